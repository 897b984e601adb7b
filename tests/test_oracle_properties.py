"""Property tests of the oracle (Hypothesis), each a fact the mathematics of
the reference pipeline fixes independently of the oracle's code:

  * positive homogeneity: bandpass(c X) = c bandpass(X) for c > 0 (the
    filters are linear and the clamp at 0 commutes with a positive scale),
    in binary64;
  * translation: shifting a frame's content by whole pixels shifts the
    bandpass, the local-maximum map and the located rows by the same amount
    wherever no border rule is reached (binary32, exactly);
  * the mirror symmetry of the filters: bandpass(flip(X)) = flip(bandpass(X))
    on frames of odd width (both boundary rules are symmetric), binary64;
  * deduplication equals the brute-force rule on point sets with many equal
    masses and points exactly at the separation;
  * the percentile threshold equals numpy.quantile on arrays with many
    repeated values and zeros.
"""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from oracle import oracle as O

SETTINGS = settings(max_examples=25, deadline=None,
                    suppress_health_check=[HealthCheck.too_slow])


def _frame(seed, h, w, n_spots, bg=100.0):
    rng = np.random.default_rng(seed)
    img = rng.poisson(bg, (h, w)).astype(np.float64)
    yy, xx = np.mgrid[0:h, 0:w]
    for _ in range(n_spots):
        cy, cx = rng.uniform(0, h - 1), rng.uniform(0, w - 1)
        amp, sig = rng.uniform(300, 3000), rng.uniform(1.2, 2.5)
        img += amp * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sig * sig))
    return np.floor(img).astype(np.float32)


@SETTINGS
@given(seed=st.integers(0, 2**31), k=st.sampled_from([3, 5, 7, 9, 11]),
       c=st.floats(0.01, 100.0), h=st.integers(12, 40), w=st.integers(12, 40))
def test_bandpass_positive_homogeneity(seed, k, c, h, w):
    img = _frame(seed, h, w, 3).astype(np.float64)
    a = O.bandpass(img * c, k, 1.0, f64=True)
    b = O.bandpass(img, k, 1.0, f64=True) * c
    assert np.allclose(a, b, rtol=1e-12, atol=1e-9 * c * max(1.0, np.abs(img).max()))


@SETTINGS
@given(seed=st.integers(0, 2**31), k=st.sampled_from([5, 9, 11]),
       h=st.integers(15, 41).filter(lambda v: v % 2 == 1))
def test_bandpass_mirror_symmetry(seed, k, h):
    img = _frame(seed, h, h, 4).astype(np.float64)
    a = O.bandpass(img[:, ::-1].copy(), k, 1.0, f64=True)
    b = O.bandpass(img, k, 1.0, f64=True)[:, ::-1]
    assert np.allclose(a, b, rtol=1e-12, atol=1e-9)
    a = O.bandpass(img[::-1, :].copy(), k, 1.0, f64=True)
    b = O.bandpass(img, k, 1.0, f64=True)[::-1, :]
    assert np.allclose(a, b, rtol=1e-12, atol=1e-9)


@SETTINGS
@given(seed=st.integers(0, 2**31), dy=st.integers(-6, 6), dx=st.integers(-6, 6),
       d=st.sampled_from([7, 9, 11]))
def test_translation_of_bandpass_maxima_and_rows(seed, dy, dx, d):
    """Content placed in the middle of a zero-padded canvas at two offsets:
    everything the border rules cannot reach moves by the offset exactly."""
    core = _frame(seed, 48, 48, 5)
    H = W = 96
    a = np.zeros((H, W), np.float32)
    b = np.zeros((H, W), np.float32)
    a[24:72, 24:72] = core
    b[24 + dy:72 + dy, 24 + dx:72 + dx] = core
    k = d
    Pa, Pb = O.bandpass(a, k, 1.0), O.bandpass(b, k, 1.0)
    m = 14   # inside this margin no filter tap reaches a frame border
    assert np.array_equal(Pa[m:H - m, m:W - m], Pb[m + dy:H - m + dy, m + dx:W - m + dx])
    p = O.Params(diameter=d, separation=d, minmass=100.0)
    ra, ca, _ = O.locate_frame(a, p)
    rb, cb, _ = O.locate_frame(b, p)
    # rows whose candidate lies well inside the canvas
    keep = lambda c: (c // W >= 20) & (c // W < H - 20) & (c % W >= 20) & (c % W < W - 20)
    ka = keep(ca) & keep(ca + dy * W + dx)
    kb = keep(cb) & keep(cb - dy * W - dx)
    assert np.array_equal(ca[ka] + dy * W + dx, cb[kb])
    assert np.array_equal(ra[ka, 1], rb[kb, 1])                  # masses: the same sums
    if ka.any():                                                # positions: u + c in binary32
        assert np.abs(ra[ka, 2] + dy - rb[kb, 2]).max() <= 1e-4
        assert np.abs(ra[ka, 3] + dx - rb[kb, 3]).max() <= 1e-4


@SETTINGS
@given(seed=st.integers(0, 2**31), n=st.integers(0, 120), sep=st.integers(1, 12),
       mass_levels=st.integers(1, 6))
def test_dedup_equals_bruteforce_with_ties(seed, n, sep, mass_levels):
    rng = np.random.default_rng(seed)
    # positions on a coarse lattice so that exact distances equal to `sep`
    # and equal masses are common
    y = rng.integers(0, 40, n).astype(np.float32)
    x = rng.integers(0, 40, n).astype(np.float32)
    mass = rng.integers(1, mass_levels + 1, n).astype(np.float32) * 100
    idx = rng.permutation(10000)[:n].astype(np.int64)
    keep = O.dedup_bruteforce(y, x, mass, idx, sep)
    # the rule of execute_gpu.rs:258-284, restated here in plain numpy
    d2 = (y[:, None] - y[None, :]) ** 2 + (x[:, None] - x[None, :]) ** 2
    near = (d2 < np.float32(sep) ** 2) & ~np.eye(n, dtype=bool)
    beats = (mass[None, :] > mass[:, None]) | ((mass[None, :] == mass[:, None]) &
                                               (idx[None, :] < idx[:, None]))
    want = ~np.any(near & beats, axis=1)
    assert np.array_equal(keep, want)


@SETTINGS
@given(seed=st.integers(0, 2**31), n=st.integers(1, 3000), pct=st.floats(0.0, 100.0),
       zeros=st.floats(0.0, 0.9), levels=st.integers(1, 50))
def test_percentile_equals_numpy_quantile(seed, n, pct, zeros, levels):
    rng = np.random.default_rng(seed)
    v = (rng.integers(0, levels, n) * 0.37).astype(np.float32)
    v[rng.random(n) < zeros] = 0.0
    got = O.percentile_threshold(v, float(np.float32(pct)))
    want = float(np.quantile(v.astype(np.float64), float(np.float32(pct)) / 100.0,
                             method="linear"))
    assert got == pytest.approx(want, rel=1e-12, abs=1e-12)
