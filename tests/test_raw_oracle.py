"""Pins the raw-signal oracle (oracle/raw.py) against what the extraction
rules of gpu-tracking/src/execute_gpu.rs:301-333 fix independently of its
code: tap counts of integer-radius disks (Gauss circle numbers), exact sums
and medians on constant and point-symmetric ramp frames, numpy's nansum and
nanmedian next to the frame border, half-away-from-zero centring and the
all-outside background."""
import numpy as np
import pytest

from oracle.raw import annulus_taps, circle_taps, median_linear, raw_signal


@pytest.mark.parametrize("r,n", [(0, 1), (1, 5), (2, 13), (3, 29), (4, 49), (5, 81), (10, 317)])
def test_circle_tap_counts(r, n):
    t = circle_taps(float(r))
    assert len(t) == n and len(set(t)) == n
    assert t == sorted(t)                      # row-major order
    assert all(i * i + j * j <= r * r for i, j in t)


def test_annulus_taps_integer_radii():
    t = annulus_taps(7.0, 3.0)
    exp = [(i, j) for i in range(-7, 8) for j in range(-7, 8) if 9 <= i * i + j * j <= 49]
    assert t == exp
    assert (0, 0) not in t and (3, 0) in t and (7, 0) in t


def test_constant_frame():
    fr = np.full((40, 40), 123.0, np.float32)
    m, med, corr = raw_signal(fr, 20.3, 19.7, 3.0, 7.0, 1.0)
    assert m == np.float32(29 * 123.0) and med == np.float32(123.0) and corr == 0.0


def test_symmetric_ramp():
    yy, xx = np.mgrid[0:50, 0:50]
    fr = (2 * yy + 3 * xx).astype(np.float32)
    m, med, corr = raw_signal(fr, 25.0, 24.0, 4.0, 9.0, 2.0)
    v0 = 2 * 25 + 3 * 24
    assert m == np.float32(49 * v0) and med == np.float32(v0) and corr == 0.0


def test_border_matches_numpy():
    rng = np.random.default_rng(2)
    fr = rng.integers(0, 4000, (30, 40)).astype(np.float32)
    for (y, x) in [(1.2, 0.6), (0.0, 39.4), (28.7, 20.0), (15.0, 15.0)]:
        cy, cx = int(np.floor(y + 0.5)), int(np.floor(x + 0.5))
        sig = [fr[cy + i, cx + j] for i in range(-3, 4) for j in range(-3, 4)
               if i * i + j * j <= 9 and 0 <= cy + i < 30 and 0 <= cx + j < 40]
        bg = [fr[cy + i, cx + j] for i in range(-8, 9) for j in range(-8, 9)
              if 16 <= i * i + j * j <= 64 and 0 <= cy + i < 30 and 0 <= cx + j < 40]
        m, med, corr = raw_signal(fr, y, x, 3.0, 8.0, 1.0)
        assert float(m) == float(np.sum(np.asarray(sig, np.float64)))
        assert float(med) == float(np.median(np.asarray(bg, np.float64)))
        assert corr == np.float32(m - np.float32(med * np.float32(29)))


def test_centre_rounds_half_away_from_zero():
    fr = np.zeros((8, 8), np.float32)
    fr[3, 3] = 7.0
    assert raw_signal(fr, 2.5, 2.5, 0.0)[0] == 7.0      # round(2.5) = 3, not 2
    assert raw_signal(fr, 3.49, 3.49, 0.0)[0] == 7.0
    assert raw_signal(fr, 3.5, 3.5, 0.0)[0] == 0.0


def test_background_all_outside():
    fr = np.ones((3, 3), np.float32)
    m, med, corr = raw_signal(fr, 1.0, 1.0, 1.0, 5.0, 3.0)
    assert m == 5.0 and np.isnan(med) and np.isnan(corr)


def test_median_interpolation():
    assert median_linear([1.0, 2.0, 4.0, 10.0]) == np.float32(3.0)
    assert median_linear([5.0, np.nan, 1.0]) == np.float32(3.0)
    assert median_linear([2.0]) == np.float32(2.0)
    assert np.isnan(median_linear([np.nan]))
    # a + f32(0.5 (b - a)): rounds once, at the half difference
    a, b = np.float32(1.0), np.float32(1.0 + 2 ** -23)
    assert median_linear([a, b]) == np.float32(a + np.float32(0.5 * (float(b) - float(a))))
