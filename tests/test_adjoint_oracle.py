"""The binary64 adjoint of the oracle against what the mathematics fixes:
central finite differences of the binary64 forward (step 1e-2 counts, the
walk and the clamp pattern unchanged), the linear-in-mass identity and the
centroid identity sum_r J_mass[r] * I[r] = mass of a linear filter."""
import numpy as np
import pytest

from oracle import oracle

H = W = 48


def _spots(seed, n=4, amp=1500.0, bg=100.0, sigma=1.6):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    img = np.full((H, W), bg)
    for _ in range(n):
        y, x = rng.uniform(12, H - 12, 2)
        img += amp * np.exp(-((yy - y) ** 2 + (xx - x) ** 2) / (2 * sigma ** 2))
    return rng.poisson(img).astype(np.float32)


def _particles(frame, p, limit=3):
    _, cand, _ = oracle.locate_frame(frame, p, f64=True)
    return [int(c) for c in cand[:limit]]


@pytest.mark.parametrize("seed,d,k", [(1, 9, 9), (2, 11, 11), (3, 7, 8)])
def test_adjoint_matches_finite_differences(seed, d, k):
    frame = _spots(seed)
    p = oracle.Params(diameter=d, separation=d, smoothing_size=k)
    parts = _particles(frame, p)
    assert parts, "no particles"
    h = 1e-2
    fd64 = frame.astype(np.float64)
    checked = 0
    for cand in parts:
        jac, (py0, px0), yxm, state = oracle.adjoint_particle(frame, p, cand, f64_forward=True)
        base, st0, na0 = oracle.particle_f64(fd64, p, cand)
        assert np.allclose(base, yxm)
        S = jac.shape[1]
        err = 0.0
        scale = np.abs(jac).reshape(3, -1).max(axis=1)
        for sy in range(S):
            for sx in range(S):
                ry, rx = py0 + sy, px0 + sx
                if not (0 <= ry < H and 0 <= rx < W):
                    continue
                f1 = fd64.copy(); f1[ry, rx] += h
                f2 = fd64.copy(); f2[ry, rx] -= h
                a, s1, n1 = oracle.particle_f64(f1, p, cand)
                b, s2, n2 = oracle.particle_f64(f2, p, cand)
                if not (np.array_equal(s1, st0) and np.array_equal(s2, st0) and n1 == na0 == n2):
                    continue  # a decision or the clamp pattern changed: not differentiable here
                fd = (a - b) / (2 * h)
                err = max(err, float(np.max(np.abs(fd - jac[:, sy, sx]) / scale)))
                checked += 1
        assert err < 1e-6, f"cand {cand}: FD vs analytic relative error {err}"
    assert checked > 100


def test_adjoint_f32_forward_close_to_f64():
    frame = _spots(5)
    p = oracle.Params(diameter=9, separation=9, smoothing_size=9)
    for cand in _particles(frame, p):
        j64, o64, yxm64, st64 = oracle.adjoint_particle(frame, p, cand, f64_forward=True)
        j32, o32, yxm32, st32 = oracle.adjoint_particle(frame, p, cand, f64_forward=False)
        if not np.array_equal(st32, st64):
            continue
        assert o32 == o64
        rel = np.abs(j32 - j64).max() / np.abs(j64).max()
        assert rel < 1e-4


def test_mass_jacobian_reproduces_mass():
    """Where the clamp is inactive the bandpass is linear, so
    mass = sum_r J_mass[r] * I[r] exactly (up to rounding)."""
    frame = _spots(7, bg=300.0)
    p = oracle.Params(diameter=9, separation=9, smoothing_size=9)
    checked = 0
    for cand in _particles(frame, p):
        jac, (py0, px0), yxm, st = oracle.adjoint_particle(frame, p, cand, f64_forward=True)
        S = jac.shape[1]
        patch = np.zeros((S, S))
        ys = slice(max(py0, 0), min(py0 + S, H)); xs = slice(max(px0, 0), min(px0 + S, W))
        patch[ys.start - py0:ys.stop - py0, xs.start - px0:xs.stop - px0] = frame[ys, xs]
        # mass is the sum of P over the disk pixels with P > 0: linear there
        m_lin = float((jac[2] * patch).sum())
        assert abs(m_lin - yxm[2]) <= 1e-9 * abs(yxm[2]) + 1e-6
        # offsets: c0 * m = sum (i - c0)... the y-row Jacobian sums to 0 against I
        assert abs(float((jac[0] * patch).sum())) < 1e-9 * abs(yxm[2])
        checked += 1
    assert checked >= 1
