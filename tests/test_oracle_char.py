"""Pins the oracle's characterisation columns before they are trusted as the
checker of the GPU's.

Two code paths of the reference produce them:
  * the shader's (centers.wgsl:64-80, `characterize`): Rg_gpu, raw_mass_gpu,
    signal_gpu, ecc_gpu over the disk of radius (d - 1) / 2 at the candidate
    pixel, divided by that disk's own mass;
  * the host's (execute_gpu.rs:165-172,286-299,346-352, `cpu_processed`):
    Rg, ecc, signal over circle_inds(d / 2) at the candidate pixel, divided
    by the WALKED mass (the mass column, Q11).
Checks, each independent of the oracle's code:
  * closed forms on hand-built frames: a ring of radius rho has Rg = rho; two
    mass points opposite each other have ecc = 1 (on either axis or a
    diagonal) and four on the axes ecc = 0, and with a centre value the
    closed forms of check_ecc; signal is the disk maximum;
    raw_mass_gpu over a constant frame c is c times the Gauss circle number
    N(r) (1, 5, 13, 29, 49, 81 for r = 0..5); the host columns divide by the
    walked mass passed in;
  * cross-column invariants on located frames with odd d (both paths then
    cover the same disk): signal == signal_gpu, Rg^2 * m_walk ==
    Rg_gpu^2 * m_cand and ecc * (m_walk - P_c) == ecc_gpu * (m_cand - P_c)
    up to summation order, where m_cand (the candidate's disk mass) and P_c
    come from the oracle's bandpass, itself pinned to the reference's
    masses.bin dump;
  * the binary32 tap tables against values fixed by the mathematics.
The mutation tests plant each plausible mistake in the oracle (on the CPU,
through orc_set_mutation) and show that at least one check fails.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

GAUSS_CIRCLE = {0: 1, 1: 5, 2: 13, 3: 29, 4: 49, 5: 81}


def _disk_frame(shape, c, pts, centre=0.0):
    P = np.zeros(shape, np.float32)
    for (i, j), val in pts:
        P[c[0] + i, c[1] + j] = val
    P[c] += centre
    return P


def _ring(rho2):
    return [(i, j) for i in range(-6, 7) for j in range(-6, 7) if i * i + j * j == rho2]


# ------------------------------------------------------------------ checks --
# Each returns a list of failure messages (empty when the check holds).

def check_ring_rg():
    out = []
    c = (20, 20)
    raw = np.ones((41, 41), np.float32)
    # d = 11: rint = band = 5; ring rho = 5 has 12 lattice points
    pts = [(q, 1.0) for q in _ring(25)]
    assert len(pts) == 12
    P = _disk_frame((41, 41), c, pts)
    cpu, gpu = O.characterise_at(P, raw, O.Params(diameter=11), *c, walked_mass=12.0)
    if abs(gpu[0] - 5.0) > 1e-6:
        out.append(f"ring rho=5 d=11: Rg_gpu {gpu[0]} != 5")
    if abs(cpu[0] - 5.0) > 1e-6:
        out.append(f"ring rho=5 d=11: Rg {cpu[0]} != 5")
    # d = 10: rint 4 (gpu disk), band 5 (cpu neighbourhood); rings rho=4
    # (4 points) and rho=5 (12 points)
    pts = [(q, 1.0) for q in _ring(16)] + [(q, 1.0) for q in _ring(25)]
    P = _disk_frame((41, 41), c, pts)
    cpu, gpu = O.characterise_at(P, raw, O.Params(diameter=10), *c, walked_mass=16.0)
    if abs(gpu[0] - 4.0) > 1e-6:
        out.append(f"rings d=10: Rg_gpu {gpu[0]} != 4 (disk radius 4 sees the inner ring only)")
    want = math.sqrt((4 * 16 + 12 * 25) / 16.0)
    if abs(cpu[0] - want) > 1e-6:
        out.append(f"rings d=10: Rg {cpu[0]} != {want} (circle_inds(5) sees both rings)")
    return out


def check_ecc():
    """The centre tap has theta = atan2(0, 0) = 0, so a centre value c adds c
    to the cosine sum (centers.wgsl:69, kernels.rs:196-205) and the
    denominator is m - c.  Pairs of mass a at angle 2*theta = pi (row axis),
    0 (column axis), pi/2 (diagonals) and four points on the axes give
    |c - 2a| / 2a, (c + 2a) / 2a, sqrt(4a^2 + c^2) / 2a and c / 4a; with
    c = 0 the pairs give 1 and the four points 0."""
    out = []
    c = (20, 20)
    raw = np.zeros((41, 41), np.float32)
    a = 3.0
    for d in (9, 11):
        p = O.Params(diameter=d)
        for cen in (0.0, 5.0):
            cases = (("row axis", [((0, 2), a), ((0, -2), a)], abs(cen - 2 * a) / (2 * a)),
                     ("column axis", [((2, 0), a), ((-2, 0), a)], (cen + 2 * a) / (2 * a)),
                     ("diagonal", [((1, 1), a), ((-1, -1), a)],
                      math.sqrt(4 * a * a + cen * cen) / (2 * a)),
                     ("anti-diagonal", [((1, -1), a), ((-1, 1), a)],
                      math.sqrt(4 * a * a + cen * cen) / (2 * a)),
                     ("four axis points", [((0, 2), a), ((0, -2), a), ((2, 0), a), ((-2, 0), a)],
                      cen / (4 * a)))
            for name, pts, want in cases:
                P = _disk_frame((41, 41), c, pts, centre=cen)
                m = float(P.sum())
                cpu, gpu = O.characterise_at(P, raw, p, *c, walked_mass=m)
                for col, got in (("ecc_gpu", gpu[3]), ("ecc", cpu[1])):
                    if abs(got - want) > 1e-6:
                        out.append(f"d={d} centre {cen} {name}: {col} {got} != {want}")
    return out


def check_signal():
    out = []
    rng = np.random.default_rng(3)
    c = (20, 20)
    for d in (9, 11):
        r = (d - 1) // 2
        P = rng.uniform(1.0, 2.0, (41, 41)).astype(np.float32)
        P[c[0] + 1, c[1] - 2] = 7.0      # the disk maximum, off centre
        P[c[0] + r + 3, c[1]] = 50.0     # outside both disks
        yy, xx = np.mgrid[-20:21, -20:21]
        dmax = P[(yy ** 2 + xx ** 2) <= r * r].max()
        cpu, gpu = O.characterise_at(P, np.zeros_like(P), O.Params(diameter=d), *c,
                                     walked_mass=float(P.sum()))
        if gpu[2] != dmax or cpu[2] != dmax:
            out.append(f"d={d}: signal {cpu[2]}, signal_gpu {gpu[2]} != disk max {dmax}")
    return out


def check_raw_mass():
    out = []
    c = (20, 20)
    rng = np.random.default_rng(4)
    P = rng.uniform(0.5, 1.5, (41, 41)).astype(np.float32)
    for d in (7, 8, 9, 10, 11):
        r = (d - 1) // 2
        cval = 3.0
        raw = np.full((41, 41), cval, np.float32)
        _, gpu = O.characterise_at(P, raw, O.Params(diameter=d), *c, walked_mass=1.0)
        if gpu[1] != cval * GAUSS_CIRCLE[r]:
            out.append(f"d={d}: raw_mass_gpu {gpu[1]} != {cval} * N({r}) = {cval * GAUSS_CIRCLE[r]}")
    return out


def check_walked_mass_divisor():
    """The host columns divide by the walked mass (execute_gpu.rs:349-350),
    the shader's by the candidate disk's own mass."""
    out = []
    c = (20, 20)
    raw = np.zeros((41, 41), np.float32)
    pts = [(q, 2.0) for q in _ring(25)] + [((0, 2), 1.0), ((0, -2), 1.0)]
    P = _disk_frame((41, 41), c, pts, centre=4.0)
    m_cand = float(P.sum())
    cpu1, gpu1 = O.characterise_at(P, raw, O.Params(diameter=11), *c, walked_mass=m_cand)
    cpu2, gpu2 = O.characterise_at(P, raw, O.Params(diameter=11), *c, walked_mass=2 * m_cand)
    if not np.array_equal(gpu1, gpu2):
        out.append("shader columns depend on the walked mass")
    if abs(cpu2[0] * math.sqrt(2) - cpu1[0]) > 1e-5 * cpu1[0]:
        out.append(f"Rg with twice the walked mass {cpu2[0]} != Rg / sqrt(2) ({cpu1[0] / math.sqrt(2)})")
    e1 = cpu1[1] * (m_cand - 4.0)
    e2 = cpu2[1] * (2 * m_cand - 4.0)
    if abs(e1 - e2) > 1e-5 * abs(e1):
        out.append(f"ecc numerator changes with the walked mass ({e1} vs {e2})")
    return out


def _frame_c1like(seed=7, size=192, n=40, sigma=1.5):
    rng = np.random.default_rng(seed)
    img = rng.poisson(100, (size, size)).astype(np.float64)
    yy, xx = np.mgrid[0:size, 0:size]
    for (cy, cx), amp in zip(rng.uniform(12, size - 12, (n, 2)), rng.uniform(500, 2000, n)):
        img += amp * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sigma ** 2))
    return np.minimum(img, 65535).astype(np.uint16).astype(np.float32)


def check_cross_columns():
    """Q11 invariants on located frames with odd d."""
    out = []
    for d, sigma, seed in ((9, 1.5, 7), (11, 2.0, 8)):
        img = _frame_c1like(seed=seed, sigma=sigma)
        p = O.Params(diameter=d, separation=d, percentile=64, cpu_processed=True,
                     characterize=True)
        rows, cand, _ = O.locate_frame(img, p)
        if len(rows) < 20:
            out.append(f"d={d}: only {len(rows)} rows")
            continue
        P = O.bandpass(img, d, 1.0)
        mc, _ = O.masses_centers(P, d)
        m_walk = rows[:, 1]
        m_cand = mc.ravel()[cand].astype(np.float64)
        pc = P.ravel()[cand].astype(np.float64)
        rg, ecc, sig = rows[:, 4], rows[:, 5], rows[:, 6]
        rg_g, sig_g, ecc_g = rows[:, 7], rows[:, 9], rows[:, 10]
        moved = m_walk != m_cand
        if moved.sum() < 3:
            out.append(f"d={d}: only {moved.sum()} rows whose walk moved (invariants untested)")
        if not np.array_equal(sig, sig_g):
            out.append(f"d={d}: signal != signal_gpu on {np.sum(sig != sig_g)} rows")
        lhs, rhs = rg ** 2 * m_walk, rg_g ** 2 * m_cand
        bad = np.abs(lhs - rhs) > 2e-6 * np.abs(rhs)
        if bad.any():
            out.append(f"d={d}: Rg^2 m_walk != Rg_gpu^2 m_cand on {bad.sum()} rows")
        lhs, rhs = ecc * (m_walk - pc), ecc_g * (m_cand - pc)
        bad = np.abs(lhs - rhs) > 1e-5 * m_cand
        if bad.any():
            out.append(f"d={d}: ecc (m_walk - P_c) != ecc_gpu (m_cand - P_c) on {bad.sum()} rows")
        # raw_mass_gpu: the input frame summed over the candidate's disk
        r = (d - 1) // 2
        H, W = img.shape
        u, v = cand // W, cand % W
        ii, jj = np.mgrid[-r:r + 1, -r:r + 1]
        mask = (ii ** 2 + jj ** 2) <= r * r
        ref = np.array([img[uu - r:uu + r + 1, vv - r:vv + r + 1][mask].astype(np.float64).sum()
                        for uu, vv in zip(u, v)])
        bad = np.abs(rows[:, 8] - ref) > 1e-6 * ref
        if bad.any():
            out.append(f"d={d}: raw_mass_gpu != disk sum of the frame on {bad.sum()} rows")
    return out


CHECKS = {
    "ring_rg": check_ring_rg,
    "ecc": check_ecc,
    "signal": check_signal,
    "raw_mass": check_raw_mass,
    "walked_mass_divisor": check_walked_mass_divisor,
    "cross_columns": check_cross_columns,
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_characterisation_check_holds(name):
    fails = CHECKS[name]()
    assert not fails, fails


# Which checks catch which planted mistake (each must be caught by these).
CAUGHT_BY = {
    "gpu_radius_band": ["ring_rg", "raw_mass"],
    "cpu_radius_rint": ["ring_rg"],
    "cpu_candidate_mass": ["walked_mass_divisor", "cross_columns"],
    "theta_not_2theta": ["ecc"],
    "ecc_full_mass": ["ecc", "cross_columns"],
    "signal_centre": ["signal"],
    "r_not_r2": ["ring_rg"],
    "middle_most_off": ["ecc", "cross_columns"],
    "raw_from_processed": ["raw_mass", "cross_columns"],
}


@pytest.mark.parametrize("mutation", sorted(O.MUTATIONS))
def test_planted_mistake_is_caught(mutation):
    assert mutation in CAUGHT_BY
    with O.mutation(mutation):
        caught = {name: CHECKS[name]() for name in CAUGHT_BY[mutation]}
    for name, fails in caught.items():
        assert fails, f"check {name} did not catch the planted mistake {mutation}"
    # and the oracle is back to normal
    assert not check_ring_rg() and not check_ecc()


# ------------------------------------------------------------- tap tables --

def test_char_tables_are_binary32():
    """kernels.rs:185-205 forms (2.*(j as f32).atan2(i as f32)).sin() in
    binary32: theta rounds to binary32 before the doubling, so at (1, 1)
    cos(2 * fl32(pi/4)) = -(2 fl32(pi/4) - pi/2) = -4.3711388e-8, not the
    binary64 cos(pi/2) = 6.1e-17; on the negative row axis sin(2 fl32(pi))
    = 2 (fl32(pi) - pi) = 1.7484555e-7."""
    r2, s2, c2, mid = O.char_tables(5)
    taps = [(i, j) for i in range(-5, 6) for j in range(-5, 6) if i * i + j * j <= 25]
    assert len(r2) == len(taps) == GAUSS_CIRCLE[5]
    assert taps[mid] == (0, 0)
    assert np.array_equal(r2, np.array([i * i + j * j for i, j in taps], np.float32))
    t = {q: k for k, q in enumerate(taps)}
    q4 = float(np.float32(math.pi / 4))
    qp = float(np.float32(math.pi))
    assert abs(c2[t[(1, 1)]] - np.float32(-(2 * q4 - math.pi / 2))) == 0
    assert abs(s2[t[(-1, 0)]] - np.float32(2 * (qp - math.pi))) == 0
    assert s2[t[(1, 0)]] == 0.0 and c2[t[(1, 0)]] == 1.0     # theta = 0 exactly
    assert c2[t[(0, 1)]] == -1.0 and s2[t[(0, 1)]] != 0.0    # 2 fl32(pi/2) != pi
    # every entry is sin / cos of the binary32 angle 2 * fl32(atan2(j, i))
    for (i, j), k in t.items():
        # (binary32 rounding of theta, doubled: <= 2 * 2^-22 for |theta| <= pi)
        ang = 2.0 * math.atan2(j, i)
        assert abs(float(s2[k]) - math.sin(ang)) <= 6e-7, (i, j)
        assert abs(float(c2[k]) - math.cos(ang)) <= 6e-7, (i, j)
