"""The linker's host subnet solver (tracking_optimizations_amd/csrc/
gt_link_solve.h) on the CPU: candidate lists built here with the binary32
tests of the GPU search (k_link_pairs: per-axis box against sy -+ r, sx -+ r,
then (0 + dy*dy) + dx*dx < r*r, ascending destination order), the solver run
by a small C++ driver (tests/link_solve_driver.cpp, built with g++), ids
compared with the C oracle linker (oracle/gt_linker.c, a restatement of
linking.rs:234-492) on the same rows.  No GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "tracking_optimizations_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("lsd") / "link_solve_driver")
    subprocess.run([gxx, "-O1", "-std=c++17", "-I", CSRC,
                    os.path.join(HERE, "link_solve_driver.cpp"), "-o", out], check=True)
    return out


def _movie_rows(seed, n_frames, n_part, size, step_sigma, blink, close_pairs):
    """[frame, mass, y, x] float32 rows of Brownian particles that blink off,
    some of them placed in close pairs so that subnets form."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, size, (n_part, 2))
    if close_pairs:
        k = n_part // 4
        pos[k:2 * k] = pos[:k] + rng.normal(0, 2.0, (k, 2))
    rows = []
    for fr in range(n_frames):
        pos = pos + rng.normal(0, step_sigma, pos.shape)
        on = rng.random(n_part) >= blink
        for (y, x) in pos[on]:
            rows.append((fr, 100.0, y, x))
    return np.asarray(rows, np.float32)


def _solver_ids(driver, rows, r, memory, tmp_path):
    r32 = np.float32(r)
    # steps (FrameSubsetter, linking.rs:34-58): a step starts where the frame
    # value differs from the previous row's, the first row compared with 0.0
    # (a first frame other than 0 leaves an empty first step)
    fr = rows[:, 0]
    starts = [i for i in range(len(fr)) if fr[i] != (fr[i - 1] if i else np.float32(0.0))]
    soff = [0] + starts + [len(fr)] if len(fr) else [0]
    # table: memory empty carried steps, an empty previous step, then the chunk
    step_off = [0] * (memory + 1) + soff
    nsteps = len(step_off) - 1
    n = len(fr)
    ys, xs = rows[:, 2].copy(), rows[:, 3].copy()
    step_of = np.zeros(n, np.int64)
    for s in range(nsteps):
        step_of[step_off[s]:step_off[s + 1]] = s
    nm = memory + 1
    off = np.zeros(n * nm + 1, np.uint32)
    cd, cdist = [], []
    r2 = r32 * r32
    for p in range(n):
        sy, sx = ys[p], xs[p]
        lo0, hi0, lo1, hi1 = sy - r32, sy + r32, sx - r32, sx + r32
        for j in range(nm):
            off[p * nm + j] = len(cd)
            ts = step_of[p] + 1 + j
            if ts >= nsteps:
                continue
            a, e = step_off[ts], step_off[ts + 1]
            dy, dx = ys[a:e], xs[a:e]
            box = ~((dy < lo0) | (dy > hi0) | (dx < lo1) | (dx > hi1))
            a0, a1 = dy - sy, dx - sx
            dist = (np.float32(0.0) + a0 * a0) + a1 * a1
            for d in np.nonzero(box & (dist < r2))[0]:
                cd.append(d)
                cdist.append(dist[d])
    off[n * nm] = len(cd)
    path = tmp_path / "in.bin"
    with open(path, "wb") as f:
        np.asarray([memory, nsteps, n, len(cd)], np.uint32).tofile(f)
        np.asarray([r32], np.float32).tofile(f)
        np.asarray(step_off, np.uint32).tofile(f)
        ys.astype(np.float32).tofile(f)
        xs.astype(np.float32).tofile(f)
        off.tofile(f)
        np.asarray(cd, np.uint32).tofile(f)
        np.asarray(cdist, np.float32).tofile(f)
    out = tmp_path / "ids.bin"
    subprocess.run([driver, str(path), str(out)], check=True)
    return np.fromfile(out, np.uint64)


@pytest.mark.parametrize("memory", [0, 1, 3])
@pytest.mark.parametrize("case", [
    dict(seed=1, n_frames=40, n_part=120, size=200.0, step_sigma=0.8, blink=0.1, close_pairs=True),
    dict(seed=2, n_frames=30, n_part=200, size=120.0, step_sigma=1.5, blink=0.05, close_pairs=True),
    dict(seed=3, n_frames=50, n_part=80, size=300.0, step_sigma=0.5, blink=0.2, close_pairs=False),
])
def test_solver_ids_equal_oracle(driver, tmp_path, memory, case):
    rows = _movie_rows(**case)
    r = 5.0
    got = _solver_ids(driver, rows, r, memory, tmp_path)
    ref = oracle.link_rows_c(rows, r, memory)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref), f"{int((got != ref).sum())} of {len(ref)} ids differ"


def _lattice_rows(seed, n_frames, side, spacing, blink):
    """[frame, mass, y, x] rows of particles on an integer lattice taking
    integer steps: squared distances are small integers, so subnets often
    hold assignments of equal total cost, and which one the solver keeps
    depends on the order it visits the subnet's destinations in."""
    rng = np.random.default_rng(seed)
    g = np.arange(side) * spacing + 10.0
    pos = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    rows = []
    for fr in range(n_frames):
        pos = pos + rng.integers(-1, 2, pos.shape)
        on = rng.random(len(pos)) >= blink
        for (y, x) in pos[on]:
            rows.append((fr, 100.0, y, x))
    return np.asarray(rows, np.float32)


@pytest.mark.parametrize("memory", [0, 1, 3])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_solver_ties_equal_oracle(driver, tmp_path, memory, seed):
    """Subnets with tied assignments: the one kept depends on the order in
    which the search visits a subnet's destinations (taking the step's
    destinations from the highest instead of the lowest changes 108-382 of
    about 900 ids here, and none of the random movies' above)."""
    rows = _lattice_rows(seed, 30, 6, 4.0, 0.15)
    got = _solver_ids(driver, rows, 2.5, memory, tmp_path)
    ref = oracle.link_rows_c(rows, 2.5, memory)
    assert np.array_equal(got, ref), f"{int((got != ref).sum())} of {len(ref)} ids differ"
