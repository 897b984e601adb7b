"""Pins the CPU linker oracle (oracle/linker.py) against what the linking
rules of gpu-tracking/src/linking.rs fix independently of its code:

  * hand-built cases with answers worked out by hand (optimal rather than
    greedy assignment, the strict ``< r*r`` rule, memory, new ids in
    destination order, FrameSubsetter's step boundaries);
  * well-separated tracks (every other particle farther than 2 r), where the
    linking is unique: ids must equal the true track ids after
    first-appearance relabelling, gaps longer than ``memory`` starting new
    tracks;
  * brute force: for dense random steps the chosen assignment must have the
    minimum total cost (summed squared distances, r*r per unmatched
    destination) over every one-to-one assignment within the radius.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from oracle.linker import Linker, frame_subsets, link_rows


def rows_of(frames):
    """[[(y, x), ...], ...] -> [frame, mass, y, x] float32 rows."""
    out = [[f, 1.0, y, x] for f, pts in enumerate(frames) for (y, x) in pts]
    return np.asarray(out, np.float32).reshape(-1, 4)


def relabel(ids):
    """First-appearance relabelling (the comparison of linking_test.ipynb)."""
    m = {}
    return np.asarray([m.setdefault(int(i), len(m)) for i in ids], np.int64)


def test_frame_subsets_steps():
    rows = np.asarray([[2, 0, 1, 1], [2, 0, 2, 2], [3, 0, 3, 3], [5, 0, 4, 4], [5, 0, 5, 5]],
                      np.float32)
    steps = list(frame_subsets(rows))
    # first frame is not 0: an empty first step, then one step per run
    assert [len(s) for s in steps] == [0, 2, 1, 2]
    assert steps[1][0] == (np.float32(1), np.float32(1))
    assert list(frame_subsets(np.zeros((0, 4), np.float32))) == []
    # a run of frame 0 at the start is the first step; a frame value that
    # comes back later starts a new step
    rows = np.asarray([[0, 0, 1, 1], [1, 0, 1, 1], [0, 0, 1, 1]], np.float32)
    assert [len(s) for s in frame_subsets(rows)] == [1, 1, 1]


def test_optimal_not_greedy():
    # dest 0 is nearer to source 0 (1.44 < 1.69), but giving it source 0
    # leaves dest 1 without a source in range (cost 9): the optimum crosses
    r = 3.0
    ids = link_rows(rows_of([[(0, 0), (0, 2.5)], [(0, 1.2), (0, -0.5)]]), r)
    assert ids.tolist() == [0, 1, 1, 0]


def test_nearer_destination_takes_the_source():
    ids = link_rows(rows_of([[(10, 10)], [(10, 12), (10, 11)]]), 3.0)
    assert ids.tolist() == [0, 1, 0]


def test_strict_radius():
    # squared distance exactly r*r is not a link (linking.rs:97, `<`)
    ids = link_rows(rows_of([[(5, 5)], [(5, 7)]]), 2.0)
    assert ids.tolist() == [0, 1]
    ids = link_rows(rows_of([[(5, 5)], [(5, 6.999)]]), 2.0)
    assert ids.tolist() == [0, 0]


def test_new_ids_in_destination_order():
    ids = link_rows(rows_of([[(10, 10), (30, 30)], [(50, 50), (10, 11), (70, 70), (30, 31)]]),
                    3.0)
    assert ids.tolist() == [0, 1, 2, 0, 3, 1]


def test_frames_without_rows_are_not_steps():
    # link_rust sees rows only: frame 2 directly follows frame 0
    frames = [[(10, 10)], [], [(10, 11)]]
    assert link_rows(rows_of(frames), 3.0, memory=0).tolist() == [0, 0]


ANCHOR = (100.0, 100.0)   # a far particle present in every frame


def _with_anchor(frames):
    return [[ANCHOR] + pts for pts in frames]


def _drop_anchor(frames, ids):
    keep = np.concatenate([[False] + [True] * len(pts) for pts in frames])
    return ids[keep].tolist()


def test_memory_bridges_gaps():
    frames = [[(10, 10)], [], [(10, 11)], [], [], [(10, 12)]]
    rows = rows_of(_with_anchor(frames))
    assert _drop_anchor(frames, link_rows(rows, 3.0, memory=0)) == [1, 2, 3]
    assert _drop_anchor(frames, link_rows(rows, 3.0, memory=1)) == [1, 1, 2]
    assert _drop_anchor(frames, link_rows(rows, 3.0, memory=2)) == [1, 1, 1]


def test_memory_entry_positions_are_where_last_seen():
    # the remembered particle is matched from its last position: (10, 10)
    # is 2.5 from (10, 12.5), in range, and 3.5 from (10, 13.5), out of range
    frames = [[(10, 10)], [], [(10, 12.5)]]
    ids = link_rows(rows_of(_with_anchor(frames)), 3.0, memory=1)
    assert _drop_anchor(frames, ids) == [1, 1]
    frames = [[(10, 10)], [], [(10, 13.5)]]
    ids = link_rows(rows_of(_with_anchor(frames)), 3.0, memory=1)
    assert _drop_anchor(frames, ids) == [1, 2]


def test_memory_entry_beside_linked_particle():
    # a particle out of range in one step stays in memory and is picked up
    # in the next, next to a particle that was linked all along
    frames = [[(10, 10), (10, 14)], [(10, 10.2)], [(10, 10.3), (10, 14.1)]]
    ids = link_rows(rows_of(frames), 3.0, memory=2)
    assert ids.tolist() == [0, 1, 0, 0, 1]


def _separated_tracks(rng, n_tracks, n_frames, r, p_off):
    """Tracks on a jittered grid with spacing 5 r, steps below r / 12 per axis."""
    g = int(np.ceil(np.sqrt(n_tracks)))
    base = np.stack(np.meshgrid(np.arange(g), np.arange(g), indexing="ij"), -1).reshape(-1, 2)
    base = base[:n_tracks] * (5 * r) + 20.0
    pos = base + rng.uniform(-0.5, 0.5, base.shape)
    frames, truth = [], []
    for f in range(n_frames):
        pos = pos + rng.uniform(-r / 12, r / 12, pos.shape)
        on = rng.random(n_tracks) >= p_off
        order = rng.permutation(n_tracks)
        pts = [(float(pos[k, 0]), float(pos[k, 1])) for k in order if on[k]]
        frames.append(pts)
        truth.append([int(k) for k in order if on[k]])
    return frames, truth


def _expected_ids(truth, memory):
    """Track ids where a gap longer than `memory` frames starts a new track."""
    last_seen, label, next_label, out = {}, {}, 0, []
    for f, ks in enumerate(truth):
        for k in ks:
            if k not in last_seen or f - last_seen[k] - 1 > memory:
                label[k] = next_label
                next_label += 1
            out.append(label[k])
        for k in ks:
            last_seen[k] = f
    return np.asarray(out)


@pytest.mark.parametrize("memory", [0, 1, 3])
def test_separated_tracks_are_recovered(memory):
    rng = np.random.default_rng(50 + memory)
    frames, truth = _separated_tracks(rng, 40, 30, 4.0, 0.2)
    ids = link_rows(rows_of(frames), 4.0, memory=memory)
    assert np.array_equal(relabel(ids), relabel(_expected_ids(truth, memory)))


def _brute_min_cost(src_pts, dst_pts, r):
    """Minimum over one-to-one assignments within the radius of
    sum(d^2) + r^2 * unmatched destinations (float64)."""
    r2 = float(np.float32(r) * np.float32(r))
    cand = []
    for d in dst_pts:
        lst = []
        for s, sp in enumerate(src_pts):
            dd = (d[0] - sp[0]) ** 2 + (d[1] - sp[1]) ** 2
            if dd < r2:
                lst.append((s, dd))
        cand.append(lst + [(None, r2)])
    best = np.inf
    for choice in itertools.product(*cand):
        used = [s for s, _ in choice if s is not None]
        if len(used) != len(set(used)):
            continue
        best = min(best, sum(c for _, c in choice))
    return best


def test_assignment_is_minimum_cost():
    rng = np.random.default_rng(7)
    r = 3.0
    checked = 0
    for trial in range(40):
        n0, n1 = rng.integers(1, 6), rng.integers(1, 6)
        src = [(float(a), float(b)) for a, b in rng.uniform(0, 6, (n0, 2)).astype(np.float32)]
        dst = [(float(a), float(b)) for a, b in rng.uniform(0, 6, (n1, 2)).astype(np.float32)]
        lk = Linker(r, 0)
        id0 = lk.advance([(np.float32(a), np.float32(b)) for a, b in src])
        id1 = lk.advance([(np.float32(a), np.float32(b)) for a, b in dst])
        by_id = {i: src[k] for k, i in enumerate(id0)}
        cost = 0.0
        r2 = float(np.float32(r) * np.float32(r))
        for k, i in enumerate(id1):
            if i in by_id:
                d = (dst[k][0] - by_id[i][0]) ** 2 + (dst[k][1] - by_id[i][1]) ** 2
                assert d < r2 + 1e-6
                cost += d
            else:
                cost += r2
        assert len(set(id1)) == len(id1)
        assert cost == pytest.approx(_brute_min_cost(src, dst, r), rel=1e-5, abs=1e-5)
        checked += 1
    assert checked == 40


def test_ids_unique_within_each_step():
    rng = np.random.default_rng(11)
    frames = [[(float(a), float(b)) for a, b in rng.uniform(0, 60, (25, 2))] for _ in range(12)]
    rows = rows_of(frames)
    ids = link_rows(rows, 4.0, memory=2)
    for f in range(len(frames)):
        sel = ids[rows[:, 0] == f]
        assert len(set(sel.tolist())) == len(sel)


def test_neighbours_match_scalar_rule():
    """The oracle's vectorised candidate search against the rule written out
    one pair at a time (box with binary32 bounds, then (0 + a*a) + b*b < r*r),
    with points on the box edges, at distance exactly r and with NaNs."""
    from oracle.linker import _neighbours
    f32 = np.float32
    rng = np.random.default_rng(3)
    r = f32(2.5)
    src = [((f32(a), f32(b)), k) for k, (a, b) in enumerate(rng.uniform(0, 12, (30, 2)))]
    dst = [(f32(a), f32(b)) for a, b in rng.uniform(0, 12, (60, 2))]
    q = src[0][0]
    dst += [(f32(q[0] + r), q[1]), (f32(q[0] - r), q[1]), (q[0], f32(q[1] + r)),
            (f32(q[0] + f32(1.5)), f32(q[1] + f32(2.0))), (f32(np.nan), q[1]),
            (q[0], f32(np.nan)), q]
    src += [((f32(np.nan), f32(3.0)), 99)]
    r2 = f32(r * r)
    got = _neighbours(src, dst, r)
    for (sp, _), lst in zip(src, got):
        exp = []
        for d, dp in enumerate(dst):
            if (dp[0] < f32(sp[0] - r) or dp[0] > f32(sp[0] + r) or dp[1] < f32(sp[1] - r)
                    or dp[1] > f32(sp[1] + r)):
                continue
            a, b = f32(dp[0] - sp[0]), f32(dp[1] - sp[1])
            dist = f32(f32(f32(0.0) + f32(a * a)) + f32(b * b))
            if dist < r2:
                exp.append((d, dist))
        assert [(d, float(v)) for d, v in lst] == [(d, float(v)) for d, v in exp]


@pytest.mark.parametrize("memory", [0, 1, 3])
def test_c_restatement_matches_python_oracle(memory):
    """oracle/gt_linker.c (the compiled CPU baseline) against oracle/linker.py
    on dense blinking Brownian tracks, edge steps included."""
    from oracle.oracle import link_rows_c
    rng = np.random.default_rng(70 + memory)
    pos = rng.uniform(0, 60, (70, 2))
    rows = []
    for f in range(25):
        on = rng.permutation(np.flatnonzero(rng.random(70) >= 0.1))
        rows += [[f + 2, 1.0, pos[k, 0], pos[k, 1]] for k in on]   # first frame is not 0
        pos = pos + rng.normal(0, 0.8, pos.shape)
    rows = np.asarray(rows, np.float32)
    for r in (2.5, 4.0):
        assert np.array_equal(link_rows_c(rows, r, memory), link_rows(rows, r, memory))
    # non-finite coordinates never link; an empty table has no ids
    bad = rows.copy()
    bad[5, 2] = np.nan
    bad[9, 3] = np.inf
    assert np.array_equal(link_rows_c(bad, 4.0, memory), link_rows(bad, 4.0, memory))
    assert link_rows_c(np.zeros((0, 4), np.float32), 4.0, memory).shape == (0,)
