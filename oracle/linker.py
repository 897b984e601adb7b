"""CPU oracle of the frame-to-frame linker (gpu-tracking/src/linking.rs).

TEST INFRASTRUCTURE ONLY: the checker for the library's gt_link.  Only tests/
may import it.  Plain Python loops, binary32 arithmetic through numpy float32
scalars, in the reference's order:

  * frame_subsets      FrameSubsetter::next (linking.rs:34-58): consecutive
                       rows with the same frame value form one step; the
                       first step is empty when the first frame is not 0;
  * Linker.advance     Linker::advance (linking.rs:228-283): memory entries
                       (unmatched sources of the last `memory` steps, oldest
                       first, ids not already present) are appended to the
                       sources after the previous step's particles;
  * link               link (linking.rs:286-492): candidate pairs within the
                       search radius (the kd-tree's per-axis box test against
                       binary32 bounds, then (dy*dy) + (dx*dx) < r*r),
                       connected subnets by depth-first search from each
                       destination in order, one-to-one pairs taken
                       directly, larger subnets solved exhaustively
                       (recurse2: minimum summed squared distance, a
                       destination left unmatched costs r*r, first strictly
                       better assignment wins), unmatched destinations get
                       new ids in destination order, unmatched sources of
                       the previous step become memory.

The order of a source's candidate destinations is the kd-tree crate's
traversal order in the reference (kd-tree 0.4.2, not vendored); here it is
ascending destination index.  It decides the depth-first visiting order and
so which of several equal-cost assignments is kept.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def frame_subsets(rows: np.ndarray, frame_col: int = 0, cols=(2, 3)):
    """FrameSubsetter (linking.rs:34-58) over a float32 row array."""
    idx, cur = 0, f32(0.0)
    n = rows.shape[0]
    while True:
        out = []
        while True:
            if idx < n:
                fr = f32(rows[idx, frame_col])
                if fr != cur:
                    cur = fr
                    break
            else:
                if out:
                    break
                return
            out.append((f32(rows[idx, cols[0]]), f32(rows[idx, cols[1]])))
            idx += 1
        yield out


def _neighbours(src, dest, radius):
    """Per source, the destinations within `radius` (ascending index), with
    the kd-tree's box test and binary32 squared distance
    (linking.rs:73-105: every coordinate inside [q - r, q + r] with binary32
    bounds, then distance = 0 + d0*d0, then + d1*d1, kept when < r*r).

    Vectorised over the destinations of a y band (numpy float32 elementwise
    operations round each operation to binary32, as the scalar rule does);
    NaN coordinates never pass the distance test, in or out of the band."""
    r = f32(radius)
    r2 = f32(r * r)
    if not dest:
        return [[] for _ in src]
    dy = np.asarray([p[0] for p in dest], np.float32)
    dx = np.asarray([p[1] for p in dest], np.float32)
    order = np.argsort(dy, kind="stable")
    sy = dy[order]
    out = []
    for (sp, _sid) in src:
        q0, q1 = f32(sp[0]), f32(sp[1])
        lo0, hi0 = f32(q0 - r), f32(q0 + r)
        lo1, hi1 = f32(q1 - r), f32(q1 + r)
        a, b = np.searchsorted(sy, lo0, "left"), np.searchsorted(sy, hi0, "right")
        if a >= b:
            out.append([])
            continue
        idx = np.sort(order[a:b])
        y, x = dy[idx], dx[idx]
        box = (x >= lo1) & (x <= hi1)
        d0 = y - q0
        d1 = x - q1
        dist = (np.float32(0.0) + d0 * d0) + d1 * d1
        keep = box & (dist < r2)
        out.append([(int(i), f32(v)) for i, v in zip(idx[keep], dist[keep])])
    return out


def link(src, dest, memory_start_idx, radius, counter):
    """link (linking.rs:286-492).  src: list of ((y, x), id); dest: list of
    (y, x).  Returns (ids of dest, unused sources, counter)."""
    s2d = _neighbours(src, dest, radius)
    d2s = [[] for _ in dest]
    for s, lst in enumerate(s2d):
        for (d, dist) in lst:
            d2s[d].append((s, dist))
    visited = [[False] * len(src), [False] * len(dest)]

    def recurse(node, path):
        # iterative form of linking.rs:316-341, same visiting order
        stack = [node]
        while stack:
            idx, kind = stack.pop()
            if visited[kind][idx]:
                continue
            visited[kind][idx] = True
            path[kind].append(idx)
            nxt = ([(d, 1) for (d, _) in s2d[idx]] if kind == 0 else
                   [(s, 0) for (s, _) in d2s[idx]])
            stack.extend(reversed(nxt))

    paths = []
    output = [None] * len(dest)
    for i in range(len(dest)):
        path = [[], []]
        recurse((i, 1), path)
        if len(path[1]) == 1:
            if len(path[0]) == 1:
                output[i] = src[path[0][0]][1]
            elif len(path[0]) > 1:
                paths.append(path)
        elif len(path[1]) > 0:
            paths.append(path)

    default = f32(f32(radius) * f32(radius))
    for path in paths:
        used = {}
        nulls = set()
        state = {"score": f32(0.0), "best": f32(np.inf)}

        def recurse2(progress):
            cur = path[1][progress]
            cands = [(src[s][1], dist) for (s, dist) in d2s[cur]] + [(None, default)]
            for (sid, dist) in cands:
                if state["score"] > state["best"]:
                    return
                state["score"] = f32(state["score"] + dist)
                if sid is not None:
                    if sid in used:
                        state["score"] = f32(state["score"] - dist)
                        continue
                    used[sid] = cur
                else:
                    nulls.add(cur)
                if progress == len(path[1]) - 1:
                    if state["score"] < state["best"]:
                        state["best"] = state["score"]
                        for k, v in used.items():
                            output[v] = k
                        for v in nulls:
                            output[v] = None
                else:
                    recurse2(progress + 1)
                if sid is not None:
                    del used[sid]
                else:
                    nulls.discard(cur)
                state["score"] = f32(state["score"] - dist)

        recurse2(0)

    ids = []
    for o in output:
        if o is None:
            ids.append(counter)
            counter += 1
        else:
            ids.append(o)
    unused = []
    if memory_start_idx is not None:
        usedset = set(ids)
        unused = [e for e in src[:memory_start_idx] if e[1] not in usedset]
    return ids, unused, counter


class Linker:
    """Linker::advance (linking.rs:228-283)."""

    def __init__(self, search_range: float, memory: int):
        self.r = f32(search_range)
        self.memory = memory
        self.prev = []
        self.prev_n = 0
        self.memvec = [[] for _ in range(memory)] if memory > 0 else None
        self.counter = 0

    def advance(self, frame):
        n = len(frame)
        mstart = None
        if self.memvec is not None:
            present = {e[1] for e in self.prev}
            for entry in (e for m in self.memvec for e in m):
                if entry[1] not in present:
                    self.prev.append(entry)
                    present.add(entry[1])
            mstart = self.prev_n
        ids, unused, self.counter = link(self.prev, frame, mstart, self.r, self.counter)
        self.prev = [((p[0], p[1]), i) for p, i in zip(frame, ids)]
        if self.memvec is not None:
            self.memvec.append(unused)
            self.memvec.pop(0)
        self.prev_n = n
        return ids


def link_rows(rows: np.ndarray, search_range: float, memory: int = 0) -> np.ndarray:
    """link_rust (lib.rs:202-212): particle id of every row of a
    [frame, mass, y, x, ...] float32 array."""
    lk = Linker(search_range, memory)
    out = []
    for step in frame_subsets(np.asarray(rows, np.float32)):
        out.extend(lk.advance(step))
    return np.asarray(out, np.uint64)
