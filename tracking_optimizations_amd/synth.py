"""Seeded synthetic movies of Gaussian spots on a Poisson background.

Spot parameters (positions, amplitudes, widths, Brownian tracks) are drawn on
the host with numpy's PCG64 generator; pixels are sampled on the device by
``gt_synth_u16`` (Philox-4x32-10 streams, PTRS Poisson sampling).  The same
(spec, seed) always gives the same movie.

Workloads (sizes and parameters of the benchmark configurations):
  * ``single_512``:  512x512, bg 100, 200 spots, sigma 1.5, amp U(500, 2000),
    spots at least 10 px apart;
  * ``movie_2048``:  2048x2048, bg 150, 5000 spots, sigma 2.0, amp
    lognormal(median 1200, s 0.5), Brownian D = 0.5 px^2/frame (step sd 1 px
    per axis), reflecting walls;
  * ``uneven``:      mixed 256x256 / 512x1024 / 2048x2048 frames (40/40/20 %),
    one spot per 800 px^2, sigma U(1.2, 2.5), bg U(50, 300) per frame;
  * ``adjoint_512``: 512x512, bg 100, 300 spots, sigma 1.5, amp U(200, 3000);
  * ``link_1024``:   1024x1024, bg 120, 3000 spots, sigma 1.8, amp lognormal
    (median 1200, s 0.5), Brownian D = 0.5, each spot off in a frame with
    probability 0.05 independently (``MovieSpec.track`` holds the spot index
    of every drawn spot).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, load


@dataclass
class MovieSpec:
    h: int
    w: int
    spots: np.ndarray      # (n_spots, 4) float32: y, x, amp, sigma
    spot_off: np.ndarray   # (n_frames + 1,) uint64
    bg: np.ndarray         # (n_frames,) float32
    seed: int
    track: np.ndarray | None = None   # (n_spots,) spot index of each drawn spot

    @property
    def n_frames(self) -> int:
        return len(self.bg)


def _dart_throw(rng, n, h, w, margin, min_dist):
    pts = []
    tries = 0
    while len(pts) < n and tries < 200 * n:
        tries += 1
        y = rng.uniform(margin, h - 1 - margin)
        x = rng.uniform(margin, w - 1 - margin)
        if all((y - a) ** 2 + (x - b) ** 2 >= min_dist ** 2 for a, b in pts):
            pts.append((y, x))
    return np.array(pts, np.float64).reshape(-1, 2)


def single_512(seed: int = 1, n_frames: int = 1) -> MovieSpec:
    rng = np.random.default_rng(seed)
    spots, off = [], [0]
    for _ in range(n_frames):
        p = _dart_throw(rng, 200, 512, 512, 5.0, 10.0)
        amp = rng.uniform(500, 2000, len(p))
        spots.append(np.column_stack([p, amp, np.full(len(p), 1.5)]))
        off.append(off[-1] + len(p))
    return MovieSpec(512, 512, np.concatenate(spots).astype(np.float32), np.array(off, np.uint64),
                     np.full(n_frames, 100.0, np.float32), seed)


def adjoint_512(seed: int = 4, n_frames: int = 256) -> MovieSpec:
    rng = np.random.default_rng(seed)
    spots, off = [], [0]
    for _ in range(n_frames):
        p = _dart_throw(rng, 300, 512, 512, 5.0, 10.0)
        amp = rng.uniform(200, 3000, len(p))
        spots.append(np.column_stack([p, amp, np.full(len(p), 1.5)]))
        off.append(off[-1] + len(p))
    return MovieSpec(512, 512, np.concatenate(spots).astype(np.float32), np.array(off, np.uint64),
                     np.full(n_frames, 100.0, np.float32), seed)


def movie_2048(seed: int = 2, n_frames: int = 1000, n_spots: int = 5000, size: int = 2048,
               sigma: float = 2.0, bg: float = 150.0, D: float = 0.5) -> MovieSpec:
    rng = np.random.default_rng(seed)
    pos = np.column_stack([rng.uniform(0, size - 1, n_spots), rng.uniform(0, size - 1, n_spots)])
    amp = np.exp(rng.normal(np.log(1200.0), 0.5, n_spots))
    step = np.sqrt(2 * D)
    frames = []
    for _ in range(n_frames):
        frames.append(np.column_stack([pos, amp, np.full(n_spots, sigma)]))
        pos = pos + rng.normal(0, step, pos.shape)
        # reflecting walls on [0, size-1]
        pos = np.where(pos < 0, -pos, pos)
        pos = np.where(pos > size - 1, 2 * (size - 1) - pos, pos)
    spots = np.concatenate(frames).astype(np.float32)
    off = np.arange(n_frames + 1, dtype=np.uint64) * n_spots
    return MovieSpec(size, size, spots, off, np.full(n_frames, bg, np.float32), seed)


def link_1024(seed: int = 5, n_frames: int = 2000, n_spots: int = 3000, size: int = 1024,
              sigma: float = 1.8, bg: float = 120.0, D: float = 0.5,
              p_off: float = 0.05) -> MovieSpec:
    """Brownian spots that blink off independently per frame."""
    rng = np.random.default_rng(seed)
    pos = np.column_stack([rng.uniform(0, size - 1, n_spots), rng.uniform(0, size - 1, n_spots)])
    amp = np.exp(rng.normal(np.log(1200.0), 0.5, n_spots))
    step = np.sqrt(2 * D)
    frames, tracks, off = [], [], [0]
    idx = np.arange(n_spots)
    for _ in range(n_frames):
        on = rng.random(n_spots) >= p_off
        frames.append(np.column_stack([pos[on], amp[on], np.full(int(on.sum()), sigma)]))
        tracks.append(idx[on])
        off.append(off[-1] + int(on.sum()))
        pos = pos + rng.normal(0, step, pos.shape)
        pos = np.where(pos < 0, -pos, pos)
        pos = np.where(pos > size - 1, 2 * (size - 1) - pos, pos)
    return MovieSpec(size, size, np.concatenate(frames).astype(np.float32),
                     np.array(off, np.uint64), np.full(n_frames, bg, np.float32), seed,
                     np.concatenate(tracks))


def uneven_shapes(seed: int = 3, n_frames: int = 4096):
    """Shapes of the uneven batch: 40 % 256x256, 40 % 512x1024, 20 % 2048x2048,
    shuffled."""
    n1 = int(round(0.4 * n_frames))
    n2 = int(round(0.4 * n_frames))
    n3 = n_frames - n1 - n2
    shapes = [(256, 256)] * n1 + [(512, 1024)] * n2 + [(2048, 2048)] * n3
    rng = np.random.default_rng(seed)
    order = rng.permutation(n_frames)
    return [shapes[i] for i in order]


def uneven_frame_spec(shape, rng) -> MovieSpec:
    h, w = shape
    n = int(round(h * w / 800.0))
    p = np.column_stack([rng.uniform(0, h - 1, n), rng.uniform(0, w - 1, n)])
    amp = np.exp(rng.normal(np.log(1200.0), 0.5, n))
    sig = rng.uniform(1.2, 2.5, n)
    bg = rng.uniform(50, 300)
    return MovieSpec(h, w, np.column_stack([p, amp, sig]).astype(np.float32),
                     np.array([0, n], np.uint64), np.array([bg], np.float32),
                     int(rng.integers(0, 2**63 - 1)))


class DeviceBuffer:
    """Device memory owned through the C ABI (no torch needed)."""

    def __init__(self, ctx: "_lib.Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(load().gt_device_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def free(self) -> None:
        if self.ptr:
            load().gt_device_free(self.ctx.handle, ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def to_host(self, dtype, shape, offset_bytes: int = 0) -> np.ndarray:
        out = np.empty(shape, dtype)
        check(load().gt_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p),
                                   ctypes.c_void_p(self.ptr + offset_bytes), out.nbytes))
        return out

    def from_host(self, arr: np.ndarray, offset_bytes: int = 0) -> None:
        a = np.ascontiguousarray(arr)
        check(load().gt_memcpy_h2d(self.ctx.handle, ctypes.c_void_p(self.ptr + offset_bytes),
                                   a.ctypes.data_as(ctypes.c_void_p), a.nbytes))


def generate(ctx: "_lib.Context", spec: MovieSpec, out: DeviceBuffer | None = None,
             offset_bytes: int = 0) -> DeviceBuffer:
    """Sample the movie on the device as uint16 (n_frames, h, w)."""
    nbytes = spec.n_frames * spec.h * spec.w * 2
    if out is None:
        out = DeviceBuffer(ctx, nbytes)
    spots = np.ascontiguousarray(spec.spots, np.float32)
    off = np.ascontiguousarray(spec.spot_off, np.uint64)
    bg = np.ascontiguousarray(spec.bg, np.float32)
    check(load().gt_synth_u16(ctx.handle, ctypes.c_void_p(out.ptr + offset_bytes), spec.n_frames,
                              spec.h, spec.w, spots.ctypes.data_as(ctypes.c_void_p),
                              off.ctypes.data_as(ctypes.c_void_p),
                              bg.ctypes.data_as(ctypes.c_void_p), spec.seed))
    return out


def generate_host(ctx: "_lib.Context", spec: MovieSpec) -> np.ndarray:
    buf = generate(ctx, spec)
    try:
        return buf.to_host(np.uint16, (spec.n_frames, spec.h, spec.w))
    finally:
        buf.free()


class UnevenMovie:
    """Device-resident uneven batch: frames of each shape live in one device
    buffer per shape; `descs` lists (ptr, h, w) in the shuffled frame order."""

    def __init__(self, ctx: "_lib.Context", n_frames: int = 4096, seed: int = 3):
        shapes = uneven_shapes(seed, n_frames)
        rng = np.random.default_rng(seed + 1)
        by_shape: dict = {}
        for i, s in enumerate(shapes):
            by_shape.setdefault(s, []).append(i)
        self.buffers = []
        self.descs = [None] * n_frames
        self.pixels = 0
        for (h, w), idxs in sorted(by_shape.items()):
            specs = [uneven_frame_spec((h, w), rng) for _ in idxs]
            spots = np.concatenate([s.spots for s in specs]).astype(np.float32)
            counts = [len(s.spots) for s in specs]
            off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
            bg = np.concatenate([s.bg for s in specs]).astype(np.float32)
            spec = MovieSpec(h, w, spots, off, bg, seed * 1000 + h * 7 + w)
            buf = generate(ctx, spec)
            self.buffers.append(buf)
            for k, fi in enumerate(idxs):
                self.descs[fi] = (buf.ptr + k * h * w * 2, h, w)
                self.pixels += h * w
        self.shapes = shapes

    def frame_descs(self):
        arr = (_lib.GtFrameDesc * len(self.descs))()
        for i, (p, h, w) in enumerate(self.descs):
            arr[i].data = p
            arr[i].h = h
            arr[i].w = w
            arr[i].frame_index = i
        return arr

    def free(self) -> None:
        for b in self.buffers:
            b.free()
