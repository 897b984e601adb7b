"""gt_locate_noalloc (no device allocation: every workspace in the library's
fixed scratch array) against gt_locate_u16 with device frames: identical
rows and candidates on two calls in a row, the second with a smaller input,
then a call that needs several scratch batches, a 2048x2048 frame, a row
capacity that is too small and a frame too large for the scratch."""
import ctypes

import numpy as np
import pytest

from tracking_optimizations_amd import _lib, synth
from tracking_optimizations_amd.gpu_tracking import make_params

pytestmark = pytest.mark.gpu

KW = dict(diameter=9, minmass=1000, separation=9, percentile=64)


@pytest.fixture(scope="module")
def ctx():
    return _lib.Context(0)


def _existing(ctx, movie, n, h, w, p):
    out = ctypes.c_void_p()
    _lib.check(_lib.load().gt_locate_u16(ctx.handle, ctypes.c_void_p(movie.ptr), n, h, w,
                                         ctypes.byref(p), _lib.GT_FRAMES_ON_DEVICE,
                                         ctypes.byref(out)))
    r = _lib.Result(out)
    return r.rows(), r.candidates()


def _noalloc(ctx, movie, n, h, w, p, cap, ncols):
    rows = synth.DeviceBuffer(ctx, max(1, cap) * ncols * 4)
    cand = synth.DeviceBuffer(ctx, max(1, cap) * 8)
    nrows = ctypes.c_uint64()
    st = _lib.load().gt_locate_noalloc(ctx.handle, ctypes.c_void_p(movie.ptr), 0, n, h, w,
                                       ctypes.byref(p), ctypes.c_void_p(rows.ptr),
                                       ctypes.c_void_p(cand.ptr), cap, ctypes.byref(nrows))
    k = min(int(nrows.value), cap)
    r = rows.to_host(np.float32, (k, ncols))
    c = cand.to_host(np.int64, (k,))
    rows.free()
    cand.free()
    return st, int(nrows.value), r, c


def _compare(ctx, spec, kw, label):
    p = make_params(**kw)
    movie = synth.generate(ctx, spec)
    n, h, w = spec.n_frames, spec.h, spec.w
    ref_rows, ref_cand = _existing(ctx, movie, n, h, w, p)
    print(f"[no-alloc] {label}: {n} x {h}x{w}, {len(ref_cand)} rows", flush=True)
    st, nrows, rows, cand = _noalloc(ctx, movie, n, h, w, p, len(ref_cand) + 100,
                                     ref_rows.shape[1])
    movie.free()
    assert st == _lib.GT_OK, _lib.load().gt_last_error()
    assert nrows == len(ref_cand)
    assert np.array_equal(cand, ref_cand), label
    assert np.array_equal(rows, ref_rows), label


def test_two_calls_match_existing_entry_point(ctx):
    _compare(ctx, synth.single_512(seed=1, n_frames=4), KW, "first call")
    small = synth.single_512(seed=2, n_frames=1)
    small.h, small.w = 256, 384          # spots beyond the crop are sampled away
    keep = (small.spots[:, 0] < 250) & (small.spots[:, 1] < 378)
    small.spots = small.spots[keep]
    small.spot_off = np.array([0, int(keep.sum())], np.uint64)
    _compare(ctx, small, KW, "second call, smaller input")


def test_several_scratch_batches(ctx):
    # about 7 MB of scratch per 512x512 frame: 40 frames take three batches
    _compare(ctx, synth.single_512(seed=3, n_frames=40), KW, "three batches")


def test_full_size_frame(ctx):
    kw = dict(diameter=11, minmass=0, separation=11, smoothing_size=11, percentile=64)
    _compare(ctx, synth.movie_2048(seed=2, n_frames=2), kw, "2048x2048")


def test_capacity_and_size_errors(ctx):
    spec = synth.single_512(seed=1, n_frames=2)
    p = make_params(**KW)
    movie = synth.generate(ctx, spec)
    ref_rows, ref_cand = _existing(ctx, movie, 2, 512, 512, p)
    st, nrows, rows, cand = _noalloc(ctx, movie, 2, 512, 512, p, 50, ref_rows.shape[1])
    assert st == _lib.GT_E_INVALID and nrows == len(ref_cand)
    assert np.array_equal(rows, ref_rows[:50]) and np.array_equal(cand, ref_cand[:50])
    movie.free()
    # a 4096x4096 frame needs more than the scratch holds: rejected before any
    # launch (the buffers are real, the frame one sized for the whole frame)
    big = synth.DeviceBuffer(ctx, 4096 * 4096 * 2)
    out = synth.DeviceBuffer(ctx, 1024)
    nrows = ctypes.c_uint64()
    st = _lib.load().gt_locate_noalloc(ctx.handle, ctypes.c_void_p(big.ptr), 0, 1, 4096, 4096,
                                       ctypes.byref(p), ctypes.c_void_p(out.ptr),
                                       ctypes.c_void_p(out.ptr), 1, ctypes.byref(nrows))
    big.free()
    out.free()
    assert st == _lib.GT_E_INVALID
    assert b"scratch" in _lib.load().gt_last_error()
