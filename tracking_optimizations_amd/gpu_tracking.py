"""Python API of the MI355X locate path, mirroring the reference package
``gpu_tracking`` (gpu-tracking/gpu_tracking/lib.py:6-106 and the PyO3 entry
points of gpu-tracking/src/lib.rs:46-200).

``batch(video, diameter, ...)`` and ``batch_file(path, diameter, ...)`` take the
reference's keyword arguments with the same defaults and return a DataFrame
with the reference's columns and dtypes (execute_gpu.rs:197-226).

Differences, all supersets:
  * ``percentile`` is implemented (the reference panics, lib.rs:73-74): when
    given, a local maximum is kept only when its bandpassed value is strictly
    above the exact ``percentile``-th percentile (linear interpolation) of all
    bandpassed pixels of its frame;
  * ``video`` may be float32 (the reference's type), uint16 or uint8, as a
    numpy array or a torch tensor (device tensors are read in place);
  * errors raise ``GtInvalid`` (a ValueError) or ``GtNotImplemented`` (a
    NotImplementedError) instead of a Rust panic.
``search_range``/``memory`` link the particles frame to frame as
execute_gpu.rs:335 does (a ``particle`` column after ``x``), and
``link_rust(array, search_range, memory)`` links a row table as lib.rs:202-212
does.  ``sig_radius``/``bg_radius``/``gap_radius`` add the raw signal columns
(execute_gpu.rs:301-333), computed on the device; ``bg_radius`` needs
``sig_radius``.  Not implemented (raise GtNotImplemented): the parameters the
reference itself rejects (maxsize, threshold, invert, topn, preprocess).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import (GT_FRAMES_ON_DEVICE, GT_PLAIN, GtError, GtInvalid, GtNotImplemented,
                   GtParams, Result, check, default_context, load)

__all__ = ["batch", "batch_rust", "batch_file", "batch_file_rust", "make_params",
           "locate", "column_names", "link_rust", "parse_ets", "read_movie", "MovieFile",
           "GtError", "GtInvalid", "GtNotImplemented"]


def make_params(diameter, minmass=None, maxsize=None, separation=None, noise_size=None,
                smoothing_size=None, threshold=None, invert=None, percentile=None, topn=None,
                preprocess=None, max_iterations=None, characterize=None, filter_close=None,
                search_range=None, memory=None, cpu_processed=None, sig_radius=None,
                bg_radius=None, gap_radius=None) -> GtParams:
    """Build gt_params from batch_rust's keyword arguments (lib.rs:77-91)."""
    if diameter is None or int(diameter) != diameter or diameter < 0:
        raise GtInvalid(_lib.GT_E_INVALID, "diameter must be a non-negative integer")
    p = GtParams()
    check(load().gt_params_default(ctypes.byref(p), int(diameter)))
    if minmass is not None:
        p.minmass = float(minmass)
    if maxsize is not None:
        p.has_maxsize, p.maxsize = 1, float(maxsize)
    if separation is not None:
        p.separation = int(separation)
    if noise_size is not None:
        p.noise_size = float(noise_size)
    if smoothing_size is not None:
        p.smoothing_size = int(smoothing_size)
    if threshold is not None:
        p.has_threshold, p.threshold = 1, float(threshold)
    if invert is not None:
        p.has_invert, p.invert = 1, int(bool(invert))
    if percentile is not None:
        p.has_percentile, p.percentile = 1, float(percentile)
    if topn is not None:
        p.has_topn, p.topn = 1, int(topn)
    if preprocess is not None:
        p.has_preprocess, p.preprocess = 1, int(bool(preprocess))
    if max_iterations is not None:
        p.max_iterations = int(max_iterations)
    if characterize is not None:
        p.characterize = int(bool(characterize))
    if filter_close is not None:
        p.filter_close = int(bool(filter_close))
    if search_range is not None:
        p.has_search_range, p.search_range = 1, float(search_range)
    if memory is not None:
        p.has_memory, p.memory = 1, int(memory)
    if cpu_processed is not None:
        p.cpu_processed = int(bool(cpu_processed))
    if sig_radius is not None:
        p.has_sig_radius, p.sig_radius = 1, float(sig_radius)
    if bg_radius is not None:
        p.has_bg_radius, p.bg_radius = 1, float(bg_radius)
        # lib.rs:91: gap_radius defaults to 0 when bg_radius is given
        p.has_gap_radius, p.gap_radius = 1, float(gap_radius if gap_radius is not None else 0.0)
    elif gap_radius is not None:
        p.has_gap_radius, p.gap_radius = 1, float(gap_radius)
    return p


def column_names(p: GtParams) -> list[tuple[str, str]]:
    names = (ctypes.c_char_p * 16)()
    dts = (ctypes.c_char_p * 16)()
    n = ctypes.c_uint32()
    check(load().gt_column_names(ctypes.byref(p), names, dts, 16, ctypes.byref(n)))
    return [(names[i].decode(), dts[i].decode()) for i in range(n.value)]


def _frames_arg(video):
    """(pointer, n, h, w, dtype, on_device, keepalive) for a 3-D movie."""
    try:
        import torch  # noqa: F401
        is_torch = type(video).__module__.startswith("torch")
    except ImportError:
        is_torch = False
    if is_torch:
        import torch
        if video.dim() == 2:
            video = video.unsqueeze(0)
        if video.dim() != 3:
            raise GtInvalid(_lib.GT_E_INVALID, "video must be (frames, height, width)")
        if not video.is_contiguous():
            raise GtInvalid(_lib.GT_E_INVALID, "Array is not standard layout")
        dt = {torch.uint16: np.uint16, torch.uint8: np.uint8, torch.float32: np.float32}.get(video.dtype)
        if dt is None:
            raise GtInvalid(_lib.GT_E_INVALID, f"unsupported dtype {video.dtype}")
        on_dev = video.is_cuda
        if not on_dev:
            arr = video.numpy()
            return arr.ctypes.data, arr.shape[0], arr.shape[1], arr.shape[2], np.dtype(dt), False, arr
        return video.data_ptr(), video.shape[0], video.shape[1], video.shape[2], np.dtype(dt), True, video
    arr = np.asarray(video)
    if arr.ndim == 2:
        arr = arr[None]
    if arr.ndim != 3:
        raise GtInvalid(_lib.GT_E_INVALID, "video must be (frames, height, width)")
    if not arr.flags.c_contiguous:
        # execute_gpu.rs:497-499 panics on non-standard layout
        raise GtInvalid(_lib.GT_E_INVALID, "Array is not standard layout")
    if arr.dtype not in _lib.DTYPE_CODES:
        if arr.dtype.kind == "f":
            raise GtInvalid(_lib.GT_E_INVALID, "float video must be float32")
        raise GtInvalid(_lib.GT_E_INVALID, f"unsupported dtype {arr.dtype}")
    return arr.ctypes.data, arr.shape[0], arr.shape[1], arr.shape[2], arr.dtype, False, arr


def locate(video, params: GtParams, plain: bool = False, device: int = 0, ctx=None):
    """Run the locate path; returns (rows float32 [N, C], candidate pixel index
    int64 [N], columns, timing dict)."""
    ptr, n, h, w, dt, on_dev, keep = _frames_arg(video)
    ctx = ctx or default_context(device)
    flags = (GT_FRAMES_ON_DEVICE if on_dev else 0) | (GT_PLAIN if plain else 0)
    fn = {np.dtype(np.uint16): load().gt_locate_u16, np.dtype(np.uint8): load().gt_locate_u8,
          np.dtype(np.float32): load().gt_locate_f32}[dt]
    out = ctypes.c_void_p()
    if on_dev:
        import torch
        torch.cuda.current_stream(video.device).synchronize()
    check(fn(ctx.handle, ctypes.c_void_p(ptr), n, h, w, ctypes.byref(params), flags, ctypes.byref(out)))
    del keep
    res = Result(out)
    return res.rows(), res.candidates(), res.columns(), ctx.timing()


def link_rust(array, search_range, memory=None, device: int = 0, ctx=None) -> np.ndarray:
    """lib.rs:202-212: particle id (uint64) of every row of a float32 table
    whose column 0 is the frame and columns 2, 3 are y, x; steps are runs of
    equal frame values (FrameSubsetter, linking.rs:12-61)."""
    arr = np.ascontiguousarray(np.asarray(array, dtype=np.float32))
    if arr.ndim != 2:
        raise GtInvalid(_lib.GT_E_INVALID, "link_rust takes a 2-d array")
    n, c = arr.shape
    ids = np.empty(n, np.uint64)
    ctx = ctx or default_context(device)
    check(load().gt_link(ctx.handle, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n, c,
                         float(search_range), int(memory or 0),
                         ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
    return ids


def batch_rust(video, diameter, minmass=None, maxsize=None, separation=None, noise_size=None,
               smoothing_size=None, threshold=None, invert=None, percentile=None, topn=None,
               preprocess=None, max_iterations=None, characterize=None, filter_close=None,
               search_range=None, memory=None, cpu_processed=None, sig_radius=None,
               bg_radius=None, gap_radius=None):
    """lib.rs:46-121: returns (float32 array [N, C], [(name, dtype), ...])."""
    p = make_params(diameter, minmass, maxsize, separation, noise_size, smoothing_size, threshold,
                    invert, percentile, topn, preprocess, max_iterations, characterize,
                    filter_close, search_range, memory, cpu_processed, sig_radius, bg_radius,
                    gap_radius)
    check(load().gt_params_validate(ctypes.byref(p), 0, 0))
    rows, _cand, cols, _t = locate(video, p)
    return rows, cols


def batch(video, diameter, minmass=None, maxsize=None, separation=None, noise_size=None,
          smoothing_size=None, threshold=None, invert=None, percentile=None, topn=None,
          preprocess=None, max_iterations=None, characterize=None, filter_close=None,
          search_range=None, memory=None, cpu_processed=None, sig_radius=None, bg_radius=None,
          gap_radius=None):
    """lib.py:6-54: locate every frame of ``video`` and return a DataFrame."""
    import pandas as pd
    arr, columns = batch_rust(video, diameter, minmass, maxsize, separation, noise_size,
                              smoothing_size, threshold, invert, percentile, topn, preprocess,
                              max_iterations, characterize, filter_close, search_range, memory,
                              cpu_processed, sig_radius, bg_radius, gap_radius)
    columns = {name: typ for name, typ in columns}
    return pd.DataFrame(arr, columns=columns).astype(columns)


class MovieFile:
    """A TIFF or ETS movie opened by the library's reader (gt_file_open)."""

    def __init__(self, path, channel: int = 0):
        self._h = ctypes.c_void_p()
        check(load().gt_file_open(os.fsencode(path), channel, ctypes.byref(self._h)))
        kind, n = ctypes.c_int(), ctypes.c_uint64()
        h, w, dt = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int()
        check(load().gt_file_info(self._h, ctypes.byref(kind), ctypes.byref(n), ctypes.byref(h),
                                  ctypes.byref(w), ctypes.byref(dt)))
        self.kind = "tiff" if kind.value == 0 else "ets"
        self.n_frames, self.shape = n.value, (h.value, w.value)
        self.dtype = np.dtype(np.uint16 if dt.value == 0 else np.uint8)

    def channels(self) -> list[int]:
        n = ctypes.c_uint32()
        check(load().gt_file_channels(self._h, None, 0, ctypes.byref(n)))
        buf = (ctypes.c_uint32 * max(n.value, 1))()
        check(load().gt_file_channels(self._h, buf, n.value, ctypes.byref(n)))
        return [int(buf[i]) for i in range(n.value)]

    def read(self, first: int = 0, count: int | None = None) -> np.ndarray:
        count = self.n_frames - first if count is None else count
        out = np.empty((count, *self.shape), self.dtype)
        check(load().gt_file_read(self._h, first, count, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def close(self) -> None:
        if self._h:
            load().gt_file_close(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_movie(path, channel: int = 0) -> np.ndarray:
    """Every frame of a TIFF (pages) or ETS (frames of `channel`) movie as an
    (n, h, w) uint16/uint8 array (decoderiter.rs)."""
    with MovieFile(path, channel) as f:
        return f.read()


def parse_ets(path) -> dict:
    """lib.rs:214-232: {channel: uint16 array (n_frames, dims[1], dims[0])}."""
    with MovieFile(path, 0xFFFFFFFF) as f:
        chans = f.channels()
    return {c: read_movie(path, c) for c in chans}


def batch_file_rust(filename, diameter, channel=None, minmass=None, maxsize=None, separation=None,
                    noise_size=None, smoothing_size=None, threshold=None, invert=None,
                    percentile=None, topn=None, preprocess=None, max_iterations=None,
                    characterize=None, filter_close=None, search_range=None, memory=None,
                    cpu_processed=None, sig_radius=None, bg_radius=None, gap_radius=None,
                    device: int = 0, ctx=None):
    """lib.rs:123-200 -> execute_file (execute_gpu.rs:510-542): decode and
    locate a TIFF or ETS movie; returns (float32 array [N, C], columns)."""
    p = make_params(diameter, minmass, maxsize, separation, noise_size, smoothing_size, threshold,
                    invert, percentile, topn, preprocess, max_iterations, characterize,
                    filter_close, search_range, memory, cpu_processed, sig_radius, bg_radius,
                    gap_radius)
    check(load().gt_params_validate(ctypes.byref(p), 0, 0))
    path = os.fsencode(filename)
    ch = 0 if channel is None else int(channel)
    MovieFile(filename, ch).close()   # reports unreadable files before a device is needed
    ctx = ctx or default_context(device)
    out = ctypes.c_void_p()
    check(load().gt_locate_file(ctx.handle, path, ch, ctypes.byref(p), 0, ctypes.byref(out)))
    res = Result(out)
    return res.rows(), res.columns()


def batch_file(path, diameter, channel=None, minmass=None, maxsize=None, separation=None,
               noise_size=None, smoothing_size=None, threshold=None, invert=None, percentile=None,
               topn=None, preprocess=None, max_iterations=None, characterize=None,
               filter_close=None, search_range=None, memory=None, cpu_processed=None,
               sig_radius=None, bg_radius=None, gap_radius=None):
    """lib.py:56-106."""
    import pandas as pd
    arr, columns = batch_file_rust(path, diameter, channel, minmass, maxsize, separation,
                                   noise_size, smoothing_size, threshold, invert, percentile, topn,
                                   preprocess, max_iterations, characterize, filter_close,
                                   search_range, memory, cpu_processed, sig_radius, bg_radius,
                                   gap_radius)
    columns = {name: typ for name, typ in columns}
    return pd.DataFrame(arr, columns=columns).astype(columns)
