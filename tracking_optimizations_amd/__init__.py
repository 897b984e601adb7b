"""MI355X (gfx950) implementation of the per-frame particle-locate path of
amfaber/tracking_optimizations' ``gpu_tracking`` crate.

Bandpass, local-maximum detection (with an exact per-frame percentile
threshold), centroid refinement, characterisation and separation
deduplication run as hand-written HIP kernels in ``libgpu_tracking.so``,
reached through a C ABI (include/gpu_tracking.h).  ``batch``,
``batch_file``, ``link_rust`` and ``parse_ets`` mirror the reference's Python
API; frame-to-frame linking and raw signal extraction are part of it.
"""
from .gpu_tracking import (GtError, GtInvalid, GtNotImplemented, MovieFile, batch, batch_file,
                           batch_file_rust, batch_rust, column_names, link_rust, locate,
                           make_params, parse_ets, read_movie)
from ._lib import Context, default_context, device_count

__all__ = ["batch", "batch_file", "batch_rust", "batch_file_rust", "make_params", "locate",
           "column_names", "link_rust", "parse_ets", "read_movie", "MovieFile", "Context",
           "default_context", "device_count", "GtError", "GtInvalid", "GtNotImplemented"]
