"""MI355X (gfx950) implementation of the per-frame particle-locate path of
amfaber/tracking_optimizations' ``gpu_tracking`` crate.

Bandpass, local-maximum detection (with an exact per-frame percentile
threshold), centroid refinement, characterisation and separation
deduplication run as hand-written HIP kernels in ``libgpu_tracking.so``,
reached through a C ABI (include/gpu_tracking.h).  ``batch`` and
``batch_file`` mirror the reference's Python API.
"""
from .gpu_tracking import (GtError, GtInvalid, GtNotImplemented, batch, batch_file,
                           batch_file_rust, batch_rust, column_names, locate, make_params)
from ._lib import Context, default_context, device_count

__all__ = ["batch", "batch_file", "batch_rust", "batch_file_rust", "make_params", "locate",
           "column_names", "Context", "default_context", "device_count", "GtError",
           "GtInvalid", "GtNotImplemented"]
