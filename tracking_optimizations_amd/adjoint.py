"""Localisation errors and gradients of the locate path.

``locate_errors(video, params)`` returns the rows of ``locate`` together with
each particle's Poisson localisation error (sigma_y, sigma_x, sigma_mass):
with the walk's discrete decisions held fixed, the Jacobian J of (y, x, mass)
with respect to the input pixels (refine and bandpass pulled back by hand on
the GPU) gives sigma^2 = sum_r J[r]^2 * lambda[r], lambda = max(I, 0).

``LocateFunction`` is a ``torch.autograd.Function``: its forward returns the
rows (frame, mass, y, x, ...) of a float32/uint16/uint8 movie on the GPU, its
backward maps d(loss)/d(mass, y, x) of every row to d(loss)/d(pixels) through
the same Jacobians (gt_adjoint_vjp, added in row order: deterministic).  The
candidate set and the walk are piecewise constant in the input, so the
gradient is the derivative of the smooth piece the input lies on.

Neither exists in the reference (execute_gpu.rs:496 returns rows only); they
extend its locate entry point.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import GT_ADJ_JACOBIAN, GT_FRAMES_ON_DEVICE, Adjoint, check, default_context, load
from .gpu_tracking import _frames_arg

__all__ = ["locate_errors", "LocateFunction", "locate_autograd"]


def _adjoint(video, params, jacobian: bool, ctx=None, device: int = 0) -> Adjoint:
    ptr, n, h, w, dt, on_dev, keep = _frames_arg(video)
    ctx = ctx or default_context(device)
    flags = (GT_FRAMES_ON_DEVICE if on_dev else 0) | (GT_ADJ_JACOBIAN if jacobian else 0)
    out = ctypes.c_void_p()
    check(load().gt_locate_adjoint(ctx.handle, ctypes.c_void_p(ptr), _lib.DTYPE_CODES[dt], n, h, w,
                                   ctypes.byref(params), flags, ctypes.byref(out)))
    del keep
    return Adjoint(out, n, h, w)


def locate_errors(video, params, ctx=None, device: int = 0, jacobian: bool = False):
    """(rows [N, C], candidates [N], sigma [N, 3] as (sigma_y, sigma_x,
    sigma_mass), Adjoint).  With jacobian=True the Adjoint also holds the
    per-row Jacobian patches (Adjoint.jacobian(), Adjoint.origin())."""
    a = _adjoint(video, params, jacobian, ctx, device)
    return a.rows(), a.candidates(), a.sigma(), a


try:
    import torch

    class LocateFunction(torch.autograd.Function):
        """rows = LocateFunction.apply(video, params): video is a CUDA tensor
        (frames, h, w) of float32 (differentiable), uint16 or uint8."""

        @staticmethod
        def forward(ctx_, video, params):
            if not video.is_cuda:
                raise _lib.GtInvalid(_lib.GT_E_INVALID, "LocateFunction needs a GPU tensor")
            video = video.contiguous()
            # the library runs on its own streams: finish torch's work first
            torch.cuda.synchronize(video.device)
            a = _adjoint(video.detach(), params, jacobian=True,
                         device=video.device.index or 0)
            ctx_.adj = a
            ctx_.shape = tuple(video.shape)
            ctx_.device = video.device
            return torch.from_numpy(a.rows()).to(video.device)

        @staticmethod
        def backward(ctx_, grad_rows):
            a = ctx_.adj
            # columns 1, 2, 3 are mass, y, x; the kernel takes (y, x, mass)
            g = grad_rows[:, [2, 3, 1]].to(torch.float32).contiguous()
            out = torch.empty(ctx_.shape, dtype=torch.float32, device=ctx_.device)
            lctx = default_context(ctx_.device.index or 0)
            torch.cuda.synchronize(ctx_.device)
            a.vjp(lctx, g.data_ptr(), out.data_ptr(), on_device=True)
            return out, None

    def locate_autograd(video, params):
        """Rows of locate as a differentiable function of the pixels."""
        return LocateFunction.apply(video, params)

except ImportError:  # torch is optional for the rest of the package
    LocateFunction = None

    def locate_autograd(video, params):
        raise _lib.GtNotImplemented(_lib.GT_E_NOT_IMPLEMENTED, "torch is not available")
