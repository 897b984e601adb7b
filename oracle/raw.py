"""CPU oracle of the raw signal extraction (gpu-tracking/src/execute_gpu.rs:
301-333, tap tables kernels.rs:149-181).

TEST INFRASTRUCTURE ONLY: the checker for the library's raw_mass,
raw_bg_median and raw_mass_corrected columns.  Only tests/ may import it.
Pure Python over numpy float32 scalars, in the reference's order:

  * circle_taps    circle_inds (kernels.rs:149-164): (i, j) over the square
                   of half-width ceil(r), row-major, kept when
                   (i*i + j*j) as f32 <= r*r (binary32);
  * annulus_taps   annulus_inds (kernels.rs:166-181): kept when
                   inner^2 <= d2 <= outer^2, inner = sig_radius + gap_radius
                   (binary32 sum);
  * raw_signal     per row: centre (round(y), round(x)) with f32 round (half
                   away from zero); taps outside the frame are NaN;
                   raw_mass = fold_axis_skipnan (sequential binary32 sum in
                   tap order, NaN skipped); raw_bg_median =
                   quantile_axis_skipnan_mut(0.5, Linear) of ndarray-stats
                   0.5.1 (a dependency absent from the reference tree; its
                   published rule: index = q (n - 1) in binary64, a and b the
                   order statistics at floor and ceil of the index,
                   a + f32(fract(index) * (b - a)), NaN for an all-NaN lane);
                   raw_mass_corrected = raw_mass - median * f32(n_sig_taps).
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32


def circle_taps(radius: float):
    r = f32(radius)
    rb = int(math.ceil(float(r)))
    r2 = f32(r * r)
    return [(i, j) for i in range(-rb, rb + 1) for j in range(-rb, rb + 1)
            if f32(i * i + j * j) <= r2]


def annulus_taps(outer: float, inner: float):
    o, n = f32(outer), f32(inner)
    rb = int(math.ceil(float(o)))
    or2, ir2 = f32(o * o), f32(n * n)
    out = []
    for i in range(-rb, rb + 1):
        for j in range(-rb, rb + 1):
            d = f32(i * i + j * j)
            if d <= or2 and d >= ir2:
                out.append((i, j))
    return out


def _round_half_away(v) -> int:
    v = float(f32(v))
    return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))


def median_linear(values):
    """ndarray-stats 0.5.1 Linear quantile at q = 0.5 of the non-NaN values."""
    vals = sorted(float(v) for v in values if not math.isnan(float(v)))
    n = len(vals)
    if n == 0:
        return f32(np.nan)
    idx = 0.5 * (n - 1)
    a, b = f32(vals[int(math.floor(idx))]), f32(vals[int(math.ceil(idx))])
    frac = idx - math.trunc(idx)
    return f32(a + f32(frac * (float(b) - float(a))))


def raw_signal(frame: np.ndarray, y, x, sig_radius: float, bg_radius=None, gap_radius=None):
    """(raw_mass, raw_bg_median, raw_mass_corrected) of one particle; the
    last two are None without bg_radius."""
    fr = np.asarray(frame, np.float32)
    H, W = fr.shape
    cy, cx = _round_half_away(y), _round_half_away(x)

    def px(i, j):
        yy, xx = cy + i, cx + j
        return fr[yy, xx] if (0 <= yy < H and 0 <= xx < W) else f32(np.nan)

    st = circle_taps(sig_radius)
    acc = f32(0.0)
    for (i, j) in st:
        v = px(i, j)
        if not np.isnan(v):
            acc = f32(acc + v)
    if bg_radius is None:
        return acc, None, None
    gap = f32(0.0 if gap_radius is None else gap_radius)
    bt = annulus_taps(bg_radius, f32(f32(sig_radius) + gap))
    med = median_linear([px(i, j) for (i, j) in bt])
    corr = f32(acc - f32(med * f32(len(st))))
    return acc, med, corr
