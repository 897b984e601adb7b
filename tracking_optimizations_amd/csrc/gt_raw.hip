// gt_raw.hip -- raw signal extraction (execute_gpu.rs:301-333) on the device.
//
// For every output row of a batch, centred on (round(y), round(x)) (f32
// round, half away from zero), on the raw input frame as float32:
//   raw_mass            sum of the signal-disk pixels (circle_inds,
//                       kernels.rs:149-164), one after the other in tap
//                       order in binary32, taps outside the frame skipped
//                       (fold_axis_skipnan);
//   raw_bg_median       median of the background-annulus pixels inside the
//                       frame (annulus_inds, kernels.rs:166-181), with the
//                       linear interpolation of ndarray-stats 0.5.1
//                       (quantile_axis_skipnan_mut, Linear): q = 0.5 of the
//                       n values in order, index 0.5 (n - 1) in binary64,
//                       a + f32(fract * (b - a)) with a, b the order
//                       statistics at floor and ceil of the index; NaN when
//                       no tap is inside;
//   raw_mass_corrected  raw_mass - raw_bg_median * n_signal_taps (binary32,
//                       no fused multiply-add).
// One wave per row: the taps are gathered into LDS by all lanes; lane 0 adds
// the signal taps in order; the background values are compacted (outside
// taps dropped), padded to a power of two with +inf and sorted by a bitonic
// network, which yields the same order statistics as any selection.
#include <math.h>

#include <algorithm>

#include "gt_internal.h"

#pragma clang fp contract(off)

namespace gt {

namespace {

__device__ __forceinline__ uint64_t lanemask_below(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ float raw_pixel(const Frame& fr, int in_type, int y, int x) {
    const size_t i = (size_t)y * fr.W + (size_t)x;
    if (in_type == IN_U16) return (float)((const uint16_t*)fr.in)[i];
    if (in_type == IN_U8) return (float)((const uint8_t*)fr.in)[i];
    return ((const float*)fr.in)[i];
}

constexpr int RAW_ROW_BLOCKS = 512;  // waves per frame, rows strided over them

// Ascending bitonic sort of 64 * VPL values held by one wave, element
// k * 64 + lane in v[k] of that lane: stages across lanes exchange through
// shuffles, stages across registers swap within the lane.  Every stage index
// is a compile-time constant once the loops are unrolled.
template <int VPL>
__device__ __forceinline__ void wave_bitonic(float (&v)[VPL], int lane) {
    constexpr int PP = 64 * VPL;
#pragma unroll
    for (int kk = 2; kk <= PP; kk <<= 1) {
#pragma unroll
        for (int j = kk >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const int jj = j >> 6;
#pragma unroll
                for (int k = 0; k < VPL; ++k) {
                    if ((k & jj) == 0) {
                        const bool up = ((k * 64 + lane) & kk) == 0;
                        const float a = v[k], c = v[k ^ jj];
                        const bool swap = up ? (a > c) : (a < c);
                        v[k] = swap ? c : a;
                        v[k ^ jj] = swap ? a : c;
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < VPL; ++k) {
                    const float x = v[k];
                    const float y = __shfl_xor(x, j, 64);
                    const bool lower = (lane & j) == 0;
                    const bool up = ((k * 64 + lane) & kk) == 0;
                    const float lo = x < y ? x : y, hi = x < y ? y : x;
                    v[k] = (lower == up) ? lo : hi;
                }
            }
        }
    }
}

template <int VPL>
__device__ __forceinline__ void sort_regs(float* s_v, int lane) {
    float v[VPL];
#pragma unroll
    for (int k = 0; k < VPL; ++k) v[k] = s_v[k * 64 + lane];
    wave_bitonic<VPL>(v, lane);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < VPL; ++k) s_v[k * 64 + lane] = v[k];
}

__global__ __launch_bounds__(64) void k_raw(Batch b, const Geom* __restrict__ g) {
    extern __shared__ float s_v[];
    const uint32_t f = blockIdx.x;
    const Frame fr = b.frames[f];
    const int H = (int)fr.H, W = (int)fr.W;
    const int lane = threadIdx.x;
    const int ns = g->raw_ns, nb = g->raw_nb, nc = g->ncols, col = g->col_raw;
    const bool with_bg = g->raw_cols == 3;
    const int8_t* tap = g->raw_tap;
    const uint64_t r0 = b.frame_row_base[f];
    const uint64_t r1 = min(r0 + (uint64_t)b.frame_rows[f], b.rows_cap);
    for (uint64_t r = r0 + blockIdx.y; r < r1; r += gridDim.y) {
        float* o = b.rows + r * (uint64_t)nc;
        const int cy = (int)roundf(o[2]), cx = (int)roundf(o[3]);
        for (int t = lane; t < ns; t += 64) {
            const int yy = cy + tap[2 * t], xx = cx + tap[2 * t + 1];
            s_v[t] = (yy >= 0 && yy < H && xx >= 0 && xx < W) ? raw_pixel(fr, b.in_type, yy, xx)
                                                              : __builtin_nanf("");
        }
        __syncthreads();
        float sig = 0.f;
        if (lane == 0) {
            for (int t = 0; t < ns; ++t) {
                const float v = s_v[t];
                if (!isnan(v)) sig = sig + v;
            }
            o[col] = sig;
        }
        __syncthreads();
        if (with_bg) {
            const int8_t* bt = tap + 2 * ns;
            int n = 0;
            for (int t0 = 0; t0 < nb; t0 += 64) {
                const int t = t0 + lane;
                float v = __builtin_nanf("");
                if (t < nb) {
                    const int yy = cy + bt[2 * t], xx = cx + bt[2 * t + 1];
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = raw_pixel(fr, b.in_type, yy, xx);
                }
                const bool ok = !isnan(v);
                const uint64_t m = __ballot(ok);
                if (ok) s_v[n + __popcll(m & lanemask_below(lane))] = v;
                n += __popcll(m);
            }
            int P = 64;
            while (P < n) P <<= 1;
            for (int t = n + lane; t < P; t += 64) s_v[t] = __builtin_inff();
            __syncthreads();
            if (P <= 512) {   // in registers (wave-uniform choice)
                if (P == 64) sort_regs<1>(s_v, lane);
                else if (P == 128) sort_regs<2>(s_v, lane);
                else if (P == 256) sort_regs<4>(s_v, lane);
                else sort_regs<8>(s_v, lane);
                __syncthreads();
            } else {
                for (int k = 2; k <= P; k <<= 1) {
                    for (int j = k >> 1; j > 0; j >>= 1) {
                        for (int i = lane; i < P; i += 64) {
                            const int l = i ^ j;
                            if (l > i) {
                                const float a = s_v[i], c = s_v[l];
                                if (((i & k) == 0) ? (a > c) : (a < c)) {
                                    s_v[i] = c;
                                    s_v[l] = a;
                                }
                            }
                        }
                        __syncthreads();
                    }
                }
            }
            if (lane == 0) {
                float med = __builtin_nanf("");
                if (n > 0) {
                    const double idx = 0.5 * (double)(n - 1);
                    const double lo = floor(idx), hi = ceil(idx);
                    const double frac = idx - trunc(idx);
                    const float a = s_v[(int)lo], c = s_v[(int)hi];
                    med = a + (float)(frac * ((double)c - (double)a));
                }
                o[col + 1] = med;
                o[col + 2] = sig - med * (float)ns;
            }
            __syncthreads();
        }
    }
}

}  // namespace

hipError_t launch_raw(const Batch& b, const Geom& g, const Geom* d_geom, hipStream_t s) {
    int P = 64;
    while (P < g.raw_nb) P <<= 1;
    const size_t lds = (size_t)std::max(g.raw_ns, g.raw_cols == 3 ? P : 0) * 4 + 16;
    hipLaunchKernelGGL(k_raw, dim3(b.nframes, RAW_ROW_BLOCKS), dim3(64), lds, s, b, d_geom);
    return hipGetLastError();
}

}  // namespace gt
