// gt_kernels.hip -- gfx950 kernels of the per-frame locate path.
//
// Pipeline per batch of frames (every stage is one launch over all frames of
// the batch; work items are (frame, tile) pairs):
//   zero      clear per-frame counters and histograms
//   bandpass  K1: u16/u8/f32 -> P (float32), optional percentile histogram
//             (preprocess_rows.wgsl + preprocess_cols.wgsl)
//   select1   percentile: histogram bin of the two order statistics
//   detect    K2: grey-dilation local maxima (max_rows.wgsl + walk_cols.wgsl
//             is_max) -> per-tile candidate slots; percentile value list
//   select2   percentile: exact order statistics -> threshold
//   refine    K3: walk + disk mass/centroid (centers.wgsl, walk_cols.wgsl
//             walk), minmass, characterisation
//   dedup     K4: separation suppression (execute_gpu.rs:258-284)
//   fscan     per-frame row offsets
//   emit      K5: rows in column_names order
//
// Arithmetic contract (identical to the CPU oracle): every `acc += a*b` of a
// WGSL shader is one fused multiply-add, every other operation one binary32
// operation in the shader's order; nothing else is contracted (this file is
// compiled with contraction off).  Divisions by the two per-call constants
// (gauss_norm, k*k) use a reciprocal with one residual correction only when
// the host has verified, for every significand, that it rounds exactly like
// IEEE division; otherwise they are IEEE divisions.
#include <algorithm>

#include "gt_internal.h"

#pragma clang fp contract(off)

namespace gt {

namespace {

constexpr uint32_t NEG_INF_BITS = 0xff800000u;

__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t n) {
    // Workgroups are dispatched round-robin over the 8 XCDs.  Give each XCD a
    // contiguous range of work items so that vertically and horizontally
    // neighbouring tiles (which share halo rows) hit the same L2.
    const uint32_t X = 8;
    const uint32_t xcd = bid % X, local = bid / X;
    const uint32_t q = n / X, r = n % X;
    return xcd * q + min(xcd, r) + local;
}

__device__ __forceinline__ uint32_t find_frame(const uint32_t* bases, uint32_t nframes,
                                               uint32_t item) {
    uint32_t lo = 0, hi = nframes;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (bases[mid] <= item) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ float cdiv(float x, float d, float rd, int fast) {
    if (fast) {
        const float q = x * rd;
        const float r = __builtin_fmaf(-q, d, x);
        return __builtin_fmaf(r, rd, q);
    }
    return x / d;
}

// Percentile histogram key of a non-negative float bit pattern.
__device__ __forceinline__ uint32_t pct_key(uint32_t bits) {
    if (bits == 0) return 0;
    const uint32_t k = (bits >> 19) + 1u;
    return k < (uint32_t)HIST_BINS ? k : (uint32_t)(HIST_BINS - 1);
}

template <typename T> struct IsFloat { static constexpr bool v = false; };
template <> struct IsFloat<float> { static constexpr bool v = true; };

template <typename T>
__device__ __forceinline__ float ld_px(const T* p, size_t i) { return (float)p[i]; }

__device__ __forceinline__ float raw_px(const void* in, int type, size_t i) {
    if (type == IN_U16) return (float)((const uint16_t*)in)[i];
    if (type == IN_U8) return (float)((const uint8_t*)in)[i];
    return ((const float*)in)[i];
}

__device__ __forceinline__ float wave_sum_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

}  // namespace

// ---------------------------------------------------------------------------
// zero: per-frame counters / histograms of a batch
// ---------------------------------------------------------------------------
__global__ void k_zero_batch(Batch b, int with_hist) {
    const uint64_t nh = with_hist ? (uint64_t)b.nframes * HIST_BINS : 0;
    const uint64_t total = nh + 2ull * b.nframes;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (i < nh) b.hist1[i] = 0;
        else if (i < nh + b.nframes) b.list_count[i - nh] = 0;
        else b.done[i - nh - b.nframes] = 0;
    }
}

hipError_t launch_zero_batch(const Batch& b, hipStream_t s) {
    const int with_hist = b.hist1 != nullptr;
    const uint64_t total = (with_hist ? (uint64_t)b.nframes * HIST_BINS : 0) + 2ull * b.nframes;
    uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 4096);
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_zero_batch, dim3(blocks), dim3(256), 0, s, b, with_hist);
    return hipGetLastError();
}

__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

hipError_t launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s) {
    uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 1024);
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K1 bandpass, streaming column strips.
//
// A work item is a 256-column strip of `seg` output rows of one frame.  The
// single wave walks down the strip: every input row (reflected at the frame
// edge) is staged once in LDS, each lane computes the row pass of its four
// columns (Gaussian with out-of-bounds taps skipped, boxcar over mirrored
// columns) and keeps the last 2R+1 row results in a register ring; once the
// ring is full the column pass produces one output row.  For integer input
// with k <= 15 the boxcar sums are exact integers in binary32, so the
// vertical boxcar is a running sum; otherwise both boxcar passes add the taps
// in the shader's order.
// ---------------------------------------------------------------------------
struct K1Params {
    float w[15];
    float gn, rgn, kk, rkk;
    int fast_gn, fast_kk;
    int seg;
    int hist;
    int vec;
};

template <int R, typename T>
__global__ __launch_bounds__(64) void k_bandpass(Batch b, K1Params kp) {
    constexpr int NT = 2 * R + 1;
    constexpr bool EXACT = !IsFloat<T>::v;
    constexpr int LB = K1_STRIP + 16;
    __shared__ __attribute__((aligned(16))) float rowbuf[2][LB];
    __shared__ uint32_t s_hist[HIST_BINS];

    const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t f = find_frame(b.k1_bases, b.nframes, item);
    const Frame fr = b.frames[f];
    const uint32_t local = item - fr.k1_base;
    const int H = (int)fr.H, W = (int)fr.W;
    const int sx = (int)(local % fr.k1_tx), sy = (int)(local / fr.k1_tx);
    const int x0 = sx * K1_STRIP;
    const int y0 = sy * kp.seg;
    const int y1 = min(y0 + kp.seg, H);
    const int t = threadIdx.x;
    const T* __restrict__ in = (const T*)fr.in;
    float* __restrict__ P = fr.P;
    const bool interior = (x0 - R >= 0) && (x0 + K1_STRIP + R <= W);
    const int cb = x0 + 4 * t;
    const bool vec_main = kp.vec && (x0 + K1_STRIP <= W);

    if (kp.hist) {
        for (int i = t; i < HIST_BINS; i += 64) s_hist[i] = 0;
    }

    float w[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) w[i] = kp.w[i];

    float rg[NT][4], rb[NT][4];
#pragma unroll
    for (int s = 0; s < NT; ++s)
#pragma unroll
        for (int q = 0; q < 4; ++q) { rg[s][q] = 0.f; rb[s][q] = 0.f; }
    float vbox[4] = {0.f, 0.f, 0.f, 0.f};
    uint32_t zc = 0;

    const int vr0 = y0 - R, vr1 = y1 + R;

    // prefetch registers for one input row: 4 main columns + (lanes 0..3)
    // 4 halo columns.
    float pm[4], ph[4];
    auto load_row = [&](int vr) {
        const int ar = vr < 0 ? -vr - 1 : (vr >= H ? 2 * H - vr - 1 : vr);
        const T* row = in + (size_t)ar * (size_t)W;
        if (vec_main) {
            if constexpr (sizeof(T) == 2) {
                const uint2 u = *(const uint2*)(row + cb);
                pm[0] = (float)(u.x & 0xffffu); pm[1] = (float)(u.x >> 16);
                pm[2] = (float)(u.y & 0xffffu); pm[3] = (float)(u.y >> 16);
            } else if constexpr (sizeof(T) == 1) {
                const uint32_t u = *(const uint32_t*)(row + cb);
                pm[0] = (float)(u & 0xffu); pm[1] = (float)((u >> 8) & 0xffu);
                pm[2] = (float)((u >> 16) & 0xffu); pm[3] = (float)(u >> 24);
            } else {
                const float4 u = *(const float4*)(row + cb);
                pm[0] = u.x; pm[1] = u.y; pm[2] = u.z; pm[3] = u.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int c = cb + q;
                if (c >= W) c = 2 * W - c - 1;
                c = max(0, min(c, W - 1));
                pm[q] = ld_px(row, (size_t)c);
            }
        }
        if (t < 4) {
            const int hb = (t < 2) ? (x0 - 8 + 4 * t) : (x0 + K1_STRIP + 4 * (t - 2));
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int c = hb + q;
                if (c < 0) c = -c - 1;
                else if (c >= W) c = 2 * W - c - 1;
                c = max(0, min(c, W - 1));
                ph[q] = ld_px(row, (size_t)c);
            }
        }
    };

    load_row(vr0);
    int parity = 0;
    for (int base = vr0; base < vr1; base += NT) {
#pragma unroll
        for (int s = 0; s < NT; ++s) {
            const int vr = base + s;
            if (vr < vr1) {  // wave-uniform
            float* buf = rowbuf[parity];
            parity ^= 1;
            *(float4*)(buf + 8 + 4 * t) = make_float4(pm[0], pm[1], pm[2], pm[3]);
            if (t < 4) {
                const int hi = (t < 2) ? (4 * t) : (8 + K1_STRIP + 4 * (t - 2));
                *(float4*)(buf + hi) = make_float4(ph[0], ph[1], ph[2], ph[3]);
            }
            if (vr + 1 < vr1) load_row(vr + 1);
            __syncthreads();

            float v[20];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const float4 x = *(const float4*)(buf + 4 * t + 4 * m);
                v[4 * m + 0] = x.x; v[4 * m + 1] = x.y; v[4 * m + 2] = x.z; v[4 * m + 3] = x.w;
            }
            const bool row_in = (vr >= 0) && (vr < H);
            float tg[4], tb[4];
            if (interior) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float gs = 0.f;
#pragma unroll
                    for (int i = -R; i <= R; ++i) gs = __builtin_fmaf(w[i + R], v[8 + q + i], gs);
                    tg[q] = gs;
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float gs = 0.f;
#pragma unroll
                    for (int i = -R; i <= R; ++i) {
                        const int c = cb + q + i;
                        if (c >= 0 && c < W) gs = __builtin_fmaf(w[i + R], v[8 + q + i], gs);
                    }
                    tg[q] = gs;
                }
            }
            if constexpr (EXACT) {
                float bx = 0.f;
#pragma unroll
                for (int i = -R; i <= R; ++i) bx = bx + v[8 + i];
                tb[0] = bx;
#pragma unroll
                for (int q = 1; q < 4; ++q) {
                    bx = (bx + v[8 + q + R]) - v[8 + q - 1 - R];
                    tb[q] = bx;
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float bx = 0.f;
#pragma unroll
                    for (int i = -R; i <= R; ++i) bx = bx + v[8 + q + i];
                    tb[q] = bx;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float g = row_in ? cdiv(tg[q], kp.gn, kp.rgn, kp.fast_gn) : 0.f;
                if constexpr (EXACT) vbox[q] = (vbox[q] - rb[s][q]) + tb[q];
                rg[s][q] = g;
                rb[s][q] = tb[q];
            }

            if (vr >= y0 + R) {
                const int o = vr - R;
                float out[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float gsum = 0.f;
#pragma unroll
                    for (int j = 0; j < NT; ++j)
                        gsum = __builtin_fmaf(w[j], rg[(s + 1 + j) % NT][q], gsum);
                    float box;
                    if constexpr (EXACT) {
                        box = vbox[q];
                    } else {
                        box = 0.f;
#pragma unroll
                        for (int j = 0; j < NT; ++j) box = box + rb[(s + 1 + j) % NT][q];
                    }
                    const float res = cdiv(gsum, kp.gn, kp.rgn, kp.fast_gn) -
                                      cdiv(box, kp.kk, kp.rkk, kp.fast_kk);
                    out[q] = (res < 0.f) ? 0.f : res;
                }
                float* prow = P + (size_t)o * (size_t)W;
                if (vec_main) {
                    *(float4*)(prow + cb) = make_float4(out[0], out[1], out[2], out[3]);
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (cb + q < W) prow[cb + q] = out[q];
                }
                if (kp.hist) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (cb + q < W) {
                            const uint32_t bits = __float_as_uint(out[q]);
                            if (bits == 0) zc++;
                            else atomicAdd(&s_hist[pct_key(bits)], 1u);
                        }
                    }
                }
            }
            }  // vr < vr1
        }
    }

    if (kp.hist) {
        __syncthreads();
        uint32_t* gh = b.hist1 + (size_t)f * HIST_BINS;
        for (int i = t; i < HIST_BINS; i += 64) {
            const uint32_t c = s_hist[i];
            if (c) atomicAdd(&gh[i], c);
        }
        const uint32_t z = wave_sum_u(zc);
        if (t == 0 && z) atomicAdd(&gh[0], z);
    }
}

// Generic bandpass for R > 7: two straightforward passes through a global
// temporary, each thread one pixel, taps in the shader's order.
__global__ void k_bandpass_rows_generic(Batch b, const Geom* __restrict__ g) {
    const uint32_t f = blockIdx.y;
    const Frame fr = b.frames[f];
    const int H = (int)fr.H, W = (int)fr.W;
    const int R = g->R;
    float* tg = fr.tmp;
    float* tb = tg + (size_t)H * W;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (size_t)H * W;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int u = (int)(idx / W), v = (int)(idx % W);
        float gs = 0.f, cs = 0.f;
        for (int i = -R; i <= R; ++i) {
            int pv = v + i;
            bool dg = true;
            if (pv < 0) { pv = -pv - 1; dg = false; }
            else if (pv >= W) { pv = 2 * W - pv - 1; dg = false; }
            const float p = raw_px(fr.in, b.in_type, (size_t)u * W + pv);
            if (dg) gs = __builtin_fmaf(g->w[i + R], p, gs);
            cs = cs + p;
        }
        tg[idx] = gs / g->gn;
        tb[idx] = cs;
    }
}

__global__ void k_bandpass_cols_generic(Batch b, const Geom* __restrict__ g, int hist) {
    __shared__ uint32_t s_hist[HIST_BINS];
    const uint32_t f = blockIdx.y;
    const Frame fr = b.frames[f];
    const int H = (int)fr.H, W = (int)fr.W;
    const int R = g->R;
    const float* tg = fr.tmp;
    const float* tb = tg + (size_t)H * W;
    if (hist)
        for (int i = threadIdx.x; i < HIST_BINS; i += blockDim.x) s_hist[i] = 0;
    __syncthreads();
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (size_t)H * W;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int u = (int)(idx / W), v = (int)(idx % W);
        float gs = 0.f, cs = 0.f;
        for (int i = -R; i <= R; ++i) {
            int pu = u + i;
            bool dg = true;
            if (pu < 0) { pu = -pu - 1; dg = false; }
            else if (pu >= H) { pu = 2 * H - pu - 1; dg = false; }
            if (dg) gs = __builtin_fmaf(g->w[i + R], tg[(size_t)pu * W + v], gs);
            cs = cs + tb[(size_t)pu * W + v];
        }
        const float res = gs / g->gn - cs / g->kk;
        const float out = res < 0.f ? 0.f : res;
        fr.P[idx] = out;
        if (hist) atomicAdd(&s_hist[pct_key(__float_as_uint(out))], 1u);
    }
    if (hist) {
        __syncthreads();
        uint32_t* gh = b.hist1 + (size_t)f * HIST_BINS;
        for (int i = threadIdx.x; i < HIST_BINS; i += blockDim.x)
            if (s_hist[i]) atomicAdd(&gh[i], s_hist[i]);
    }
}

template <int R, typename T>
static hipError_t launch_bp_t(const Batch& b, const K1Params& kp, hipStream_t s) {
    hipLaunchKernelGGL((k_bandpass<R, T>), dim3(b.k1_items), dim3(64), 0, s, b, kp);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_bp_r(int R, const Batch& b, const K1Params& kp, hipStream_t s) {
    switch (R) {
        case 0: return launch_bp_t<0, T>(b, kp, s);
        case 1: return launch_bp_t<1, T>(b, kp, s);
        case 2: return launch_bp_t<2, T>(b, kp, s);
        case 3: return launch_bp_t<3, T>(b, kp, s);
        case 4: return launch_bp_t<4, T>(b, kp, s);
        case 5: return launch_bp_t<5, T>(b, kp, s);
        case 6: return launch_bp_t<6, T>(b, kp, s);
        case 7: return launch_bp_t<7, T>(b, kp, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_bandpass(const Batch& b, const Geom& g, const Geom* d_geom, int seg_rows,
                           hipStream_t s) {
    const int hist = b.hist1 != nullptr && g.has_pct;
    if (g.R <= 7) {
        K1Params kp;
        for (int i = 0; i < 15; ++i) kp.w[i] = (i < 2 * g.R + 1) ? g.w[i] : 0.f;
        kp.gn = g.gn; kp.rgn = g.rgn; kp.kk = g.kk; kp.rkk = g.rkk;
        kp.fast_gn = g.fastdiv_gn; kp.fast_kk = g.fastdiv_kk;
        kp.seg = seg_rows;
        kp.hist = hist;
        kp.vec = b.vec;
        const int it = b.in_type;
        if (it == IN_U16) return launch_bp_r<uint16_t>(g.R, b, kp, s);
        if (it == IN_U8) return launch_bp_r<uint8_t>(g.R, b, kp, s);
        return launch_bp_r<float>(g.R, b, kp, s);
    }
    return hipErrorInvalidValue;  // generic path is launched by the host with a temp
}

hipError_t launch_bandpass_generic(const Batch& b, const Geom* d_geom, uint32_t max_px,
                                   int hist, hipStream_t s) {
    const uint32_t bx = (uint32_t)std::min<uint64_t>((max_px + 255) / 256, 2048);
    hipLaunchKernelGGL(k_bandpass_rows_generic, dim3(bx, b.nframes), dim3(256), 0, s, b, d_geom);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bandpass_cols_generic, dim3(bx, b.nframes), dim3(256), 0, s, b, d_geom,
                       hist);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// select1: histogram bins holding the order statistics lo and hi
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pct_ranks(uint64_t n, float percentile, uint64_t& lo,
                                          uint64_t& hi, double& frac) {
    const double q = (double)percentile / 100.0;
    const double h = (double)(n - 1) * q;
    const double fl = floor(h);
    lo = (uint64_t)fl;
    if (lo > n - 1) lo = n - 1;
    hi = (lo + 1 < n) ? lo + 1 : n - 1;
    frac = h - fl;
}

__global__ __launch_bounds__(256) void k_select1(Batch b, float percentile) {
    __shared__ uint32_t s_part[256];
    const uint32_t f = blockIdx.x;
    const Frame fr = b.frames[f];
    const uint64_t n = (uint64_t)fr.H * fr.W;
    uint64_t lo, hi;
    double frac;
    pct_ranks(n, percentile, lo, hi, frac);
    const uint32_t* hist = b.hist1 + (size_t)f * HIST_BINS;
    const int t = threadIdx.x;
    constexpr int PER = HIST_BINS / 256;
    uint32_t c[PER];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) { c[i] = hist[t * PER + i]; sum += c[i]; }
    s_part[t] = sum;
    __syncthreads();
    // exclusive prefix of the 256 partial sums (serial per thread is cheap
    // enough here: 256 adds from LDS)
    uint64_t before = 0;
    for (int i = 0; i < t; ++i) before += s_part[i];
    uint64_t cum = before;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const uint64_t nxt = cum + c[i];
        if (lo >= cum && lo < nxt) {
            b.selstate[f * 4 + 0] = t * PER + i;
            b.selstate[f * 4 + 1] = (int32_t)(lo - cum);
        }
        if (hi >= cum && hi < nxt) {
            b.selstate[f * 4 + 2] = t * PER + i;
            b.selstate[f * 4 + 3] = (int32_t)(hi - cum);
        }
        cum = nxt;
    }
}

hipError_t launch_select1(const Batch& b, const Geom& g, hipStream_t s) {
    hipLaunchKernelGGL(k_select1, dim3(b.nframes), dim3(256), 0, s, b, g.percentile);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K2 detect: local maxima of the dilation window (ties kept), candidate
// slots in raster order within each 64x64 tile.
// ---------------------------------------------------------------------------
struct DetParams {
    int a;              // window start offset
    int kd;             // window size (runtime copy for the generic kernel)
    int band;
    int diw;            // disk inside window (zero-valued maxima cannot yield rows)
    int has_pct;
};

// y[i] = max(x[i .. i+K-1]) for i in [0, N) (van Herk / Gil-Werman): the
// input is cut into blocks of K; a window is the suffix maximum of one block
// joined with the prefix maximum of the next.  All indices are compile-time.
template <int N, int K>
__device__ __forceinline__ void sliding_max(const float (&x)[N + K - 1], float (&y)[N]) {
    constexpr int L = N + K - 1;
    float S[N];   // suffix max to the end of the block, positions 0..N-1
    float Pm[L];  // prefix max from the start of the block
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        const int be = ((j / K) + 1) * K - 1 < L - 1 ? ((j / K) + 1) * K - 1 : L - 1;
        if (j == be) {
            S[j] = x[j];
        } else if (j + 1 < N) {
            S[j] = fmaxf(x[j], S[j + 1]);
        } else {
            float m = x[j];
#pragma unroll
            for (int q = j + 1; q <= be; ++q) m = fmaxf(m, x[q]);
            S[j] = m;
        }
    }
#pragma unroll
    for (int j = 0; j < L; ++j) Pm[j] = (j % K == 0) ? x[j] : fmaxf(Pm[j - 1], x[j]);
#pragma unroll
    for (int i = 0; i < N; ++i) y[i] = fmaxf(S[i], Pm[i + K - 1]);
}

template <int KD>
__global__ __launch_bounds__(256) void k_detect(Batch b, DetParams dp) {
    constexpr int N = 16;
    constexpr int LR = TILE_H + KD - 1;
    constexpr int LC = TILE_W + KD - 1;
    constexpr int LS = ((LC + 3) / 4) * 4 + 4;
    constexpr int NIN = N + KD - 1;
    constexpr int NV4 = (NIN + 3) / 4;
    __shared__ __attribute__((aligned(16))) float L[LR * LS];
    __shared__ __attribute__((aligned(16))) float RM[LR * TILE_W];
    __shared__ uint64_t s_mask[TILE_H];
    __shared__ uint32_t s_off[TILE_H];

    const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t f = find_frame(b.tile_bases, b.nframes, item);
    const Frame fr = b.frames[f];
    const uint32_t local = item - fr.tile_base;
    const int H = (int)fr.H, W = (int)fr.W;
    const int x0 = (int)(local % fr.tx) * TILE_W;
    const int y0 = (int)(local / fr.tx) * TILE_H;
    const int tid = threadIdx.x;
    const int a = dp.a;
    const float* __restrict__ P = fr.P;

    // stage P over the tile plus the window halo (-inf outside the frame)
    for (int idx = tid; idx < LR * LC; idx += 256) {
        const int r = idx / LC, j = idx - r * LC;
        const int u = y0 - a + r, v = x0 - a + j;
        float val = __uint_as_float(NEG_INF_BITS);
        if (u >= 0 && u < H && v >= 0 && v < W) val = P[(size_t)u * W + v];
        L[r * LS + j] = val;
    }
    __syncthreads();

    // row pass: RM[r][c] = max(L[r][c .. c+KD-1])
    for (int run = tid; run < LR * 4; run += 256) {
        const int r = run >> 2, c0 = (run & 3) * N;
        float x[NIN];
        const float* src = L + r * LS + c0;
#pragma unroll
        for (int m = 0; m < NV4; ++m) {
            const float4 q = *(const float4*)(src + 4 * m);
            if (4 * m + 0 < NIN) x[4 * m + 0] = q.x;
            if (4 * m + 1 < NIN) x[4 * m + 1] = q.y;
            if (4 * m + 2 < NIN) x[4 * m + 2] = q.z;
            if (4 * m + 3 < NIN) x[4 * m + 3] = q.w;
        }
        float y[N];
        sliding_max<N, KD>(x, y);
        float* dst = RM + r * TILE_W + c0;
#pragma unroll
        for (int m = 0; m < N / 4; ++m)
            *(float4*)(dst + 4 * m) = make_float4(y[4 * m], y[4 * m + 1], y[4 * m + 2], y[4 * m + 3]);
    }
    __syncthreads();

    // column pass + candidate test: lane = column, wave = 16-row run
    const int c = tid & 63, k = tid >> 6;
    const int v = x0 + c;
    float x[NIN];
#pragma unroll
    for (int m = 0; m < NIN; ++m) x[m] = RM[(k * N + m) * TILE_W + c];
    float cm[N];
    sliding_max<N, KD>(x, cm);

    int32_t bin_lo = -1, bin_hi = -1;
    if (dp.has_pct) {
        bin_lo = b.selstate[f * 4 + 0];
        bin_hi = b.selstate[f * 4 + 2];
    }
    uint32_t* list = b.list + fr.list_base;
    uint32_t cbits = 0;
#pragma unroll
    for (int m = 0; m < N; ++m) {
        const int r = k * N + m;
        const int u = y0 + r;
        const float center = L[(r + a) * LS + (c + a)];
        const bool inb = (u < H) && (v < W);
        const bool interior = (u - dp.band >= 0) && (u + dp.band < H) &&
                              (v - dp.band >= 0) && (v + dp.band < W);
        const bool cand = inb && interior && (center >= cm[m]) &&
                          !(dp.diw && center == 0.f);
        const uint64_t msk = __ballot(cand);
        if (c == 0) s_mask[r] = msk;
        cbits |= (cand ? 1u : 0u) << m;
        if (dp.has_pct) {
            const uint32_t bits = __float_as_uint(center);
            const int key = (int)pct_key(bits);
            const bool match = inb && key != 0 && (key == bin_lo || key == bin_hi);
            const uint64_t mm = __ballot(match);
            if (mm) {
                uint32_t basei = 0;
                if (c == 0) basei = atomicAdd(&b.list_count[f], (uint32_t)__popcll(mm));
                basei = __shfl(basei, 0, 64);
                if (match) list[basei + __popcll(mm & lanemask_lt(c))] = bits;
            }
        }
    }
    __syncthreads();
    if (tid < 64) {
        const uint32_t cnt = (uint32_t)__popcll(s_mask[tid]);
        const uint32_t inc = wave_incl_scan(cnt, tid);
        s_off[tid] = inc - cnt;
        if (tid == 63) {
            b.tile_count[item] = inc;
            if (inc) atomicAdd((unsigned long long*)b.cand_total, (unsigned long long)inc);
        }
    }
    __syncthreads();
    uint16_t* slots = b.slots + (size_t)item * TILE_PX;
#pragma unroll
    for (int m = 0; m < N; ++m) {
        if (cbits & (1u << m)) {
            const int r = k * N + m;
            const uint32_t rank = s_off[r] + (uint32_t)__popcll(s_mask[r] & lanemask_lt(c));
            slots[rank] = (uint16_t)(r * TILE_W + c);
        }
    }
}

// Generic detect for window sizes without a specialisation: straightforward
// window loops over the staged tile.
__global__ __launch_bounds__(256) void k_detect_generic(Batch b, DetParams dp) {
    extern __shared__ float dyn[];
    __shared__ uint64_t s_mask[TILE_H];
    __shared__ uint32_t s_off[TILE_H];
    const int KD = dp.kd;
    const int LR = TILE_H + max(KD, 1) - 1, LC = TILE_W + max(KD, 1) - 1;
    float* L = dyn;                  // LR * LC
    float* RM = dyn + LR * LC;       // LR * TILE_W
    const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t f = find_frame(b.tile_bases, b.nframes, item);
    const Frame fr = b.frames[f];
    const uint32_t local = item - fr.tile_base;
    const int H = (int)fr.H, W = (int)fr.W;
    const int x0 = (int)(local % fr.tx) * TILE_W;
    const int y0 = (int)(local / fr.tx) * TILE_H;
    const int tid = threadIdx.x;
    const int a = dp.a;
    const float* P = fr.P;
    for (int idx = tid; idx < LR * LC; idx += 256) {
        const int r = idx / LC, j = idx - r * LC;
        const int u = y0 - a + r, v = x0 - a + j;
        float val = __uint_as_float(NEG_INF_BITS);
        if (u >= 0 && u < H && v >= 0 && v < W) val = P[(size_t)u * W + v];
        L[r * LC + j] = val;
    }
    __syncthreads();
    for (int idx = tid; idx < LR * TILE_W; idx += 256) {
        const int r = idx / TILE_W, c = idx % TILE_W;
        float m = __uint_as_float(NEG_INF_BITS);
        for (int j = 0; j < KD; ++j) m = fmaxf(m, L[r * LC + c + j]);
        RM[idx] = m;
    }
    __syncthreads();
    const int c = tid & 63, k = tid >> 6;
    const int v = x0 + c;
    int32_t bin_lo = -1, bin_hi = -1;
    if (dp.has_pct) {
        bin_lo = b.selstate[f * 4 + 0];
        bin_hi = b.selstate[f * 4 + 2];
    }
    uint32_t* list = b.list + fr.list_base;
    uint32_t cbits = 0;
    for (int m = 0; m < 16; ++m) {
        const int r = k * 16 + m;
        const int u = y0 + r;
        float cmx = __uint_as_float(NEG_INF_BITS);
        for (int i = 0; i < KD; ++i) cmx = fmaxf(cmx, RM[(r + i) * TILE_W + c]);
        const float center = L[(r + a) * LC + (c + a)];
        const bool inb = (u < H) && (v < W);
        const bool interior = (u - dp.band >= 0) && (u + dp.band < H) &&
                              (v - dp.band >= 0) && (v + dp.band < W);
        const bool cand = inb && interior && (center >= cmx) && !(dp.diw && center == 0.f);
        const uint64_t msk = __ballot(cand);
        if (c == 0) s_mask[r] = msk;
        cbits |= (cand ? 1u : 0u) << m;
        if (dp.has_pct) {
            const uint32_t bits = __float_as_uint(center);
            const int key = (int)pct_key(bits);
            const bool match = inb && key != 0 && (key == bin_lo || key == bin_hi);
            const uint64_t mm = __ballot(match);
            if (mm) {
                uint32_t basei = 0;
                if (c == 0) basei = atomicAdd(&b.list_count[f], (uint32_t)__popcll(mm));
                basei = __shfl(basei, 0, 64);
                if (match) list[basei + __popcll(mm & lanemask_lt(c))] = bits;
            }
        }
    }
    __syncthreads();
    if (tid < 64) {
        const uint32_t cnt = (uint32_t)__popcll(s_mask[tid]);
        const uint32_t inc = wave_incl_scan(cnt, tid);
        s_off[tid] = inc - cnt;
        if (tid == 63) {
            b.tile_count[item] = inc;
            if (inc) atomicAdd((unsigned long long*)b.cand_total, (unsigned long long)inc);
        }
    }
    __syncthreads();
    uint16_t* slots = b.slots + (size_t)item * TILE_PX;
    for (int m = 0; m < 16; ++m) {
        if (cbits & (1u << m)) {
            const int r = k * 16 + m;
            const uint32_t rank = s_off[r] + (uint32_t)__popcll(s_mask[r] & lanemask_lt(c));
            slots[rank] = (uint16_t)(r * TILE_W + c);
        }
    }
}

template <int KD>
static hipError_t launch_det_t(const Batch& b, const DetParams& dp, hipStream_t s) {
    hipLaunchKernelGGL(k_detect<KD>, dim3(b.tiles), dim3(256), 0, s, b, dp);
    return hipGetLastError();
}

hipError_t launch_detect(const Batch& b, const Geom& g, const Geom* d_geom, hipStream_t s) {
    (void)d_geom;
    DetParams dp;
    dp.a = g.a;
    dp.kd = g.kd;
    dp.band = g.band;
    dp.diw = g.disk_in_window;
    dp.has_pct = g.has_pct;
    switch (g.kd) {
        case 4: return launch_det_t<4>(b, dp, s);
        case 5: return launch_det_t<5>(b, dp, s);
        case 7: return launch_det_t<7>(b, dp, s);
        case 8: return launch_det_t<8>(b, dp, s);
        case 9: return launch_det_t<9>(b, dp, s);
        case 11: return launch_det_t<11>(b, dp, s);
        case 12: return launch_det_t<12>(b, dp, s);
        case 14: return launch_det_t<14>(b, dp, s);
        case 15: return launch_det_t<15>(b, dp, s);
        case 16: return launch_det_t<16>(b, dp, s);
        case 18: return launch_det_t<18>(b, dp, s);
        case 19: return launch_det_t<19>(b, dp, s);
        case 21: return launch_det_t<21>(b, dp, s);
        default: break;
    }
    const int kd1 = max(g.kd, 1);
    const size_t shm = sizeof(float) * ((size_t)(TILE_H + kd1 - 1) * (TILE_W + kd1 - 1) +
                                        (size_t)(TILE_H + kd1 - 1) * TILE_W);
    if (shm > 60 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_detect_generic, dim3(b.tiles), dim3(256), shm, s, b, dp);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// select2: exact order statistics from the value list, then the threshold
// ---------------------------------------------------------------------------
__device__ uint32_t select_in_list(const uint32_t* list, uint32_t cnt, int bin, int rank,
                                   uint32_t* s_hist /* >= 2048 */, uint32_t* s_res) {
    const int t = threadIdx.x, nt = blockDim.x;
    // level A: bits [18:8]
    for (int i = t; i < 2048; i += nt) s_hist[i] = 0;
    __syncthreads();
    for (uint32_t i = t; i < cnt; i += nt) {
        const uint32_t bits = list[i];
        if ((int)pct_key(bits) == bin) atomicAdd(&s_hist[(bits >> 8) & 2047u], 1u);
    }
    __syncthreads();
    if (t == 0) {
        uint32_t cum = 0;
        int sa = 0;
        for (int i = 0; i < 2048; ++i) {
            if ((uint32_t)rank < cum + s_hist[i]) { sa = i; break; }
            cum += s_hist[i];
        }
        s_res[0] = (uint32_t)sa;
        s_res[1] = (uint32_t)rank - cum;
    }
    __syncthreads();
    const uint32_t sa = s_res[0], ra = s_res[1];
    // level B: bits [7:0]
    for (int i = t; i < 256; i += nt) s_hist[i] = 0;
    __syncthreads();
    for (uint32_t i = t; i < cnt; i += nt) {
        const uint32_t bits = list[i];
        if ((int)pct_key(bits) == bin && ((bits >> 8) & 2047u) == sa)
            atomicAdd(&s_hist[bits & 255u], 1u);
    }
    __syncthreads();
    if (t == 0) {
        uint32_t cum = 0;
        uint32_t sb = 0;
        for (int i = 0; i < 256; ++i) {
            if (ra < cum + s_hist[i]) { sb = (uint32_t)i; break; }
            cum += s_hist[i];
        }
        s_res[2] = ((uint32_t)(bin - 1) << 19) | (sa << 8) | sb;
    }
    __syncthreads();
    const uint32_t v = s_res[2];
    __syncthreads();
    return v;
}

__global__ __launch_bounds__(1024) void k_select2(Batch b, float percentile) {
    __shared__ uint32_t s_hist[2048];
    __shared__ uint32_t s_res[4];
    const uint32_t f = blockIdx.x;
    const Frame fr = b.frames[f];
    const uint64_t n = (uint64_t)fr.H * fr.W;
    uint64_t lo, hi;
    double frac;
    pct_ranks(n, percentile, lo, hi, frac);
    const int bin_lo = b.selstate[f * 4 + 0], r_lo = b.selstate[f * 4 + 1];
    const int bin_hi = b.selstate[f * 4 + 2], r_hi = b.selstate[f * 4 + 3];
    const uint32_t* list = b.list + fr.list_base;
    const uint32_t cnt = b.list_count[f];
    float vlo = 0.f, vhi = 0.f;
    if (bin_lo != 0) vlo = __uint_as_float(select_in_list(list, cnt, bin_lo, r_lo, s_hist, s_res));
    if (hi == lo) vhi = vlo;
    else if (bin_hi != 0) vhi = __uint_as_float(select_in_list(list, cnt, bin_hi, r_hi, s_hist, s_res));
    if (threadIdx.x == 0) b.thr[f] = (double)vlo + frac * ((double)vhi - (double)vlo);
}

hipError_t launch_select2(const Batch& b, const Geom& g, hipStream_t s) {
    hipLaunchKernelGGL(k_select2, dim3(b.nframes), dim3(1024), 0, s, b, g.percentile);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K3 refine: one lane per candidate, one wave per tile.
// ---------------------------------------------------------------------------

// centers.wgsl:23-109 at one pixel (binary32, taps in shader order).
__device__ __forceinline__ void disk_eval(const float* __restrict__ P, int H, int W, int band,
                                          int rint, int u, int v, float& m, float& c0,
                                          float& c1) {
    if (u - band < 0 || u + band >= H || v - band < 0 || v + band >= W) {
        m = 0.f; c0 = 0.f; c1 = 0.f;
        return;
    }
    float s0 = 0.f, s1 = 0.f, ms = 0.f;
    for (int i = -rint; i <= rint; ++i) {
        int hw = 0;
        while ((hw + 1) * (hw + 1) + i * i <= rint * rint) ++hw;
        const float* row = P + (size_t)(u + i) * W + v;
        const float fi = (float)i;
        for (int j = -hw; j <= hw; ++j) {
            const float d = row[j];
            s0 = __builtin_fmaf(d, fi, s0);
            s1 = __builtin_fmaf(d, (float)j, s1);
            ms = ms + d;
        }
    }
    m = ms;
    c0 = s0 / ms;
    c1 = s1 / ms;
}

template <int RI>
__device__ __forceinline__ void disk_eval_t(const float* __restrict__ P, int H, int W, int band,
                                            int u, int v, float& m, float& c0, float& c1) {
    if (u - band < 0 || u + band >= H || v - band < 0 || v + band >= W) {
        m = 0.f; c0 = 0.f; c1 = 0.f;
        return;
    }
    float s0 = 0.f, s1 = 0.f, ms = 0.f;
#pragma unroll
    for (int i = -RI; i <= RI; ++i) {
        const float* row = P + (size_t)(u + i) * W + v;
#pragma unroll
        for (int j = -RI; j <= RI; ++j) {
            if (i * i + j * j > RI * RI) continue;
            const float d = row[j];
            s0 = __builtin_fmaf(d, (float)i, s0);
            s1 = __builtin_fmaf(d, (float)j, s1);
            ms = ms + d;
        }
    }
    m = ms;
    c0 = s0 / ms;
    c1 = s1 / ms;
}

template <int RI>
__device__ __forceinline__ void disk_any(const float* P, int H, int W, int band, int rint, int u,
                                         int v, float& m, float& c0, float& c1) {
    if constexpr (RI > 0) disk_eval_t<RI>(P, H, W, band, u, v, m, c0, c1);
    else disk_eval(P, H, W, band, rint, u, v, m, c0, c1);
}

template <int RI>
__global__ __launch_bounds__(256) void k_refine(Batch b, const Geom* __restrict__ g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t item = blockIdx.x * 4 + wave;
    if (item >= b.tiles) return;
    const uint32_t f = find_frame(b.tile_bases, b.nframes, item);
    const Frame fr = b.frames[f];
    const uint32_t local = item - fr.tile_base;
    const int H = (int)fr.H, W = (int)fr.W;
    const int x0 = (int)(local % fr.tx) * TILE_W;
    const int y0 = (int)(local / fr.tx) * TILE_H;
    const uint32_t cnt = b.tile_count[item];
    const size_t slot0 = (size_t)item * TILE_PX;
    const float* __restrict__ P = fr.P;
    const int band = g->band, rint = g->rint;
    const uint32_t maxit = g->max_iterations;
    const float minmass = g->minmass;
    const double thr = g->has_pct ? b.thr[f] : 0.0;
    const float sthr = 0.6f;
    float* rec = b.rec;
    const uint64_t cap = b.slot_cap;

    for (uint32_t i = lane; i < cnt; i += 64) {
        const uint32_t l = b.slots[slot0 + i];
        const int u0 = y0 + (int)(l / TILE_W), v0 = x0 + (int)(l % TILE_W);
        bool valid = true;
        if (g->has_pct) valid = (double)P[(size_t)u0 * W + v0] > thr;
        float y = 0.f, x = 0.f, mass = 0.f;
        if (valid) {
            int u = u0, v = v0;
            float c0 = 0.f, c1 = 0.f, m = 0.f;
            if (maxit == 0) {
                float t0, t1;
                disk_any<RI>(P, H, W, band, rint, 0, 0, m, t0, t1);
            }
            for (uint32_t it = 0; it < maxit; ++it) {
                disk_any<RI>(P, H, W, band, rint, u, v, m, c0, c1);
                bool changed = false;
                if (c0 > sthr && u < H - 1) { u += 1; changed = true; }
                else if (c0 < -sthr && u > 0) { u -= 1; changed = true; }
                if (c1 > sthr && v < W - 1) { v += 1; changed = true; }
                else if (c1 < -sthr && v > 0) { v -= 1; changed = true; }
                if (!changed) break;
            }
            y = (float)u + c0;
            x = (float)v + c1;
            mass = m;
            valid = (mass > minmass) && (mass != 0.f);
        }
        rec[REC_MASS * cap + slot0 + i] = valid ? mass : 0.f;
        rec[REC_Y * cap + slot0 + i] = y;
        rec[REC_X * cap + slot0 + i] = x;
        if (valid && g->cpu_processed) {
            // execute_gpu.rs:165-172,286-299,346-352 at the candidate pixel:
            // products with the host tables, each summed in ndarray's
            // eight-accumulator order (full chunks of 8, pairwise fold, then
            // the remainder in order).
            const int r = g->band;
            const int nt = g->ntap_cpu;
            const int nfull = nt - (nt % 8);
            float pr[8], ps[8], pc[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) { pr[q] = 0.f; ps[q] = 0.f; pc[q] = 0.f; }
            float sig = 0.f, mid = 0.f;
            int tap = 0;
            for (int di = -r; di <= r; ++di) {
                for (int dj = -r; dj <= r; ++dj) {
                    if (di * di + dj * dj > r * r) continue;
                    const float d = P[(size_t)(u0 + di) * W + (v0 + dj)];
                    sig = (tap == 0 || d > sig) ? d : sig;
                    if (di == 0 && dj == 0) mid = d;
                    if (tap < nfull) {
                        const int q = tap & 7;
                        pr[q] = pr[q] + d * g->r2_cpu[tap];
                        ps[q] = ps[q] + d * g->sin_cpu[tap];
                        pc[q] = pc[q] + d * g->cos_cpu[tap];
                    }
                    ++tap;
                }
            }
            float rg = 0.f, ss = 0.f, cc = 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                rg = rg + (pr[q] + pr[q + 4]);
                ss = ss + (ps[q] + ps[q + 4]);
                cc = cc + (pc[q] + pc[q + 4]);
            }
            tap = 0;
            for (int di = -r; di <= r; ++di) {
                for (int dj = -r; dj <= r; ++dj) {
                    if (di * di + dj * dj > r * r) continue;
                    if (tap >= nfull) {
                        const float d = P[(size_t)(u0 + di) * W + (v0 + dj)];
                        rg = rg + d * g->r2_cpu[tap];
                        ss = ss + d * g->sin_cpu[tap];
                        cc = cc + d * g->cos_cpu[tap];
                    }
                    ++tap;
                }
            }
            ss = ss * ss;
            cc = cc * cc;
            const float ecc = sqrtf(ss + cc);
            rec[REC_RG * cap + slot0 + i] = sqrtf(rg / mass);
            rec[REC_ECC * cap + slot0 + i] = ecc / (mass - mid);
            rec[REC_SIGNAL * cap + slot0 + i] = sig;
        }
        if (valid && g->characterize) {
            // centers.wgsl:64-80 at the candidate pixel
            float o4[4] = {0.f, 0.f, 0.f, 0.f};
            if (!(u0 - band < 0 || u0 + band >= H || v0 - band < 0 || v0 + band >= W)) {
                float ms = 0.f, rg = 0.f, rawm = 0.f, sg = 0.f, es = 0.f, ec = 0.f;
                int tap = 0;
                for (int di = -rint; di <= rint; ++di)
                    for (int dj = -rint; dj <= rint; ++dj) {
                        if (di * di + dj * dj > rint * rint) continue;
                        const size_t idx = (size_t)(u0 + di) * W + (v0 + dj);
                        const float d = P[idx];
                        ms = ms + d;
                        rg = __builtin_fmaf(d, (float)(di * di + dj * dj), rg);
                        rawm = rawm + raw_px(fr.in, b.in_type, idx);
                        sg = (d > sg) ? d : sg;
                        es = __builtin_fmaf(d, g->sin2_gpu[tap], es);
                        ec = __builtin_fmaf(d, g->cos2_gpu[tap], ec);
                        ++tap;
                    }
                es = es * es;
                ec = ec * ec;
                o4[3] = sqrtf(es + ec) / (ms - P[(size_t)u0 * W + v0]);
                o4[0] = sqrtf(rg / ms);
                o4[1] = rawm;
                o4[2] = sg;
            }
            rec[REC_RG_GPU * cap + slot0 + i] = o4[0];
            rec[REC_RAW_GPU * cap + slot0 + i] = o4[1];
            rec[REC_SIG_GPU * cap + slot0 + i] = o4[2];
            rec[REC_ECC_GPU * cap + slot0 + i] = o4[3];
        }
    }
}

hipError_t launch_refine(const Batch& b, const Geom& g, const Geom* d_geom, hipStream_t s) {
    const dim3 grid((b.tiles + 3) / 4);
    switch (g.rint) {
        case 2: hipLaunchKernelGGL(k_refine<2>, grid, dim3(256), 0, s, b, d_geom); break;
        case 3: hipLaunchKernelGGL(k_refine<3>, grid, dim3(256), 0, s, b, d_geom); break;
        case 4: hipLaunchKernelGGL(k_refine<4>, grid, dim3(256), 0, s, b, d_geom); break;
        case 5: hipLaunchKernelGGL(k_refine<5>, grid, dim3(256), 0, s, b, d_geom); break;
        case 6: hipLaunchKernelGGL(k_refine<6>, grid, dim3(256), 0, s, b, d_geom); break;
        case 7: hipLaunchKernelGGL(k_refine<7>, grid, dim3(256), 0, s, b, d_geom); break;
        default: hipLaunchKernelGGL(k_refine<0>, grid, dim3(256), 0, s, b, d_geom); break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K4 dedup: one wave per tile; neighbours are read tile by tile with
// wave-uniform loads.  The last wave of a frame scans its tiles' keep counts.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dedup(Batch b, const Geom* __restrict__ g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t item = blockIdx.x * 4 + wave;
    if (item >= b.tiles) return;
    const uint32_t f = find_frame(b.tile_bases, b.nframes, item);
    const Frame fr = b.frames[f];
    const uint32_t local = item - fr.tile_base;
    const int W = (int)fr.W;
    const int tx = (int)(local % fr.tx), ty = (int)(local / fr.tx);
    const uint32_t cnt = b.tile_count[item];
    const size_t slot0 = (size_t)item * TILE_PX;
    const uint64_t cap = b.slot_cap;
    const float* rmass = b.rec + REC_MASS * cap;
    const float* ry = b.rec + REC_Y * cap;
    const float* rx = b.rec + REC_X * cap;
    const float rad = g->sep;
    const float r2 = rad * rad;
    const int rty = g->dedup_rty, rtx = g->dedup_rtx;
    uint32_t kept = 0;

    for (uint32_t ib = 0; ib < cnt; ib += 64) {
        const uint32_t i = ib + lane;
        const bool own = i < cnt;
        float m = 0.f, y = 0.f, x = 0.f;
        int64_t idx = 0;
        if (own) {
            m = rmass[slot0 + i];
            y = ry[slot0 + i];
            x = rx[slot0 + i];
            const uint32_t l = b.slots[slot0 + i];
            idx = (int64_t)(ty * TILE_H + (int)(l / TILE_W)) * W + (tx * TILE_W + (int)(l % TILE_W));
        }
        bool keep = own && (m != 0.f);
        if (__ballot(keep)) {
            for (int ny = max(0, ty - rty); ny <= min((int)fr.ty - 1, ty + rty); ++ny) {
                for (int nx = max(0, tx - rtx); nx <= min((int)fr.tx - 1, tx + rtx); ++nx) {
                    const uint32_t it2 = fr.tile_base + (uint32_t)(ny * (int)fr.tx + nx);
                    const uint32_t c2 = b.tile_count[it2];
                    const size_t s2 = (size_t)it2 * TILE_PX;
                    for (uint32_t j = 0; j < c2; ++j) {
                        const float nm = rmass[s2 + j];
                        if (nm == 0.f) continue;
                        const float nyv = ry[s2 + j], nxv = rx[s2 + j];
                        const float dy = nyv - y, dx = nxv - x;
                        float d = dy * dy;
                        d = d + dx * dx;
                        if (d < r2 && !(it2 == item && j == i)) {
                            bool drop = m < nm;
                            if (!drop && m == nm) {
                                const uint32_t l2 = b.slots[s2 + j];
                                const int64_t nidx =
                                    (int64_t)(ny * TILE_H + (int)(l2 / TILE_W)) * W +
                                    (nx * TILE_W + (int)(l2 % TILE_W));
                                drop = idx > nidx;
                            }
                            if (drop) keep = false;
                        }
                    }
                }
            }
        }
        if (own) b.keep[slot0 + i] = keep ? 1 : 0;
        kept += (uint32_t)__popcll(__ballot(keep));
    }
    if (lane == 0) b.tile_keep[item] = kept;

    // last wave of this frame: exclusive scan of the frame's tile keep counts
    __threadfence();
    uint32_t old = 0;
    if (lane == 0) old = atomicAdd(&b.done[f], 1u);
    old = __shfl(old, 0, 64);
    const uint32_t ntiles = fr.tx * fr.ty;
    if (old == ntiles - 1) {
        __threadfence();
        uint32_t carry = 0;
        for (uint32_t base = 0; base < ntiles; base += 64) {
            const uint32_t k = base + lane;
            const uint32_t c =
                k < ntiles ? __hip_atomic_load(&b.tile_keep[fr.tile_base + k], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)
                           : 0u;
            const uint32_t inc = wave_incl_scan(c, lane);
            if (k < ntiles) b.tile_row_off[fr.tile_base + k] = carry + inc - c;
            carry += __shfl(inc, 63, 64);
        }
        if (lane == 0) b.frame_rows[f] = carry;
    }
}

hipError_t launch_dedup(const Batch& b, const Geom& g, const Geom* d_geom, hipStream_t s) {
    (void)g;
    hipLaunchKernelGGL(k_dedup, dim3((b.tiles + 3) / 4), dim3(256), 0, s, b, d_geom);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// fscan: frame row bases (continuing the running count across batches)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_frame_scan(Batch b) {
    __shared__ uint64_t s_part[256];
    const int t = threadIdx.x;
    const uint32_t n = b.nframes;
    const uint32_t per = (n + 255) / 256;
    uint64_t sum = 0;
    for (uint32_t i = 0; i < per; ++i) {
        const uint32_t k = t * per + i;
        if (k < n) sum += b.frame_rows[k];
    }
    s_part[t] = sum;
    __syncthreads();
    uint64_t before = *b.row_carry;
    for (int i = 0; i < t; ++i) before += s_part[i];
    for (uint32_t i = 0; i < per; ++i) {
        const uint32_t k = t * per + i;
        if (k < n) {
            b.frame_row_base[k] = before;
            before += b.frame_rows[k];
        }
    }
    __syncthreads();
    if (t == 255) *b.row_carry = before;
}

hipError_t launch_frame_scan(const Batch& b, hipStream_t s) {
    hipLaunchKernelGGL(k_frame_scan, dim3(1), dim3(256), 0, s, b);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K5 emit: rows [frame, mass, y, x, (Rg, ecc, signal), (Rg_gpu, raw_mass_gpu,
// signal_gpu, ecc_gpu)] in frame, tile, raster-within-tile order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_emit(Batch b, const Geom* __restrict__ g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t item = blockIdx.x * 4 + wave;
    if (item >= b.tiles) return;
    const uint32_t f = find_frame(b.tile_bases, b.nframes, item);
    const Frame fr = b.frames[f];
    const uint32_t local = item - fr.tile_base;
    const int W = (int)fr.W;
    const int tx = (int)(local % fr.tx), ty = (int)(local / fr.tx);
    const uint32_t cnt = b.tile_count[item];
    const size_t slot0 = (size_t)item * TILE_PX;
    const uint64_t cap = b.slot_cap;
    const int nc = g->ncols;
    uint64_t row = b.frame_row_base[f] + b.tile_row_off[item];
    for (uint32_t ib = 0; ib < cnt; ib += 64) {
        const uint32_t i = ib + lane;
        const bool k = (i < cnt) && b.keep[slot0 + i];
        const uint64_t msk = __ballot(k);
        if (k) {
            const uint64_t r = row + (uint64_t)__popcll(msk & lanemask_lt(lane));
            if (r < b.rows_cap) {
                float* o = b.rows + r * (uint64_t)nc;
                const size_t s = slot0 + i;
                o[0] = (float)fr.frame_index;
                o[1] = b.rec[REC_MASS * cap + s];
                o[2] = b.rec[REC_Y * cap + s];
                o[3] = b.rec[REC_X * cap + s];
                int c = 4;
                if (g->cpu_processed) {
                    o[c++] = b.rec[REC_RG * cap + s];
                    o[c++] = b.rec[REC_ECC * cap + s];
                    o[c++] = b.rec[REC_SIGNAL * cap + s];
                }
                if (g->characterize) {
                    o[c++] = b.rec[REC_RG_GPU * cap + s];
                    o[c++] = b.rec[REC_RAW_GPU * cap + s];
                    o[c++] = b.rec[REC_SIG_GPU * cap + s];
                    o[c++] = b.rec[REC_ECC_GPU * cap + s];
                }
                const uint32_t l = b.slots[s];
                b.rows_cand[r] = (int64_t)(ty * TILE_H + (int)(l / TILE_W)) * W +
                                 (tx * TILE_W + (int)(l % TILE_W));
            }
        }
        row += (uint64_t)__popcll(msk);
    }
}

hipError_t launch_emit(const Batch& b, const Geom& g, const Geom* d_geom, hipStream_t s) {
    (void)g;
    hipLaunchKernelGGL(k_emit, dim3((b.tiles + 3) / 4), dim3(256), 0, s, b, d_geom);
    return hipGetLastError();
}

}  // namespace gt
