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
#include <type_traits>

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
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh;
         i += (uint64_t)gridDim.x * blockDim.x)
        b.hist1[i] = 0;
}

hipError_t launch_zero_batch(const Batch& b, hipStream_t s) {
    const int with_hist = b.hist1 != nullptr;
    const uint64_t total = with_hist ? (uint64_t)b.nframes * HIST_BINS : 0;
    uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 4096);
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_zero_batch, dim3(blocks), dim3(256), 0, s, b, with_hist);
    return hipGetLastError();
}

// Per-item tables of a batch from its per-frame bases: frame blockIdx.y
// writes the frame index of each of its bandpass items and tiles and the
// top-left tile of each of its 2x2 dedup blocks (the frame's blocks in
// raster order).
__global__ void k_fill_tables(Batch b, const uint32_t* __restrict__ dbases) {
    const uint32_t f = blockIdx.y;
    const uint32_t k1a = b.k1_bases[f], k1n = b.k1_bases[f + 1] - k1a;
    const uint32_t ta = b.tile_bases[f], tn = b.tile_bases[f + 1] - ta;
    const uint32_t da = dbases[f], dn = dbases[f + 1] - da;
    const uint32_t tx = b.frames[f].tx;
    const uint32_t bx = (tx + 1) / 2;   // dedup blocks per block row
    uint32_t* k1f = (uint32_t*)b.k1_frame;
    uint32_t* tf = (uint32_t*)b.tile_frame;
    uint32_t* dt = (uint32_t*)b.dblock_tile;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k1n + tn + dn;
         i += gridDim.x * blockDim.x) {
        if (i < k1n) {
            k1f[k1a + i] = f;
        } else if (i < k1n + tn) {
            tf[ta + (i - k1n)] = f;
        } else {
            const uint32_t j = i - k1n - tn;
            dt[da + j] = ta + 2 * (j / bx) * tx + 2 * (j % bx);
        }
    }
}

hipError_t launch_fill_tables(const Batch& b, const uint32_t* dblock_bases,
                              uint32_t max_frame_work, hipStream_t s) {
    const uint32_t bx = std::min<uint32_t>((max_frame_work + 255) / 256, 64);
    hipLaunchKernelGGL(k_fill_tables, dim3(bx, b.nframes), dim3(256), 0, s, b, dblock_bases);
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

// Plain device copy: 16-byte vectors, grid-stride, byte tail.  Also the kernel
// of known traffic (n bytes read, n written) for scaling byte counters.
__global__ void k_copy(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n16,
                       uint8_t* __restrict__ dtail, const uint8_t* __restrict__ stail,
                       uint32_t ntail) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
    if (blockIdx.x == 0 && threadIdx.x < ntail) dtail[threadIdx.x] = stail[threadIdx.x];
}

hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s) {
    const uint64_t n16 = bytes / 16;
    const uint32_t ntail = (uint32_t)(bytes % 16);
    uint32_t blocks = (uint32_t)std::min<uint64_t>((n16 + 255) / 256, 8192);
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, s, (uint4*)dst, (const uint4*)src,
                       n16, (uint8_t*)dst + n16 * 16, (const uint8_t*)src + n16 * 16, ntail);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K1 bandpass, streaming column strips.
//
// A work item is a 256-column strip of `seg` output rows of one frame.  The
// single wave walks down the strip: every input row (reflected at the top and
// bottom edges) is staged once in LDS, each lane computes the row pass of its
// four columns and keeps the last 2R+1 row results in a register ring; once
// the ring is full the column pass produces one output row.  Input rows are
// fetched 2R+1 rows ahead with 8-byte (u16) / 4-byte (u8) vector loads of
// 4-column chunks.  With W % 4 == 0 no chunk straddles a frame edge, so a
// chunk outside the frame is the mirror image of an in-frame chunk: it is
// loaded from there and reversed when staged.  Two staged copies of the row
// implement the two boundary rules: out-of-frame columns are zero in the
// Gaussian copy (taps skipped, Q3) and mirrored in the boxcar copy (Q4).
//
// Integer input with k <= 15 keeps every boxcar sum an exact integer below
// 2^24 in binary32, so the horizontal boxcar slides and the vertical boxcar is
// a running sum.  FAST selects the corrected-reciprocal division by the two
// per-call divisors (host-verified to equal IEEE division).
// ---------------------------------------------------------------------------
// Pointers fetched from the frame table are generic; marking them global
// lets the compiler emit global (not flat) memory instructions, whose waits
// do not also drain the LDS counter.  Only the device pass knows the
// address space.
#if defined(__HIP_DEVICE_COMPILE__)
#define GLOBAL __attribute__((address_space(1)))
#define SHARED __attribute__((address_space(3)))
#else
#define GLOBAL
#define SHARED
#endif

struct K1Params {
    float w[17];
    float gn, rgn, kk, rkk;
    int seg;
    int hist;
};

template <typename T> struct Chunk4;
template <> struct Chunk4<uint16_t> { uint2 v; };
template <> struct Chunk4<uint8_t> { uint32_t v; };
template <> struct Chunk4<float> { float4 v; };

// base: the frame (wave-uniform); off: byte offset, 32-bit (frames hold
// fewer than 2^30 pixels), so the load takes the frame as a scalar base
template <typename T>
__device__ __forceinline__ Chunk4<T> ld_chunk(const GLOBAL char* base, uint32_t off) {
    const GLOBAL char* p = base + off;
    Chunk4<T> c;
    if constexpr (IsFloat<T>::v) c.v = *(const GLOBAL float4*)p;
    else if constexpr (sizeof(T) == 2) c.v = *(const GLOBAL uint2*)p;
    else c.v = *(const GLOBAL uint32_t*)p;
    return c;
}

typedef float f4v __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ f4v unpack_chunk(const Chunk4<T>& c) {
    f4v r;
    if constexpr (IsFloat<T>::v) {
        r.x = c.v.x; r.y = c.v.y; r.z = c.v.z; r.w = c.v.w;
    } else if constexpr (sizeof(T) == 2) {
        r.x = (float)(c.v.x & 0xffffu); r.y = (float)(c.v.x >> 16);
        r.z = (float)(c.v.y & 0xffffu); r.w = (float)(c.v.y >> 16);
    } else {
        r.x = (float)(c.v & 0xffu); r.y = (float)((c.v >> 8) & 0xffu);
        r.z = (float)((c.v >> 16) & 0xffu); r.w = (float)(c.v >> 24);
    }
    return r;
}

// Where the 4-column chunk starting at column c (c % 4 == 0) comes from:
// its own columns, or (outside the frame) the mirrored in-frame chunk, read
// in reverse.  Chunks that no output can reach are clamped into the frame.
__device__ __forceinline__ void chunk_src(int c, int W, int& src, bool& rev, bool& oob) {
    if (c < 0) { src = -c - 4; rev = true; oob = true; }
    else if (c >= W) { src = 2 * W - c - 4; rev = true; oob = true; }
    else { src = c; rev = false; oob = false; }
    src = max(0, min(src, W - 4));
}

__device__ __forceinline__ float cdiv_t(float x, float d, float rd, bool fast) {
    if (fast) {
        const float q = x * rd;
        const float r = __builtin_fmaf(-q, d, x);
        return __builtin_fmaf(r, rd, q);
    }
    return x / d;
}

// Five 16-byte LDS reads kept whole (volatile): otherwise the compiler
// drops unused lanes and splits them into 4-byte reads, which at a 16-byte
// lane stride collide on the LDS banks.
__device__ __forceinline__ void lds_ld20(const float* p, float (&v)[20]) {
    const volatile SHARED f4v* q = (const volatile SHARED f4v*)p;
    const f4v a0 = q[0], a1 = q[1], a2 = q[2], a3 = q[3], a4 = q[4];
    v[0] = a0.x;  v[1] = a0.y;  v[2] = a0.z;  v[3] = a0.w;
    v[4] = a1.x;  v[5] = a1.y;  v[6] = a1.z;  v[7] = a1.w;
    v[8] = a2.x;  v[9] = a2.y;  v[10] = a2.z; v[11] = a2.w;
    v[12] = a3.x; v[13] = a3.y; v[14] = a3.z; v[15] = a3.w;
    v[16] = a4.x; v[17] = a4.y; v[18] = a4.z; v[19] = a4.w;
}

// K1 layout: lane t owns the 4-column chunk at column x0 - 8 + 4t; lanes
// 2..61 produce output (a 240-column strip), lanes 0, 1, 62, 63 only load
// the 8-column halos on either side.
constexpr int K1_LANE0 = 2;
constexpr int K1_LANES = K1_STRIP / 4;
static_assert(K1_LANES + 2 * K1_LANE0 == 64, "strip + halo must fill one wave");
constexpr int K1_ROWBUF = 256;   // floats per staged row

// Three waves per SIMD fit without spills up to R = 5.
template <int R, typename T, bool FAST>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 64),
                          amdgpu_waves_per_eu(R <= 5 ? 3 : 2)))
void k_bandpass(Batch b, K1Params kp) {
    constexpr int NT = 2 * R + 1;
    static_assert(R <= 8, "taps must stay inside the staged halo");
    // float input: the boxcar sums are not exact, so they are formed tap by
    // tap in the shader's order (from +0, as preprocess_*.wgsl do) instead
    // of as running sums, and the clamp keeps NaN as the shader's select does
    constexpr bool FLT = IsFloat<T>::v;
    // per parity: boxcar copy (mirrored) then Gaussian copy (zero outside)
    __shared__ __attribute__((aligned(16))) float buf[2][2][K1_ROWBUF];
    // two 16-bit bins per word (word i: bins 2i + 1 and 2i + 2; zeros are
    // counted apart); a work item adds at most K1_STRIP * seg < 65536 to one
    // bin (checked by the host)
    __shared__ uint32_t s_hist[HIST_BINS / 2];

    const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t f = b.k1_frame[item];
    const Frame fr = b.frames[f];
    const uint32_t local = item - fr.k1_base;
    const int H = (int)fr.H, W = (int)fr.W;
    const int sx = (int)(local % fr.k1_tx), sy = (int)(local / fr.k1_tx);
    const int x0 = sx * K1_STRIP;
    const int y0 = sy * kp.seg;
    const int y1 = min(y0 + kp.seg, H);
    const int t = threadIdx.x;
    const GLOBAL T* in = (const GLOBAL T*)fr.in;
    GLOBAL float* P = (GLOBAL float*)fr.P;
    // some chunk of this strip (output or halo) lies outside the frame
    const bool border = (x0 - 8 < 0) || (x0 + K1_STRIP + 8 > W);
    const int cc = x0 - 8 + 4 * t;
    int csrc;
    bool crev, coob;
    chunk_src(cc, W, csrc, crev, coob);
    const bool outl = t >= K1_LANE0 && t < K1_LANE0 + K1_LANES && cc < W;
    const uint32_t cboff = (uint32_t)csrc * (uint32_t)sizeof(T);
    const uint32_t pcoff = 4u * (uint32_t)cc;   // used by output lanes only
    // first staged column this lane reads (halo lanes read in range, unused)
    const int rbase = min(max(4 * t - 8, 0), K1_ROWBUF - 20);

    if (kp.hist) {
        for (int i = t; i < HIST_BINS / 2; i += 64) s_hist[i] = 0;
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

    const int vr0 = y0 - R, vr1 = y1 + R;
    const GLOBAL char* inb = (const GLOBAL char*)in;
    // byte offset of row vr, reflected at the top and bottom edges
    auto row_off = [&](int vr) -> uint32_t {
        const int ar = vr < 0 ? -vr - 1 : (vr >= H ? 2 * H - vr - 1 : vr);
        return (uint32_t)ar * (uint32_t)W * (uint32_t)sizeof(T);
    };

    // prefetch: row vr is loaded PF rows ahead into slot (vr - vr0) % NT,
    // so only PF + 1 slots are live at a time
    // (float chunks take 4 registers each: 3 rows ahead)
    constexpr int PFMAX = FLT ? 3 : 4;
    constexpr int PF = NT < PFMAX ? NT : PFMAX;
    Chunk4<T> pm[NT];
#pragma unroll
    for (int s = 0; s < PF; ++s) pm[s] = ld_chunk<T>(inb, row_off(min(vr0 + s, vr1 - 1)) + cboff);

    uint32_t zc = 0;  // zero outputs of this work item (wave-uniform)
    const uint64_t outl_m = __builtin_amdgcn_ballot_w64(outl);
    // the row loop, specialised on whether any chunk of the strip lies
    // outside the frame (mirrored / zeroed staging) or none does
    auto rows = [&](auto border_c) {
        constexpr bool BORDER = decltype(border_c)::value;
        int parity = 0;
        // whole groups of NT rows: the last group runs past vr1 (those rows
        // load row vr1 - 1 and emit nothing) so that no branch surrounds a
        // row; one there makes the compiler wait for every load and store in
        // flight before each row instead of only for the row's own chunk
        for (int base = vr0; base < vr1; base += NT) {
#pragma unroll
            for (int s = 0; s < NT; ++s) {
                const int vr = base + s;
                {
                float* bb_ = buf[parity][0];
                float* bg_ = buf[parity][1];
                parity ^= 1;
                // rows above and below the frame are zero in the Gaussian copy
                // (skipped taps) and mirrored in the boxcar copy
                const bool row_in = (vr >= 0) && (vr < H);
                {
                    f4v m = unpack_chunk<T>(pm[s]);
                    if constexpr (BORDER) {
                        if (crev) m = m.wzyx;
                        *(f4v*)(bb_ + 4 * t) = m;
                        if (coob || !row_in) m = f4v{0.f, 0.f, 0.f, 0.f};
                        *(f4v*)(bg_ + 4 * t) = m;
                    } else {
                        // every chunk of an interior strip is its own, so the
                        // two copies agree on rows inside the frame and the
                        // Gaussian copy of a row outside it is zero: one copy
                        // is staged
                        *(f4v*)(bb_ + 4 * t) = m;
                    }
                }
                pm[(s + PF) % NT] = ld_chunk<T>(inb, row_off(min(vr + PF, vr1 - 1)) + cboff);
                __syncthreads();

                // boxcar row pass on the mirrored copy (integer sums: exact)
                float v[20];
                lds_ld20(bb_ + rbase, v);
                float tb[4];
                if constexpr (FLT) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        float bx = 0.f;
#pragma unroll
                        for (int i = -R; i <= R; ++i) bx = bx + v[8 + q + i];
                        tb[q] = bx;
                    }
                } else {
                    float bx = v[8 - R];
#pragma unroll
                    for (int i = -R + 1; i <= R; ++i) bx = bx + v[8 + i];
                    tb[0] = bx;
#pragma unroll
                    for (int q = 1; q < 4; ++q) {
                        bx = (bx + v[8 + q + R]) - v[8 + q - 1 - R];
                        tb[q] = bx;
                    }
                }
                // Gaussian row pass: zeros outside the frame (skipped taps).
                // A row outside the frame sums to +0 (FMAs of +0 into +0)
                // and so divides to +0: interior strips skip it.
                auto gauss = [&](const float (&u)[20], float (&tg)[4]) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        float gs = 0.f;
#pragma unroll
                        for (int i = -R; i <= R; ++i) gs = __builtin_fmaf(w[i + R], u[8 + q + i], gs);
                        tg[q] = cdiv_t(gs, kp.gn, kp.rgn, FAST);
                    }
                };
                float tg[4];
                if constexpr (BORDER) {
                    float u[20];
                    lds_ld20(bg_ + rbase, u);
                    gauss(u, tg);
                } else if (row_in) {  // wave-uniform
                    gauss(v, tg);
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q) tg[q] = 0.f;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float gq = tg[q];
                    if constexpr (!FLT) vbox[q] = (vbox[q] - rb[s][q]) + tb[q];
                    rg[s][q] = gq;
                    rb[s][q] = tb[q];
                }

                if (vr >= y0 + R && vr < vr1) {  // wave-uniform
                    const int o = vr - R;
                    float out[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        float gsum = 0.f;
#pragma unroll
                        for (int j = 0; j < NT; ++j)
                            gsum = __builtin_fmaf(w[j], rg[(s + 1 + j) % NT][q], gsum);
                        if constexpr (FLT) {
                            float bsum = 0.f;
#pragma unroll
                            for (int j = 0; j < NT; ++j) bsum = bsum + rb[(s + 1 + j) % NT][q];
                            const float res = cdiv_t(gsum, kp.gn, kp.rgn, FAST) -
                                              cdiv_t(bsum, kp.kk, kp.rkk, FAST);
                            out[q] = res < 0.f ? 0.f : res;
                        } else {
                            const float res = cdiv_t(gsum, kp.gn, kp.rgn, FAST) -
                                              cdiv_t(vbox[q], kp.kk, kp.rkk, FAST);
                            out[q] = __builtin_fmaxf(res, 0.f);   // res is finite, never -0
                        }
                    }
                    if (outl) {
                        // the row offset goes through a scalar register, so
                        // each store adds one uniform value to this lane's
                        // column offset
                        const uint32_t ro = (uint32_t)__builtin_amdgcn_readfirstlane(
                            (int)(4u * (uint32_t)o * (uint32_t)W));
                        *(GLOBAL float4*)((GLOBAL char*)P + (ro + pcoff)) =
                            make_float4(out[0], out[1], out[2], out[3]);
                    }
                    if (kp.hist) {  // wave-uniform
                        // bin (bits >> 19) + 1 for positive values (finite and
                        // >= +0, so no clamp is needed): word bits >> 20, half
                        // bit 19; zero pixels (most of a frame) are counted by
                        // a ballot over the output lanes (one compare serves
                        // both) instead of contending for one word
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            uint32_t bits = __float_as_uint(out[q]);
                            const bool pos = bits != 0u;
                            zc += (uint32_t)__popcll(~__builtin_amdgcn_ballot_w64(pos) & outl_m);
                            // float input can give +inf or NaN: the last bin
                            // takes them (bin 4095: word 2047, low half)
                            if constexpr (FLT) bits = min(bits, 0x7ff7ffffu);
                            if (outl && pos)
                                atomicAdd(&s_hist[bits >> 20], 1u << ((bits >> 15) & 16u));
                        }
                    }
                }
                }
            }
        }
    };
    if (border) rows(std::true_type{});
    else rows(std::false_type{});

    if (kp.hist) {
        __syncthreads();
        uint32_t* gh = b.hist1 + (size_t)f * HIST_BINS;
        for (int i = t; i < HIST_BINS / 2; i += 64) {
            const uint32_t c = s_hist[i];
            if (c & 0xffffu) atomicAdd(&gh[2 * i + 1], c & 0xffffu);
            if ((c >> 16) && 2 * i + 2 < HIST_BINS) atomicAdd(&gh[2 * i + 2], c >> 16);
        }
        if (t == 0 && zc) atomicAdd(&gh[0], zc);
    }
}

// Generic bandpass for R > 8 (and rows whose width is not a multiple of 4): two straightforward passes through a global
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
static hipError_t launch_bp_t(const Batch& b, const K1Params& kp, bool fast, hipStream_t s) {
    if (fast) hipLaunchKernelGGL((k_bandpass<R, T, true>), dim3(b.k1_items), dim3(64), 0, s, b, kp);
    else hipLaunchKernelGGL((k_bandpass<R, T, false>), dim3(b.k1_items), dim3(64), 0, s, b, kp);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_bp_r(int R, const Batch& b, const K1Params& kp, bool fast, hipStream_t s) {
    switch (R) {
        case 1: return launch_bp_t<1, T>(b, kp, fast, s);
        case 2: return launch_bp_t<2, T>(b, kp, fast, s);
        case 3: return launch_bp_t<3, T>(b, kp, fast, s);
        case 4: return launch_bp_t<4, T>(b, kp, fast, s);
        case 5: return launch_bp_t<5, T>(b, kp, fast, s);
        case 6: return launch_bp_t<6, T>(b, kp, fast, s);
        case 7: return launch_bp_t<7, T>(b, kp, fast, s);
        case 8: return launch_bp_t<8, T>(b, kp, fast, s);
        default: return hipErrorInvalidValue;
    }
}

bool bandpass_fast_path(const Batch& b, const Geom& g) {
    return g.R >= 1 && g.R <= 8 && b.vec &&
           (b.in_type == IN_U16 || b.in_type == IN_U8 || b.in_type == IN_F32);
}

hipError_t launch_bandpass(const Batch& b, const Geom& g, const Geom* d_geom, int seg_rows,
                           hipStream_t s) {
    (void)d_geom;
    if (!bandpass_fast_path(b, g)) return hipErrorInvalidValue;
    K1Params kp;
    for (int i = 0; i < 17; ++i) kp.w[i] = (i < 2 * g.R + 1) ? g.w[i] : 0.f;
    kp.gn = g.gn; kp.rgn = g.rgn; kp.kk = g.kk; kp.rkk = g.rkk;
    kp.seg = seg_rows;
    kp.hist = b.hist1 != nullptr && g.has_pct;
    // the LDS histogram keeps 16-bit bins
    if (kp.hist && (int64_t)K1_STRIP * seg_rows >= 65536) return hipErrorInvalidValue;
    const bool fast = g.fastdiv_gn && g.fastdiv_kk;
    if (b.in_type == IN_U16) return launch_bp_r<uint16_t>(g.R, b, kp, fast, s);
    if (b.in_type == IN_F32) return launch_bp_r<float>(g.R, b, kp, fast, s);
    return launch_bp_r<uint8_t>(g.R, b, kp, fast, s);
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
// slots in raster order within each TILE_W x TILE_H tile.
// ---------------------------------------------------------------------------
struct DetParams {
    int a;              // window start offset
    int kd;             // window size (runtime copy for the generic kernel)
    int band;
    int diw;            // disk inside window (zero-valued maxima cannot yield rows)
    int has_pct;
    int walk_ri;        // disk radius of the walk run inside detect (0: none)
    uint32_t maxit;
    float minmass;
    uint64_t slot_cap;
};

// y[i] = max(x[i .. i+K-1]) for i in [0, N) (van Herk / Gil-Werman): the
// input is cut into blocks of K; a window is the suffix maximum of one block
// joined with the prefix maximum of the next.  All indices are compile-time.
//
// The detect kernels run it on float bit patterns as signed integers: their
// values are finite floats >= +0 (bandpass output, never -0) or -inf (outside
// the frame), for which signed integer order is float order and equal values
// have equal bits.  Integer max needs no NaN canonicalisation.
template <int N, int K>
__device__ __forceinline__ void sliding_max(const int32_t (&x)[N + K - 1], int32_t (&y)[N]) {
    constexpr int L = N + K - 1;
    int32_t S[N];   // suffix max to the end of the block, positions 0..N-1
    int32_t Pm[L];  // prefix max from the start of the block
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        const int be = ((j / K) + 1) * K - 1 < L - 1 ? ((j / K) + 1) * K - 1 : L - 1;
        if (j == be) {
            S[j] = x[j];
        } else if (j + 1 < N) {
            S[j] = max(x[j], S[j + 1]);
        } else {
            int32_t m = x[j];
#pragma unroll
            for (int q = j + 1; q <= be; ++q) m = max(m, x[q]);
            S[j] = m;
        }
    }
#pragma unroll
    for (int j = 0; j < L; ++j) Pm[j] = (j % K == 0) ? x[j] : max(Pm[j - 1], x[j]);
#pragma unroll
    for (int i = 0; i < N; ++i) y[i] = max(S[i], Pm[i + K - 1]);
}

// Centroid walk of the tile's candidates (centers.wgsl:23-109 and the walk
// of walk_cols.wgsl, as in k_refine): candidate k of the tile (raster
// order) walks in lane k % 256 of the workgroup, so a typical tile (a few
// dozen candidates) occupies part of one wave.  Disks inside the staged area
// are summed from LDS with the taps in the shader's order, others from P.
template <int RI, int LR, int LC, int LS>
__device__ __forceinline__ void walk_tile(const Batch& b, const DetParams& dp,
                                          const float* __restrict__ L,
                                          const GLOBAL uint16_t* __restrict__ slots,
                                          uint32_t total, uint32_t item,
                                          const GLOBAL float* __restrict__ P, int H, int W,
                                          int x0, int y0, int a, int tid) {
    constexpr int BW = 2 * RI + 1;
    const float sthr = 0.6f;
    const int band = dp.band;
    const size_t slot0 = (size_t)item * TILE_PX;
    const uint64_t cap = dp.slot_cap;
    float* rec = b.rec;
    for (uint32_t k = (uint32_t)tid; k < total; k += 256) {
        const uint32_t l = slots[k];  // written by this workgroup before the barrier
        const int lr0 = (int)(l / TILE_W), lc0 = (int)(l % TILE_W);
        int u = y0 + lr0, v = x0 + lc0;
        float m = 0.f, c0 = 0.f, c1 = 0.f;
        for (uint32_t it = 0; it < dp.maxit; ++it) {
            if (u - band < 0 || u + band >= H || v - band < 0 || v + band >= W) {
                m = 0.f; c0 = 0.f; c1 = 0.f;
            } else {
                // top-left corner of the disk's box in the staged area
                const int br = u - y0 + a - RI, bc = v - x0 + a - RI;
                float s0 = 0.f, s1 = 0.f, ms = 0.f;
                // a tap with a zero coefficient leaves its sum unchanged:
                // fma(d, 0, s) = s exactly (d is finite and s never -0,
                // since the sums start at +0), so those FMAs are skipped
                if (br >= 0 && br + BW <= LR && bc >= 0 && bc + BW <= LC) {
                    const float* box = L + br * LS + bc;
#pragma unroll
                    for (int di = -RI; di <= RI; ++di)
#pragma unroll
                        for (int dj = -RI; dj <= RI; ++dj) {
                            if (di * di + dj * dj > RI * RI) continue;
                            const float d = box[(di + RI) * LS + (dj + RI)];
                            if (di != 0) s0 = __builtin_fmaf(d, (float)di, s0);
                            if (dj != 0) s1 = __builtin_fmaf(d, (float)dj, s1);
                            ms = ms + d;
                        }
                } else {
#pragma unroll
                    for (int di = -RI; di <= RI; ++di) {
                        const GLOBAL float* row = P + (size_t)(u + di) * W + v;
#pragma unroll
                        for (int dj = -RI; dj <= RI; ++dj) {
                            if (di * di + dj * dj > RI * RI) continue;
                            const float d = row[dj];
                            if (di != 0) s0 = __builtin_fmaf(d, (float)di, s0);
                            if (dj != 0) s1 = __builtin_fmaf(d, (float)dj, s1);
                            ms = ms + d;
                        }
                    }
                }
                m = ms;
                c0 = s0 / ms;
                c1 = s1 / ms;
            }
            bool changed = false;
            if (c0 > sthr && u < H - 1) { u += 1; changed = true; }
            else if (c0 < -sthr && u > 0) { u -= 1; changed = true; }
            if (c1 > sthr && v < W - 1) { v += 1; changed = true; }
            else if (c1 < -sthr && v > 0) { v -= 1; changed = true; }
            if (!changed) break;
        }
        const bool valid = (m > dp.minmass) && (m != 0.f);
        rec[REC_MASS * cap + slot0 + k] = valid ? m : 0.f;
        rec[REC_Y * cap + slot0 + k] = (float)u + c0;
        rec[REC_X * cap + slot0 + k] = (float)v + c1;
        if (dp.has_pct) rec[REC_CENTER * cap + slot0 + k] = L[(lr0 + a) * LS + (lc0 + a)];
    }
}

// Bit-pattern range [base, base + width) of percentile histogram bin k
// (width 0 for bin 0 and for "no bin"): a value is listed when
// bits - base < width, unsigned.
__device__ __forceinline__ void pct_bin_range(int32_t k, uint32_t& base, uint32_t& width) {
    if (k <= 0) { base = 0; width = 0; return; }
    base = (uint32_t)(k - 1) << 19;
    width = (k >= HIST_BINS - 1) ? 0x80000000u - base : (1u << 19);
}

// One workgroup per 48x64 tile (TILE_W x TILE_H).  Lane c of every wave owns
// staged column c, frame column x0 - a + c: the 64 lanes hold the tile's 48
// columns and the dilation window's halo (KD - 1 <= 16 columns), so the
// column maxima come out in the lanes that own the columns, and the row
// window over them is formed by lane shifts (DPP and ds_bpermute).  The
// staged tile ((64 + KD - 1) x 64 floats: 19.5 KB for KD = 15) is never
// overwritten, so the walk reads it as staged; wave w tests output rows
// 16w .. 16w + 15.  With about 20 KB of LDS and 59 VGPRs, seven workgroups
// share a CU (the row ballots' scalar registers set that).
//
// RI > 0: the centroid walk of every candidate (refine) runs here too, on
// the staged tile: each candidate walks in one lane, reading its disk from
// LDS (from P when the walk leaves the staged area).  Every candidate walks;
// the percentile test on its start value is applied in dedup, which reads
// that value from the REC_CENTER plane.
template <int KD, int RI>
__global__ __launch_bounds__(256) void k_detect(Batch b, DetParams dp) {
    constexpr int N = 16;
    constexpr int A = (KD - 1) / 2;   // window start trunc(KD / 2 - 0.5), as dp.a
    constexpr int LR = TILE_H + KD - 1;
    constexpr int LS = 64;            // staged columns: one per lane
    constexpr int NIN = N + KD - 1;
    // largest power of two <= KD: the span the doubling steps reach
    constexpr int P2 = KD >= 16 ? 16 : KD >= 8 ? 8 : KD >= 4 ? 4 : KD >= 2 ? 2 : 1;
    static_assert(TILE_W + KD - 1 <= LS, "tile and window halo must fit the 64 lanes");
    __shared__ __attribute__((aligned(16))) float L[LR * LS];
    // candidates per tile row, then (in place) the row's first slot
    __shared__ uint32_t s_rowcnt[TILE_H];
    __shared__ uint32_t s_lcnt, s_total;

    const int tid = threadIdx.x;
    const int c = tid & 63;
    const int kw = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t f = b.tile_frame[item];
    const Frame fr = b.frames[f];
    const uint32_t local = item - fr.tile_base;
    const int H = (int)fr.H, W = (int)fr.W;
    const int x0 = (int)(local % fr.tx) * TILE_W;
    const int y0 = (int)(local / fr.tx) * TILE_H;
    const int v = x0 - A + c;         // this lane's frame column
    const GLOBAL float* P = (const GLOBAL float*)fr.P;
    if (tid == 0) s_lcnt = 0;

    // Staging: wave kw fills rows kw, kw + 4, ...; every row is one 256-byte
    // load across the wave.  Tiles whose staged area lies inside the frame
    // (most of them) load with one 32-bit byte offset per lane, advanced by
    // four rows per load (the frame base stays scalar); the others clamp
    // their indices into the frame (every load in bounds) and store -inf
    // outside it.  All loads are issued before any store.
    {
        constexpr int RPW = (LR + 3) / 4;
        constexpr int QF = RPW - 1;
        const GLOBAL char* Pb = (const GLOBAL char*)P;
        const bool inner = (x0 - A >= 0) && (x0 - A + LS <= W) && (y0 - A >= 0) && (y0 - A + LR <= H);
        float fv[RPW];
        if (inner) {  // workgroup-uniform
            // rows kw + 4q, q < QF, always exist; the last group only for the
            // first waves
            const bool last = kw + 4 * QF < LR;  // wave-uniform
            const uint32_t step = 16u * (uint32_t)W;
            const uint32_t o = 4u * (uint32_t)((y0 - A + kw) * W + v);
#pragma unroll
            for (int q = 0; q < QF; ++q) fv[q] = *(const GLOBAL float*)(Pb + (o + (uint32_t)q * step));
            if (last) fv[QF] = *(const GLOBAL float*)(Pb + (o + (uint32_t)QF * step));
            float* d = L + kw * LS + c;
#pragma unroll
            for (int q = 0; q < QF; ++q) d[4 * q * LS] = fv[q];
            if (last) d[4 * QF * LS] = fv[QF];
        } else {
            const uint32_t oc = 4u * (uint32_t)min(max(v, 0), W - 1);
            const bool cin = v >= 0 && v < W;
#pragma unroll
            for (int q = 0; q < RPW; ++q) {
                const int r = min(kw + 4 * q, LR - 1);  // repeated last row: same values
                const int u = y0 - A + r;
                const uint32_t rb = 4u * (uint32_t)(min(max(u, 0), H - 1) * W);
                fv[q] = *(const GLOBAL float*)(Pb + (rb + oc));
            }
#pragma unroll
            for (int q = 0; q < RPW; ++q) {
                const int r = min(kw + 4 * q, LR - 1);
                const int u = y0 - A + r;
                const bool ur = u >= 0 && u < H;  // wave-uniform
                L[r * LS + c] = (ur && cin) ? fv[q] : __uint_as_float(NEG_INF_BITS);
            }
        }
    }
    __syncthreads();

    // Column pass: this lane's column over the wave's 16 output rows and the
    // window's rows below them; the max filter works on the bit patterns
    // (see sliding_max).  x[m + A] is the centre of output row kw * 16 + m.
    const int32_t* Li = (const int32_t*)L;
    int32_t x[NIN];
#pragma unroll
    for (int m = 0; m < NIN; ++m) x[m] = Li[(kw * N + m) * LS + c];
    int32_t wm[N];
    sliding_max<N, KD>(x, wm);
    // Row window by lane shifts: after the doubling steps wm[m] in lane j is
    // the maximum over lanes j .. j + P2 - 1; the window of the output in
    // lane c is lanes c - A .. c - A + KD - 1, the union of the spans that
    // start at c - A and at c - A + KD - P2.  Lanes past 63 wrap around and
    // feed only lanes that hold no output (c <= A + TILE_W - 1 reads at
    // most lane A + TILE_W - 1 - A + KD - 1 <= 63).
    // Shifts by 1 and 2 are DPP wave shifts (lane j reads lane j + 1; lane
    // 63 reads 0, and lanes past the last output's window feed no output),
    // folded into the max; longer shifts go through ds_bpermute (measured:
    // a shift by 4 as four DPP moves is slower).
#pragma unroll
    for (int s2 = 1; s2 < P2; s2 *= 2) {
        if (s2 <= 2) {
#pragma unroll
            for (int m = 0; m < N; ++m) {
                int32_t t = wm[m];
#pragma unroll
                for (int q = 0; q < s2; ++q) t = __builtin_amdgcn_mov_dpp(t, 0x130, 0xf, 0xf, true);
                wm[m] = max(wm[m], t);
            }
        } else {
            const int src = 4 * ((c + s2) & 63);
#pragma unroll
            for (int m = 0; m < N; ++m) wm[m] = max(wm[m], __builtin_amdgcn_ds_bpermute(src, wm[m]));
        }
    }
    {
        constexpr int D0 = -A, D1 = KD - P2 - A;
        const int src0 = 4 * ((c + D0) & 63), src1 = 4 * ((c + D1) & 63);
#pragma unroll
        for (int m = 0; m < N; ++m) {
            const int32_t w0 = D0 == 0 ? wm[m] : __builtin_amdgcn_ds_bpermute(src0, wm[m]);
            const int32_t w1 = D1 == D0 ? w0
                               : D1 == 0 ? wm[m]
                                         : __builtin_amdgcn_ds_bpermute(src1, wm[m]);
            wm[m] = max(w0, w1);
        }
    }

    // candidates: the ballots of the wave's 16 rows stay in scalar registers
    // (bitwise tests: no short-circuit branches); output lanes are
    // A .. A + TILE_W - 1
    const bool olane = (c >= A) & (c < A + TILE_W);
    const uint64_t colmask = __builtin_amdgcn_ballot_w64(olane & (v < W) & (v - dp.band >= 0) &
                                                         (v + dp.band < W));
    const bool diw = dp.diw != 0;
    // rows m of the wave's run that can hold a candidate: band <= u0 + m <= H - 1 - band
    const int u0 = y0 + kw * N;
    const int rlo = max(dp.band - u0, 0), rhi = min(H - 1 - dp.band - u0, N - 1);
    const uint32_t rowbits = rhi >= rlo ? ((2u << rhi) - (1u << rlo)) : 0u;
    uint64_t bal[N];
    uint32_t mycnt = 0;
#pragma unroll
    for (int m = 0; m < N; ++m) {
        const uint64_t ge = __builtin_amdgcn_ballot_w64(x[m + A] >= wm[m]);
        const uint64_t nz = diw ? __builtin_amdgcn_ballot_w64(x[m + A] != 0) : ~0ull;
        bal[m] = (rowbits >> m) & 1u ? (colmask & ge & nz) : 0ull;
        mycnt = (c == m) ? (uint32_t)__popcll(bal[m]) : mycnt;
    }
    if (c < N) s_rowcnt[kw * N + c] = mycnt;

    if (dp.has_pct) {
        // values in the two histogram bins holding the order statistics go to
        // the tile's list (any order), each pixel from the one tile whose
        // output lanes hold it; out-of-frame centres are -inf and fall
        // outside every range.  The two order statistics have adjacent
        // ranks, so no value lies strictly between their bins: one range
        // from the start of the lower bin to the end of the upper one holds
        // exactly the two bins.
        uint32_t lb, lw, hb, hw;
        pct_bin_range(b.selstate[f * 4 + 0], lb, lw);
        pct_bin_range(b.selstate[f * 4 + 2], hb, hw);
        const uint32_t base = lw ? lb : hb;
        const uint32_t width = olane ? (hw ? (hb + hw - base) : lw) : 0u;
        uint32_t mbits = 0;
#pragma unroll
        for (int m = N - 1; m >= 0; --m) {
            const uint32_t bits = (uint32_t)x[m + A];
            mbits = 2u * mbits + ((bits - base < width) ? 1u : 0u);
        }
        if (__ballot(mbits != 0u)) {  // wave-uniform
            const uint32_t n = (uint32_t)__popc(mbits);
            const uint32_t inc = wave_incl_scan(n, c);
            const uint32_t tot = __shfl(inc, 63, 64);
            uint32_t lbase = 0;
            if (c == 0) lbase = atomicAdd(&s_lcnt, tot);
            lbase = __builtin_amdgcn_readfirstlane(lbase);
            uint32_t j = lbase + inc - n;
            GLOBAL uint32_t* list = (GLOBAL uint32_t*)b.list + (size_t)item * TILE_PX;
#pragma unroll
            for (int m = 0; m < N; ++m) {
                if (mbits & (1u << m)) list[j++] = (uint32_t)x[m + A];
            }
        }
    }
    __syncthreads();
    if (tid < 64) {
        const uint32_t cnt = s_rowcnt[tid];
        const uint32_t inc = wave_incl_scan(cnt, tid);
        s_rowcnt[tid] = inc - cnt;
        if (tid == 63) {
            b.tile_count[item] = inc;
            s_total = inc;
            if (dp.has_pct) b.tile_lcount[item] = s_lcnt;
        }
    }
    __syncthreads();
    GLOBAL uint16_t* slots = (GLOBAL uint16_t*)b.slots + (size_t)item * TILE_PX;
#pragma unroll
    for (int m = 0; m < N; ++m) {
        if (bal[m] == 0) continue;  // wave-uniform: most row pieces hold no candidate
        const int r = kw * N + m;
        if ((bal[m] >> c) & 1ull) {
            const uint32_t rank =
                s_rowcnt[r] + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal[m] >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)bal[m], 0u));
            slots[rank] = (uint16_t)(r * TILE_W + (c - A));
        }
    }

    if constexpr (RI > 0) {
        const uint32_t total = s_total;
        if (total == 0) return;  // workgroup-uniform
        __syncthreads();         // slots visible
        walk_tile<RI, LR, LS, LS>(b, dp, L, slots, total, item, P, H, W, x0, y0, A, tid);
    }
}

// Detect for dilation windows too wide for the lanes of k_detect (KD 18 to
// 26, separation 13 to 19), row first: one workgroup per 48x64 tile, the tile
// and its window halo staged in LDS (all loads issued before the first
// store); the row maxima overwrite the staged tile in place, so one tile
// needs (64+KD-1) x (48+KD+3) floats of LDS and six workgroups share a CU.
// Lane c holds output column c (lanes 48..63 hold none in the column pass).
//
// RI > 0: the centroid walk of every candidate (refine) runs here too, on
// the staged tile: after the candidate test the tile's columns are staged
// again from P (the row maxima overwrote them) and each candidate walks in
// one lane, reading its disk from LDS (from P when the walk leaves the
// staged area).  Every candidate walks; the percentile test on its start
// value is applied in dedup, which reads that value from the REC_CENTER
// plane.
template <int KD, int RI>
__global__ __launch_bounds__(256) void k_detect_rows(Batch b, DetParams dp) {
    constexpr int N = 16;
    constexpr int LR = TILE_H + KD - 1;
    constexpr int LC = TILE_W + KD - 1;
    constexpr int LS = ((LC + 3) / 4) * 4 + 4;
    constexpr int NIN = N + KD - 1;
    constexpr int NV4 = (NIN + 3) / 4;
    constexpr int RUNS = TILE_W / N;              // 16-column runs of an output row
    constexpr int NRUN = (LR * RUNS + 255) / 256;
    static_assert(LC - 64 <= 64, "halo wider than one extra column pass");
    static_assert(LC >= 64, "the first staging pass fills 64 columns");
    __shared__ __attribute__((aligned(16))) float L[LR * LS];
    // candidates per tile row, then (in place) the row's first slot
    __shared__ uint32_t s_rowcnt[TILE_H];
    __shared__ uint32_t s_lcnt, s_total;

    const int tid = threadIdx.x;
    const int a = dp.a;
    const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t f = b.tile_frame[item];
    const Frame fr = b.frames[f];
    const uint32_t local = item - fr.tile_base;
    const int H = (int)fr.H, W = (int)fr.W;
    const int x0 = (int)(local % fr.tx) * TILE_W;
    const int y0 = (int)(local / fr.tx) * TILE_H;
    const GLOBAL float* P = (const GLOBAL float*)fr.P;
    if (tid == 0) s_lcnt = 0;

    // Staging: wave w fills rows w, w+4, ...: lane = column (plus a second
    // pass over the KD-1 halo columns).  Indices are clamped into the frame
    // so every load is in bounds; out-of-frame cells become -inf.  All loads
    // are issued before any store.
    // Tiles whose staged area lies inside the frame (most of them) skip the
    // -inf selects.  Byte offsets are 32-bit (frames hold < 2^30 pixels), so
    // the loads take the frame base as a scalar address.
    const bool inner = (x0 - a >= 0) && (x0 - a + LC <= W) && (y0 - a >= 0) && (y0 - a + LR <= H);
    const GLOBAL char* Pb = (const GLOBAL char*)P;
    auto stage_tile = [&](bool halo) {
        constexpr int RPW = (LR + 3) / 4;
        const int lane = tid & 63;
        const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
        const int v1 = x0 - a + lane, v2 = v1 + 64;
        const bool h2 = halo && lane < LC - 64;
        if (inner) {  // workgroup-uniform
            // every staged row lies inside the frame: one byte offset per
            // lane, advanced by four rows per load, and LDS stores at
            // constant offsets from one address (rows wv + 4q, q < RPW - 1,
            // always exist; the last group only for the first waves)
            constexpr int QF = RPW - 1;
            const bool last = wv + 4 * QF < LR;  // wave-uniform
            const uint32_t step = 16u * (uint32_t)W;
            const uint32_t o = 4u * (uint32_t)((y0 - a + wv) * W + v1);
            float f1[RPW], f2[RPW];
#pragma unroll
            for (int q = 0; q < QF; ++q) f1[q] = *(const GLOBAL float*)(Pb + (o + (uint32_t)q * step));
            if (last) f1[QF] = *(const GLOBAL float*)(Pb + (o + (uint32_t)QF * step));
            if (h2) {
#pragma unroll
                for (int q = 0; q < QF; ++q)
                    f2[q] = *(const GLOBAL float*)(Pb + (o + 256u + (uint32_t)q * step));
                if (last) f2[QF] = *(const GLOBAL float*)(Pb + (o + 256u + (uint32_t)QF * step));
            }
            float* d = L + wv * LS + lane;
#pragma unroll
            for (int q = 0; q < QF; ++q) d[4 * q * LS] = f1[q];
            if (last) d[4 * QF * LS] = f1[QF];
            if (h2) {
#pragma unroll
                for (int q = 0; q < QF; ++q) d[4 * q * LS + 64] = f2[q];
                if (last) d[4 * QF * LS + 64] = f2[QF];
            }
            return;
        }
        const uint32_t o1 = 4u * (uint32_t)min(max(v1, 0), W - 1);
        const uint32_t o2 = 4u * (uint32_t)min(max(v2, 0), W - 1);
        float f1[RPW], f2[RPW];
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            const int r = min(wv + 4 * q, LR - 1);  // repeated last row: same values
            const int u = y0 - a + r;
            const uint32_t rb = 4u * (uint32_t)(min(max(u, 0), H - 1) * W);
            f1[q] = *(const GLOBAL float*)(Pb + (rb + o1));
            if (h2) f2[q] = *(const GLOBAL float*)(Pb + (rb + o2));
        }
        const bool c1 = v1 >= 0 && v1 < W;
        const bool c2 = v2 >= 0 && v2 < W;
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            const int r = min(wv + 4 * q, LR - 1);
            const int u = y0 - a + r;
            const bool ur = u >= 0 && u < H;  // wave-uniform
            L[r * LS + lane] = (ur && c1) ? f1[q] : __uint_as_float(NEG_INF_BITS);
            if (h2) L[r * LS + 64 + lane] = (ur && c2) ? f2[q] : __uint_as_float(NEG_INF_BITS);
        }
    };
    stage_tile(true);
    __syncthreads();

    // the max filter works on the bit patterns (see sliding_max)
    const int32_t* Li = (const int32_t*)L;
    // centre values of this thread's 16 outputs, read before the overwrite
    // (lanes past the tile read a column inside it; they hold no output)
    const int c = tid & 63;
    const int cc = min(c, TILE_W - 1);
    const int kw = __builtin_amdgcn_readfirstlane(tid >> 6);
    int32_t center[N];
#pragma unroll
    for (int m = 0; m < N; ++m) center[m] = Li[(kw * N + m + a) * LS + (cc + a)];

    // row pass: RM[r][c] = max(L[r][c .. c+KD-1]), written over L[r][0..47]
    int32_t y[NRUN][N];
#pragma unroll
    for (int k2 = 0; k2 < NRUN; ++k2) {
        const int run = tid + k2 * 256;
        if (run < LR * RUNS) {
            const int r = run / RUNS, c0 = (run % RUNS) * N;
            int32_t x[NIN];
            const int32_t* src = Li + r * LS + c0;
#pragma unroll
            for (int m = 0; m < NV4; ++m) {
                const int4 q = *(const int4*)(src + 4 * m);
                if (4 * m + 0 < NIN) x[4 * m + 0] = q.x;
                if (4 * m + 1 < NIN) x[4 * m + 1] = q.y;
                if (4 * m + 2 < NIN) x[4 * m + 2] = q.z;
                if (4 * m + 3 < NIN) x[4 * m + 3] = q.w;
            }
            sliding_max<N, KD>(x, y[k2]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < NRUN; ++k2) {
        const int run = tid + k2 * 256;
        if (run < LR * RUNS) {
            const int r = run / RUNS, c0 = (run % RUNS) * N;
            int32_t* dst = (int32_t*)L + r * LS + c0;
#pragma unroll
            for (int m = 0; m < N / 4; ++m)
                *(int4*)(dst + 4 * m) =
                    make_int4(y[k2][4 * m], y[k2][4 * m + 1], y[k2][4 * m + 2], y[k2][4 * m + 3]);
        }
    }
    __syncthreads();

    // column pass + candidate test: lane = column, wave = 16-row run
    const int v = x0 + c;
    int32_t cm[N];
    {
        int32_t x[NIN];
#pragma unroll
        for (int m = 0; m < NIN; ++m) x[m] = Li[(kw * N + m) * LS + cc];
        sliding_max<N, KD>(x, cm);
    }

    // candidates: the ballots of the wave's 16 rows stay in scalar registers
    // (bitwise tests: no short-circuit branches)
    // (lane masks combined in scalar registers: one vector compare each)
    const uint64_t colmask = __builtin_amdgcn_ballot_w64((c < TILE_W) & (v < W) &
                                                         (v - dp.band >= 0) & (v + dp.band < W));
    const bool diw = dp.diw != 0;
    // rows m of the wave's run that can hold a candidate: band <= u0 + m <= H - 1 - band
    const int u0 = y0 + kw * N;
    const int rlo = max(dp.band - u0, 0), rhi = min(H - 1 - dp.band - u0, N - 1);
    const uint32_t rowbits = rhi >= rlo ? ((2u << rhi) - (1u << rlo)) : 0u;
    uint64_t bal[N];
    uint32_t mycnt = 0;
#pragma unroll
    for (int m = 0; m < N; ++m) {
        const uint64_t ge = __builtin_amdgcn_ballot_w64(center[m] >= cm[m]);
        const uint64_t nz = diw ? __builtin_amdgcn_ballot_w64(center[m] != 0) : ~0ull;
        bal[m] = (rowbits >> m) & 1u ? (colmask & ge & nz) : 0ull;
        mycnt = (c == m) ? (uint32_t)__popcll(bal[m]) : mycnt;
    }
    if (c < N) s_rowcnt[kw * N + c] = mycnt;

    if (dp.has_pct) {
        // values in the two histogram bins holding the order statistics go to
        // the tile's list (any order); out-of-frame centres are -inf and
        // fall outside every range
        // The two order statistics have adjacent ranks, so no value lies
        // strictly between their bins: one range from the start of the lower
        // bin to the end of the upper one holds exactly the two bins.
        uint32_t lb, lw, hb, hw;
        pct_bin_range(b.selstate[f * 4 + 0], lb, lw);
        pct_bin_range(b.selstate[f * 4 + 2], hb, hw);
        const uint32_t base = lw ? lb : hb;
        const uint32_t width = c < TILE_W ? (hw ? (hb + hw - base) : lw) : 0u;
        uint32_t mbits = 0;
#pragma unroll
        for (int m = N - 1; m >= 0; --m) {
            const uint32_t bits = (uint32_t)center[m];
            mbits = 2u * mbits + ((bits - base < width) ? 1u : 0u);
        }
        if (__ballot(mbits != 0u)) {  // wave-uniform
            const uint32_t n = (uint32_t)__popc(mbits);
            const uint32_t inc = wave_incl_scan(n, c);
            const uint32_t tot = __shfl(inc, 63, 64);
            uint32_t base = 0;
            if (c == 0) base = atomicAdd(&s_lcnt, tot);
            base = __builtin_amdgcn_readfirstlane(base);
            uint32_t j = base + inc - n;
            GLOBAL uint32_t* list = (GLOBAL uint32_t*)b.list + (size_t)item * TILE_PX;
#pragma unroll
            for (int m = 0; m < N; ++m) {
                if (mbits & (1u << m)) list[j++] = (uint32_t)center[m];
            }
        }
    }
    __syncthreads();
    if (tid < 64) {
        const uint32_t cnt = s_rowcnt[tid];
        const uint32_t inc = wave_incl_scan(cnt, tid);
        s_rowcnt[tid] = inc - cnt;
        if (tid == 63) {
            b.tile_count[item] = inc;
            s_total = inc;
            if (dp.has_pct) b.tile_lcount[item] = s_lcnt;
        }
    }
    __syncthreads();
    GLOBAL uint16_t* slots = (GLOBAL uint16_t*)b.slots + (size_t)item * TILE_PX;
#pragma unroll
    for (int m = 0; m < N; ++m) {
        if (bal[m] == 0) continue;  // wave-uniform: most 64-pixel row pieces hold no candidate
        const int r = kw * N + m;
        if ((bal[m] >> c) & 1ull) {
            const uint32_t rank =
                s_rowcnt[r] + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal[m] >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)bal[m], 0u));
            slots[rank] = (uint16_t)(r * TILE_W + c);
        }
    }

    if constexpr (RI > 0) {
        const uint32_t total = s_total;
        if (total == 0) return;  // workgroup-uniform
        __syncthreads();         // column pass done with L; slots visible
        stage_tile(false);       // columns 0..63 again hold P
        __syncthreads();
        walk_tile<RI, LR, LC, LS>(b, dp, L, slots, total, item, P, H, W, x0, y0, a, tid);
    }
}

// Generic detect for window sizes without a specialisation: straightforward
// window loops over the staged tile.
__global__ __launch_bounds__(256) void k_detect_generic(Batch b, DetParams dp) {
    extern __shared__ float dyn[];
    __shared__ uint64_t s_mask[TILE_H];
    __shared__ uint32_t s_off[TILE_H];
    __shared__ uint32_t s_lcnt;
    if (threadIdx.x == 0) s_lcnt = 0;
    const int KD = dp.kd;
    const int LR = TILE_H + max(KD, 1) - 1, LC = TILE_W + max(KD, 1) - 1;
    float* L = dyn;                  // LR * LC
    float* RM = dyn + LR * LC;       // LR * TILE_W
    const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t f = b.tile_frame[item];
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
    const int c = tid & 63, k = tid >> 6;   // lanes c >= TILE_W hold no output
    const int v = x0 + c;
    int32_t bin_lo = -1, bin_hi = -1;
    if (dp.has_pct) {
        bin_lo = b.selstate[f * 4 + 0];
        bin_hi = b.selstate[f * 4 + 2];
    }
    uint32_t* list = b.list + (size_t)item * TILE_PX;
    uint32_t cbits = 0;
    for (int m = 0; m < 16; ++m) {
        const int r = k * 16 + m;
        const int u = y0 + r;
        const bool inb = (c < TILE_W) && (u < H) && (v < W);
        float cmx = __uint_as_float(NEG_INF_BITS), center = __uint_as_float(NEG_INF_BITS);
        if (c < TILE_W) {
            for (int i = 0; i < KD; ++i) cmx = fmaxf(cmx, RM[(r + i) * TILE_W + c]);
            center = L[(r + a) * LC + (c + a)];
        }
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
                if (c == 0) basei = atomicAdd(&s_lcnt, (uint32_t)__popcll(mm));
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
            if (dp.has_pct) b.tile_lcount[item] = s_lcnt;
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

// Dilation windows too wide for the staged tile of k_detect_generic (its
// (64 + kd - 1)^2 + (64 + kd - 1) * 64 floats exceed 60 KB from kd = 61, i.e.
// separation 44): the row maxima of max_rows.wgsl are written to a plane
// (the frame's tmp buffer) by one kernel, and the column test of
// walk_cols.wgsl:23-34 reads that plane in a second.  Each output takes kd
// loads (coalesced across the wave, reused through the caches), as the
// shaders do.
__global__ __launch_bounds__(256) void k_dilate_rows(Batch b, int KD, int a) {
    const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t f = b.tile_frame[item];
    const Frame fr = b.frames[f];
    const uint32_t local = item - fr.tile_base;
    const int H = (int)fr.H, W = (int)fr.W;
    const int x0 = (int)(local % fr.tx) * TILE_W;
    const int y0 = (int)(local / fr.tx) * TILE_H;
    const int c = threadIdx.x & 63, k = threadIdx.x >> 6;
    const int v = x0 + c;
    if (c >= TILE_W || v >= W) return;
    const float* P = fr.P;
    float* RM = fr.tmp;
    const int j0 = max(0, a - v), j1 = min(KD, W + a - v);   // in-bounds taps v - a + j
    for (int m = 0; m < 16; ++m) {
        const int u = y0 + k * 16 + m;
        if (u >= H) break;
        const float* row = P + (size_t)u * W;
        float mx = __uint_as_float(NEG_INF_BITS);
        for (int j = j0; j < j1; ++j) mx = fmaxf(mx, row[v - a + j]);
        RM[(size_t)u * W + v] = mx;
    }
}

__global__ __launch_bounds__(256) void k_detect_plane(Batch b, DetParams dp) {
    __shared__ uint64_t s_mask[TILE_H];
    __shared__ uint32_t s_off[TILE_H];
    __shared__ uint32_t s_lcnt;
    if (threadIdx.x == 0) s_lcnt = 0;
    __syncthreads();
    const int KD = dp.kd, a = dp.a;
    const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t f = b.tile_frame[item];
    const Frame fr = b.frames[f];
    const uint32_t local = item - fr.tile_base;
    const int H = (int)fr.H, W = (int)fr.W;
    const int x0 = (int)(local % fr.tx) * TILE_W;
    const int y0 = (int)(local / fr.tx) * TILE_H;
    const int tid = threadIdx.x, c = tid & 63, k = tid >> 6;
    const int v = x0 + c;
    const float* P = fr.P;
    const float* RM = fr.tmp;
    int32_t bin_lo = -1, bin_hi = -1;
    if (dp.has_pct) {
        bin_lo = b.selstate[f * 4 + 0];
        bin_hi = b.selstate[f * 4 + 2];
    }
    uint32_t* list = b.list + (size_t)item * TILE_PX;
    uint32_t cbits = 0;
    for (int m = 0; m < 16; ++m) {
        const int r = k * 16 + m;
        const int u = y0 + r;
        const bool inb = (c < TILE_W) && (u < H) && (v < W);
        float center = 0.f;
        bool cand = false;
        if (inb) {
            center = P[(size_t)u * W + v];
            const int i0 = max(0, a - u), i1 = min(KD, H + a - u);   // in-bounds rows u - a + i
            float cmx = __uint_as_float(NEG_INF_BITS);
            for (int i = i0; i < i1; ++i) cmx = fmaxf(cmx, RM[(size_t)(u - a + i) * W + v]);
            const bool interior = (u - dp.band >= 0) && (u + dp.band < H) &&
                                  (v - dp.band >= 0) && (v + dp.band < W);
            cand = interior && (center >= cmx) && !(dp.diw && center == 0.f);
        }
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
                if (c == 0) basei = atomicAdd(&s_lcnt, (uint32_t)__popcll(mm));
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
            if (dp.has_pct) b.tile_lcount[item] = s_lcnt;
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

static bool detect_specialised(int kd);

static size_t detect_generic_lds(int kd) {
    const int kd1 = max(kd, 1);
    return sizeof(float) * ((size_t)(TILE_H + kd1 - 1) * (TILE_W + kd1 - 1) +
                            (size_t)(TILE_H + kd1 - 1) * TILE_W);
}

bool detect_needs_plane(const Geom& g) {
    return !detect_specialised(g.kd) && detect_generic_lds(g.kd) > 60 * 1024;
}

template <int KD>
static hipError_t launch_det_t(const Batch& b, const DetParams& dp, hipStream_t s) {
    if (dp.a != (KD - 1) / 2) return hipErrorInvalidValue;   // the kernel's window start
    switch (dp.walk_ri) {
        case 0: hipLaunchKernelGGL((k_detect<KD, 0>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 1: hipLaunchKernelGGL((k_detect<KD, 1>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 2: hipLaunchKernelGGL((k_detect<KD, 2>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 3: hipLaunchKernelGGL((k_detect<KD, 3>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 4: hipLaunchKernelGGL((k_detect<KD, 4>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 5: hipLaunchKernelGGL((k_detect<KD, 5>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 6: hipLaunchKernelGGL((k_detect<KD, 6>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 7: hipLaunchKernelGGL((k_detect<KD, 7>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int KD>
static hipError_t launch_det_rows_t(const Batch& b, const DetParams& dp, hipStream_t s) {
    switch (dp.walk_ri) {
        case 0: hipLaunchKernelGGL((k_detect_rows<KD, 0>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 1: hipLaunchKernelGGL((k_detect_rows<KD, 1>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 2: hipLaunchKernelGGL((k_detect_rows<KD, 2>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 3: hipLaunchKernelGGL((k_detect_rows<KD, 3>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 4: hipLaunchKernelGGL((k_detect_rows<KD, 4>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 5: hipLaunchKernelGGL((k_detect_rows<KD, 5>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 6: hipLaunchKernelGGL((k_detect_rows<KD, 6>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 7: hipLaunchKernelGGL((k_detect_rows<KD, 7>), dim3(b.tiles), dim3(256), 0, s, b, dp); break;
        case 8:  // disk radius 8 (diameters 17, 18) walks here from window 22 on
            if constexpr (KD >= 22) {
                hipLaunchKernelGGL((k_detect_rows<KD, 8>), dim3(b.tiles), dim3(256), 0, s, b, dp);
                break;
            } else {
                return hipErrorInvalidValue;
            }
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

static bool detect_specialised(int kd) {
    switch (kd) {
        case 4: case 5: case 7: case 8: case 9: case 11: case 12: case 14: case 15: case 16:
        case 18: case 19: case 21: case 22: case 24: case 25: case 26:
            return true;
        default:
            return false;
    }
}

bool walk_in_detect_supported(const Geom& g) {
    const int max_ri = g.kd >= 22 ? 8 : 7;
    return detect_specialised(g.kd) && g.rint >= 1 && g.rint <= max_ri && g.max_iterations >= 1 &&
           !g.characterize && !g.cpu_processed;
}

hipError_t launch_detect(const Batch& b, const Geom& g, const Geom* d_geom, hipStream_t s) {
    (void)d_geom;
    DetParams dp;
    dp.a = g.a;
    dp.kd = g.kd;
    dp.band = g.band;
    dp.diw = g.disk_in_window;
    dp.has_pct = g.has_pct;
    dp.walk_ri = g.walk_in_detect ? g.rint : 0;
    dp.maxit = g.max_iterations;
    dp.minmass = g.minmass;
    dp.slot_cap = b.slot_cap;
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
        case 18: return launch_det_rows_t<18>(b, dp, s);
        case 19: return launch_det_rows_t<19>(b, dp, s);
        case 21: return launch_det_rows_t<21>(b, dp, s);
        case 22: return launch_det_rows_t<22>(b, dp, s);
        case 24: return launch_det_rows_t<24>(b, dp, s);
        case 25: return launch_det_rows_t<25>(b, dp, s);
        case 26: return launch_det_rows_t<26>(b, dp, s);
        default: break;
    }
    if (detect_needs_plane(g)) {
        hipLaunchKernelGGL(k_dilate_rows, dim3(b.tiles), dim3(256), 0, s, b, g.kd, g.a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_detect_plane, dim3(b.tiles), dim3(256), 0, s, b, dp);
        return hipGetLastError();
    }
    const size_t shm = detect_generic_lds(g.kd);
    hipLaunchKernelGGL(k_detect_generic, dim3(b.tiles), dim3(256), shm, s, b, dp);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// select2: exact order statistics from the value list, then the threshold
// ---------------------------------------------------------------------------
// Exact order statistics lo and hi of one frame by radix selection over the
// per-tile value lists: level A counts bits [18:8] of the values in each
// target bin, level B bits [7:0] of the values sharing the chosen level-A
// digit.  Both targets share every pass; each thread walks whole tile lists
// with 16-byte loads.
__global__ __launch_bounds__(1024) void k_select2(Batch b, float percentile) {
    __shared__ uint32_t s_ha[2][2048];
    __shared__ uint32_t s_res[2][3];
    __shared__ uint32_t s_wsum[2][16];
    const uint32_t f = blockIdx.x;
    const Frame fr = b.frames[f];
    const uint64_t n = (uint64_t)fr.H * fr.W;
    uint64_t lo, hi;
    double frac;
    pct_ranks(n, percentile, lo, hi, frac);
    const int t = threadIdx.x;
    if (t < 6) s_res[t / 3][t % 3] = 0;
    const int bin[2] = {b.selstate[f * 4 + 0], b.selstate[f * 4 + 2]};
    const int rank[2] = {b.selstate[f * 4 + 1], b.selstate[f * 4 + 3]};
    const bool need[2] = {bin[0] != 0, hi != lo && bin[1] != 0};
    const uint32_t ntiles = fr.tx * fr.ty;
    for (int pass = 0; pass < 2; ++pass) {
        const int nb = pass == 0 ? 2048 : 256;
        for (int i = t; i < 2 * 2048; i += 1024) s_ha[i / 2048][i % 2048] = 0;
        __syncthreads();
        const uint32_t sa0 = pass ? s_res[0][0] : 0u, sa1 = pass ? s_res[1][0] : 0u;
        for (uint32_t tl = t; tl < ntiles; tl += 1024) {
            const uint32_t it = fr.tile_base + tl;
            const uint32_t cnt = b.tile_lcount[it];
            const GLOBAL uint4* lst = (const GLOBAL uint4*)(b.list + (size_t)it * TILE_PX);
            for (uint32_t i4 = 0; i4 < (cnt + 3) / 4; ++i4) {
                const uint4 q = lst[i4];
                const uint32_t vals[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (i4 * 4 + e >= cnt) break;
                    const uint32_t bits = vals[e];
                    const int key = (int)pct_key(bits);
                    const uint32_t da = (bits >> 8) & 2047u;
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        if (!need[s] || key != bin[s]) continue;
                        if (pass == 0) atomicAdd(&s_ha[s][da], 1u);
                        else if (da == (s ? sa1 : sa0)) atomicAdd(&s_ha[s][bits & 255u], 1u);
                    }
                }
            }
        }
        __syncthreads();
        // Selection: a workgroup-wide exclusive prefix sum over the bins
        // (thread t holds bins 2t, 2t+1 of 2048, or bin t of 256); the thread
        // whose range holds the wanted rank picks the bin.
        const uint32_t want0 = pass == 0 ? (uint32_t)rank[0] : s_res[0][1];
        const uint32_t want1 = pass == 0 ? (uint32_t)rank[1] : s_res[1][1];
        uint32_t c0a, c0b, c1a, c1b;
        if (nb == 2048) {
            c0a = s_ha[0][2 * t]; c0b = s_ha[0][2 * t + 1];
            c1a = s_ha[1][2 * t]; c1b = s_ha[1][2 * t + 1];
        } else {
            c0a = t < 256 ? s_ha[0][t] : 0u; c0b = 0u;
            c1a = t < 256 ? s_ha[1][t] : 0u; c1b = 0u;
        }
        const int lane = t & 63, wv = t >> 6;
        const uint32_t l0 = c0a + c0b, l1 = c1a + c1b;
        const uint32_t i0 = wave_incl_scan(l0, lane), i1 = wave_incl_scan(l1, lane);
        if (lane == 63) { s_wsum[0][wv] = i0; s_wsum[1][wv] = i1; }
        __syncthreads();
        uint32_t e0 = i0 - l0, e1 = i1 - l1;
        for (int q = 0; q < wv; ++q) { e0 += s_wsum[0][q]; e1 += s_wsum[1][q]; }
        const int per = nb == 2048 ? 2 : 1;
#pragma unroll
        for (int sidx = 0; sidx < 2; ++sidx) {
            const uint32_t want = sidx ? want1 : want0;
            const uint32_t ex = sidx ? e1 : e0, ca = sidx ? c1a : c0a, cb = sidx ? c1b : c0b;
            if (need[sidx] && want >= ex && want < ex + ca + cb) {
                const bool first = want < ex + ca;
                const uint32_t sel = (uint32_t)(per * t) + (first ? 0u : 1u);
                const uint32_t below = first ? ex : ex + ca;
                if (pass == 0) {
                    s_res[sidx][0] = sel;
                    s_res[sidx][1] = want - below;
                } else {
                    s_res[sidx][2] = ((uint32_t)(bin[sidx] - 1) << 19) | (s_res[sidx][0] << 8) | sel;
                }
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        const float vlo = need[0] ? __uint_as_float(s_res[0][2]) : 0.f;
        const float vhi = (hi == lo) ? vlo : (need[1] ? __uint_as_float(s_res[1][2]) : 0.f);
        b.thr[f] = (double)vlo + frac * ((double)vhi - (double)vlo);
    }
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

// Characterisation columns of one row, at the candidate (start) pixel:
// execute_gpu.rs:165-172,286-299,346-352 (cpu_processed: Rg, ecc, signal
// with the walked mass) and centers.wgsl:64-80 (characterize: Rg_gpu,
// raw_mass_gpu, signal_gpu, ecc_gpu with the candidate pixel's own disk).
__device__ __forceinline__ void characterise(const Batch& b, const Geom* __restrict__ g,
                                          const float* __restrict__ P, const void* in, int H,
                                          int W, int u0, int v0, float mass, size_t s) {
    const uint64_t cap = b.slot_cap;
    const int band = g->band, rint = g->rint;
        if (g->cpu_processed) {
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
            b.rec[REC_RG * cap + s] = sqrtf(rg / mass);
            b.rec[REC_ECC * cap + s] = ecc / (mass - mid);
            b.rec[REC_SIGNAL * cap + s] = sig;
        }
        if (g->characterize) {
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
                        rawm = rawm + raw_px(in, b.in_type, idx);
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
            b.rec[REC_RG_GPU * cap + s] = o4[0];
            b.rec[REC_RAW_GPU * cap + s] = o4[1];
            b.rec[REC_SIG_GPU * cap + s] = o4[2];
            b.rec[REC_ECC_GPU * cap + s] = o4[3];
        }
}

// Candidates of RT_TILES consecutive tiles are packed over the 256 lanes of a
// workgroup (a tile holds ~16 candidates on average; one wave per tile would
// leave three quarters of the lanes idle while issuing the same number of
// scattered loads).
constexpr int RT_TILES = 16;

struct RefTile {
    const float* P;
    const void* in;
    int x0, y0, H, W;
    uint32_t f, cnt, pref;
};

template <int RI>
__global__ __launch_bounds__(256) void k_refine(Batch b, const Geom* __restrict__ g) {
    __shared__ RefTile s_t[RT_TILES];
    __shared__ uint32_t s_total;
    const uint32_t item0 = blockIdx.x * RT_TILES;
    const int nt = (int)min<uint32_t>(RT_TILES, b.tiles - item0);
    const int tid = threadIdx.x;
    if (tid < 64) {
        uint32_t cnt = 0;
        RefTile rt;
        if (tid < nt) {
            const uint32_t item = item0 + tid;
            const uint32_t f = b.tile_frame[item];
            const Frame fr = b.frames[f];
            const uint32_t local = item - fr.tile_base;
            rt.P = fr.P;
            rt.in = fr.in;
            rt.x0 = (int)(local % fr.tx) * TILE_W;
            rt.y0 = (int)(local / fr.tx) * TILE_H;
            rt.H = (int)fr.H;
            rt.W = (int)fr.W;
            rt.f = f;
            cnt = b.tile_count[item];
            rt.cnt = cnt;
        }
        const uint32_t inc = wave_incl_scan(cnt, tid);
        if (tid < nt) {
            rt.pref = inc - cnt;
            s_t[tid] = rt;
        }
        if (tid == 63) s_total = inc;
    }
    __syncthreads();
    const uint32_t total = s_total;
    const int band = g->band, rint = g->rint;
    const uint32_t maxit = g->max_iterations;
    const float minmass = g->minmass;
    const float sthr = 0.6f;
    float* rec = b.rec;
    const uint64_t cap = b.slot_cap;

    for (uint32_t gi = tid; gi < total; gi += 256) {
        int j = 0;
#pragma unroll
        for (int q = 1; q < RT_TILES; ++q)
            if (q < nt && s_t[q].pref <= gi) j = q;
        const RefTile& rt = s_t[j];
        const uint32_t i = gi - rt.pref;
        const uint32_t item = item0 + (uint32_t)j;
        const size_t slot0 = (size_t)item * TILE_PX;
        const int H = rt.H, W = rt.W;
        const float* __restrict__ P = rt.P;
        const double thr = g->has_pct ? b.thr[rt.f] : 0.0;
        const int x0 = rt.x0, y0 = rt.y0;
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
        if (valid && (g->cpu_processed || g->characterize))
            characterise(b, g, P, rt.in, H, W, u0, v0, mass, slot0 + i);
    }
}

// Refinement with LDS-staged disk boxes (1 <= RI <= 7, max_iterations >= 1).
// Lanes hold candidates (16 consecutive tiles per workgroup).  Every walk
// step needs the disk around the lane's current pixel; instead of each lane
// issuing (2RI+1)^2 scattered 4-byte loads, the wave stages the
// (2RI+1)x(2RI+1) boxes of up to 16 lanes at a time with row-coalesced loads
// (two load instructions per box for RI = 5), then each owner lane sums its
// disk from LDS in the shader's tap order.
template <int RI>
__global__ __launch_bounds__(256) void k_refine_staged(Batch b, const Geom* __restrict__ g) {
    constexpr int BW = 2 * RI + 1;
    constexpr int BOX = BW * BW;
    constexpr int G = 16;
    constexpr int NE = (BOX + 63) / 64;
    __shared__ RefTile s_t[RT_TILES];
    __shared__ uint32_t s_total;
    __shared__ float s_box[4][G][BOX];
    const uint32_t item0 = blockIdx.x * RT_TILES;
    const int nt = (int)min<uint32_t>(RT_TILES, b.tiles - item0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < 64) {
        uint32_t cnt = 0;
        RefTile rt;
        if (tid < nt) {
            const uint32_t item = item0 + tid;
            const uint32_t f = b.tile_frame[item];
            const Frame fr = b.frames[f];
            const uint32_t local = item - fr.tile_base;
            rt.P = fr.P;
            rt.in = fr.in;
            rt.x0 = (int)(local % fr.tx) * TILE_W;
            rt.y0 = (int)(local / fr.tx) * TILE_H;
            rt.H = (int)fr.H;
            rt.W = (int)fr.W;
            rt.f = f;
            cnt = b.tile_count[item];
            rt.cnt = cnt;
        }
        const uint32_t inc = wave_incl_scan(cnt, tid);
        if (tid < nt) {
            rt.pref = inc - cnt;
            s_t[tid] = rt;
        }
        if (tid == 63) s_total = inc;
    }
    __syncthreads();
    const uint32_t total = s_total;
    const int band = g->band;
    const uint32_t maxit = g->max_iterations;
    const float minmass = g->minmass;
    const int has_pct = g->has_pct;
    const float sthr = 0.6f;
    const uint64_t cap = b.slot_cap;

    for (uint32_t gb = 0; gb < total; gb += 256) {
        const uint32_t gi = gb + tid;
        const bool has = gi < total;
        int j = 0;
#pragma unroll
        for (int q = 1; q < RT_TILES; ++q)
            if (q < nt && s_t[q].pref <= gi) j = q;
        const RefTile& rt = s_t[j];
        const uint32_t i = has ? gi - rt.pref : 0;
        const size_t slot = (size_t)(item0 + (uint32_t)j) * TILE_PX + i;
        const int H = rt.H, W = rt.W;
        int u0 = 0, v0 = 0;
        bool active = false;
        if (has) {
            const uint32_t l = b.slots[slot];
            u0 = rt.y0 + (int)(l / TILE_W);
            v0 = rt.x0 + (int)(l % TILE_W);
            active = true;
            if (has_pct) active = (double)rt.P[(size_t)u0 * W + v0] > b.thr[rt.f];
        }
        const bool started = active;
        int u = u0, v = v0;
        float c0 = 0.f, c1 = 0.f, m = 0.f;
        uint32_t it = 0;
        while (__ballot(active)) {
            uint64_t pend = __ballot(active);
            while (pend) {
                int src[G];
                uint64_t gm = 0;
#pragma unroll
                for (int q = 0; q < G; ++q) {
                    src[q] = pend ? __ffsll((long long)pend) - 1 : -1;
                    if (pend) {
                        gm |= 1ull << src[q];
                        pend &= pend - 1;
                    }
                }
                float fb[G][NE];
#pragma unroll
                for (int q = 0; q < G; ++q) {
                    if (src[q] >= 0) {
                        const int uq = __shfl(u, src[q], 64), vq = __shfl(v, src[q], 64);
                        const int jq = __shfl(j, src[q], 64);
                        const RefTile& tq = s_t[jq];
                        if (uq - band >= 0 && uq + band < tq.H && vq - band >= 0 && vq + band < tq.W) {
                            const GLOBAL float* Pq = (const GLOBAL float*)tq.P;
#pragma unroll
                            for (int e = 0; e < NE; ++e) {
                                const int el = lane + 64 * e;
                                if (el < BOX) {
                                    const int r = el / BW, c = el - r * BW;
                                    fb[q][e] = Pq[(size_t)(uq - RI + r) * tq.W + (vq - RI + c)];
                                }
                            }
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < G; ++q) {
                    if (src[q] >= 0) {
#pragma unroll
                        for (int e = 0; e < NE; ++e) {
                            const int el = lane + 64 * e;
                            if (el < BOX) s_box[wave][q][el] = fb[q][e];
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if ((gm >> lane) & 1ull) {
                    const int sl = (int)__popcll(gm & lanemask_lt(lane));
                    if (u - band < 0 || u + band >= H || v - band < 0 || v + band >= W) {
                        m = 0.f; c0 = 0.f; c1 = 0.f;
                    } else {
                        const float* box = s_box[wave][sl];
                        float s0 = 0.f, s1 = 0.f, ms = 0.f;
#pragma unroll
                        for (int di = -RI; di <= RI; ++di)
#pragma unroll
                            for (int dj = -RI; dj <= RI; ++dj) {
                                if (di * di + dj * dj > RI * RI) continue;
                                const float d = box[(di + RI) * BW + (dj + RI)];
                                s0 = __builtin_fmaf(d, (float)di, s0);
                                s1 = __builtin_fmaf(d, (float)dj, s1);
                                ms = ms + d;
                            }
                        m = ms;
                        c0 = s0 / ms;
                        c1 = s1 / ms;
                    }
                    bool changed = false;
                    if (c0 > sthr && u < H - 1) { u += 1; changed = true; }
                    else if (c0 < -sthr && u > 0) { u -= 1; changed = true; }
                    if (c1 > sthr && v < W - 1) { v += 1; changed = true; }
                    else if (c1 < -sthr && v > 0) { v -= 1; changed = true; }
                    ++it;
                    if (!changed || it >= maxit) active = false;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        if (has) {
            const float y = (float)u + c0, x = (float)v + c1;
            const bool valid = started && (m > minmass) && (m != 0.f);
            b.rec[REC_MASS * cap + slot] = valid ? m : 0.f;
            b.rec[REC_Y * cap + slot] = started ? y : 0.f;
            b.rec[REC_X * cap + slot] = started ? x : 0.f;
            if (valid && (g->cpu_processed || g->characterize))
                characterise(b, g, rt.P, rt.in, H, W, u0, v0, m, slot);
        }
    }
}

hipError_t launch_refine(const Batch& b, const Geom& g, const Geom* d_geom, hipStream_t s) {
    if (g.walk_in_detect) return hipSuccess;  // done by detect
    const dim3 grid((b.tiles + RT_TILES - 1) / RT_TILES);
    if (g.max_iterations >= 1) {
        switch (g.rint) {
            case 1: hipLaunchKernelGGL(k_refine_staged<1>, grid, dim3(256), 0, s, b, d_geom); return hipGetLastError();
            case 2: hipLaunchKernelGGL(k_refine_staged<2>, grid, dim3(256), 0, s, b, d_geom); return hipGetLastError();
            case 3: hipLaunchKernelGGL(k_refine_staged<3>, grid, dim3(256), 0, s, b, d_geom); return hipGetLastError();
            case 4: hipLaunchKernelGGL(k_refine_staged<4>, grid, dim3(256), 0, s, b, d_geom); return hipGetLastError();
            case 5: hipLaunchKernelGGL(k_refine_staged<5>, grid, dim3(256), 0, s, b, d_geom); return hipGetLastError();
            case 6: hipLaunchKernelGGL(k_refine_staged<6>, grid, dim3(256), 0, s, b, d_geom); return hipGetLastError();
            case 7: hipLaunchKernelGGL(k_refine_staged<7>, grid, dim3(256), 0, s, b, d_geom); return hipGetLastError();
            default: break;
        }
    }
    hipLaunchKernelGGL(k_refine<0>, grid, dim3(256), 0, s, b, d_geom);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K4 dedup: one wave per block of 2x2 tiles (the block's tiles share one
// chain of dependent loads and one staged neighbourhood).  The valid records
// (mass != 0) of the tiles that can hold a neighbour, and that lie within
// `separation` of the bounding box of the block's own valid positions, are
// staged in LDS as 16-byte entries.  Each lane then tests its own candidate
// against every staged record (execute_gpu.rs:258-284: drop when a record
// strictly within `separation` has a larger mass, or an equal mass and a
// lower raster index).  Neighbourhoods larger than the staging area are
// staged and tested in rounds.  Outputs stay per tile: keep flags, kept
// count and kept rows per pixel row.
// ---------------------------------------------------------------------------
constexpr int DD_CAP = 256;   // staged records per wave and round
constexpr int DD_CELLS_1D = 16;
constexpr int DD_CELLS = DD_CELLS_1D * DD_CELLS_1D;   // cells of a staged neighbourhood

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
void k_dedup(Batch b, const Geom* __restrict__ g) {
    // staged records, one 16-byte entry each: mass, y, x, raster index (bits)
    __shared__ float4 s_rec[4][DD_CAP];
    __shared__ uint32_t s_kr[4][4][TILE_H];   // kept rows per block tile and pixel row
    // staged records are ordered by square cells (row-major, at most 16 x
    // 16): s_perm lists them cell by cell, cell c starting at s_cell[wave][c]
    // (s_cell[wave][DD_CELLS] = staged count)
    __shared__ uint16_t s_cell[4][DD_CELLS + 1];
    __shared__ uint32_t s_ccur[4][DD_CELLS / 2];   // 16-bit counts, then fill cursors
    __shared__ uint8_t s_perm[4][DD_CAP];
    static_assert(DD_CAP <= 256, "record numbers are bytes");
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t item = blockIdx.x * 4 + wave;
    if (item >= b.dblocks) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) s_kr[wave][q][lane] = 0;
    const uint64_t cap = b.slot_cap;
    const GLOBAL float* rmass = (const GLOBAL float*)(b.rec + REC_MASS * cap);
    const GLOBAL float* ry = (const GLOBAL float*)(b.rec + REC_Y * cap);
    const GLOBAL float* rx = (const GLOBAL float*)(b.rec + REC_X * cap);
    const GLOBAL uint16_t* slots = (const GLOBAL uint16_t*)b.slots;
    const float rad = g->sep;
    const float r2 = rad * rad;
    const int rty = g->dedup_rty, rtx = g->dedup_rtx;

    const uint32_t top = b.dblock_tile[item];   // the block's top-left tile
    const uint32_t f = b.tile_frame[top];
    const Frame fr = b.frames[f];
    const uint32_t local = top - fr.tile_base;
    const int W = (int)fr.W;
    const int tx = (int)(local % fr.tx), ty = (int)(local / fr.tx);
    const bool hx = tx + 1 < (int)fr.tx, hy = ty + 1 < (int)fr.ty;
    // tile q = 2 dy + dx of the block (absent ones hold no candidate)
    const bool has[4] = {true, hx, hy, hx && hy};
    uint32_t tq[4], cq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        tq[q] = top + (uint32_t)((q >> 1) * (int)fr.tx + (q & 1));
        cq[q] = has[q] ? b.tile_count[tq[q]] : 0u;
    }
    const uint32_t e1 = cq[0], e2 = e1 + cq[1], e3 = e2 + cq[2], own_n = e3 + cq[3];
    // own record j of the block: its tile q and slot
    auto own_q = [&](uint32_t j) { return j < e1 ? 0 : (j < e2 ? 1 : (j < e3 ? 2 : 3)); };
    auto own_slot = [&](uint32_t j, int q) {
        const uint32_t first = q == 0 ? 0u : (q == 1 ? e1 : (q == 2 ? e2 : e3));
        return (size_t)tq[q] * TILE_PX + (j - first);
    };
    // when the walk ran in detect, every candidate walked and the percentile
    // test on its start value (refine's `P > thr`) is applied here
    const bool use_thr = g->walk_in_detect && g->has_pct;
    const double thr = use_thr ? b.thr[f] : 0.0;
    const GLOBAL float* rcen = (const GLOBAL float*)(b.rec + REC_CENTER * cap);
    // a record's start value (percentile test), loaded beside its other
    // fields so that a record costs one round trip to memory
    auto cen_of = [&](size_t s) { return use_thr ? rcen[s] : 0.f; };
    auto ok = [&](float mass, float cen) {
        return (mass != 0.f) & (!use_thr | ((double)cen > thr));
    };
    const int y_lo = max(0, ty - rty), y_hi = min((int)fr.ty - 1, ty + 1 + rty);
    const int x_lo = max(0, tx - rtx), x_hi = min((int)fr.tx - 1, tx + 1 + rtx);
    const int nxr = x_hi - x_lo + 1;
    const int ntl = (y_hi - y_lo + 1) * nxr;

    // bounding box of the block's own valid positions
    float bymin = INFINITY, bymax = -INFINITY, bxmin = INFINITY, bxmax = -INFINITY;
    for (uint32_t j = lane; j < own_n; j += 64) {
        const size_t sl = own_slot(j, own_q(j));
        const float m0 = rmass[sl], c0 = cen_of(sl), y = ry[sl], x = rx[sl];
        if (ok(m0, c0)) {
            bymin = fminf(bymin, y); bymax = fmaxf(bymax, y);
            bxmin = fminf(bxmin, x); bxmax = fmaxf(bxmax, x);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        bymin = fminf(bymin, __shfl_xor(bymin, o, 64));
        bymax = fmaxf(bymax, __shfl_xor(bymax, o, 64));
        bxmin = fminf(bxmin, __shfl_xor(bxmin, o, 64));
        bxmax = fmaxf(bxmax, __shfl_xor(bxmax, o, 64));
    }
    if (!(bymin <= bymax)) {  // no valid own record
        for (uint32_t j = lane; j < own_n; j += 64) b.keep[own_slot(j, own_q(j))] = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (!has[q]) continue;
            if (lane == 0) b.tile_keep[tq[q]] = 0;
            b.rowk[(size_t)tq[q] * TILE_H + lane] = 0;
        }
        return;
    }
    // wave-uniform from here on: keep them in scalar registers
    auto uni = [](float v) {
        return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
    };
    const float margin = rad + 1.0f;
    bymin = uni(bymin - margin); bymax = uni(bymax + margin);
    bxmin = uni(bxmin - margin); bxmax = uni(bxmax + margin);
    // cells at least rad + 1 wide: a record strictly within `separation` of
    // a position lies in that position's cell or one of its eight neighbours
    // (cell coordinates are monotone; their rounding is far below one cell)
    const float ch = fmaxf(rad + 1.0f, fmaxf(bymax - bymin, bxmax - bxmin) * (1.0f / 15.0f));
    const float cinv = uni(1.0f / ch);
    const int ncy = min(DD_CELLS_1D, (int)((bymax - bymin) * cinv) + 1);
    const int ncx = min(DD_CELLS_1D, (int)((bxmax - bxmin) * cinv) + 1);
    auto cell_y = [&](float yy) { return min(ncy - 1, max(0, (int)((yy - bymin) * cinv))); };
    auto cell_x = [&](float xx) { return min(ncx - 1, max(0, (int)((xx - bxmin) * cinv))); };

    // The neighbourhood is staged in rounds of at most DD_CAP records: the
    // neighbour tiles are taken 64 at a time (lane q holds tile q's count),
    // their records enumerated 64 at a time (the loads of a round are
    // independent) and filtered by the bounding box; when the next 64 could
    // overflow the staging area, the own candidates are tested against what
    // is staged and staging resumes.  When the whole neighbourhood fits in
    // one round it is staged once and serves every batch of own candidates.
    uint32_t S = 0;
    bool resident = false;
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    uint32_t kept[4] = {0, 0, 0, 0};
    for (uint32_t ib = 0; ib < own_n; ib += 64) {
        const uint32_t j = ib + lane;
        const bool own = j < own_n;
        const int q = own ? own_q(j) : 0;
        float m = 0.f, y = 0.f, x = 0.f;
        int32_t idx = -1;
        size_t sl = 0;
        float cen = 0.f;
        if (own) {
            sl = own_slot(j, q);
            m = rmass[sl];
            y = ry[sl];
            x = rx[sl];
            cen = cen_of(sl);
            const uint32_t l = slots[sl];
            idx = ((ty + (q >> 1)) * TILE_H + (int)(l / TILE_W)) * W +
                  ((tx + (q & 1)) * TILE_W + (int)(l % TILE_W));
        }
        const bool valid = own && ok(m, cen);
        bool drop = false;
        // Counting sort of the n staged records by cell into s_perm: count,
        // scan (lane l owns cells 4l .. 4l + 3), then place through per-cell
        // cursors (the cell is recomputed from the staged position, so
        // nothing is held across the passes).
        auto cell_of = [&](uint32_t k) {
            const float4 rr = s_rec[wave][k];
            return cell_y(rr.y) * ncx + cell_x(rr.z);
        };
        // Counts and cursors are 16-bit halves of 32-bit words (at most
        // DD_CAP records).
        auto cell_add = [&](int c) {
            const uint32_t sh = 16u * (uint32_t)(c & 1);
            return (atomicAdd(&s_ccur[wave][c >> 1], 1u << sh) >> sh) & 0xffffu;
        };
        auto cell_sort = [&](uint32_t n) {
            constexpr int PC = DD_CELLS / 64;   // cells per lane in the scan
            static_assert(PC % 2 == 0, "a lane's cells fill whole words");
#pragma unroll
            for (int i = 0; i < PC / 2; ++i) s_ccur[wave][(PC / 2) * lane + i] = 0;
            wave_sync();
            for (uint32_t k = lane; k < n; k += 64) cell_add(cell_of(k));
            wave_sync();
            uint32_t c[PC], tot = 0;
#pragma unroll
            for (int i = 0; i < PC / 2; ++i) {
                const uint32_t w2 = s_ccur[wave][(PC / 2) * lane + i];
                c[2 * i] = w2 & 0xffffu;
                c[2 * i + 1] = w2 >> 16;
                tot += c[2 * i] + c[2 * i + 1];
            }
            uint32_t run = wave_incl_scan(tot, lane) - tot;
#pragma unroll
            for (int i = 0; i < PC / 2; ++i) {
                const uint32_t r0 = run, r1 = run + c[2 * i];
                s_cell[wave][PC * lane + 2 * i] = (uint16_t)r0;
                s_cell[wave][PC * lane + 2 * i + 1] = (uint16_t)r1;
                s_ccur[wave][(PC / 2) * lane + i] = r0 | (r1 << 16);
                run = r1 + c[2 * i + 1];
            }
            if (lane == 63) s_cell[wave][DD_CELLS] = (uint16_t)run;
            wave_sync();
            for (uint32_t k = lane; k < n; k += 64) s_perm[wave][cell_add(cell_of(k))] = (uint8_t)k;
            wave_sync();
        };
        // every lane tests its own candidate against the records of its cell
        // and the eight neighbouring cells, one contiguous run of cells per
        // cell row (execute_gpu.rs:258-284).  A lane's own record never
        // drops it (equal mass, equal raster index), and distinct candidates
        // have distinct raster indices, so no self test is needed.
        auto test_cells = [&]() {
            if (!valid) return;
            const int cy0 = cell_y(y), cx0 = cell_x(x);
            const int cxa = max(cx0 - 1, 0), cxb = min(cx0 + 1, ncx - 1);
            for (int cy = max(cy0 - 1, 0); cy <= min(cy0 + 1, ncy - 1); ++cy) {
                const uint32_t lo = s_cell[wave][cy * ncx + cxa];
                const uint32_t hi = s_cell[wave][cy * ncx + cxb + 1];
                for (uint32_t k = lo; k < hi; ++k) {
                    const float4 rr = s_rec[wave][s_perm[wave][k]];
                    const float dy = rr.y - y, dx = rr.z - x;
                    float d = dy * dy;
                    d = d + dx * dx;
                    const int32_t nidx = __float_as_int(rr.w);
                    drop |= d < r2 && (m < rr.x || (m == rr.x && idx > nidx));
                }
            }
        };
        if (__ballot(valid)) {  // wave-uniform
            if (resident) {
                test_cells();
            } else {
                bool single = true;   // one round held the whole neighbourhood
                for (int g0 = 0; g0 < ntl; g0 += 64) {
                    const int ng = min(64, ntl - g0);
                    uint32_t myc = 0, myit = 0;
                    if (lane < ng) {
                        const int qq = g0 + lane;
                        const int ny = y_lo + qq / nxr, nx = x_lo + qq % nxr;
                        myit = fr.tile_base + (uint32_t)(ny * (int)fr.tx + nx);
                        myc = b.tile_count[myit];
                    }
                    const uint32_t inc = wave_incl_scan(myc, lane);
                    const uint32_t exq = inc - myc;
                    const uint32_t T = __shfl(inc, 63, 64);
                    uint32_t jb = 0;
                    do {
                        S = 0;
                        for (; jb < T; jb += 64) {
                            const uint32_t jj = jb + lane;
                            // tile of record jj: the last qq whose first record is <= jj
                            int qt = 0;
                            for (int qq = 1; qq < ng; ++qq)  // wave-uniform: scalar read of lane qq
                                if ((uint32_t)__builtin_amdgcn_readlane((int)exq, qq) <= jj) qt = qq;
                            const uint32_t it2 = __shfl(myit, qt, 64);
                            const uint32_t i2 = jj - __shfl(exq, qt, 64);
                            const size_t s2 = (size_t)it2 * TILE_PX + i2;
                            bool take = false;
                            float nm = 0.f, yy = 0.f, xx = 0.f, cn = 0.f;
                            uint32_t l2 = 0;
                            if (jj < T) {   // every field of the record in one round trip
                                nm = rmass[s2];
                                yy = ry[s2];
                                xx = rx[s2];
                                cn = cen_of(s2);
                                l2 = slots[s2];
                                take = ok(nm, cn) & (yy >= bymin) & (yy <= bymax) & (xx >= bxmin) &
                                       (xx <= bxmax);
                            }
                            const uint64_t tm = __ballot(take);
                            const uint32_t nt = (uint32_t)__popcll(tm);
                            if (S + nt > (uint32_t)DD_CAP) break;   // test, then resume at jb
                            if (take) {
                                const uint32_t r = S + (uint32_t)__popcll(tm & lanemask_lt(lane));
                                const int qg = g0 + qt;
                                const int ny = y_lo + qg / nxr, nx = x_lo + qg % nxr;
                                const int32_t ri = (ny * TILE_H + (int)(l2 / TILE_W)) * W +
                                                   (nx * TILE_W + (int)(l2 % TILE_W));
                                s_rec[wave][r] = make_float4(nm, yy, xx, __int_as_float(ri));
                            }
                            S += nt;
                        }
                        wave_sync();
                        cell_sort(S);
                        test_cells();
                        if (jb < T || g0 + 64 < ntl) single = false;
                        wave_sync();   // every lane's reads done before the next round's writes
                    } while (jb < T);
                }
                resident = single;
            }
        }
        const bool keep = valid && !drop;
        if (own) b.keep[sl] = keep ? 1 : 0;
        if (keep) {
            const int row = idx / W - (ty + (q >> 1)) * TILE_H;
            atomicAdd(&s_kr[wave][q][row], 1u);
        }
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) kept[qq] += (uint32_t)__popcll(__ballot(keep && q == qq));
    }
    // kept rows of each tile in pixel rows <= r (the emit order is raster)
    wave_sync();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (!has[q]) continue;
        if (lane == 0) b.tile_keep[tq[q]] = kept[q];
        b.rowk[(size_t)tq[q] * TILE_H + lane] = wave_incl_scan(s_kr[wave][q][lane], lane);
    }
}

hipError_t launch_dedup(const Batch& b, const Geom& g, const Geom* d_geom, hipStream_t s) {
    (void)g;
    hipLaunchKernelGGL(k_dedup, dim3((b.dblocks + 3) / 4), dim3(256), 0, s, b, d_geom);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// tile scan: one workgroup per frame turns its tiles' keep counts into row
// offsets and sums its candidates and rows
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tile_scan(Batch b) {
    constexpr int PT = 8;   // tiles per thread in the one-round path
    __shared__ uint32_t s_part[4];
    __shared__ uint32_t s_cand[4];
    const uint32_t f = blockIdx.x;
    const Frame fr = b.frames[f];
    const uint32_t ntiles = fr.tx * fr.ty;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (ntiles <= 256 * PT) {  // workgroup-uniform
        // thread t takes tiles t * PT .. t * PT + PT - 1: every load of the
        // frame in one round, then one workgroup scan of the thread sums
        const uint32_t k0 = (uint32_t)t * PT;
        uint32_t kc[PT], cc[PT];
#pragma unroll
        for (int i = 0; i < PT; ++i) {
            const bool in = k0 + i < ntiles;
            kc[i] = in ? b.tile_keep[fr.tile_base + k0 + i] : 0u;
            cc[i] = in ? b.tile_count[fr.tile_base + k0 + i] : 0u;
        }
        uint32_t sum = 0, cands = 0;
#pragma unroll
        for (int i = 0; i < PT; ++i) { sum += kc[i]; cands += cc[i]; }
        const uint32_t inc = wave_incl_scan(sum, lane);
        cands = wave_sum_u(cands);
        if (lane == 63) s_part[wave] = inc;
        if (lane == 0) s_cand[wave] = cands;
        __syncthreads();
        uint32_t before = inc - sum;
        for (int w = 0; w < wave; ++w) before += s_part[w];
#pragma unroll
        for (int i = 0; i < PT; ++i) {
            if (k0 + i < ntiles) b.tile_row_off[fr.tile_base + k0 + i] = before;
            before += kc[i];
        }
        if (t == 0) {
            b.frame_rows[f] = s_part[0] + s_part[1] + s_part[2] + s_part[3];
            b.frame_cands[f] = s_cand[0] + s_cand[1] + s_cand[2] + s_cand[3];
        }
        return;
    }
    uint32_t carry = 0, cands = 0;
    for (uint32_t base = 0; base < ntiles; base += 256) {
        const uint32_t k = base + t;
        const uint32_t c = k < ntiles ? b.tile_keep[fr.tile_base + k] : 0u;
        cands += k < ntiles ? b.tile_count[fr.tile_base + k] : 0u;
        const uint32_t inc = wave_incl_scan(c, lane);
        if (lane == 63) s_part[wave] = inc;
        __syncthreads();
        uint32_t before = carry;
        for (int w = 0; w < wave; ++w) before += s_part[w];
        if (k < ntiles) b.tile_row_off[fr.tile_base + k] = before + inc - c;
        carry += s_part[0] + s_part[1] + s_part[2] + s_part[3];
        __syncthreads();
    }
    cands = wave_sum_u(cands);
    if (lane == 0) s_cand[wave] = cands;
    __syncthreads();
    if (t == 0) {
        b.frame_rows[f] = carry;
        b.frame_cands[f] = s_cand[0] + s_cand[1] + s_cand[2] + s_cand[3];
    }
}

// Raster order within a frame (get_work scans masses in raster order,
// execute_gpu.rs:108-195): a kept row at pixel row r of tile t in tile row tr
// comes after every kept row of the earlier tile rows, of tiles left of t in
// pixel rows <= r and of tiles right of t in pixel rows < r; the rows of tile
// t itself are in raster order already.  With K[t][r] = kept rows of tile t
// in pixel rows <= r (from dedup):
//   rowoff[t][r] = base(tr) + sum_{t' < t} K[t'][r] + sum_{t' > t} K[t'][r - 1].
// One wave per tile row, lane = pixel row.  Tile rows of up to 48 tiles
// (frames up to 2304 columns) load every K once, in one round, and take
// K[t][r - 1] from the lane below; wider ones load in chunks of 32 tiles.
__global__ __launch_bounds__(256) void k_row_scan(Batch b) {
    constexpr int CH = 32;
    constexpr int C1 = 48;
    const uint32_t f = blockIdx.x;
    const Frame fr = b.frames[f];
    const int lane = threadIdx.x & 63;
    const uint32_t tr = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (tr >= fr.ty) return;
    const uint32_t t0 = fr.tile_base + tr * fr.tx;
    const uint32_t base = b.tile_row_off[t0];
    const uint32_t* K = b.rowk + (size_t)t0 * TILE_H + lane;
    if (fr.tx <= C1) {  // wave-uniform
        uint32_t kle[C1], klt[C1];
#pragma unroll
        for (int j = 0; j < C1; ++j) kle[j] = (uint32_t)j < fr.tx ? K[(size_t)j * TILE_H] : 0u;
        uint32_t total = 0;
#pragma unroll
        for (int j = 0; j < C1; ++j) {
            const uint32_t up = __shfl_up(kle[j], 1, 64);
            klt[j] = lane > 0 ? up : 0u;
            total += klt[j];
        }
        uint32_t pre = 0, cum = 0;
#pragma unroll
        for (int j = 0; j < C1; ++j) {
            if ((uint32_t)j < fr.tx) {
                cum += klt[j];
                b.rowoff[(size_t)(t0 + j) * TILE_H + lane] = base + pre + (total - cum);
                pre += kle[j];
            }
        }
        return;
    }
    uint32_t total_lt = 0;
    for (uint32_t q0 = 0; q0 < fr.tx; q0 += CH) {
        uint32_t v[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j)
            v[j] = (q0 + j < fr.tx && lane > 0) ? K[(size_t)(q0 + j) * TILE_H - 1] : 0u;
#pragma unroll
        for (int j = 0; j < CH; ++j) total_lt += v[j];
    }
    uint32_t pre_le = 0, cum_lt = 0;
    for (uint32_t q0 = 0; q0 < fr.tx; q0 += CH) {
        uint32_t kle[CH], klt[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const bool in = q0 + j < fr.tx;
            kle[j] = in ? K[(size_t)(q0 + j) * TILE_H] : 0u;
            klt[j] = (in && lane > 0) ? K[(size_t)(q0 + j) * TILE_H - 1] : 0u;
        }
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            if (q0 + j < fr.tx) {
                cum_lt += klt[j];
                b.rowoff[(size_t)(t0 + q0 + j) * TILE_H + lane] = base + pre_le + (total_lt - cum_lt);
                pre_le += kle[j];
            }
        }
    }
}

hipError_t launch_tile_scan(const Batch& b, uint32_t max_ty, hipStream_t s) {
    hipLaunchKernelGGL(k_tile_scan, dim3(b.nframes), dim3(256), 0, s, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_row_scan, dim3(b.nframes, (max_ty + 3) / 4), dim3(256), 0, s, b);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// fscan: frame row bases (continuing the running count across batches)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_frame_scan(Batch b) {
    __shared__ uint64_t s_part[256];
    __shared__ uint64_t s_cand[256];
    const int t = threadIdx.x;
    const uint32_t n = b.nframes;
    const uint32_t per = (n + 255) / 256;
    uint64_t sum = 0, cs = 0;
    for (uint32_t i = 0; i < per; ++i) {
        const uint32_t k = t * per + i;
        if (k < n) {
            sum += b.frame_rows[k];
            cs += b.frame_cands[k];
        }
    }
    s_part[t] = sum;
    s_cand[t] = cs;
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
    if (t == 255) {
        *b.row_carry = before;
        uint64_t c = 0;
        for (int i = 0; i < 256; ++i) c += s_cand[i];
        *b.cand_total += c;
    }
}

hipError_t launch_frame_scan(const Batch& b, hipStream_t s) {
    hipLaunchKernelGGL(k_frame_scan, dim3(1), dim3(256), 0, s, b);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K5 emit: rows [frame, mass, y, x, (Rg, ecc, signal), (raw_mass, raw_bg_median,
// raw_mass_corrected: left for k_raw), (Rg_gpu, raw_mass_gpu, signal_gpu,
// ecc_gpu)] in frame order, raster order of the candidate pixel
// within a frame (row positions from k_tile_scan).
// ---------------------------------------------------------------------------
// EMIT_TPW tiles per wave, EMIT_LPT lanes each: a tile holds about a dozen
// candidates, and a wave's time is mostly its chain of dependent loads, so
// fewer, fuller waves finish sooner.
constexpr int EMIT_TPW = 4;
constexpr int EMIT_LPT = 64 / EMIT_TPW;
__global__ __launch_bounds__(256) void k_emit(Batch b, const Geom* __restrict__ g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int part = lane / EMIT_LPT, hl = lane % EMIT_LPT;
    const uint64_t hmask = (~0ull >> (64 - EMIT_LPT)) << (part * EMIT_LPT);
    const uint32_t item = (blockIdx.x * 4 + wave) * EMIT_TPW + part;
    if (__ballot(item < b.tiles) == 0) return;  // wave-uniform
    const bool tile_in = item < b.tiles;
    const uint32_t it = tile_in ? item : 0u;   // loads stay in bounds
    const uint32_t f = b.tile_frame[it];
    const Frame fr = b.frames[f];
    const uint32_t local = it - fr.tile_base;
    const int W = (int)fr.W;
    const int tx = (int)(local % fr.tx), ty = (int)(local / fr.tx);
    const uint32_t cnt = tile_in ? b.tile_count[it] : 0u;
    const size_t slot0 = (size_t)it * TILE_PX;
    const uint64_t cap = b.slot_cap;
    const int nc = g->ncols;
    const uint64_t fbase = b.frame_row_base[f];
    uint32_t before = 0;  // kept rows of this tile before the chunk
    for (uint32_t ib = 0; ib < cnt; ib += EMIT_LPT) {
        const uint32_t i = ib + (uint32_t)hl;
        // the keep flag, the slot and the position record of every candidate
        // of the chunk are loaded together (all in bounds for i < cnt); only
        // the row offset waits for the slot
        const bool in = i < cnt;
        const size_t s = slot0 + (in ? i : 0u);
        const bool kf = b.keep[s];
        const uint32_t l = b.slots[s];
        const float rm = b.rec[REC_MASS * cap + s];
        const float ryv = b.rec[REC_Y * cap + s];
        const float rxv = b.rec[REC_X * cap + s];
        const bool k = in && kf;
        const uint64_t msk = __ballot(k) & hmask;
        if (k) {
            const uint64_t r = fbase + b.rowoff[(size_t)it * TILE_H + l / TILE_W] + before +
                               (uint64_t)__popcll(msk & lanemask_lt(lane));
            if (r < b.rows_cap) {
                float* o = b.rows + r * (uint64_t)nc;
                o[0] = (float)fr.frame_index;
                o[1] = rm;
                o[2] = ryv;
                o[3] = rxv;
                int c = 4;
                if (g->cpu_processed) {
                    o[c++] = b.rec[REC_RG * cap + s];
                    o[c++] = b.rec[REC_ECC * cap + s];
                    o[c++] = b.rec[REC_SIGNAL * cap + s];
                }
                c += g->raw_cols;   // raw signal columns, written by k_raw
                if (g->characterize) {
                    o[c++] = b.rec[REC_RG_GPU * cap + s];
                    o[c++] = b.rec[REC_RAW_GPU * cap + s];
                    o[c++] = b.rec[REC_SIG_GPU * cap + s];
                    o[c++] = b.rec[REC_ECC_GPU * cap + s];
                }
                b.rows_cand[r] = (int64_t)(ty * TILE_H + (int)(l / TILE_W)) * W +
                                 (tx * TILE_W + (int)(l % TILE_W));
            }
        }
        before += (uint32_t)__popcll(msk);
    }
}

hipError_t launch_emit(const Batch& b, const Geom& g, const Geom* d_geom, hipStream_t s) {
    (void)g;
    hipLaunchKernelGGL(k_emit, dim3((b.tiles + 4 * EMIT_TPW - 1) / (4 * EMIT_TPW)), dim3(256), 0, s,
                       b, d_geom);
    return hipGetLastError();
}

}  // namespace gt
