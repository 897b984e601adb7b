// gt_synth.hip -- seeded synthetic uint16 movies generated on the device.
//
// Pixel (f, y, x) = min(Poisson(lambda), 65535) with
//   lambda = bg[f] + sum_i amp_i * exp(-((y - y_i)^2 + (x - x_i)^2) / (2 sigma_i^2))
// over spots within 5 sigma (point sampled at pixel centres).  Random numbers
// come from Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed by
// (seed, frame) with the pixel index as counter, so every pixel is an
// independent, reproducible stream.  Poisson variates use Hoermann's
// transformed rejection with squeeze, PTRS (Insurance: Mathematics and
// Economics 12, 1993), for lambda >= 10 and sequential inversion below.
#include <algorithm>

#include "gt_internal.h"

namespace gt {

namespace {

struct Philox {
    uint32_t k0, k1;
    uint32_t c0, c1, c2, c3;
    __device__ uint4 next() {
        uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
        uint32_t a = k0, b = k1;
#pragma unroll
        for (int r = 0; r < 10; ++r) {
            const uint32_t lo0 = 0xD2511F53u * x0;
            const uint32_t hi0 = __umulhi(0xD2511F53u, x0);
            const uint32_t lo1 = 0xCD9E8D57u * x2;
            const uint32_t hi1 = __umulhi(0xCD9E8D57u, x2);
            x0 = hi1 ^ x1 ^ a;
            x1 = lo1;
            x2 = hi0 ^ x3 ^ b;
            x3 = lo0;
            a += 0x9E3779B9u;
            b += 0xBB67AE85u;
        }
        c2 += 1;  // next draw of this pixel's stream
        return make_uint4(x0, x1, x2, x3);
    }
};

__device__ __forceinline__ double u01(uint32_t hi, uint32_t lo) {
    // 53-bit uniform in [0, 1)
    const uint64_t m = ((uint64_t)(hi >> 5) << 26) | (uint64_t)(lo >> 6);
    return (double)m * (1.0 / 9007199254740992.0);
}

__device__ uint32_t poisson(double lam, Philox& rng) {
    if (!(lam > 0.0)) return 0;
    if (lam < 10.0) {
        const uint4 r = rng.next();
        const double U = u01(r.x, r.y);
        double p = exp(-lam), F = p;
        uint32_t k = 0;
        while (U > F && k < 1000) {
            ++k;
            p *= lam / (double)k;
            F += p;
        }
        return k;
    }
    const double slam = sqrt(lam);
    const double loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < 64; ++it) {
        const uint4 r = rng.next();
        const double U = u01(r.x, r.y) - 0.5;
        const double V = u01(r.z, r.w);
        const double us = 0.5 - fabs(U);
        const double kd = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return kd < 0.0 ? 0u : (uint32_t)kd;
        if (kd < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <=
            -lam + kd * loglam - lgamma(kd + 1.0))
            return (uint32_t)kd;
    }
    // 64 consecutive rejections has probability far below 1e-60; fall back
    // to the rounded mean so the kernel always terminates.
    return (uint32_t)floor(lam + 0.5);
}

}  // namespace

__global__ __launch_bounds__(256) void k_synth(uint16_t* __restrict__ out, uint64_t nframes,
                                               uint32_t h, uint32_t w,
                                               const float4* __restrict__ spots,
                                               const uint32_t* __restrict__ cell_off,
                                               uint32_t cx, uint32_t cy, uint32_t cell,
                                               const float* __restrict__ bg, uint64_t seed,
                                               uint64_t frame0) {
    const uint64_t px = (uint64_t)h * w;
    const uint64_t total = px * nframes;
    const uint32_t ncell = cx * cy;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t f = g / px;
        const uint64_t p = g - f * px;
        const int u = (int)(p / w), v = (int)(p % w);
        const uint32_t* co = cell_off + f * (ncell + 1);
        float lam = bg[f];
        const int ccy = u / (int)cell, ccx = v / (int)cell;
        for (int dy = -1; dy <= 1; ++dy) {
            const int yy = ccy + dy;
            if (yy < 0 || yy >= (int)cy) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int xx = ccx + dx;
                if (xx < 0 || xx >= (int)cx) continue;
                const uint32_t c = (uint32_t)yy * cx + (uint32_t)xx;
                for (uint32_t s = co[c]; s < co[c + 1]; ++s) {
                    const float4 sp = spots[s];
                    const float ddy = (float)u - sp.x, ddx = (float)v - sp.y;
                    const float d2 = ddy * ddy + ddx * ddx;
                    const float s2 = sp.w * sp.w;
                    if (d2 <= 25.0f * s2) lam += sp.z * expf(-d2 / (2.0f * s2));
                }
            }
        }
        const uint64_t fid = frame0 + f;
        Philox rng;
        rng.k0 = (uint32_t)seed ^ (uint32_t)(fid * 0x9E3779B97F4A7C15ull);
        rng.k1 = (uint32_t)(seed >> 32) ^ (uint32_t)((fid * 0x9E3779B97F4A7C15ull) >> 32);
        rng.c0 = (uint32_t)p;
        rng.c1 = (uint32_t)(p >> 32);
        rng.c2 = 0;
        rng.c3 = 0x5eed;
        const uint32_t k = poisson((double)lam, rng);
        out[g] = (uint16_t)(k > 65535u ? 65535u : k);
    }
}

hipError_t launch_synth(uint16_t* frames, uint64_t nframes, uint32_t h, uint32_t w,
                        const float4* spots, const uint32_t* cell_off, uint32_t cells_x,
                        uint32_t cells_y, uint32_t cell, const float* bg, uint64_t seed,
                        uint64_t frame0, hipStream_t s) {
    const uint64_t total = (uint64_t)h * w * nframes;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_synth, dim3(blocks ? blocks : 1), dim3(256), 0, s, frames, nframes, h, w,
                       spots, cell_off, cells_x, cells_y, cell, bg, seed, frame0);
    return hipGetLastError();
}

}  // namespace gt
