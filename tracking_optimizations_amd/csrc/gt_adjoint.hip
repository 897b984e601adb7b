// gt_adjoint.hip -- gfx950 adjoint of the locate path (refine + bandpass).
//
// For every output row the particle's forward state is recomputed from the
// batch's bandpass output P with the forward's own arithmetic: the walk from
// the row's candidate pixel gives the last evaluated disk (u_e, v_e), its
// mass m and offsets c = s / m.  With the walk decisions held fixed,
//   y = u_final + c0,  x = v_final + c1,  mass = m,
// so on the disk pixels q
//   dy/dP[q] = (i_q - c0) / m,  dx/dP[q] = (j_q - c1) / m,  dmass/dP[q] = 1,
// zero where the bandpass clamp is active (P[q] == 0).  The bandpass
//   P[q] = max(0, G[q] / gn^2 - B[q] / k^2)
// is linear in the input where P > 0: G is the separable Gaussian with taps
// outside the frame skipped (weight w_i w_j at offset (j, i)), B the boxcar
// over the (2R+1)^2 window with one mirror reflection (a pixel reached twice
// by reflection counts twice).  The Jacobian of (y, x, mass) with respect to
// the input pixels is therefore the disk adjoint pulled back through the
// transposed filters; its support is the S x S patch, S = 2 (rint + R) + 1,
// centred on (u_e, v_e).  The Poisson localisation error of each output is
//   sigma^2 = sum_r J[r]^2 * lambda[r],  lambda[r] = max(I[r], 0).
//
// The vector-Jacobian product used by the Python autograd wrapper adds the
// patches of a frame's rows in row order, one workgroup per frame, so the
// gradient is deterministic.
#include <algorithm>

#include "gt_internal.h"

#pragma clang fp contract(off)

namespace gt {

namespace {

__device__ __forceinline__ float in_px(const void* in, int type, size_t i) {
    if (type == IN_U16) return (float)((const uint16_t*)in)[i];
    if (type == IN_U8) return (float)((const uint8_t*)in)[i];
    return ((const float*)in)[i];
}

// Boxcar multiplicity: how many offsets d in [-R, R] map input row (or
// column) q + d onto r under the single mirror reflection -i-1 / 2n-i-1.
__device__ __forceinline__ float box_count(int q, int r, int R, int n) {
    float c = 0.f;
    if (abs(r - q) <= R) c += 1.f;           // direct
    if (abs(-r - 1 - q) <= R) c += 1.f;      // reflected at the low edge
    if (abs(2 * n - r - 1 - q) <= R) c += 1.f;  // reflected at the high edge
    return c;
}

constexpr int ADJ_TAPS = 1024;  // disk taps ((2 rint + 1)^2 <= 1024: rint <= 15)

__global__ __launch_bounds__(64) void k_adjoint(Batch b, const Geom* __restrict__ g, AdjOut ao,
                                               uint32_t frame_pos0) {
    __shared__ float s_a[3][ADJ_TAPS];
    __shared__ int16_t s_ti[ADJ_TAPS], s_tj[ADJ_TAPS];
    __shared__ float s_w[33];
    const int lane = threadIdx.x;
    const uint32_t f = blockIdx.x;
    const Frame fr = b.frames[f];
    const int H = (int)fr.H, W = (int)fr.W;
    const float* __restrict__ P = fr.P;
    const int band = g->band, rint = g->rint, R = g->R;
    const uint32_t maxit = g->max_iterations;
    const float gn2 = g->gn * g->gn, kk = g->kk;
    const int S = 2 * (rint + R) + 1;
    const uint64_t base = b.frame_row_base[f];
    const uint32_t nrows = b.frame_rows[f];
    const float sthr = 0.6f;

    if (lane < 2 * R + 1) s_w[lane] = g->w[lane];
    // disk tap offsets in the shader's order
    int ntap = 0;
    for (int i = -rint; i <= rint; ++i)
        for (int j = -rint; j <= rint; ++j) {
            if (i * i + j * j > rint * rint) continue;
            if (ntap % 64 == lane) { s_ti[ntap] = (int16_t)i; s_tj[ntap] = (int16_t)j; }
            ++ntap;
        }
    __syncthreads();

    for (uint32_t rl = blockIdx.y; rl < nrows; rl += gridDim.y) {
        const uint64_t row = base + rl;
        if (row >= b.rows_cap) break;
        const int64_t pix = b.rows_cand[row];
        const int u0 = (int)(pix / W), v0 = (int)(pix % W);
        // forward walk, identical arithmetic to the refine stage
        int u = u0, v = v0, ue = u0, ve = v0;
        float m = 0.f, c0 = 0.f, c1 = 0.f;
        for (uint32_t it = 0; it < maxit; ++it) {
            ue = u;
            ve = v;
            if (u - band < 0 || u + band >= H || v - band < 0 || v + band >= W) {
                m = 0.f; c0 = 0.f; c1 = 0.f;
            } else {
                float s0 = 0.f, s1 = 0.f, ms = 0.f;
                for (int t = 0; t < ntap; ++t) {
                    const int i = s_ti[t], j = s_tj[t];
                    const float d = P[(size_t)(u + i) * W + (v + j)];
                    s0 = __builtin_fmaf(d, (float)i, s0);
                    s1 = __builtin_fmaf(d, (float)j, s1);
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
            if (!changed) break;
        }
        const int py0 = ue - rint - R, px0 = ve - rint - R;
        if (m == 0.f || ue - band < 0 || ue + band >= H || ve - band < 0 || ve + band >= W) {
            // not an output row of the forward (mass 0): nothing to pull back
            if (lane == 0) {
                for (int o = 0; o < 3; ++o) ao.sigma[row * 3 + o] = 0.f;
                ao.origin[row * 2 + 0] = py0;
                ao.origin[row * 2 + 1] = px0;
                ao.rowpos[row] = (int32_t)(frame_pos0 + f);
            }
            if (ao.jac)
                for (int p = lane; p < 3 * S * S; p += 64) ao.jac[(size_t)row * 3 * S * S + p] = 0.f;
            continue;
        }
        // disk adjoint (zero where the clamp is active)
        __syncthreads();
        const float rm = 1.f / m;
        for (int t = lane; t < ntap; t += 64) {
            const int i = s_ti[t], j = s_tj[t];
            const float on = P[(size_t)(ue + i) * W + (ve + j)] > 0.f ? 1.f : 0.f;
            s_a[0][t] = on * (((float)i - c0) * rm);
            s_a[1][t] = on * (((float)j - c1) * rm);
            s_a[2][t] = on;
        }
        __syncthreads();
        // pull back through the transposed bandpass, one patch pixel per lane
        double var[3] = {0.0, 0.0, 0.0};
        for (int p = lane; p < S * S; p += 64) {
            const int ry = py0 + p / S, rx = px0 + p % S;
            float jo[3] = {0.f, 0.f, 0.f};
            if (ry >= 0 && ry < H && rx >= 0 && rx < W) {
                for (int t = 0; t < ntap; ++t) {
                    const int qy = ue + s_ti[t], qx = ve + s_tj[t];
                    const int dy = ry - qy, dx = rx - qx;
                    const float wg = (abs(dy) <= R && abs(dx) <= R)
                                         ? s_w[dy + R] * s_w[dx + R] / gn2 : 0.f;
                    const float wb = box_count(qy, ry, R, H) * box_count(qx, rx, R, W) / kk;
                    const float kq = wg - wb;
#pragma unroll
                    for (int o = 0; o < 3; ++o) jo[o] = __builtin_fmaf(s_a[o][t], kq, jo[o]);
                }
                const double lam = (double)fmaxf(in_px(fr.in, b.in_type, (size_t)ry * W + rx), 0.f);
#pragma unroll
                for (int o = 0; o < 3; ++o) var[o] += (double)jo[o] * (double)jo[o] * lam;
            }
            if (ao.jac) {
#pragma unroll
                for (int o = 0; o < 3; ++o)
                    ao.jac[((size_t)row * 3 + o) * (size_t)(S * S) + p] = jo[o];
            }
        }
#pragma unroll
        for (int o = 0; o < 3; ++o)
            for (int s = 32; s > 0; s >>= 1) var[o] += __shfl_xor(var[o], s, 64);
        if (lane == 0) {
            // output order: y, x, mass
            ao.sigma[row * 3 + 0] = (float)sqrt(var[0]);
            ao.sigma[row * 3 + 1] = (float)sqrt(var[1]);
            ao.sigma[row * 3 + 2] = (float)sqrt(var[2]);
            ao.origin[row * 2 + 0] = py0;
            ao.origin[row * 2 + 1] = px0;
            ao.rowpos[row] = (int32_t)(frame_pos0 + f);
        }
    }
}

// The same adjoint with the filters applied separably.  One wave per row:
//  * the walk is replayed with the forward's own arithmetic (taps summed in
//    the shader's order by every lane), the disk's D x D box of P staged in
//    LDS by all lanes first (D = 2 rint + 1);
//  * the disk adjoint A_o (o = y, x, mass) lives on that box (0 off the
//    disk and where the clamp is active);
//  * the Gaussian part of the pullback, sum_q A(q) w[ry - qy] w[rx - qx] /
//    gn^2, is a column pass T[i][c] = sum_j w[c - j] A[i][j] then a row pass
//    sum_i w[r - i] T[i][c]; the boxcar part the same with the mirror
//    multiplicities box_count (a table per patch row and column).
// Sums are reassociated against the per-pair form of k_adjoint (results
// agree to binary32 rounding).  Dynamic LDS, sized from the geometry.
__global__ __launch_bounds__(64) void k_adjoint_sep(Batch b, const Geom* __restrict__ g,
                                                   AdjOut ao, uint32_t frame_pos0) {
    extern __shared__ float sh[];
    const int lane = threadIdx.x;
    const uint32_t f = blockIdx.x;
    const Frame fr = b.frames[f];
    const int H = (int)fr.H, W = (int)fr.W;
    const float* __restrict__ P = fr.P;
    const int band = g->band, rint = g->rint, R = g->R;
    const uint32_t maxit = g->max_iterations;
    const float gn2 = g->gn * g->gn, kk = g->kk;
    const int D = 2 * rint + 1, S = 2 * (rint + R) + 1, NW = 2 * R + 1;
    const uint64_t base = b.frame_row_base[f];
    const uint32_t nrows = b.frame_rows[f];
    const float sthr = 0.6f;
    // LDS: box[D*D] A[3][D*D] TG[3][D*S] TB[3][D*S] BX[S*D] BY[S*D] w[NW]
    float* box = sh;
    float* A = box + D * D;
    float* TG = A + 3 * D * D;
    float* TB = TG + 3 * D * S;
    float* BX = TB + 3 * D * S;
    float* BY = BX + S * D;
    float* w = BY + S * D;
    for (int k = lane; k < NW; k += 64) w[k] = g->w[k];
    __syncthreads();

    for (uint32_t rl = blockIdx.y; rl < nrows; rl += gridDim.y) {
        const uint64_t row = base + rl;
        if (row >= b.rows_cap) break;
        const int64_t pix = b.rows_cand[row];
        const int u0 = (int)(pix / W), v0 = (int)(pix % W);
        // forward walk (refine's arithmetic, every lane alike)
        int u = u0, v = v0, ue = u0, ve = v0;
        float m = 0.f, c0 = 0.f, c1 = 0.f;
        bool staged = false;   // box holds P around (ue, ve)
        for (uint32_t it = 0; it < maxit; ++it) {
            ue = u;
            ve = v;
            if (u - band < 0 || u + band >= H || v - band < 0 || v + band >= W) {
                m = 0.f; c0 = 0.f; c1 = 0.f;
                staged = false;
            } else {
                __syncthreads();
                for (int e = lane; e < D * D; e += 64)
                    box[e] = P[(size_t)(u + e / D - rint) * W + (v + e % D - rint)];
                __syncthreads();
                staged = true;
                float s0 = 0.f, s1 = 0.f, ms = 0.f;
                for (int i = -rint; i <= rint; ++i)
                    for (int j = -rint; j <= rint; ++j) {
                        if (i * i + j * j > rint * rint) continue;
                        const float d = box[(i + rint) * D + (j + rint)];
                        s0 = __builtin_fmaf(d, (float)i, s0);
                        s1 = __builtin_fmaf(d, (float)j, s1);
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
            if (!changed) break;
        }
        const int py0 = ue - rint - R, px0 = ve - rint - R;
        if (m == 0.f || !staged) {
            // not an output row of the forward (mass 0): nothing to pull back
            if (lane == 0) {
                for (int o = 0; o < 3; ++o) ao.sigma[row * 3 + o] = 0.f;
                ao.origin[row * 2 + 0] = py0;
                ao.origin[row * 2 + 1] = px0;
                ao.rowpos[row] = (int32_t)(frame_pos0 + f);
            }
            if (ao.jac)
                for (int p = lane; p < 3 * S * S; p += 64) ao.jac[(size_t)row * 3 * S * S + p] = 0.f;
            continue;
        }
        // disk adjoint on the box; boxcar multiplicity tables
        const float rm = 1.f / m;
        for (int e = lane; e < D * D; e += 64) {
            const int i = e / D - rint, j = e % D - rint;
            float a0 = 0.f, a1 = 0.f, a2 = 0.f;
            if (i * i + j * j <= rint * rint) {
                const float on = box[e] > 0.f ? 1.f : 0.f;
                a0 = on * (((float)i - c0) * rm);
                a1 = on * (((float)j - c1) * rm);
                a2 = on;
            }
            A[e] = a0;
            A[D * D + e] = a1;
            A[2 * D * D + e] = a2;
        }
        for (int e = lane; e < S * D; e += 64) {
            const int c = e / D, j = e % D;
            BX[e] = box_count(ve + j - rint, px0 + c, R, W);
            BY[e] = box_count(ue + j - rint, py0 + c, R, H);
        }
        __syncthreads();
        // column pass: TG_o[i][c] = sum_j w[c - j] A_o[i][j], TB_o with BX
        for (int e = lane; e < D * S; e += 64) {
            const int i = e / S, c = e % S;
            float tg[3] = {0.f, 0.f, 0.f}, tb[3] = {0.f, 0.f, 0.f};
            for (int j = 0; j < D; ++j) {
                const int k = c - j;
                const float wg = (k >= 0 && k < NW) ? w[k] : 0.f;
                const float wb = BX[c * D + j];
#pragma unroll
                for (int o = 0; o < 3; ++o) {
                    const float a = A[o * D * D + i * D + j];
                    tg[o] = __builtin_fmaf(wg, a, tg[o]);
                    tb[o] = __builtin_fmaf(wb, a, tb[o]);
                }
            }
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                TG[o * D * S + e] = tg[o];
                TB[o * D * S + e] = tb[o];
            }
        }
        __syncthreads();
        // row pass and the Poisson variance, one patch pixel per lane
        double var[3] = {0.0, 0.0, 0.0};
        for (int p = lane; p < S * S; p += 64) {
            const int r = p / S, c = p % S;
            const int ry = py0 + r, rx = px0 + c;
            float jo[3] = {0.f, 0.f, 0.f};
            if (ry >= 0 && ry < H && rx >= 0 && rx < W) {
                float sg[3] = {0.f, 0.f, 0.f}, sb[3] = {0.f, 0.f, 0.f};
                for (int i = 0; i < D; ++i) {
                    const int k = r - i;
                    const float wg = (k >= 0 && k < NW) ? w[k] : 0.f;
                    const float wb = BY[r * D + i];
#pragma unroll
                    for (int o = 0; o < 3; ++o) {
                        sg[o] = __builtin_fmaf(wg, TG[o * D * S + i * S + c], sg[o]);
                        sb[o] = __builtin_fmaf(wb, TB[o * D * S + i * S + c], sb[o]);
                    }
                }
#pragma unroll
                for (int o = 0; o < 3; ++o) jo[o] = sg[o] / gn2 - sb[o] / kk;
                const double lam = (double)fmaxf(in_px(fr.in, b.in_type, (size_t)ry * W + rx), 0.f);
#pragma unroll
                for (int o = 0; o < 3; ++o) var[o] += (double)jo[o] * (double)jo[o] * lam;
            }
            if (ao.jac) {
#pragma unroll
                for (int o = 0; o < 3; ++o)
                    ao.jac[((size_t)row * 3 + o) * (size_t)(S * S) + p] = jo[o];
            }
        }
#pragma unroll
        for (int o = 0; o < 3; ++o)
            for (int sft = 32; sft > 0; sft >>= 1) var[o] += __shfl_xor(var[o], sft, 64);
        if (lane == 0) {
            ao.sigma[row * 3 + 0] = (float)sqrt(var[0]);
            ao.sigma[row * 3 + 1] = (float)sqrt(var[1]);
            ao.sigma[row * 3 + 2] = (float)sqrt(var[2]);
            ao.origin[row * 2 + 0] = py0;
            ao.origin[row * 2 + 1] = px0;
            ao.rowpos[row] = (int32_t)(frame_pos0 + f);
        }
        __syncthreads();   // the next row overwrites the tables
    }
}

// grad[f] += sum over the frame's rows, in row order, of
// gy * J_y + gx * J_x + gm * J_mass on each row's patch.
__global__ __launch_bounds__(256) void k_adjoint_vjp(const float* __restrict__ jac,
                                                    const int32_t* __restrict__ origin,
                                                    const uint64_t* __restrict__ frame_rows,
                                                    const float* __restrict__ grad_rows,
                                                    float* __restrict__ grad, int S, int H,
                                                    int W) {
    const uint32_t f = blockIdx.x;
    float* gf = grad + (size_t)f * H * W;
    const int SS = S * S;
    for (uint64_t row = frame_rows[f]; row < frame_rows[f + 1]; ++row) {
        const float gy = grad_rows[row * 3 + 0], gx = grad_rows[row * 3 + 1],
                    gm = grad_rows[row * 3 + 2];
        const int py0 = origin[row * 2 + 0], px0 = origin[row * 2 + 1];
        const float* jr = jac + row * 3 * (size_t)SS;
        for (int p = threadIdx.x; p < SS; p += blockDim.x) {
            const int ry = py0 + p / S, rx = px0 + p % S;
            if (ry < 0 || ry >= H || rx < 0 || rx >= W) continue;
            float v = gy * jr[p];
            v = __builtin_fmaf(gx, jr[SS + p], v);
            v = __builtin_fmaf(gm, jr[2 * SS + p], v);
            gf[(size_t)ry * W + rx] += v;
        }
        __syncthreads();  // rows in order: overlapping patches add deterministically
    }
}

__global__ void k_fill_f32(float* p, uint64_t n, float v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

int adjoint_patch(const Geom& g) { return 2 * (g.rint + g.R) + 1; }

bool adjoint_supported(const Geom& g) {
    return g.rint >= 0 && (2 * g.rint + 1) * (2 * g.rint + 1) <= ADJ_TAPS && g.R <= 16;
}

size_t adjoint_sep_lds(const Geom& g) {
    const size_t D = 2 * (size_t)g.rint + 1, S = (size_t)adjoint_patch(g), NW = 2 * (size_t)g.R + 1;
    return 4 * (D * D * 4 + 6 * D * S + 2 * S * D + NW);
}

hipError_t launch_adjoint(const Batch& b, const Geom& g, const Geom* d_geom, const AdjOut& ao,
                          uint32_t frame_pos0, hipStream_t s) {
    const size_t lds = adjoint_sep_lds(g);
    if (lds <= 48 * 1024)
        hipLaunchKernelGGL(k_adjoint_sep, dim3(b.nframes, 64), dim3(64), lds, s, b, d_geom, ao,
                           frame_pos0);
    else
        hipLaunchKernelGGL(k_adjoint, dim3(b.nframes, 64), dim3(64), 0, s, b, d_geom, ao,
                           frame_pos0);
    return hipGetLastError();
}

hipError_t launch_adjoint_vjp(const float* jac, const int32_t* origin, const uint64_t* frame_rows,
                              uint32_t nframes, const float* grad_rows, float* grad, int S, int H,
                              int W, hipStream_t s) {
    const uint64_t n = (uint64_t)nframes * H * W;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_f32, dim3(std::max(blocks, 1u)), dim3(256), 0, s, grad, n, 0.f);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nframes == 0) return e;
    hipLaunchKernelGGL(k_adjoint_vjp, dim3(nframes), dim3(256), 0, s, jac, origin, frame_rows,
                       grad_rows, grad, S, H, W);
    return hipGetLastError();
}

}  // namespace gt
