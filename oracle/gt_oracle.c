/*
 * gt_oracle.c -- CPU oracle for the per-frame locate path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
 * tests/, the smoke entry point and bench.py's CPU-baseline leg load it.
 * See gt_oracle.h for the arithmetic contract and the reference files it
 * restates.
 *
 * Build (oracle/Makefile): gcc -O3 -march=native -ffp-contract=off -fopenmp
 * so that no multiply-add is contracted except where fma() is written.
 */
#include "gt_oracle.h"

#include <alloca.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* gpu_setup.rs:45-60 (dilation size (2*sep/sqrt(2)) as u32 in binary32),
 * max_rows.wgsl:17 (window start trunc(kd/2 - 0.5)), centers.wgsl:24,99. */
int orc_geometry(const orc_params* p, orc_geom* g) {
    if (!p || !g) return -1;
    if (p->diameter < 1 || p->smoothing_size < 1) return -1;
    g->k = (int32_t)p->smoothing_size;
    g->R = g->k / 2;
    g->band = (int32_t)(p->diameter / 2);
    g->rint = ((int32_t)p->diameter - 1) / 2;
    float t = 2.0f * (float)p->separation;
    t = t / sqrtf(2.0f);
    g->kd = (int32_t)(uint32_t)t;
    g->a = (int32_t)((float)g->kd / 2.0f - 0.5f);
    g->r2 = (float)g->rint * (float)g->rint;
    return 0;
}

int32_t orc_ncols(const orc_params* p) {
    return 4 + (p->cpu_processed ? 3 : 0) + (p->characterize ? 4 : 0);
}

/* Threads over the rows of one frame in the bandpass and local-maximum
 * passes (orc_set_threads); 1: serial. */
static int orc_threads = 1;

int orc_set_threads(int n) {
    int old = orc_threads;
    orc_threads = n > 0 ? n : 1;
    return old;
}

/* Planted mistake of the characterisation (orc_set_mutation): ORC_MUT_NONE
 * everywhere but in the tests that show their checks catch each one. */
static int orc_mut = ORC_MUT_NONE;

int orc_set_mutation(int m) {
    int old = orc_mut;
    orc_mut = (m >= 0 && m < ORC_MUT_COUNT) ? m : ORC_MUT_NONE;
    return old;
}

/* Disk radius of the characterize columns (centers.wgsl:24: (d - 1) / 2) and
 * of the cpu_processed columns (execute_gpu.rs:381: circle_inds(d / 2)). */
static int32_t char_radius_gpu(const orc_geom* g) {
    return orc_mut == ORC_MUT_GPU_RADIUS_BAND ? g->band : g->rint;
}
static int32_t char_radius_cpu(const orc_geom* g) {
    return orc_mut == ORC_MUT_CPU_RADIUS_RINT ? g->rint : g->band;
}

int32_t orc_char_tables(int32_t r, float* r2, float* sin2, float* cos2, int32_t cap,
                        int32_t* mid) {
    if (r < 0) return -1;
    int32_t t = 0, m = -1;
    for (int32_t i = -r; i <= r; i++)
        for (int32_t j = -r; j <= r; j++) {
            if (i * i + j * j > r * r) continue;
            if (t >= cap) return -1;
            /* kernels.rs:190,201: (2.*(y as f32).atan2(x as f32)).sin() with
             * x the row offset i and y the column offset j, in binary32 */
            const float th = atan2f((float)j, (float)i);
            const float a = (orc_mut == ORC_MUT_THETA) ? th : 2.0f * th;
            const float q = (float)(i * i + j * j);
            if (r2) r2[t] = (orc_mut == ORC_MUT_R_NOT_R2) ? sqrtf(q) : q;
            if (sin2) sin2[t] = sinf(a);
            if (cos2) cos2[t] = cosf(a);
            if (i == 0 && j == 0) m = t;
            t++;
        }
    if (mid) *mid = m;
    return t;
}

#define REAL float
#define SFX(name) name##_f32
#define FMA(a, b, c) fmaf((a), (b), (c))
#define SQRT(x) sqrtf(x)
#include "gt_oracle_impl.inc"
#undef REAL
#undef SFX
#undef FMA
#undef SQRT

#define REAL double
#define SFX(name) name##_f64
#define FMA(a, b, c) fma((a), (b), (c))
#define SQRT(x) sqrt(x)
#include "gt_oracle_impl.inc"
#undef REAL
#undef SFX
#undef FMA
#undef SQRT

int orc_gauss_weights(int32_t k, float noise_size, float* w, float* norm) {
    if (k < 1) return -1;
    weights_f32(k / 2, noise_size, w, norm);
    return 0;
}

int orc_bandpass_f32(const float* in, int32_t H, int32_t W, int32_t k,
                     float noise_size, float* out) {
    return bandpass_f32(in, H, W, k, noise_size, out);
}

int orc_bandpass_f64(const double* in, int32_t H, int32_t W, int32_t k,
                     double noise_size, double* out) {
    return bandpass_f64(in, H, W, k, noise_size, out);
}

int orc_masses_centers_f32(const float* P, int32_t H, int32_t W,
                           uint32_t diameter, float* masses, float* centers) {
    orc_params p;
    memset(&p, 0, sizeof(p));
    p.diameter = diameter;
    p.smoothing_size = 1;
    orc_geom g;
    if (orc_geometry(&p, &g) != 0) return -1;
    for (int32_t u = 0; u < H; u++)
        for (int32_t v = 0; v < W; v++) {
            float m, c0, c1;
            disk_f32(P, H, W, &g, u, v, &m, &c0, &c1);
            size_t i = (size_t)u * W + v;
            masses[i] = m;
            centers[2 * i] = c0;
            centers[2 * i + 1] = c1;
        }
    return 0;
}

int orc_is_max_map(const float* P, int32_t H, int32_t W, uint32_t separation,
                   uint8_t* out) {
    orc_params p;
    memset(&p, 0, sizeof(p));
    p.diameter = 1;
    p.smoothing_size = 1;
    p.separation = separation;
    orc_geom g;
    if (orc_geometry(&p, &g) != 0) return -1;
    is_max_map_f32(P, H, W, &g, out);
    return 0;
}

double orc_percentile_threshold(const float* P, size_t n, float percentile) {
    if (n == 0) return NAN;
    return threshold_f32(P, n, percentile);
}

int orc_locate_frame(const float* frame, int32_t H, int32_t W,
                     const orc_params* p, int32_t frame_index, int32_t f64,
                     double* rows, int64_t* cand, int64_t max_rows,
                     int64_t* n_rows, double* threshold_out) {
    if (!f64)
        return locate_f32(frame, H, W, p, frame_index, rows, cand, max_rows,
                          n_rows, threshold_out);
    size_t n = (size_t)H * W;
    double* fd = (double*)malloc(sizeof(double) * n);
    if (!fd) return -2;
    for (size_t i = 0; i < n; i++) fd[i] = (double)frame[i];
    int rc = locate_f64(fd, H, W, p, frame_index, rows, cand, max_rows, n_rows,
                        threshold_out);
    free(fd);
    return rc;
}

int orc_characterise_at(const float* P, const float* raw, int32_t H, int32_t W,
                        const orc_params* p, int32_t u, int32_t v, float walked_mass,
                        float* cpu3, float* gpu4) {
    orc_geom g;
    if (!P || !raw || orc_geometry(p, &g) != 0) return -1;
    if (u < 0 || u >= H || v < 0 || v >= W) return -1;
    const int32_t rc = char_radius_cpu(&g), rg = char_radius_gpu(&g);
    const int32_t cap_c = (2 * rc + 1) * (2 * rc + 1), cap_g = (2 * rg + 1) * (2 * rg + 1);
    float* r2t = (float*)malloc(sizeof(float) * (size_t)cap_c * 3);
    float* tg = (float*)malloc(sizeof(float) * (size_t)cap_g * 2);
    if (!r2t || !tg) { free(r2t); free(tg); return -2; }
    float* sint = r2t + cap_c;
    float* cost = sint + cap_c;
    int32_t mid = 0;
    const int32_t ntc = orc_char_tables(rc, r2t, sint, cost, cap_c, &mid);
    const int32_t ntg = orc_char_tables(rg, NULL, tg, tg + cap_g, cap_g, NULL);
    if (ntc < 0 || ntg < 0) { free(r2t); free(tg); return -1; }
    if (cpu3) char_cpu_f32(P, H, W, &g, r2t, sint, cost, ntc, walked_mass, u, v, cpu3);
    if (gpu4) char_gpu_f32(P, raw, H, W, &g, tg, tg + cap_g, u, v, gpu4);
    free(r2t);
    free(tg);
    return 0;
}

int orc_candidates_frame(const float* frame, int32_t H, int32_t W,
                         const orc_params* p, int64_t* cand, float* y, float* x,
                         float* mass, int64_t max_n, int64_t* n) {
    orc_geom g;
    if (orc_geometry(p, &g) != 0) return -1;
    int64_t cnt = 0;
    rec_f32* rs = candidates_f32(frame, H, W, p, &g, &cnt, NULL);
    if (!rs) return -2;
    for (int64_t i = 0; i < cnt && i < max_n; i++) {
        cand[i] = rs[i].cand;
        y[i] = rs[i].y;
        x[i] = rs[i].x;
        mass[i] = rs[i].mass;
    }
    free(rs);
    *n = cnt;
    return cnt > max_n ? -3 : 0;
}

int orc_dedup_bruteforce(const float* y, const float* x, const float* mass,
                         const int64_t* idx, int64_t n, uint32_t separation,
                         uint8_t* keep) {
    const float rad = (float)separation;
    const float r2 = rad * rad;
    for (int64_t i = 0; i < n; i++) {
        keep[i] = 1;
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            float dy = y[j] - y[i], dx = x[j] - x[i];
            float d = dy * dy;
            d = d + dx * dx;
            if (!(d < r2)) continue;
            if (mass[i] < mass[j] || (mass[i] == mass[j] && idx[i] > idx[j])) {
                keep[i] = 0;
                break;
            }
        }
    }
    return 0;
}

/* ---- adjoint (no reference counterpart: checks the GPU adjoint) ---------- */

/* Boxcar multiplicity of input index r in the window of q under the single
 * mirror reflection (preprocess_rows.wgsl / preprocess_cols.wgsl boxcar). */
static double box_count(int32_t q, int32_t r, int32_t R, int32_t n) {
    double c = 0.0;
    for (int32_t d = -R; d <= R; d++) {
        int32_t t = q + d;
        if (t < 0) t = -t - 1;
        else if (t >= n) t = 2 * n - t - 1;
        if (t == r) c += 1.0;
    }
    return c;
}

int orc_adjoint_particle(const float* frame, int32_t H, int32_t W, const orc_params* p,
                         int32_t f64_forward, int64_t cand, double* jac, int32_t* S_out,
                         int32_t* origin, double* yxm, int32_t* state) {
    orc_geom g;
    if (orc_geometry(p, &g) != 0) return -1;
    const size_t npx = (size_t)H * W;
    const int32_t R = g.R, r = g.rint, S = 2 * (r + R) + 1;
    double* Pd = (double*)malloc(sizeof(double) * npx);
    double* w = (double*)malloc(sizeof(double) * (size_t)(2 * R + 1));
    if (!Pd || !w) { free(Pd); free(w); return -2; }
    double gn, m, c0, c1;
    int32_t st[4];
    const int32_t u0 = (int32_t)(cand / W), v0 = (int32_t)(cand % W);
    if (f64_forward) {
        double* fd = (double*)malloc(sizeof(double) * npx);
        if (!fd) { free(Pd); free(w); return -2; }
        for (size_t i = 0; i < npx; i++) fd[i] = (double)frame[i];
        if (bandpass_f64(fd, H, W, g.k, (double)p->noise_size, Pd) != 0) {
            free(fd); free(Pd); free(w); return -1;
        }
        free(fd);
        weights_f64(R, (double)p->noise_size, w, &gn);
        walk_state_f64(Pd, H, W, &g, p->max_iterations, u0, v0, st, &m, &c0, &c1);
    } else {
        float* Pf = (float*)malloc(sizeof(float) * npx);
        float* wf = (float*)malloc(sizeof(float) * (size_t)(2 * R + 1));
        if (!Pf || !wf) { free(Pf); free(wf); free(Pd); free(w); return -2; }
        if (bandpass_f32(frame, H, W, g.k, p->noise_size, Pf) != 0) {
            free(Pf); free(wf); free(Pd); free(w); return -1;
        }
        float gnf, mf, c0f, c1f;
        weights_f32(R, p->noise_size, wf, &gnf);
        walk_state_f32(Pf, H, W, &g, p->max_iterations, u0, v0, st, &mf, &c0f, &c1f);
        for (size_t i = 0; i < npx; i++) Pd[i] = (double)Pf[i];
        for (int32_t i = 0; i < 2 * R + 1; i++) w[i] = (double)wf[i];
        gn = (double)gnf; m = (double)mf; c0 = (double)c0f; c1 = (double)c1f;
        free(Pf); free(wf);
    }
    const int32_t ue = st[0], ve = st[1];
    if (state) for (int q = 0; q < 4; q++) state[q] = st[q];
    if (yxm) { yxm[0] = (double)st[2] + c0; yxm[1] = (double)st[3] + c1; yxm[2] = m; }
    const int32_t py0 = ue - r - R, px0 = ve - r - R;
    if (origin) { origin[0] = py0; origin[1] = px0; }
    if (S_out) *S_out = S;
    const double gn2 = gn * gn, kk = (double)(g.k * g.k);
    if (jac) {
        memset(jac, 0, sizeof(double) * 3 * (size_t)S * S);
        for (int32_t i = -r; i <= r; i++)
            for (int32_t j = -r; j <= r; j++) {
                if (i * i + j * j > r * r) continue;
                const int32_t qy = ue + i, qx = ve + j;
                if (!(Pd[(size_t)qy * W + qx] > 0.0)) continue;   /* clamp active */
                const double a[3] = {((double)i - c0) / m, ((double)j - c1) / m, 1.0};
                for (int32_t sy = 0; sy < S; sy++) {
                    const int32_t ry = py0 + sy;
                    if (ry < 0 || ry >= H) continue;
                    const double wy = abs(ry - qy) <= R ? w[ry - qy + R] : 0.0;
                    const double by = box_count(qy, ry, R, H);
                    for (int32_t sx = 0; sx < S; sx++) {
                        const int32_t rx = px0 + sx;
                        if (rx < 0 || rx >= W) continue;
                        const double wx = abs(rx - qx) <= R ? w[rx - qx + R] : 0.0;
                        const double k = wy * wx / gn2 - by * box_count(qx, rx, R, W) / kk;
                        for (int o = 0; o < 3; o++)
                            jac[((size_t)o * S + sy) * S + sx] += a[o] * k;
                    }
                }
            }
    }
    free(Pd); free(w);
    return 0;
}

int orc_particle_f64(const double* frame, int32_t H, int32_t W, const orc_params* p,
                     int64_t cand, double* yxm, int32_t* state, int32_t* n_active) {
    orc_geom g;
    if (orc_geometry(p, &g) != 0) return -1;
    const size_t npx = (size_t)H * W;
    double* Pd = (double*)malloc(sizeof(double) * npx);
    if (!Pd) return -2;
    if (bandpass_f64(frame, H, W, g.k, (double)p->noise_size, Pd) != 0) { free(Pd); return -1; }
    double m, c0, c1;
    int32_t st[4];
    walk_state_f64(Pd, H, W, &g, p->max_iterations, (int32_t)(cand / W), (int32_t)(cand % W), st,
                   &m, &c0, &c1);
    yxm[0] = (double)st[2] + c0;
    yxm[1] = (double)st[3] + c1;
    yxm[2] = m;
    if (state) for (int q = 0; q < 4; q++) state[q] = st[q];
    if (n_active) {
        int32_t na = 0;
        for (int32_t i = -g.rint; i <= g.rint; i++)
            for (int32_t j = -g.rint; j <= g.rint; j++) {
                if (i * i + j * j > g.rint * g.rint) continue;
                if (Pd[(size_t)(st[0] + i) * W + (st[1] + j)] > 0.0) na++;
            }
        *n_active = na;
    }
    free(Pd);
    return 0;
}
