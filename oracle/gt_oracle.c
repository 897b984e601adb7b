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
