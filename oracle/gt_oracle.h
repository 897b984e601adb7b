/*
 * gt_oracle.h -- CPU oracle for the per-frame locate path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against.  Only tests/, the smoke entry point and
 * the CPU-baseline leg of bench.py may load it; the product path never does.
 *
 * It restates, in plain C, the reference's locate pipeline
 * (gpu-tracking/src/shaders/ WGSL files, gpu-tracking/src/gpu_setup.rs:45-60,
 * gpu-tracking/src/execute_gpu.rs:108-369, gpu-tracking/src/kernels.rs:134-205)
 * plus the percentile rule of
 * emily_tracking/work_in_progress/torch_tracking/torch_tracking_v2.py:271-301.
 *
 * Arithmetic contract of the binary32 ("faithful") mode:
 *   - every `acc += a * b` inside a WGSL shader is one fused multiply-add
 *     (a shader compiler may contract it; the oracle pins the contracted form);
 *   - every other shader operation is one IEEE-754 binary32 operation, in the
 *     reference's loop order;
 *   - host-side Rust arithmetic (kd-tree distances, CPU characterisation) is
 *     evaluated literally, with no contraction;
 *   - Gaussian tap weights are float(exp(double(-(i*i) / (2*noise_size))))
 *     with the argument formed in binary32, computed on the host.
 * The binary64 mode runs the same algorithm in double precision and exists to
 * state the binary32 error of each output.
 */
#ifndef GT_ORACLE_H
#define GT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    uint32_t diameter;        /* odd or even; disk radius (d-1)/2, band d/2 */
    float minmass;            /* strict `mass > minmass` */
    uint32_t separation;      /* dilation size (2*sep/sqrt2) and dedup radius */
    float noise_size;         /* used as sigma^2 of the Gaussian (Q1) */
    uint32_t smoothing_size;  /* k taps of both 1-D filters */
    uint32_t max_iterations;  /* walk cap, default 10 */
    int32_t has_percentile;   /* 0: no percentile test (reference default) */
    float percentile;         /* in [0, 100] */
    int32_t characterize;     /* adds Rg_gpu, raw_mass_gpu, signal_gpu, ecc_gpu */
    int32_t cpu_processed;    /* adds Rg, ecc, signal */
    int32_t literal;          /* 1: no exact short-cuts (plateau/band skip) */
} orc_params;

/* Derived kernel geometry (gpu_setup.rs:45-60, max_rows.wgsl:17). */
typedef struct orc_geom {
    int32_t k;          /* smoothing taps */
    int32_t R;          /* k / 2 */
    int32_t band;       /* diameter / 2: centers.wgsl:99 border band */
    int32_t rint;       /* (diameter - 1) / 2: disk radius, centers.wgsl:24 */
    int32_t kd;         /* dilation size */
    int32_t a;          /* window start offset: window is [-a, kd-1-a] */
    float r2;           /* float(rint)^2 */
} orc_geom;

int orc_geometry(const orc_params* p, orc_geom* g);

/* Gaussian weights (k of them) and their binary32 sequential sum. */
int orc_gauss_weights(int32_t k, float noise_size, float* w, float* norm);

/* Bandpass: preprocess_rows.wgsl + preprocess_cols.wgsl. */
int orc_bandpass_f32(const float* in, int32_t H, int32_t W, int32_t k,
                     float noise_size, float* out);
int orc_bandpass_f64(const double* in, int32_t H, int32_t W, int32_t k,
                     double noise_size, double* out);

/* centers.wgsl main for every pixel: masses[H*W] and centers[H*W*2]
 * (row offset, column offset) interleaved. */
int orc_masses_centers_f32(const float* P, int32_t H, int32_t W,
                           uint32_t diameter, float* masses, float* centers);

/* max_rows.wgsl + walk_cols.wgsl is_max for every pixel (1 = candidate). */
int orc_is_max_map(const float* P, int32_t H, int32_t W, uint32_t separation,
                   uint8_t* out);

/* Exact percentile threshold of all n values (linear interpolation, f64). */
double orc_percentile_threshold(const float* P, size_t n, float percentile);

/* Full locate of one frame (binary32 faithful mode when f64 == 0).
 * rows: n_rows x ncols row-major doubles, columns in column_names order
 * (execute_gpu.rs:197-226) restricted to frame, mass, y, x, [Rg, ecc,
 * signal], [Rg_gpu, raw_mass_gpu, signal_gpu, ecc_gpu].
 * cand: candidate (start) pixel index of each row.  Rows are in raster
 * order of the candidate pixel.  Pass rows == NULL to only count. */
int orc_locate_frame(const float* frame, int32_t H, int32_t W,
                     const orc_params* p, int32_t frame_index, int32_t f64,
                     double* rows, int64_t* cand, int64_t max_rows,
                     int64_t* n_rows, double* threshold_out);

int32_t orc_ncols(const orc_params* p);

/* Characterisation tap tables (kernels.rs:171-205 r2/sin/cos_in_circle; the
 * same sin/cos of 2*atan2(j, i) as centers.wgsl:67-69): one entry per tap of
 * the disk of integer radius r, row-major over (i, j), every value formed in
 * binary32 as the reference forms it ((2.*(j as f32).atan2(i as f32)).sin()).
 * mid: index of the (0, 0) tap (circle_inds' middle_most, kernels.rs:134-147).
 * Returns the tap count, or -1 when r < 0 or the table exceeds cap. */
int32_t orc_char_tables(int32_t r, float* r2, float* sin2, float* cos2, int32_t cap,
                        int32_t* mid);

/* The characterisation columns of one candidate pixel (u, v) of a bandpassed
 * frame P (raw: the input frame) in binary32:
 *   cpu3 = Rg, ecc, signal (execute_gpu.rs:165-172,286-299,346-352: the
 *          neighbourhood over circle_inds(diameter/2) at the candidate,
 *          divided by the walked mass `walked_mass`);
 *   gpu4 = Rg_gpu, raw_mass_gpu, signal_gpu, ecc_gpu (centers.wgsl:64-80:
 *          the disk of radius (diameter-1)/2 at the candidate, divided by
 *          that disk's own mass; zero in the border band). */
int orc_characterise_at(const float* P, const float* raw, int32_t H, int32_t W,
                        const orc_params* p, int32_t u, int32_t v, float walked_mass,
                        float* cpu3, float* gpu4);

/* Threads over the rows of one frame (bandpass and local-maximum passes);
 * default 1.  Results do not depend on it.  Returns the previous value. */
int orc_set_threads(int n);

/* Planted mistakes, for the tests that show the characterisation checks
 * catch them (never set outside those tests).  Returns the previous value. */
enum {
    ORC_MUT_NONE = 0,
    ORC_MUT_GPU_RADIUS_BAND = 1,  /* gpu columns over diameter/2, not (d-1)/2 */
    ORC_MUT_CPU_RADIUS_RINT = 2,  /* cpu columns over (d-1)/2, not diameter/2 */
    ORC_MUT_CPU_CAND_MASS = 3,    /* cpu Rg/ecc divided by the candidate's mass */
    ORC_MUT_THETA = 4,            /* sin/cos of theta instead of 2*theta */
    ORC_MUT_ECC_FULL_MASS = 5,    /* ecc divided by m, not m - P[centre] */
    ORC_MUT_SIGNAL_CENTRE = 6,    /* signal = P[centre] instead of the max */
    ORC_MUT_R_NOT_R2 = 7,         /* Rg weights |r| instead of r^2 */
    ORC_MUT_MIDDLE = 8,           /* middle_most one tap off */
    ORC_MUT_RAW_FROM_P = 9,       /* raw_mass_gpu summed over P, not the frame */
    ORC_MUT_COUNT = 10
};
int orc_set_mutation(int m);

/* Pre-dedup candidate records of one frame (after walk and minmass), raster
 * order: cand, y, x, mass.  Used to check deduplication in isolation. */
int orc_candidates_frame(const float* frame, int32_t H, int32_t W,
                         const orc_params* p, int64_t* cand, float* y, float* x,
                         float* mass, int64_t max_n, int64_t* n);

/* Deduplication (execute_gpu.rs:258-284), brute force O(n^2).
 * keep[i] = 1 if point i survives. */
int orc_dedup_bruteforce(const float* y, const float* x, const float* mass,
                         const int64_t* idx, int64_t n, uint32_t separation,
                         uint8_t* keep);

/* Adjoint of one particle started at candidate pixel `cand` (no reference
 * counterpart; checks the GPU adjoint).  Forward state (bandpass, walk) in
 * binary32 (f64_forward == 0) or binary64; the Jacobian itself in binary64:
 * jac[3][S][S] = d(y, x, mass)/d(input pixel) on the patch whose corner is
 * origin[0..1], S = 2*(rint + smoothing_size/2) + 1, walk decisions fixed.
 * yxm: forward (y, x, mass); state: last evaluated disk centre (row, col)
 * and final position (row, col). */
int orc_adjoint_particle(const float* frame, int32_t H, int32_t W, const orc_params* p,
                         int32_t f64_forward, int64_t cand, double* jac, int32_t* S_out,
                         int32_t* origin, double* yxm, int32_t* state);

/* Binary64 forward (bandpass + walk) of one particle from a binary64 frame,
 * for finite differences; n_active counts the disk pixels with P > 0. */
int orc_particle_f64(const double* frame, int32_t H, int32_t W, const orc_params* p,
                     int64_t cand, double* yxm, int32_t* state, int32_t* n_active);

/* link_rust (gpu-tracking/src/lib.rs:202-212, linking.rs) restated in C
 * (gt_linker.c): ids[i] = particle id of row i of a row-major n x ncols
 * float32 table (column 0 frame, 2 y, 3 x).  0, or -1 when out of memory. */
int orc_link_rows(const float* rows, int64_t n, int32_t ncols, float r, int64_t memory,
                  uint64_t* ids);

#ifdef __cplusplus
}
#endif
#endif
