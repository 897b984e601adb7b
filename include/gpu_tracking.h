/*
 * gpu_tracking.h -- C ABI of the MI355X (gfx950) locate library
 * (libgpu_tracking.so).
 *
 * The ABI replaces the per-frame locate path of the reference crate
 * `gpu_tracking` (gpu-tracking/src/execute_gpu.rs, gpu-tracking/src/lib.rs).
 * Every entry point names the reference interface it stands in for.
 *
 * Conventions
 *   - Plain C types only.  Nothing unwinds across the ABI: every call returns a
 *     gt_status; gt_last_error() gives the calling thread's last message.
 *   - A gt_context owns one HIP device, its stream, its cached device
 *     workspaces and compiled-shape caches.  A context is not thread-safe:
 *     use one context per host thread.
 *   - Frames are row-major H x W.  Host pointers are read during the call;
 *     device pointers (GT_FRAMES_ON_DEVICE) must stay valid until it returns.
 *   - Results are float32 rows in the reference's column order
 *     (execute_gpu.rs:197-226), one row per located feature.
 */
#ifndef GPU_TRACKING_H
#define GPU_TRACKING_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GT_ABI_VERSION 1

typedef enum gt_status {
    GT_OK = 0,
    GT_E_INVALID = 1,          /* bad argument or parameter combination */
    GT_E_NOT_IMPLEMENTED = 2,  /* parameter the reference also rejects */
    GT_E_NO_DEVICE = 3,        /* no HIP device / runtime unavailable */
    GT_E_HIP = 4,              /* a HIP runtime call failed */
    GT_E_OOM = 5,              /* host or device allocation failed */
    GT_E_INTERNAL = 6
} gt_status;

/* Mirror of `TrackingParams` (execute_gpu.rs:16-38), field for field and in
 * the same order.  An `Option<T>` field of the reference becomes
 * `has_<name>` + value.  The six fields that lib.rs:73-74 rejects when passed
 * from Python (maxsize, threshold, invert, percentile, topn, preprocess)
 * carry a `has_<name>` flag meaning "explicitly set"; setting any of them
 * gives GT_E_NOT_IMPLEMENTED, except `percentile`, which this library
 * implements (exact per-frame quantile of the bandpassed frame, candidate
 * kept when its value is strictly above it). */
typedef struct gt_params {
    uint32_t diameter;
    float minmass;
    uint8_t has_maxsize;
    float maxsize;
    uint32_t separation;
    float noise_size;
    uint32_t smoothing_size;
    uint8_t has_threshold;
    float threshold;
    uint8_t has_invert;
    uint8_t invert;
    uint8_t has_percentile;
    float percentile;
    uint8_t has_topn;
    uint32_t topn;
    uint8_t has_preprocess;
    uint8_t preprocess;
    uint32_t max_iterations;
    uint8_t characterize;
    uint8_t filter_close;
    uint8_t has_search_range;
    float search_range;
    uint8_t has_memory;
    uint64_t memory;
    uint8_t cpu_processed;
    uint8_t has_sig_radius;
    float sig_radius;
    uint8_t has_bg_radius;
    float bg_radius;
    uint8_t has_gap_radius;
    float gap_radius;
} gt_params;

/* One frame of an uneven batch (replaces one item of the frame iterator fed
 * to execute_gpu, execute_gpu.rs:372). */
typedef struct gt_frame_desc {
    const void* data;       /* H x W pixels, dtype given by the call */
    uint32_t h, w;
    int64_t frame_index;    /* value written to the `frame` column */
} gt_frame_desc;

typedef struct gt_context gt_context;
typedef struct gt_result gt_result;

/* Flags of the locate calls. */
#define GT_FRAMES_ON_DEVICE  0x1u  /* frame pointers are HIP device pointers */
#define GT_PLAIN             0x2u  /* per-frame pipeline: launch + readback per frame */
#define GT_RESULT_ON_DEVICE  0x4u  /* leave rows in device memory (see gt_result_device_data) */
#define GT_ADJ_JACOBIAN      0x8u  /* gt_locate_adjoint: keep the Jacobian patches */

/* Per-call device timing of the last locate call on a context, measured with
 * HIP events on the context's stream. */
typedef struct gt_timing {
    double device_ms;       /* first kernel start .. last kernel end */
    double total_ms;        /* host wall time of the call */
    uint64_t frames;
    uint64_t pixels;
    uint64_t candidates;    /* local maxima that entered refinement */
    uint64_t rows;          /* rows returned */
    uint64_t launches;      /* kernel launches issued */
} gt_timing;

/* --- parameters (no device needed) ------------------------------------- */

/* Python defaults of batch_rust (lib.rs:77-91): minmass 0, separation
 * diameter+1, noise_size 1, smoothing_size diameter, max_iterations 10,
 * characterize false, filter_close true, cpu_processed false; all optional
 * fields unset. */
gt_status gt_params_default(gt_params* p, uint32_t diameter);

/* Rust `Default` of TrackingParams (execute_gpu.rs:40-65). */
gt_status gt_params_rust_default(gt_params* p);

/* Checks a parameter set for frames of h x w (h = w = 0: shape-free checks).
 * Returns GT_E_NOT_IMPLEMENTED for what lib.rs:73-74 rejects. */
gt_status gt_params_validate(const gt_params* p, uint32_t h, uint32_t w);

/* column_names (execute_gpu.rs:197-226): names and dtypes ("int"/"float")
 * of the result columns for this parameter set.  Pointers are static. */
gt_status gt_column_names(const gt_params* p, const char** names,
                          const char** dtypes, uint32_t cap, uint32_t* ncols);

const char* gt_last_error(void);

/* Test hook; no counterpart in the reference, which panics when an
 * allocation fails (execute_gpu.rs:497-499, gpu_setup.rs:201,208): from now
 * on, the after-th device allocation (kind 0: the library's workspaces and
 * result buffers) or host result allocation (kind 1) made by the library
 * fails, once, as if out of memory, so that callers can check that the call
 * returns GT_E_OOM and the context keeps working.  after = 0 disarms.
 * Process-wide. */
gt_status gt_debug_inject_alloc_failure(int kind, uint32_t after);
const char* gt_version(void);

/* Number of visible HIP devices (0 when there is none; never fails). */
int gt_device_count(void);

/* --- context ----------------------------------------------------------- */

/* Replaces gpu_setup::setup_state (gpu_setup.rs:191), which the reference
 * runs per call; here the device state persists across calls. */
gt_status gt_context_create(int device, gt_context** out);
gt_status gt_context_destroy(gt_context* ctx);
gt_status gt_context_timing(const gt_context* ctx, gt_timing* out);
/* Frames per launch of the batched path (0 = automatic). */
gt_status gt_context_set_batch(gt_context* ctx, uint32_t frames_per_launch);
/* 1: the batched path's launch sequence is captured as a HIP graph and
 * replayed while calls repeat with the same inputs, parameters and buffers
 * (a device staging buffer refilled between calls, a benchmark loop).  The
 * graph holds no memset node.  0 (default): launches are issued directly. */
gt_status gt_context_set_graphs(gt_context* ctx, int enable);

/* Rows per bandpass work item (16..272, default 128); clears tuned shapes. */
gt_status gt_context_set_launch_shape(gt_context* ctx, uint32_t seg_rows);

/* Autotuner: times the batched path on n_frames synthetic uint16 frames of
 * h x w for every launch shape (rows per bandpass work item 64/128/256,
 * frames per launch automatic/16/32, the centroid walk inside detect or in
 * its own kernel where both are possible) and keeps the fastest in the
 * context for calls whose frames are all h x w with the same filter
 * geometry.  Results do not depend on the shape. */
typedef struct gt_tuning {
    uint32_t seg_rows;            /* rows per bandpass work item */
    uint32_t frames_per_launch;   /* 0 = as many as the workspace holds */
    double device_ms_per_frame;   /* of the chosen shape */
    uint32_t fused_walk;          /* 1: walk inside detect, 0: separate refine kernel */
} gt_tuning;
gt_status gt_context_autotune(gt_context* ctx, const gt_params* p, uint32_t h, uint32_t w,
                              uint32_t n_frames, gt_tuning* out);

/* 1 (default): where the geometry allows it, the centroid walk runs inside
 * the detect kernel on the tile it has staged; 0: a separate refine kernel
 * walks every candidate.  Results are identical either way. */
gt_status gt_context_set_fused_walk(gt_context* ctx, int enable);

/* Per-stage kernel timing (HIP events recorded around every launch on the
 * context's stream).  Stages, in pipeline order: */
#define GT_STAGE_ZERO 0      /* clear per-frame counters */
#define GT_STAGE_BANDPASS 1  /* preprocess_rows.wgsl + preprocess_cols.wgsl */
#define GT_STAGE_SELECT1 2   /* percentile: histogram bin of the quantile */
#define GT_STAGE_DETECT 3    /* max_rows.wgsl + walk_cols.wgsl is_max */
#define GT_STAGE_SELECT2 4   /* percentile: exact order statistics */
#define GT_STAGE_REFINE 5    /* walk + centers.wgsl at visited pixels */
#define GT_STAGE_DEDUP 6     /* execute_gpu.rs:258-284 */
#define GT_STAGE_SCAN 7      /* row offsets */
#define GT_STAGE_EMIT 8      /* rows in column_names order */
#define GT_NSTAGES 9
/* enable != 0 turns event timing on for the following locate calls. */
gt_status gt_context_set_profiling(gt_context* ctx, int enable);
/* Sum of kernel durations (ms) and launch counts per stage since profiling
 * was last enabled; arrays of GT_NSTAGES entries. */
gt_status gt_context_stage_times(const gt_context* ctx, double* ms, uint64_t* launches);

/* --- locate ------------------------------------------------------------- */

/* execute_ndarray (execute_gpu.rs:496) / batch_rust (lib.rs:46) for a
 * C-contiguous (n_frames, h, w) uint16 movie.  uint16 -> float32 is exact,
 * so results equal those for the same frames as float32. */
gt_status gt_locate_u16(gt_context* ctx, const uint16_t* frames,
                        uint64_t n_frames, uint32_t h, uint32_t w,
                        const gt_params* p, uint32_t flags, gt_result** out);

/* Same for uint8 frames (the reference converts u8 pages to f32,
 * decoderiter.rs:31-35). */
gt_status gt_locate_u8(gt_context* ctx, const uint8_t* frames,
                       uint64_t n_frames, uint32_t h, uint32_t w,
                       const gt_params* p, uint32_t flags, gt_result** out);

/* Same for float32 frames, the reference's own input type (lib.rs:50).
 * Pixel values must be finite. */
gt_status gt_locate_f32(gt_context* ctx, const float* frames,
                        uint64_t n_frames, uint32_t h, uint32_t w,
                        const gt_params* p, uint32_t flags, gt_result** out);

/* Frames of different shapes in one batched launch sequence; dtype 0 =
 * uint16, 1 = float32, 2 = uint8.  Rows are grouped by descriptor order. */
gt_status gt_locate_batch(gt_context* ctx, const gt_frame_desc* frames,
                          uint64_t n_frames, int dtype, const gt_params* p,
                          uint32_t flags, gt_result** out);

/* No counterpart in the reference.  gt_locate_u16 / gt_locate_u8 /
 * gt_locate_f32 with GT_FRAMES_ON_DEVICE | GT_RESULT_ON_DEVICE, for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the call allocates no device memory and
 * calls no allocating library routine.  Every workspace (frame tables,
 * bandpassed frames, candidate slots and records, counters) lives in one
 * fixed device array of GT_NOALLOC_SCRATCH_BYTES per device, held from
 * library load until the process ends; frames are located in batches that
 * fit it, and a single frame that needs more fails with GT_E_INVALID (about
 * 113 MB for a 2048 x 2048 frame at d = 11 with percentile; the
 * characterisation columns need more per frame).
 *   frames   device pointer, n_frames x h x w of dtype (0 u16, 1 f32, 2 u8);
 *   rows     device pointer, rows_capacity x ncols float32 (column_names
 *            order), written row-major;
 *   cand     device pointer, rows_capacity int64 candidate pixel indices;
 *   n_rows   rows found.  When it exceeds rows_capacity the call returns
 *            GT_E_INVALID and only the first rows_capacity rows are written.
 * Rows equal those of gt_locate_* for the same input.  search_range is not
 * supported here (the linker allocates).  Calls on one device share its
 * scratch: calls that overlap in time on the same device (two threads, two
 * contexts) race; calls in sequence do not. */
#define GT_NOALLOC_SCRATCH_BYTES (128u << 20)
gt_status gt_locate_noalloc(gt_context* ctx, const void* frames, int dtype, uint64_t n_frames,
                            uint32_t h, uint32_t w, const gt_params* p, float* rows,
                            int64_t* cand, uint64_t rows_capacity, uint64_t* n_rows);

/* --- results ------------------------------------------------------------ */

uint64_t gt_result_rows(const gt_result* r);
uint32_t gt_result_ncols(const gt_result* r);
/* Host rows, row-major n_rows x ncols float32 (NULL with GT_RESULT_ON_DEVICE). */
const float* gt_result_data(const gt_result* r);
/* Candidate (start) pixel index y*w + x of each row, host int64 (NULL with
 * GT_RESULT_ON_DEVICE).  Not a reference column: a key for comparisons. */
const int64_t* gt_result_candidates(const gt_result* r);
/* Device rows (GT_RESULT_ON_DEVICE only): valid until the next locate call
 * on the same context or gt_result_free. */
const float* gt_result_device_data(const gt_result* r);
gt_status gt_result_columns(const gt_result* r, const char** names,
                            const char** dtypes, uint32_t cap, uint32_t* ncols);
void gt_result_free(gt_result* r);

/* --- adjoint: localisation error and gradients ------------------------- */

/* Adjoint of the locate path (refine and bandpass) for the rows of one
 * batched locate of n_frames frames of h x w (dtype 0 u16, 1 f32, 2 u8).
 * No counterpart in the reference, which has no gradients; it extends
 * execute_ndarray (execute_gpu.rs:496) for callers that propagate Poisson
 * shot noise into per-particle localisation errors or differentiate through
 * locate.  For every output row, with the walk's discrete decisions held
 * fixed, it forms the Jacobian of (y, x, mass) with respect to the input
 * pixels on an S x S patch (S = 2 * ((diameter-1)/2 + smoothing_size/2) + 1)
 * centred on the particle's last evaluated disk, and the Poisson errors
 * sigma = sqrt(sum_r J[r]^2 * max(I[r], 0)).  Flags: GT_FRAMES_ON_DEVICE,
 * GT_ADJ_JACOBIAN (keep the patches, required by gt_adjoint_vjp).  The rows
 * are those gt_locate_* returns for the same input. */
typedef struct gt_adjoint gt_adjoint;
gt_status gt_locate_adjoint(gt_context* ctx, const void* frames, int dtype, uint64_t n_frames,
                            uint32_t h, uint32_t w, const gt_params* p, uint32_t flags,
                            gt_adjoint** out);
uint64_t gt_adjoint_rows(const gt_adjoint* a);
uint32_t gt_adjoint_ncols(const gt_adjoint* a);
uint32_t gt_adjoint_patch(const gt_adjoint* a);
/* Forward rows (as gt_result_data) and candidate pixels. */
const float* gt_adjoint_data(const gt_adjoint* a);
const int64_t* gt_adjoint_candidates(const gt_adjoint* a);
/* n_rows x 3: sigma_y, sigma_x, sigma_mass. */
const float* gt_adjoint_sigma(const gt_adjoint* a);
/* n_rows x 2: frame row and column of each patch's corner. */
const int32_t* gt_adjoint_origin(const gt_adjoint* a);
/* Copies the patches, n_rows x 3 (y, x, mass) x S x S floats. */
gt_status gt_adjoint_jacobian(const gt_adjoint* a, float* out, uint64_t capacity_floats);
/* Vector-Jacobian product: grad_frames (n_frames x h x w) = sum over rows of
 * grad_rows[row] . J[row] (grad_rows: n_rows x 3, d/dy, d/dx, d/dmass), added
 * in row order (deterministic).  GT_FRAMES_ON_DEVICE: both are device
 * pointers.  Needs GT_ADJ_JACOBIAN at gt_locate_adjoint. */
gt_status gt_adjoint_vjp(gt_context* ctx, const gt_adjoint* a, const float* grad_rows,
                         float* grad_frames, uint32_t flags);
void gt_adjoint_free(gt_adjoint* a);

/* --- movie files (gpu-tracking/src/decoderiter.rs, execute_gpu.rs:510-542) -- */

typedef struct gt_file gt_file;

/* Opens a TIFF (.tif/.tiff: the pages, decoderiter.rs:9-71) or ETS (.ets:
 * the frames of `channel`, MinimalETSParser, decoderiter.rs:77-260) movie
 * and reads its frame table; channel is ignored for TIFF, and UINT32_MAX
 * opens an ETS file for gt_file_channels only (no frames).  No device
 * needed.  TIFF: classic or BigTIFF, 8/16-bit unsigned grey strips,
 * uncompressed, LZW, Deflate or PackBits, optional horizontal predictor.
 * ETS: uint16 data. */
gt_status gt_file_open(const char* path, uint32_t channel, gt_file** out);
/* kind 0 TIFF, 1 ETS; frame shape h x w as stored (ETS: h = dims[1],
 * w = dims[0], the layout parse_ets returns); dtype 0 u16, 2 u8. */
gt_status gt_file_info(const gt_file* f, int* kind, uint64_t* n_frames, uint32_t* h,
                       uint32_t* w, int* dtype);
/* ETS channel numbers present in the file (TIFF: none). */
gt_status gt_file_channels(const gt_file* f, uint32_t* channels, uint32_t cap, uint32_t* n);
/* Decodes frames [first, first + count) into dst (host memory,
 * count x h x w elements of the file's dtype). */
gt_status gt_file_read(const gt_file* f, uint64_t first, uint64_t count, void* dst);
void gt_file_close(gt_file* f);

/* execute_file (execute_gpu.rs:510-542, behind batch_file_rust, lib.rs:
 * 123-200): decode + locate a movie file.  Frames are read in chunks of
 * about 256 MiB into pinned host buffers, the next chunk being read while
 * the current one is located; with search_range the rows are linked once
 * over all frames.  ETS frames are located with the ETS header's
 * (dims[0], dims[1]) as (rows, columns), as execute_file does.  flags: 0 or
 * GT_PLAIN. */
gt_status gt_locate_file(gt_context* ctx, const char* path, uint32_t channel,
                         const gt_params* p, uint32_t flags, gt_result** out);

/* --- linking (gpu-tracking/src/linking.rs) --------------------------------- */

/* Statistics of the last link on a context. */
typedef struct gt_link_stats {
    uint64_t steps;         /* Linker::advance calls */
    uint64_t pairs;         /* (source, destination) pairs within the radius, all ages */
    uint64_t subnets;       /* subnets solved by exhaustive assignment */
    uint64_t max_subnet;    /* destinations of the largest subnet */
    uint64_t ids;           /* particle ids handed out */
    double search_ms;       /* host wall time of the GPU candidate search */
    double solve_ms;        /* host wall time of the subnet solver */
} gt_link_stats;

/* link_rust (gpu-tracking/src/lib.rs:202-212 -> FrameSubsetter, linking.rs:
 * 12-61 -> linker_all, linking.rs:570-580): ids[i] = particle id of row i of
 * a row-major n_rows x ncols float32 table in host memory, column 0 = frame,
 * columns 2 and 3 = y and x.  A step is a run of equal frame values (a first
 * frame other than 0 gives an empty first step); search_range and memory as
 * Linker::new (linking.rs:211).  Candidates within the radius are found on
 * the device, subnets are solved on the host. */
gt_status gt_link(gt_context* ctx, const float* rows, uint64_t n_rows, uint32_t ncols,
                  float search_range, uint64_t memory, uint64_t* ids);

/* Statistics of the last gt_link or locate call with search_range. */
gt_status gt_context_link_stats(const gt_context* ctx, gt_link_stats* out);

/* --- synthetic movies (device side, for tests and benchmarks) ------------ */

/* Fills n_frames x h x w uint16 device frames with
 * min(Poisson(bg[f] + sum_i amp_i exp(-d^2 / (2 sigma_i^2))), 65535), spots
 * truncated at 5 sigma, Poisson sampled with Hoermann's PTRS (lambda >= 10)
 * or sequential inversion (lambda < 10) from Philox-4x32-10 streams keyed by
 * (seed, frame).  spots: host array of n_spots x 4 floats (y, x, amp, sigma)
 * grouped by frame with frame f's spots at [spot_off[f], spot_off[f+1]). */
gt_status gt_synth_u16(gt_context* ctx, uint16_t* d_frames, uint64_t n_frames,
                       uint32_t h, uint32_t w, const float* spots,
                       const uint64_t* spot_off, const float* bg, uint64_t seed);

/* Device memory helpers for callers without another allocator. */
gt_status gt_device_alloc(gt_context* ctx, uint64_t bytes, void** out);
gt_status gt_device_free(gt_context* ctx, void* p);
gt_status gt_memcpy_h2d(gt_context* ctx, void* dst, const void* src, uint64_t bytes);
gt_status gt_memcpy_d2h(gt_context* ctx, void* dst, const void* src, uint64_t bytes);
/* Device-to-device copy by one plain copy kernel (16-byte aligned pointers);
 * returns when the copy is done. */
gt_status gt_memcpy_d2d(gt_context* ctx, void* dst, const void* src, uint64_t bytes);
gt_status gt_synchronize(gt_context* ctx);

#ifdef __cplusplus
}
#endif
#endif
