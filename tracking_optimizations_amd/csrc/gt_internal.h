// gt_internal.h -- device-side descriptors and launch entry points shared by
// the HIP kernels (gt_kernels.hip) and the host orchestration (gt_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "gt_link_solve.h"

namespace gt {

// Output tile of the detect stage (K2).  It is also the grouping unit of
// candidates: each tile owns TILE_H*TILE_W candidate slots, refinement,
// deduplication and row emission all work tile by tile.  48 columns: with
// the dilation window's halo (up to 16 columns) a tile's staged columns are
// the 64 lanes of a wave.
constexpr int TILE_H = 64;
constexpr int TILE_W = 48;
constexpr int TILE_PX = TILE_H * TILE_W;

// Bandpass (K1) column strip: 64 lanes x 4 columns.
constexpr int K1_STRIP = 240;   // output columns per bandpass work item

// Raw signal extraction: taps of the signal disk and of the background
// annulus (each at most RAW_MAX_TAPS, offsets within +-127).
constexpr int RAW_MAX_TAPS = 4096;

// Percentile histogram: 4096 bins over the float32 bit pattern
// (bin 0 = exact zero, bin b >= 1 = bits >> 19 == b - 1).
constexpr int HIST_BINS = 4096;

enum InType : int { IN_U16 = 0, IN_F32 = 1, IN_U8 = 2 };

// Record planes stored per candidate slot.
enum RecPlane : int {
    REC_MASS = 0, REC_Y = 1, REC_X = 2,
    REC_RG = 3, REC_ECC = 4, REC_SIGNAL = 5,                 // cpu_processed
    REC_RG_GPU = 6, REC_RAW_GPU = 7, REC_SIG_GPU = 8, REC_ECC_GPU = 9,  // characterize
    REC_NPLANES = 10,
    // start value P[u0, v0] of each candidate, stored only when the walk
    // runs in detect (characterisation off, so this plane is free)
    REC_CENTER = 3
};

struct Frame {
    const void* in;        // H*W input pixels (device)
    float* P;              // H*W bandpassed frame
    float* tmp;            // 2*H*W row-pass temporary (generic bandpass), then the
                           // row-max plane of wide dilation windows
    uint64_t slot_base;    // first candidate slot of this frame
    uint32_t H, W;
    uint32_t k1_tx, k1_ty; // bandpass strips x segments
    uint32_t tx, ty;       // detect tiles
    uint32_t tile_base;    // first detect tile of this frame in the batch
    uint32_t k1_base;      // first bandpass item of this frame
    int64_t frame_index;
};

// Parameters derived once per call (gpu_setup.rs:45-60 plus host tables).
struct Geom {
    int32_t k;             // smoothing taps parameter
    int32_t R;             // k / 2
    int32_t band;          // diameter / 2
    int32_t rint;          // (diameter - 1) / 2
    int32_t kd;            // dilation size
    int32_t a;             // window start offset
    uint32_t max_iterations;
    float minmass;
    float sep;             // dedup radius
    int32_t has_pct;
    float percentile;
    int32_t characterize;
    int32_t cpu_processed;
    int32_t disk_in_window;
    int32_t fastdiv_gn;    // division by gauss_norm via corrected reciprocal
    int32_t fastdiv_kk;    // division by k*k via corrected reciprocal
    float w[33];           // Gaussian weights, taps -R..R (R <= 16)
    float gn, rgn;         // gauss_norm and its rounded reciprocal
    float kk, rkk;         // float(k*k) and its rounded reciprocal
    int32_t dedup_rty, dedup_rtx;  // tile radius of the dedup neighbourhood
    int32_t ncols;         // output columns
    int32_t ntap_gpu;      // disk taps (characterize tables)
    int32_t ntap_cpu;      // circle taps (cpu_processed tables)
    int32_t walk_in_detect;  // centroid walk fused into detect (no refine launch)
    int32_t raw_cols;      // 0, 1 (raw_mass) or 3 (+ raw_bg_median, raw_mass_corrected)
    int32_t col_raw;       // output column of raw_mass
    int32_t raw_ns, raw_nb;  // signal and background tap counts
    int8_t raw_tap[4 * RAW_MAX_TAPS];  // (dy, dx) of the signal taps, then the background taps
    float sin2_gpu[1024], cos2_gpu[1024];                // disk radius rint
    float r2_cpu[1024], sin_cpu[1024], cos_cpu[1024];    // circle radius band
};

struct Batch {
    const Frame* frames;   // device array [nframes]
    const uint32_t* k1_bases;   // [nframes + 1]
    const uint32_t* tile_bases; // [nframes + 1]
    const uint32_t* k1_frame;   // [k1_items] frame of each bandpass item
    const uint32_t* tile_frame; // [tiles] frame of each detect tile
    const uint32_t* dblock_tile;  // [dblocks] top-left tile of each 2x2 block of tiles (dedup)
    uint32_t nframes;
    uint32_t k1_items;
    uint32_t tiles;
    uint32_t dblocks;
    int in_type;           // InType
    int vec;               // W % 4 == 0 and aligned frame bases: vector loads
    // per-tile arrays [tiles]
    uint32_t* tile_count;
    uint32_t* tile_keep;
    uint32_t* tile_row_off;
    uint32_t* rowk;        // [tiles * 64] kept rows of the tile in pixel rows <= r
    uint32_t* rowoff;      // [tiles * 64] frame-relative first row of pixel row r
    uint32_t* tile_lcount; // percentile list entries of the tile
    // per-frame arrays [nframes]
    uint32_t* hist1;       // [nframes * HIST_BINS]
    uint32_t* list;        // percentile candidates (bit patterns), TILE_PX per tile
    int32_t* selstate;     // [nframes * 4]: bin_lo, rank_lo, bin_hi, rank_hi
    double* thr;           // [nframes]
    uint32_t* frame_rows;  // [nframes]
    uint32_t* frame_cands; // [nframes] candidates entering refinement
    uint64_t* frame_row_base;  // [nframes]
    // candidate slots and records
    uint16_t* slots;       // [tiles * TILE_PX]
    float* rec;            // [REC_NPLANES * slot_cap]
    uint8_t* keep;         // [slot_cap]
    uint64_t slot_cap;
    // output
    uint64_t* cand_total;  // device running candidate count (statistics)
    uint64_t* row_carry;   // device running row count across batches
    float* rows;
    int64_t* rows_cand;
    uint64_t rows_cap;
};

// Adjoint outputs, indexed by output row (gt_adjoint.hip).
struct AdjOut {
    float* sigma;          // [rows_cap * 3] sigma_y, sigma_x, sigma_mass
    int32_t* origin;       // [rows_cap * 2] patch origin (row, column)
    int32_t* rowpos;       // [rows_cap] position of the row's frame in the call
    float* jac;            // [rows_cap * 3 * S * S] Jacobian patches, or null
};

// Launchers (gt_kernels.hip).  All launch on `s`; return hipSuccess or the
// first launch error.
hipError_t launch_zero_batch(const Batch& b, hipStream_t s);
// The streaming bandpass handles uint8, uint16 and float32 input with 1 <= R <= 8, W % 4 == 0
// and aligned frame bases (b.vec); everything else takes the generic kernels.
bool bandpass_fast_path(const Batch& b, const Geom& g);
hipError_t launch_bandpass(const Batch& b, const Geom& g, const Geom* d_geom,
                           int seg_rows, hipStream_t s);
hipError_t launch_bandpass_generic(const Batch& b, const Geom* d_geom, uint32_t max_px,
                                   int hist, hipStream_t s);
hipError_t launch_select1(const Batch& b, const Geom& g, hipStream_t s);
bool walk_in_detect_supported(const Geom& g);
// Dilation windows too wide for a staged tile: detect writes the row maxima
// to each frame's tmp plane (Frame::tmp, H*W floats at least) first.
bool detect_needs_plane(const Geom& g);
hipError_t launch_detect(const Batch& b, const Geom& g, const Geom* d_geom,
                         hipStream_t s);
hipError_t launch_select2(const Batch& b, const Geom& g, hipStream_t s);
// Expands the batch's bases (first bandpass item, tile and dedup block of
// every frame) into b.k1_frame, b.tile_frame and b.dblock_tile.
hipError_t launch_fill_tables(const Batch& b, const uint32_t* dblock_bases,
                              uint32_t max_frame_work, hipStream_t s);
hipError_t launch_refine(const Batch& b, const Geom& g, const Geom* d_geom,
                         hipStream_t s);
hipError_t launch_dedup(const Batch& b, const Geom& g, const Geom* d_geom,
                        hipStream_t s);
hipError_t launch_tile_scan(const Batch& b, uint32_t max_ty, hipStream_t s);
hipError_t launch_frame_scan(const Batch& b, hipStream_t s);
hipError_t launch_emit(const Batch& b, const Geom& g, const Geom* d_geom,
                       hipStream_t s);
hipError_t launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s);

// Raw signal and background of every output row (gt_raw.hip).
hipError_t launch_raw(const Batch& b, const Geom& g, const Geom* d_geom, hipStream_t s);

// Adjoint (gt_adjoint.hip).
int adjoint_patch(const Geom& g);
bool adjoint_supported(const Geom& g);
hipError_t launch_adjoint(const Batch& b, const Geom& g, const Geom* d_geom, const AdjOut& ao,
                          uint32_t frame_pos0, hipStream_t s);
hipError_t launch_adjoint_vjp(const float* jac, const int32_t* origin, const uint64_t* frame_rows,
                              uint32_t nframes, const float* grad_rows, float* grad, int S, int H,
                              int W, hipStream_t s);

// Linker (gt_link.hip).
struct LinkStats {
    uint64_t steps = 0, pairs = 0, subnets = 0, max_subnet = 0, ids = 0;
    double search_ms = 0, solve_ms = 0;   // candidate search, subnet solver
};
// Particle ids of the n = step_off.back() positions (ys, xs: host arrays),
// step s being positions [step_off[s], step_off[s + 1]).  On failure `err`
// says why.
// Device buffers and pinned tables kept between calls (one per context).
struct LinkCache;
LinkCache* link_cache_create();
void link_cache_destroy(LinkCache* c);
// A stream of the cache's own (default priority), created on first use with
// the calling thread's device current.
hipError_t link_cache_stream(LinkCache* c, hipStream_t* out);
// Linker state between chunks of steps (LinkState, gt_link_solve.h).
LinkState* link_state_create(uint32_t memory);
void link_state_destroy(LinkState* s);
// Links the steps of one chunk (soff: chunk-local step offsets) after the
// steps the state has seen; ids of the chunk's rows go to ids.
hipError_t link_chunk(hipStream_t st, LinkCache* cache, LinkState* state, const float* ys,
                      const float* xs, const std::vector<uint32_t>& soff, float search_range,
                      uint64_t* ids, LinkStats* stats, std::string& err);
hipError_t link_steps(hipStream_t st, LinkCache* cache, const float* ys, const float* xs,
                      const std::vector<uint32_t>& step_off, float search_range,
                      uint32_t memory, uint64_t* ids, LinkStats* stats, std::string& err);
// link_rust: steps are FrameSubsetter runs of column 0, y = column 2, x = 3.
hipError_t link_rows(hipStream_t st, LinkCache* cache, const float* rows, uint64_t n_rows,
                     uint32_t ncols, float search_range, uint32_t memory, uint64_t* ids,
                     LinkStats* stats, std::string& err);

// Synthetic frames (gt_synth.hip).
hipError_t launch_synth(uint16_t* frames, uint64_t nframes, uint32_t h,
                        uint32_t w, const float4* spots, const uint32_t* cell_off,
                        uint32_t cells_x, uint32_t cells_y, uint32_t cell,
                        const float* bg, uint64_t seed, uint64_t frame0,
                        hipStream_t s);

}  // namespace gt
