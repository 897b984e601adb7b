// gt_api.hip -- host side of libgpu_tracking.so: parameter handling, the
// persistent device context, batch planning and the C ABI of
// include/gpu_tracking.h.
#include <algorithm>
#include <cstdlib>
#include <new>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpu_tracking.h"
#include "gt_internal.h"

using namespace gt;

// Scratch of gt_locate_noalloc: every workspace of that entry point lives in
// this array, which each device holds from module load until the process
// ends.  Host code never passes `scratch` itself (its name is a host-side
// stand-in, not a device address): gt_context_create takes the device
// address with hipGetSymbolAddress, with the device current, and the entry
// point uses that.
static __device__ __align__(256) unsigned char scratch[GT_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string t_err;

gt_status fail(gt_status s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
gt_status fail(gt_status s, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    t_err = buf;
    return s;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? GT_E_OOM : GT_E_HIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);              \
    } while (0)

#define GT_TRY(expr)                  \
    do {                              \
        gt_status s_ = (expr);        \
        if (s_ != GT_OK) return s_;   \
    } while (0)

// C entry points turn C++ exceptions (host allocation failures, thread
// creation) into status codes instead of letting them cross the C ABI.
template <class F>
gt_status guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(GT_E_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(GT_E_INTERNAL, "%s", e.what());
    } catch (...) {
        return fail(GT_E_INTERNAL, "unknown exception");
    }
}

// gt_debug_inject_alloc_failure: countdowns to a failing device (0) or host
// result (1) allocation; 0 = disarmed.
std::atomic<uint32_t> g_fail_after[2];

// True when this allocation of `kind` is the one that must fail.
bool injected_failure(int kind) {
    uint32_t c = g_fail_after[kind].load();
    while (c != 0) {
        if (g_fail_after[kind].compare_exchange_weak(c, c - 1)) return c == 1;
    }
    return false;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n == 0) n = 256;
        if (p && n <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        n = (n + 255) & ~(size_t)255;
        if (injected_failure(0)) return hipErrorOutOfMemory;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        else p = nullptr;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return (T*)p; }
};

// ---- exact division check -------------------------------------------------

__attribute__((target("fma"))) bool corrected_div_exact_fma(float d) {
    const float rd = 1.0f / d;
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        const uint32_t bits = 0x3f800000u | m;
        float x;
        std::memcpy(&x, &bits, 4);
        const float q = x * rd;
        const float r = __builtin_fmaf(-q, d, x);
        const float q2 = __builtin_fmaf(r, rd, q);
        if (q2 != x / d) return false;
    }
    return true;
}

bool corrected_div_exact_soft(float d) {
    const float rd = 1.0f / d;
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        const uint32_t bits = 0x3f800000u | m;
        float x;
        std::memcpy(&x, &bits, 4);
        const float q = x * rd;
        const float r = std::fma(-q, d, x);
        const float q2 = std::fma(r, rd, q);
        if (q2 != x / d) return false;
    }
    return true;
}

// Division x / d replaced by q = x*rd; r = fma(-q, d, x); q' = fma(r, rd, q)
// with rd = RN(1/d) is correctly rounded for every x exactly when it is for
// every significand in [1, 2) (the three steps commute with scaling by powers
// of two while no result leaves the normal range).  Checked exhaustively,
// cached per divisor.
bool corrected_div_exact(float d) {
    static std::mutex mu;
    static std::map<uint32_t, bool> cache;
    uint32_t key;
    std::memcpy(&key, &d, 4);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    if (!(d > 0.f) || !std::isfinite(d) || !std::isnormal(d)) return false;
    const bool ok = __builtin_cpu_supports("fma") ? corrected_div_exact_fma(d)
                                                  : corrected_div_exact_soft(d);
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = ok;
    return ok;
}

// ---- parameters -----------------------------------------------------------

gt_status check_params(const gt_params* p, uint32_t h, uint32_t w) {
    if (!p) return fail(GT_E_INVALID, "params is NULL");
    // lib.rs:73-74: not_implemented!(maxsize, threshold, invert, percentile, topn, preprocess)
    if (p->has_maxsize) return fail(GT_E_NOT_IMPLEMENTED, "maxsize is not implemented");
    if (p->has_threshold) return fail(GT_E_NOT_IMPLEMENTED, "threshold is not implemented");
    if (p->has_invert) return fail(GT_E_NOT_IMPLEMENTED, "invert is not implemented");
    if (p->has_topn) return fail(GT_E_NOT_IMPLEMENTED, "topn is not implemented");
    if (p->has_preprocess) return fail(GT_E_NOT_IMPLEMENTED, "preprocess is not implemented");
    // execute_gpu.rs:404-406,440: the background needs the signal radius (the
    // reference declares the background columns but fills them only then)
    if (p->has_bg_radius && !p->has_sig_radius)
        return fail(GT_E_INVALID, "bg_radius needs sig_radius");
    if (p->has_sig_radius && !std::isfinite(p->sig_radius))
        return fail(GT_E_INVALID, "sig_radius must be finite");
    if (p->has_bg_radius && (!std::isfinite(p->bg_radius) ||
                             (p->has_gap_radius && !std::isfinite(p->gap_radius))))
        return fail(GT_E_INVALID, "bg_radius and gap_radius must be finite");
    if (p->diameter < 1) return fail(GT_E_INVALID, "diameter must be >= 1");
    if (p->smoothing_size < 1) return fail(GT_E_INVALID, "smoothing_size must be >= 1");
    if (p->smoothing_size > 33) return fail(GT_E_INVALID, "smoothing_size must be <= 33");
    if (!(p->noise_size > 0.f) || !std::isfinite(p->noise_size))
        return fail(GT_E_INVALID, "noise_size must be positive and finite");
    if (!std::isfinite(p->minmass)) return fail(GT_E_INVALID, "minmass must be finite");
    if (p->has_percentile && !(p->percentile >= 0.f && p->percentile <= 100.f))
        return fail(GT_E_INVALID, "percentile must be in [0, 100]");
    if ((p->characterize || p->cpu_processed) && p->diameter > 31)
        return fail(GT_E_INVALID, "characterize/cpu_processed need diameter <= 31");
    if (p->max_iterations > 4096) return fail(GT_E_INVALID, "max_iterations must be <= 4096");
    if (p->separation > 4096) return fail(GT_E_INVALID, "separation must be <= 4096");
    if (h || w) {
        const uint32_t R = p->smoothing_size / 2;
        if (h < 1 || w < 1) return fail(GT_E_INVALID, "frame shape must be positive");
        if (h < R || w < R)
            return fail(GT_E_INVALID,
                        "frame %ux%u is smaller than smoothing_size/2 = %u (single mirror "
                        "reflection of the boxcar would leave the frame)",
                        h, w, R);
        // kernels address a frame's float32 plane with 32-bit byte offsets
        if ((uint64_t)h * w >= (1ull << 30)) return fail(GT_E_INVALID, "frame too large");
    }
    return GT_OK;
}

int32_t raw_cols_of(const gt_params* p) {
    return p->has_sig_radius ? (p->has_bg_radius ? 3 : 1) : 0;
}

int32_t ncols_of(const gt_params* p) {
    return 4 + (p->cpu_processed ? 3 : 0) + raw_cols_of(p) + (p->characterize ? 4 : 0);
}

// circle_inds (kernels.rs:149-164) and annulus_inds (kernels.rs:166-181):
// offsets (i, j), i = row, in row-major order over the bounding square,
// with the squared integer distance compared in binary32.
void circle_taps(float radius, std::vector<std::pair<int, int>>& out) {
    const int rb = (int)std::ceil(radius);
    const float r2 = radius * radius;
    for (int i = -rb; i <= rb; ++i)
        for (int j = -rb; j <= rb; ++j)
            if ((float)(i * i + j * j) <= r2) out.push_back({i, j});
}

void annulus_taps(float outer, float inner, std::vector<std::pair<int, int>>& out) {
    const int rb = (int)std::ceil(outer);
    const float or2 = outer * outer, ir2 = inner * inner;
    for (int i = -rb; i <= rb; ++i)
        for (int j = -rb; j <= rb; ++j) {
            const float d = (float)(i * i + j * j);
            if (d <= or2 && d >= ir2) out.push_back({i, j});
        }
}

// Derived geometry and host tables (gpu_setup.rs:45-60, kernels.rs:134-205).
gt_status make_geom(const gt_params* p, bool integer_input, Geom& g) {
    std::memset(&g, 0, sizeof(g));
    g.k = (int32_t)p->smoothing_size;
    g.R = g.k / 2;
    g.band = (int32_t)(p->diameter / 2);
    g.rint = ((int32_t)p->diameter - 1) / 2;
    {
        volatile float t = 2.0f * (float)p->separation;
        t = t / std::sqrt(2.0f);
        g.kd = (int32_t)(uint32_t)t;
        volatile float hk = (float)g.kd / 2.0f;
        hk = hk - 0.5f;
        g.a = (int32_t)hk;
    }
    g.max_iterations = p->max_iterations;
    g.minmass = p->minmass;
    g.sep = (float)p->separation;
    g.has_pct = p->has_percentile ? 1 : 0;
    g.percentile = p->percentile;
    g.characterize = p->characterize ? 1 : 0;
    g.cpu_processed = p->cpu_processed ? 1 : 0;
    g.raw_cols = raw_cols_of(p);
    g.col_raw = 4 + (p->cpu_processed ? 3 : 0);
    if (g.raw_cols) {
        if (std::ceil(std::fabs(p->sig_radius)) > 127.f ||
            (p->has_bg_radius && std::ceil(std::fabs(p->bg_radius)) > 127.f))
            return fail(GT_E_INVALID, "sig_radius and bg_radius must be at most 127");
        std::vector<std::pair<int, int>> st, bt;
        circle_taps(p->sig_radius, st);
        if (p->has_bg_radius) {
            // lib.rs:91: gap_radius defaults to 0 when bg_radius is given
            const float gap = p->has_gap_radius ? p->gap_radius : 0.f;
            annulus_taps(p->bg_radius, p->sig_radius + gap, bt);
        }
        if (st.size() > (size_t)RAW_MAX_TAPS || bt.size() > (size_t)RAW_MAX_TAPS)
            return fail(GT_E_INVALID, "sig_radius/bg_radius give more than %d taps", RAW_MAX_TAPS);
        g.raw_ns = (int32_t)st.size();
        g.raw_nb = (int32_t)bt.size();
        int k = 0;
        for (const auto& t : st) {
            g.raw_tap[k++] = (int8_t)t.first;
            g.raw_tap[k++] = (int8_t)t.second;
        }
        for (const auto& t : bt) {
            g.raw_tap[k++] = (int8_t)t.first;
            g.raw_tap[k++] = (int8_t)t.second;
        }
    }
    g.disk_in_window = (g.rint <= g.a) && (g.rint <= g.kd - 1 - g.a);
    // Gaussian weights (preprocess_rows.wgsl:44): argument in binary32,
    // exponential evaluated in binary64 on the host and rounded once.
    float gn = 0.f, wmin = 1.f;
    for (int i = -g.R; i <= g.R; ++i) {
        const float x = (float)i;
        volatile float arg = -(x * x);
        arg = arg / (2.0f * p->noise_size);
        const float wi = (float)std::exp((double)arg);
        g.w[i + g.R] = wi;
        gn = gn + wi;
        if (wi > 0.f) wmin = std::min(wmin, wi);
    }
    g.gn = gn;
    g.rgn = 1.0f / gn;
    g.kk = (float)(g.k * g.k);
    g.rkk = 1.0f / g.kk;
    g.fastdiv_gn = integer_input && wmin >= std::ldexp(1.0f, -40) && corrected_div_exact(gn);
    g.fastdiv_kk = integer_input && corrected_div_exact(g.kk);
    // dedup neighbourhood in tiles: start pixels of two rows closer than
    // `separation` differ by less than sep + 2*(max_iterations + rint) per axis
    const double D = std::ceil((double)p->separation) + 2.0 * ((double)p->max_iterations + g.rint) + 1.0;
    g.dedup_rty = (int32_t)(D / TILE_H) + 1;
    g.dedup_rtx = (int32_t)(D / TILE_W) + 1;
    g.ncols = ncols_of(p);
    // characterise tables (centers.wgsl:67-69 sin/cos of 2*atan2(j, i)), in
    // binary32 as kernels.rs:185-205 forms them: (2.*(j as f32).atan2(i as
    // f32)).sin() with the row offset i and the column offset j
    int t = 0;
    for (int i = -g.rint; i <= g.rint; ++i)
        for (int j = -g.rint; j <= g.rint; ++j) {
            if (i * i + j * j > g.rint * g.rint) continue;
            if (t >= 1024) return fail(GT_E_INVALID, "diameter too large for characterisation");
            const float a = 2.0f * atan2f((float)j, (float)i);
            g.sin2_gpu[t] = sinf(a);
            g.cos2_gpu[t] = cosf(a);
            ++t;
        }
    g.ntap_gpu = t;
    // cpu_processed tables over circle_inds(diameter/2) (kernels.rs:134-205)
    t = 0;
    for (int i = -g.band; i <= g.band; ++i)
        for (int j = -g.band; j <= g.band; ++j) {
            if (i * i + j * j > g.band * g.band) continue;
            if (t >= 1024) return fail(GT_E_INVALID, "diameter too large for characterisation");
            const float a = 2.0f * atan2f((float)j, (float)i);
            g.r2_cpu[t] = (float)(i * i + j * j);
            g.sin_cpu[t] = sinf(a);
            g.cos_cpu[t] = cosf(a);
            ++t;
        }
    g.ntap_cpu = t;
    return GT_OK;
}

const char* const kColFrame = "frame";

}  // namespace

// ---- context --------------------------------------------------------------

// Device buffers of one in-flight batch.  Consecutive batches alternate
// between two workspaces and two streams so that one batch's
// compute-bound bandpass overlaps the next batch's latency-bound stages.
struct Workspace {
    DevBuf in_stage, P, tmp, list, slots, rec, keep, tile3, perframe;
    void release() {
        DevBuf* b[] = {&in_stage, &P, &tmp, &list, &slots, &rec, &keep, &tile3, &perframe};
        for (DevBuf* x : b) x->release();
    }
};

constexpr int kStreams = 2;

struct gt_context {
    int device = 0;
    hipStream_t stream = nullptr;            // = streams[0]
    hipStream_t streams[kStreams] = {nullptr, nullptr};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_join = nullptr;
    hipEvent_t ev_scan[kStreams] = {nullptr, nullptr};
    gt_timing last{};
    uint32_t batch_frames = 0;   // 0 = automatic
    size_t workspace_limit = (size_t)6 << 30;   // per workspace
    int seg_rows = 128;
    // rows per bandpass work item of the plain per-frame path unless the
    // caller sets the launch shape: one frame per launch needs shorter
    // segments to occupy the chip (2048x2048 uint16, d = 11: 32 rows
    // measured 0.139 ms per frame against 0.181 ms with 128)
    int plain_seg_rows = 32;
    bool seg_set = false;   // gt_context_set_launch_shape was called
    int fuse_walk = 1;     // run the centroid walk inside detect where supported
    // batched launch sequence captured as a graph and replayed while the
    // call's inputs, parameters and buffers are unchanged
    int graphs = 0;
    // (call key, executable graph, launches) of recent calls, most recent
    // last: host-staged calls alternate their chunks between two pinned
    // buffers, so consecutive chunks have different keys
    struct GraphEntry {
        std::vector<uint64_t> key;
        hipGraphExec_t exec;
        uint64_t launches;
    };
    static constexpr size_t kGraphCache = 4;
    std::vector<GraphEntry> gcache;
    void drop_graphs() {
        for (auto& e : gcache) (void)hipGraphExecDestroy(e.exec);
        gcache.clear();
    }
    // autotuned (rows per bandpass work item, frames per launch) per frame
    // shape and filter geometry
    struct Tuned {
        int seg;
        uint32_t batch;
        int fuse;
    };
    std::map<std::array<int64_t, 6>, Tuned> tuned;
    // per-stage event timing (forces one stream so that kernels do not overlap)
    int profiling = 0;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
    double stage_ms[GT_NSTAGES] = {0};
    uint64_t stage_launches[GT_NSTAGES] = {0};
    DevBuf geom, frames_tab, bases, rows, rows_cand, counters;
    Workspace ws[kStreams];
    LinkStats last_link;
    LinkCache* link_cache = nullptr;   // linker buffers, created on first use
    // pinned host buffers of gt_locate_file (two chunks), kept for reuse
    void* pinned[2] = {nullptr, nullptr};
    size_t pinned_bytes = 0;
    // gt_locate_noalloc: this device's address of `scratch` (null when the
    // symbol lookup failed) and the host frame tables of one batch, sized at
    // creation so that the entry point itself allocates nothing
    void* scratch = nullptr;
    static constexpr size_t kNaFrames = 1024, kNaBases = 16384;
    std::vector<Frame> na_ftab;
    std::vector<uint32_t> na_bases;
    hipEvent_t ev_na = nullptr;
};

// Host result buffers.  Blocks of 1 MiB and more are kept when a result is
// freed (up to 256 MiB in all) and handed to the next result that fits, so
// calls in a loop write their rows into memory whose pages are already
// mapped instead of faulting in fresh pages on every call.
class HostBlockCache {
  public:
    static void* get(size_t bytes) {
        if (bytes >= kMin) {
            std::lock_guard<std::mutex> lk(mu());
            auto& v = blocks();
            size_t best = v.size();
            for (size_t i = 0; i < v.size(); ++i)
                if (v[i].second >= bytes && v[i].second <= 2 * bytes &&
                    (best == v.size() || v[i].second < v[best].second))
                    best = i;
            if (best != v.size()) {
                void* p = v[best].first;
                total() -= v[best].second;
                v.erase(v.begin() + (std::ptrdiff_t)best);
                return p;
            }
        }
        if (injected_failure(1)) throw std::bad_alloc();
        void* p = std::malloc(bytes ? bytes : 1);
        if (!p) throw std::bad_alloc();
        return p;
    }
    // bytes: the size the block was requested with (a reused block may be larger)
    static void put(void* p, size_t bytes) {
        if (!p) return;
        if (bytes < kMin || bytes > kMax) {
            std::free(p);
            return;
        }
        std::lock_guard<std::mutex> lk(mu());
        auto& v = blocks();
        v.emplace_back(p, bytes);
        total() += bytes;
        while (total() > kMax) {   // drop the oldest
            total() -= v.front().second;
            std::free(v.front().first);
            v.erase(v.begin());
        }
    }

  private:
    static constexpr size_t kMin = (size_t)1 << 20, kMax = (size_t)256 << 20;
    // never destroyed: results may be freed during process teardown
    static std::mutex& mu() { static std::mutex* m = new std::mutex; return *m; }
    static std::vector<std::pair<void*, size_t>>& blocks() {
        static auto* v = new std::vector<std::pair<void*, size_t>>;
        return *v;
    }
    static size_t& total() { static size_t t = 0; return t; }
};

template <class T>
struct HostAlloc {
    using value_type = T;
    HostAlloc() = default;
    template <class U>
    HostAlloc(const HostAlloc<U>&) {}
    T* allocate(size_t n) { return (T*)HostBlockCache::get(n * sizeof(T)); }
    void deallocate(T* p, size_t n) { HostBlockCache::put(p, n * sizeof(T)); }
    template <class U>
    bool operator==(const HostAlloc<U>&) const { return true; }
    template <class U>
    bool operator!=(const HostAlloc<U>&) const { return false; }
};
template <class T>
using HostVec = std::vector<T, HostAlloc<T>>;

struct gt_result {
    uint64_t rows = 0;
    uint32_t ncols = 0;
    HostVec<float> data;
    HostVec<int64_t> cand;
    const float* ddata = nullptr;
    std::vector<std::string> names, dtypes;
};

// Adjoint of one locate call: the forward rows plus, per row, the Poisson
// localisation errors, the Jacobian patch origin and (on request) the
// Jacobian patches, which stay in device memory for gt_adjoint_vjp.
struct gt_adjoint {
    gt_result* fwd = nullptr;
    int device = 0;
    uint32_t S = 0;
    uint64_t nframes = 0;
    uint32_t h = 0, w = 0;     // frame shape (0 when the call had mixed shapes)
    bool jacobian = false;
    HostVec<float> sigma;
    HostVec<int32_t> origin, rowpos;
    DevBuf d_jac, d_origin, d_frame_rows;
    ~gt_adjoint() {
        delete fwd;
        d_jac.release();
        d_origin.release();
        d_frame_rows.release();
    }
};

namespace {

struct FrameIn {
    const void* ptr;
    uint32_t h, w;
    int64_t fi;
};

size_t elem_size(int in_type) { return in_type == IN_U16 ? 2 : (in_type == IN_U8 ? 1 : 4); }

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

struct BatchPlan {
    size_t b0, b1;
};

struct FrameSizes {
    uint64_t px;
    uint32_t k1_tx, k1_ty, k1_items;
    uint32_t tx, ty, tiles;
};

std::array<int64_t, 6> tune_key(uint32_t h, uint32_t w, const Geom& g) {
    return {(int64_t)h, (int64_t)w, g.R, g.kd, g.rint, g.has_pct};
}

FrameSizes sizes_of(uint32_t h, uint32_t w, int seg) {
    FrameSizes s;
    s.px = (uint64_t)h * w;
    s.k1_tx = (w + K1_STRIP - 1) / K1_STRIP;
    s.k1_ty = (h + seg - 1) / seg;
    s.k1_items = s.k1_tx * s.k1_ty;
    s.tx = (w + TILE_W - 1) / TILE_W;
    s.ty = (h + TILE_H - 1) / TILE_H;
    s.tiles = s.tx * s.ty;
    return s;
}

int rec_planes(const Geom& g) {
    if (g.characterize) return REC_NPLANES;
    if (g.cpu_processed) return REC_SIGNAL + 1;
    if (g.walk_in_detect && g.has_pct) return REC_CENTER + 1;
    return REC_X + 1;
}

struct Need {
    uint64_t px = 0, tiles = 0, k1 = 0, in_bytes = 0;
    size_t nframes = 0;
};

size_t workspace_bytes(const Need& n, const Geom& g, bool tmp_planes) {
    size_t b = 0;
    b += al256(n.px * 4);                                   // P
    if (tmp_planes) b += al256(n.px * 8);                   // tmp
    if (g.has_pct) b += al256(n.tiles * TILE_PX * 4);       // list
    b += al256(n.tiles * TILE_PX * 2);                      // slots
    b += al256(n.tiles * TILE_PX * 4 * (uint64_t)rec_planes(g));  // rec
    b += al256(n.tiles * TILE_PX);                          // keep
    b += al256(n.in_bytes);                                 // input staging
    return b;
}

hipEvent_t take_event(gt_context* ctx) {
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Runs one stage launch; with profiling on, brackets it with events.
template <class F>
hipError_t stage(gt_context* ctx, hipStream_t st, int id, F&& launch) {
    if (!ctx->profiling) return launch();
    hipEvent_t a = take_event(ctx), b = take_event(ctx);
    if (!a || !b) return hipErrorOutOfMemory;
    hipError_t e = hipEventRecord(a, st);
    if (e != hipSuccess) return e;
    e = launch();
    if (e != hipSuccess) return e;
    e = hipEventRecord(b, st);
    ctx->ev_used.push_back({id, {a, b}});
    return e;
}

// After a stream synchronisation: fold the recorded event pairs into the
// per-stage totals and recycle the events.
void collect_stage_times(gt_context* ctx) {
    for (auto& u : ctx->ev_used) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, u.second.first, u.second.second) == hipSuccess) {
            ctx->stage_ms[u.first] += ms;
            ctx->stage_launches[u.first] += 1;
        }
        ctx->ev_pool.push_back(u.second.first);
        ctx->ev_pool.push_back(u.second.second);
    }
    ctx->ev_used.clear();
}

// Device buffers an adjoint request fills while the batches' P is resident.
struct AdjReq {
    bool jacobian = false;
    DevBuf sigma, origin, rowpos, jac;
    ~AdjReq() {
        sigma.release();
        origin.release();
        rowpos.release();
        jac.release();
    }
};

// Linking inside locate (execute_gpu.rs:335, 395-396): Linker::advance runs
// once per frame, frames without particles included, and every row gets the
// particle id as a float column right after x (execute_gpu.rs:351).
gt_status link_result(gt_context* ctx, const std::vector<FrameIn>& frames, const gt_params* p,
                      gt_result* res) {
    const uint32_t nc = res->ncols;
    const uint64_t n = res->rows;
    std::vector<uint32_t> step_off(frames.size() + 1, 0);
    std::vector<float> ys(n), xs(n);
    uint64_t r = 0;
    for (size_t f = 0; f < frames.size(); ++f) {
        const float fv = (float)frames[f].fi;
        for (; r < n && res->data[r * nc] == fv; ++r) {
            ys[r] = res->data[r * nc + 2];
            xs[r] = res->data[r * nc + 3];
        }
        step_off[f + 1] = (uint32_t)r;
    }
    if (r != n) return fail(GT_E_INTERNAL, "link: rows are not in frame order");
    std::vector<uint64_t> ids(n);
    const uint32_t mem = p->has_memory ? (uint32_t)std::min<uint64_t>(p->memory, UINT32_MAX) : 0;
    std::string err;
    if (!ctx->link_cache) ctx->link_cache = link_cache_create();
    const hipError_t e = link_steps(ctx->streams[0], ctx->link_cache, ys.data(), xs.data(), step_off,
                                    p->search_range, mem, ids.data(), &ctx->last_link, err);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? GT_E_OOM
                    : e == hipErrorInvalidValue ? GT_E_INVALID : GT_E_HIP, "%s", err.c_str());
    HostVec<float> out(n * (nc + 1));
    for (uint64_t i = 0; i < n; ++i) {
        const float* src = res->data.data() + i * nc;
        float* dst = out.data() + i * (nc + 1);
        for (uint32_t c = 0; c < 4; ++c) dst[c] = src[c];
        dst[4] = (float)ids[i];
        for (uint32_t c = 4; c < nc; ++c) dst[c + 1] = src[c];
    }
    res->data.swap(out);
    res->ncols = nc + 1;
    return GT_OK;
}

// Linking beside locate: chunks of located rows are linked in order on a
// worker thread (its own stream and buffers) while the next chunk is being
// located; Linker::advance runs once per frame, as in link_result.
struct LinkPipeline {
    gt_context* ctx;
    const gt_params* p;
    LinkState* state;
    std::future<std::pair<gt_status, std::string>> tail;
    LinkStats total{};
    // the result rows (particle column at index 4, `stride` floats per row),
    // grown by the caller under `mu`; the worker writes each chunk's ids
    // there under the same lock
    HostVec<float>* data;
    std::mutex* mu;
    uint32_t stride;
    LinkPipeline(gt_context* c, const gt_params* prm, uint64_t n_frames, HostVec<float>* d,
                 std::mutex* m, uint32_t st)
        : ctx(c), p(prm), data(d), mu(m), stride(st) {
        const uint64_t mem = prm->has_memory ? prm->memory : 0;
        state = link_state_create((uint32_t)std::min<uint64_t>(mem, std::max<uint64_t>(n_frames, 1)));
        if (!ctx->link_cache) ctx->link_cache = link_cache_create();
    }
    ~LinkPipeline() {
        if (tail.valid()) tail.wait();
        link_state_destroy(state);
    }
    // rows: the chunk's rows (nc columns, frame order), frames: its frames,
    // r0: the chunk's first row in the result
    gt_status submit(const std::vector<FrameIn>& frames, const float* rows, uint64_t nrows,
                     uint32_t nc, uint64_t r0) {
        std::vector<uint32_t> so(frames.size() + 1, 0);
        std::vector<float> ys(nrows), xs(nrows);
        uint64_t r = 0;
        for (size_t f = 0; f < frames.size(); ++f) {
            const float fv = (float)frames[f].fi;
            for (; r < nrows && rows[r * nc] == fv; ++r) {
                ys[r] = rows[r * nc + 2];
                xs[r] = rows[r * nc + 3];
            }
            so[f + 1] = (uint32_t)r;
        }
        if (r != nrows) return fail(GT_E_INTERNAL, "link: rows are not in frame order");
        auto prev = std::move(tail);
        tail = std::async(std::launch::async,
                          [this, prev = std::move(prev), ys = std::move(ys), xs = std::move(xs),
                           so = std::move(so), nrows, r0]() mutable -> std::pair<gt_status, std::string> {
                              if (prev.valid()) {
                                  auto pr = prev.get();
                                  if (pr.first != GT_OK) return pr;
                              }
                              if (hipSetDevice(ctx->device) != hipSuccess)
                                  return {GT_E_HIP, "link: hipSetDevice failed"};
                              hipStream_t st = nullptr;
                              std::string err;
                              hipError_t e = link_cache_stream(ctx->link_cache, &st);
                              LinkStats ls;
                              std::vector<uint64_t> ids(nrows);
                              if (e == hipSuccess)
                                  e = link_chunk(st, ctx->link_cache, state, ys.data(), xs.data(), so,
                                                 p->search_range, ids.data(), &ls, err);
                              if (e != hipSuccess)
                                  return {e == hipErrorOutOfMemory ? GT_E_OOM
                                          : e == hipErrorInvalidValue ? GT_E_INVALID : GT_E_HIP,
                                          err.empty() ? hipGetErrorString(e) : err};
                              total.steps += ls.steps;
                              total.pairs += ls.pairs;
                              total.subnets += ls.subnets;
                              total.max_subnet = std::max(total.max_subnet, ls.max_subnet);
                              total.ids = ls.ids;
                              total.search_ms += ls.search_ms;
                              total.solve_ms += ls.solve_ms;
                              {
                                  std::lock_guard<std::mutex> lk(*mu);
                                  float* d = data->data() + r0 * stride + 4;
                                  for (uint64_t i = 0; i < nrows; ++i) d[i * stride] = (float)ids[i];
                              }
                              return {GT_OK, {}};
                          });
        return GT_OK;
    }
    // waits for the last chunk (every id is then in place)
    gt_status finish() {
        if (tail.valid()) {
            auto r = tail.get();
            if (r.first != GT_OK) return fail(r.first, "%s", r.second.c_str());
        }
        ctx->last_link = total;
        return GT_OK;
    }
};

gt_status run_locate(gt_context* ctx, const std::vector<FrameIn>& frames, int in_type,
                     const gt_params* p, uint32_t flags, gt_result** out,
                     AdjReq* adj = nullptr) {
    const auto t_start = std::chrono::steady_clock::now();
    if (!ctx) return fail(GT_E_INVALID, "context is NULL");
    if (!out) return fail(GT_E_INVALID, "out is NULL");
    *out = nullptr;
    GT_TRY(check_params(p, 0, 0));
    for (const auto& f : frames) {
        GT_TRY(check_params(p, f.h, f.w));
        if (!f.ptr) return fail(GT_E_INVALID, "frame pointer is NULL");
    }
    HIP_TRY(hipSetDevice(ctx->device));
    const bool on_device = flags & GT_FRAMES_ON_DEVICE;
    const bool plain = flags & GT_PLAIN;
    const bool result_on_device = flags & GT_RESULT_ON_DEVICE;
    const bool integer_input = in_type != IN_F32;

    Geom g;
    GT_TRY(make_geom(p, integer_input, g));
    g.walk_in_detect = ctx->fuse_walk && walk_in_detect_supported(g) ? 1 : 0;
    // autotuned launch shape for calls whose frames all share one shape
    int seg = plain && !ctx->seg_set ? ctx->plain_seg_rows : ctx->seg_rows;
    uint32_t batch_frames = ctx->batch_frames;
    if (!frames.empty() && ctx->batch_frames == 0 && !plain) {
        bool uniform = true;
        for (const auto& f : frames) uniform = uniform && f.h == frames[0].h && f.w == frames[0].w;
        if (uniform) {
            auto it = ctx->tuned.find(tune_key(frames[0].h, frames[0].w, g));
            if (it != ctx->tuned.end()) {
                seg = it->second.seg;
                batch_frames = it->second.batch;
                if (!it->second.fuse) g.walk_in_detect = 0;
            }
        }
    }
    if (p->has_search_range) {
        if (result_on_device || adj)
            return fail(GT_E_NOT_IMPLEMENTED,
                        "search_range needs host results (no GT_RESULT_ON_DEVICE, no adjoint)");
        for (size_t f = 1; f < frames.size(); ++f)
            if ((float)frames[f].fi == (float)frames[f - 1].fi)
                return fail(GT_E_INVALID, "search_range needs consecutive frames with distinct indices");
    }
    if (adj && !adjoint_supported(g))
        return fail(GT_E_NOT_IMPLEMENTED, "adjoint needs (2*rint+1)^2 <= 1024 and smoothing_size/2 <= 16");
    const size_t es = elem_size(in_type);
    // vector loads need W % 4 == 0 and frame bases aligned to 4 pixels
    bool vec = true;
    for (const auto& f : frames) {
        if (f.w % 4 != 0) vec = false;
        if (on_device && ((uintptr_t)f.ptr) % (4 * es) != 0) vec = false;
    }
    bool generic;
    {
        Batch probe;
        std::memset(&probe, 0, sizeof(probe));
        probe.in_type = in_type;
        probe.vec = vec ? 1 : 0;
        generic = !bandpass_fast_path(probe, g);
    }
    // per-frame tmp planes: the generic bandpass's row pass, or the row
    // maxima of a dilation window too wide for detect's staged tile
    const bool tmp_planes = generic || detect_needs_plane(g);

    // owned here until it is handed to the caller: every error return frees it
    std::unique_ptr<gt_result> res(new gt_result());
    res->ncols = (uint32_t)g.ncols;
    {
        const char* names[16];
        const char* dtypes[16];
        uint32_t nc = 0;
        gt_column_names(p, names, dtypes, 16, &nc);
        for (uint32_t i = 0; i < nc; ++i) {
            res->names.push_back(names[i]);
            res->dtypes.push_back(dtypes[i]);
        }
    }
    ctx->last = gt_timing{};
    ctx->last.frames = frames.size();
    if (frames.empty()) {
        *out = res.release();
        return GT_OK;
    }

    // ---- batch plan -------------------------------------------------------
    const size_t max_batch = plain ? 1 : (batch_frames ? batch_frames : 4096);
    std::vector<BatchPlan> plan;
    Need maxneed;
    {
        size_t b0 = 0;
        while (b0 < frames.size()) {
            Need n;
            size_t b1 = b0;
            while (b1 < frames.size() && (b1 - b0) < max_batch) {
                const FrameSizes fs = sizes_of(frames[b1].h, frames[b1].w, seg);
                Need m = n;
                m.px += fs.px;
                m.tiles += fs.tiles;
                m.k1 += fs.k1_items;
                m.in_bytes += on_device ? 0 : al256(fs.px * es);
                m.nframes += 1;
                if (b1 > b0 && workspace_bytes(m, g, tmp_planes) > ctx->workspace_limit) break;
                n = m;
                ++b1;
            }
            plan.push_back({b0, b1});
            maxneed.px = std::max(maxneed.px, n.px);
            maxneed.tiles = std::max(maxneed.tiles, n.tiles);
            maxneed.k1 = std::max(maxneed.k1, n.k1);
            maxneed.in_bytes = std::max(maxneed.in_bytes, n.in_bytes);
            maxneed.nframes = std::max(maxneed.nframes, n.nframes);
            b0 = b1;
        }
    }
    uint64_t total_px = 0;
    for (const auto& f : frames) total_px += (uint64_t)f.h * f.w;
    ctx->last.pixels = total_px;

    // ---- workspaces -------------------------------------------------------
    const int nstreams = (plain || plan.size() < 2 || ctx->profiling) ? 1 : kStreams;
    const uint64_t slot_cap = maxneed.tiles * TILE_PX;
    const size_t nfb = maxneed.nframes;
    // per-frame: hist (4096 u32), frame_rows, frame_cands, spare, selstate(4), thr(double), row_base(u64)
    const size_t pf_hist = al256(nfb * HIST_BINS * 4);
    const size_t pf_small = al256(nfb * 4);
    const size_t pf_sel = al256(nfb * 16);
    const size_t pf_8 = al256(nfb * 8);
    HIP_TRY(ctx->geom.ensure(sizeof(Geom)));
    for (int si = 0; si < nstreams; ++si) {
        Workspace& w = ctx->ws[si];
        HIP_TRY(w.P.ensure(maxneed.px * 4));
        if (tmp_planes) HIP_TRY(w.tmp.ensure(maxneed.px * 8));
        if (g.has_pct) HIP_TRY(w.list.ensure(slot_cap * 4));
        HIP_TRY(w.slots.ensure(slot_cap * 2));
        HIP_TRY(w.rec.ensure(slot_cap * 4 * (uint64_t)rec_planes(g)));
        HIP_TRY(w.keep.ensure(slot_cap));
        if (!on_device) HIP_TRY(w.in_stage.ensure(maxneed.in_bytes));
        HIP_TRY(w.tile3.ensure(al256(maxneed.tiles * 4) * 4 + 2 * al256(maxneed.tiles * 4 * TILE_H)));
        HIP_TRY(w.perframe.ensure(pf_hist + 3 * pf_small + pf_sel + 2 * pf_8));
    }
    HIP_TRY(ctx->counters.ensure(256));

    // rows capacity: one row per 16 pixels (per 256 when Jacobian patches
    // ride along), grown on overflow
    uint64_t rows_cap = std::max<uint64_t>(4096, total_px / (adj && adj->jacobian ? 256 : 16));
    if (!(adj && adj->jacobian) && ctx->rows.bytes >= rows_cap * 4 * g.ncols &&
        ctx->rows_cand.bytes >= rows_cap * 8) {
        rows_cap = std::min<uint64_t>(ctx->rows.bytes / (4 * (uint64_t)g.ncols), ctx->rows_cand.bytes / 8);
    }

    // ---- frame tables -------------------------------------------------------
    // The host builds the frame table and, per batch, the small base arrays
    // (first bandpass item, first tile and first dedup block of every frame:
    // 3 x (nb + 1) entries); k_fill_tables expands them on the device into the
    // per-item tables (frame of every bandpass item and tile, top-left tile of
    // every dedup block), so a call uploads kilobytes, not megabytes.
    std::vector<Frame> ftab(frames.size());
    std::vector<uint32_t> small;              // per batch: k1, tile and dedup-block bases
    std::vector<size_t> small_off(plan.size());
    std::vector<uint64_t> big_off(plan.size());   // per batch: k1_frame, tile_frame, dblock_tile
    std::vector<uint32_t> k1_items_of(plan.size()), tiles_of(plan.size()), dblocks_of(plan.size());
    uint32_t max_frame_work = 1;              // largest bandpass items + tiles + blocks of a frame
    uint64_t big_total = 0;
    for (size_t bi = 0; bi < plan.size(); ++bi) {
        const auto& bp = plan[bi];
        Workspace& wsb = ctx->ws[bi % nstreams];
        const size_t nb = bp.b1 - bp.b0;
        uint64_t pxo = 0, inoff = 0;
        uint32_t k1b = 0, tb = 0, db = 0;
        small_off[bi] = small.size();
        small.resize(small.size() + 3 * (nb + 1));
        uint32_t* k1v = small.data() + small_off[bi];
        uint32_t* tbv = k1v + nb + 1;
        uint32_t* dbv = tbv + nb + 1;
        for (size_t fi = bp.b0; fi < bp.b1; ++fi) {
            const FrameSizes fs = sizes_of(frames[fi].h, frames[fi].w, seg);
            Frame& F = ftab[fi];
            std::memset(&F, 0, sizeof(F));
            if (on_device) {
                F.in = frames[fi].ptr;
            } else {
                F.in = (const char*)wsb.in_stage.p + inoff;
                inoff += al256(fs.px * es);
            }
            F.P = wsb.P.as<float>() + pxo;
            F.tmp = tmp_planes ? wsb.tmp.as<float>() + 2 * pxo : nullptr;
            F.H = frames[fi].h;
            F.W = frames[fi].w;
            F.k1_tx = fs.k1_tx;
            F.k1_ty = fs.k1_ty;
            F.tx = fs.tx;
            F.ty = fs.ty;
            F.tile_base = tb;
            F.k1_base = k1b;
            F.slot_base = (uint64_t)tb * TILE_PX;
            F.frame_index = frames[fi].fi;
            const uint32_t nblk = ((fs.ty + 1) / 2) * ((fs.tx + 1) / 2);
            k1v[fi - bp.b0] = k1b;
            tbv[fi - bp.b0] = tb;
            dbv[fi - bp.b0] = db;
            max_frame_work = std::max<uint32_t>(max_frame_work, fs.k1_items + fs.tiles + nblk);
            k1b += fs.k1_items;
            tb += fs.tiles;
            db += nblk;
            pxo += fs.px;
        }
        k1v[nb] = k1b;
        tbv[nb] = tb;
        dbv[nb] = db;
        k1_items_of[bi] = k1b;
        tiles_of[bi] = tb;
        dblocks_of[bi] = db;
        big_off[bi] = big_total;
        big_total += (uint64_t)k1b + tb + db;
    }
    HIP_TRY(ctx->frames_tab.ensure(ftab.size() * sizeof(Frame)));
    HIP_TRY(ctx->bases.ensure((small.size() + big_total) * 4));
    HIP_TRY(hipMemcpy(ctx->frames_tab.p, ftab.data(), ftab.size() * sizeof(Frame), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->bases.p, small.data(), small.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->geom.p, &g, sizeof(Geom), hipMemcpyHostToDevice));
    const Geom* d_geom = ctx->geom.as<const Geom>();

    uint64_t launches = 0;
    const int S = adjoint_patch(g);
    for (int attempt = 0; attempt < 2; ++attempt) {
        HIP_TRY(ctx->rows.ensure(rows_cap * 4 * (uint64_t)g.ncols));
        HIP_TRY(ctx->rows_cand.ensure(rows_cap * 8));
        AdjOut ao{};
        if (adj) {
            HIP_TRY(adj->sigma.ensure(rows_cap * 12));
            HIP_TRY(adj->origin.ensure(rows_cap * 8));
            HIP_TRY(adj->rowpos.ensure(rows_cap * 4));
            ao.sigma = adj->sigma.as<float>();
            ao.origin = adj->origin.as<int32_t>();
            ao.rowpos = adj->rowpos.as<int32_t>();
            if (adj->jacobian) {
                HIP_TRY(adj->jac.ensure(rows_cap * 12 * (uint64_t)S * S));
                ao.jac = adj->jac.as<float>();
            }
        }
        uint64_t* counters = ctx->counters.as<uint64_t>();  // [0] cand_total, [1] row_carry
        res->data.clear();
        res->cand.clear();
        launches = 0;

        hipStream_t s0 = ctx->streams[0];
        uint64_t host_rows_done = 0;
        auto enqueue = [&]() -> gt_status {
        HIP_TRY(launch_fill_u64(counters, 2, 0, s0));
        ++launches;
        if (nstreams > 1) {
            HIP_TRY(hipEventRecord(ctx->ev_join, s0));
            HIP_TRY(hipStreamWaitEvent(ctx->streams[1], ctx->ev_join, 0));
        }
        for (size_t bi = 0; bi < plan.size(); ++bi) {
            const auto& bp = plan[bi];
            const int si = (int)(bi % nstreams);
            hipStream_t st = ctx->streams[si];
            Workspace& w = ctx->ws[si];
            const uint32_t nb = (uint32_t)(bp.b1 - bp.b0);
            Batch b;
            std::memset(&b, 0, sizeof(b));
            b.frames = ctx->frames_tab.as<const Frame>() + bp.b0;
            b.k1_bases = ctx->bases.as<const uint32_t>() + small_off[bi];
            b.tile_bases = b.k1_bases + nb + 1;
            b.nframes = nb;
            b.k1_items = k1_items_of[bi];
            b.tiles = tiles_of[bi];
            const uint32_t* dblock_bases = b.tile_bases + nb + 1;
            b.k1_frame = ctx->bases.as<const uint32_t>() + small.size() + big_off[bi];
            b.tile_frame = b.k1_frame + b.k1_items;
            b.dblock_tile = b.tile_frame + b.tiles;
            b.dblocks = dblocks_of[bi];
            b.in_type = in_type;
            b.vec = vec ? 1 : 0;
            char* t3 = w.tile3.as<char>();
            const size_t t3s = al256(maxneed.tiles * 4);
            b.tile_count = (uint32_t*)t3;
            b.tile_keep = (uint32_t*)(t3 + t3s);
            b.tile_row_off = (uint32_t*)(t3 + 2 * t3s);
            b.tile_lcount = (uint32_t*)(t3 + 3 * t3s);
            b.rowk = (uint32_t*)(t3 + 4 * t3s);
            b.rowoff = (uint32_t*)(t3 + 4 * t3s + al256(maxneed.tiles * 4 * TILE_H));
            char* pf = w.perframe.as<char>();
            b.hist1 = g.has_pct ? (uint32_t*)pf : nullptr;
            pf += pf_hist;
            b.list = g.has_pct ? w.list.as<uint32_t>() : nullptr;
            b.frame_rows = (uint32_t*)pf; pf += pf_small;
            b.frame_cands = (uint32_t*)pf; pf += pf_small;
            pf += pf_small;
            b.selstate = (int32_t*)pf; pf += pf_sel;
            b.thr = (double*)pf; pf += pf_8;
            b.frame_row_base = (uint64_t*)pf;
            b.slots = w.slots.as<uint16_t>();
            b.rec = w.rec.as<float>();
            b.keep = w.keep.as<uint8_t>();
            b.slot_cap = slot_cap;
            b.cand_total = counters;
            b.row_carry = counters + 1;
            b.rows = ctx->rows.as<float>();
            b.rows_cand = ctx->rows_cand.as<int64_t>();
            b.rows_cap = rows_cap;

            if (!on_device) {
                size_t off = 0;
                for (size_t fi = bp.b0; fi < bp.b1; ++fi) {
                    const size_t nbytes = (size_t)frames[fi].h * frames[fi].w * es;
                    HIP_TRY(hipMemcpyAsync((char*)w.in_stage.p + off, frames[fi].ptr, nbytes,
                                           hipMemcpyHostToDevice, st));
                    off += al256(nbytes);
                }
            }
            HIP_TRY(stage(ctx, st, GT_STAGE_ZERO, [&] {
                const hipError_t e = launch_fill_tables(b, dblock_bases, max_frame_work, st);
                return e != hipSuccess ? e : launch_zero_batch(b, st);
            }));
            launches += 2;
            if (generic) {
                uint32_t maxpx = 0;
                for (size_t fi = bp.b0; fi < bp.b1; ++fi)
                    maxpx = std::max<uint32_t>(maxpx, frames[fi].h * frames[fi].w);
                HIP_TRY(stage(ctx, st, GT_STAGE_BANDPASS, [&] {
                    return launch_bandpass_generic(b, d_geom, maxpx, g.has_pct, st);
                }));
                launches += 2;
            } else {
                HIP_TRY(stage(ctx, st, GT_STAGE_BANDPASS,
                              [&] { return launch_bandpass(b, g, d_geom, seg, st); }));
                ++launches;
            }
            if (g.has_pct) {
                HIP_TRY(stage(ctx, st, GT_STAGE_SELECT1, [&] { return launch_select1(b, g, st); }));
                ++launches;
            }
            HIP_TRY(stage(ctx, st, GT_STAGE_DETECT, [&] { return launch_detect(b, g, d_geom, st); }));
            ++launches;
            if (g.has_pct) {
                HIP_TRY(stage(ctx, st, GT_STAGE_SELECT2, [&] { return launch_select2(b, g, st); }));
                ++launches;
            }
            HIP_TRY(stage(ctx, st, GT_STAGE_REFINE, [&] { return launch_refine(b, g, d_geom, st); }));
            HIP_TRY(stage(ctx, st, GT_STAGE_DEDUP, [&] { return launch_dedup(b, g, d_geom, st); }));
            uint32_t max_ty = 1;
            for (size_t fi = bp.b0; fi < bp.b1; ++fi)
                max_ty = std::max(max_ty, ftab[fi].ty);
            HIP_TRY(stage(ctx, st, GT_STAGE_SCAN, [&] { return launch_tile_scan(b, max_ty, st); }));
            // row bases continue the previous batch's count: order the scans
            if (nstreams > 1 && bi > 0)
                HIP_TRY(hipStreamWaitEvent(st, ctx->ev_scan[(bi - 1) % nstreams], 0));
            HIP_TRY(stage(ctx, st, GT_STAGE_SCAN, [&] { return launch_frame_scan(b, st); }));
            if (nstreams > 1) HIP_TRY(hipEventRecord(ctx->ev_scan[si], st));
            HIP_TRY(stage(ctx, st, GT_STAGE_EMIT, [&] { return launch_emit(b, g, d_geom, st); }));
            launches += 5;
            if (g.raw_cols) {
                HIP_TRY(launch_raw(b, g, d_geom, st));
                ++launches;
            }
            if (adj) {
                HIP_TRY(launch_adjoint(b, g, d_geom, ao, (uint32_t)bp.b0, st));
                ++launches;
            }

            if (plain && result_on_device) {
                // per-frame completion, as the reference's get_work waits per frame
                HIP_TRY(hipStreamSynchronize(st));
            } else if (plain) {
                // per-frame readback, as the reference's get_work does
                uint64_t carry = 0;
                HIP_TRY(hipMemcpyAsync(&carry, counters + 1, 8, hipMemcpyDeviceToHost, st));
                HIP_TRY(hipStreamSynchronize(st));
                const uint64_t upto = std::min(carry, rows_cap);
                if (upto > host_rows_done) {
                    const uint64_t n = upto - host_rows_done;
                    res->data.resize(upto * g.ncols);
                    res->cand.resize(upto);
                    HIP_TRY(hipMemcpyAsync(res->data.data() + host_rows_done * g.ncols,
                                           ctx->rows.as<float>() + host_rows_done * g.ncols,
                                           n * g.ncols * 4, hipMemcpyDeviceToHost, st));
                    HIP_TRY(hipMemcpyAsync(res->cand.data() + host_rows_done,
                                           ctx->rows_cand.as<int64_t>() + host_rows_done, n * 8,
                                           hipMemcpyDeviceToHost, st));
                    HIP_TRY(hipStreamSynchronize(st));
                    host_rows_done = upto;
                }
            }
        }
        if (nstreams > 1) {
            HIP_TRY(hipEventRecord(ctx->ev_join, ctx->streams[1]));
            HIP_TRY(hipStreamWaitEvent(s0, ctx->ev_join, 0));
        }
        return GT_OK;
        };  // enqueue

        // Replay a captured launch sequence when this call repeats one of the
        // last kGraphCache calls (same inputs, parameters, launch shape and
        // buffers); capture it otherwise.  Any capture failure falls back to launching
        // directly and turns graphs off for the context.
        bool launched = false;
        if (ctx->graphs && !plain && !ctx->profiling && !adj) {
            std::vector<uint64_t> key;
            for (const auto& f : frames) {
                key.push_back((uint64_t)(uintptr_t)f.ptr);
                key.push_back(((uint64_t)f.h << 32) | f.w);
                key.push_back((uint64_t)f.fi);
            }
            const uint64_t* pw = (const uint64_t*)p;
            for (size_t i = 0; i < sizeof(gt_params) / 8; ++i) key.push_back(pw[i]);
            key.push_back(((uint64_t)flags << 32) | (uint32_t)in_type);
            key.push_back(rows_cap);
            key.push_back(((uint64_t)seg << 32) | (uint32_t)nstreams);
            key.push_back(((uint64_t)batch_frames << 32) | (uint32_t)g.walk_in_detect);
            const DevBuf* bufs[] = {&ctx->geom, &ctx->frames_tab, &ctx->bases, &ctx->rows,
                                    &ctx->rows_cand, &ctx->counters};
            for (const DevBuf* d : bufs) key.push_back((uint64_t)(uintptr_t)d->p);
            for (int si = 0; si < nstreams; ++si) {
                const Workspace& w = ctx->ws[si];
                const DevBuf* wb[] = {&w.in_stage, &w.P, &w.tmp, &w.list, &w.slots,
                                      &w.rec, &w.keep, &w.tile3, &w.perframe};
                for (const DevBuf* d : wb) key.push_back((uint64_t)(uintptr_t)d->p);
            }
            size_t hit = ctx->gcache.size();
            for (size_t i = 0; i < ctx->gcache.size(); ++i)
                if (ctx->gcache[i].key == key) hit = i;
            if (hit == ctx->gcache.size()) {
                hipGraphExec_t ex = nullptr;
                if (hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed) == hipSuccess) {
                    const gt_status st = enqueue();
                    hipGraph_t gr = nullptr;
                    const hipError_t e = hipStreamEndCapture(s0, &gr);
                    if (!(st == GT_OK && e == hipSuccess && gr &&
                          hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) == hipSuccess))
                        ex = nullptr;
                    if (gr) (void)hipGraphDestroy(gr);
                }
                if (ex) {
                    if (ctx->gcache.size() == gt_context::kGraphCache) {
                        (void)hipGraphExecDestroy(ctx->gcache.front().exec);
                        ctx->gcache.erase(ctx->gcache.begin());
                    }
                    ctx->gcache.push_back({key, ex, launches});
                    hit = ctx->gcache.size() - 1;
                } else {
                    ctx->graphs = 0;
                    (void)hipGetLastError();
                }
            }
            if (hit < ctx->gcache.size()) {
                // most recent last
                std::rotate(ctx->gcache.begin() + (std::ptrdiff_t)hit,
                            ctx->gcache.begin() + (std::ptrdiff_t)hit + 1, ctx->gcache.end());
                const auto& e = ctx->gcache.back();
                HIP_TRY(hipEventRecord(ctx->ev0, s0));
                HIP_TRY(hipGraphLaunch(e.exec, s0));
                launches = e.launches;
                launched = true;
            }
        }
        if (!launched) {
            launches = 0;
            HIP_TRY(hipEventRecord(ctx->ev0, s0));
            GT_TRY(enqueue());
        }
        HIP_TRY(hipEventRecord(ctx->ev1, s0));
        uint64_t cnt[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(cnt, counters, 16, hipMemcpyDeviceToHost, s0));
        HIP_TRY(hipStreamSynchronize(s0));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last.device_ms = ms;
        collect_stage_times(ctx);
        ctx->last.candidates = cnt[0];
        const uint64_t total_rows = cnt[1];
        if (total_rows > rows_cap) {
            rows_cap = total_rows + 1024;
            continue;  // rerun with enough room (results are deterministic)
        }
        res->rows = total_rows;
        if (result_on_device) {
            res->ddata = ctx->rows.as<float>();
        } else if (!plain) {
            res->data.resize(total_rows * g.ncols);
            res->cand.resize(total_rows);
            if (total_rows) {
                HIP_TRY(hipMemcpy(res->data.data(), ctx->rows.p, total_rows * g.ncols * 4,
                                  hipMemcpyDeviceToHost));
                HIP_TRY(hipMemcpy(res->cand.data(), ctx->rows_cand.p, total_rows * 8,
                                  hipMemcpyDeviceToHost));
            }
        }
        break;
    }
    if (p->has_search_range) GT_TRY(link_result(ctx, frames, p, res.get()));
    ctx->last.rows = res->rows;
    ctx->last.launches = launches;
    ctx->last.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    *out = res.release();
    return GT_OK;
}

}  // namespace

namespace gt {
gt_status set_last_error(gt_status s, const std::string& msg) { return fail(s, "%s", msg.c_str()); }
}  // namespace gt

// ---- C ABI ----------------------------------------------------------------

extern "C" {

const char* gt_last_error(void) { return t_err.c_str(); }

gt_status gt_debug_inject_alloc_failure(int kind, uint32_t after) {
    if (kind != 0 && kind != 1) return fail(GT_E_INVALID, "kind %d (0 device, 1 host)", kind);
    g_fail_after[kind].store(after);
    return GT_OK;
}

const char* gt_version(void) { return "gpu_tracking-mi355x 0.1.0 (gfx950, abi 1)"; }

int gt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

gt_status gt_params_default(gt_params* p, uint32_t diameter) {
    if (!p) return fail(GT_E_INVALID, "params is NULL");
    std::memset(p, 0, sizeof(*p));
    // lib.rs:77-91
    p->diameter = diameter;
    p->minmass = 0.f;
    p->maxsize = INFINITY;
    p->separation = diameter + 1;
    p->noise_size = 1.f;
    p->smoothing_size = diameter;
    p->threshold = 1.f / 255.f;
    p->invert = 0;
    p->percentile = 64.f;
    p->topn = UINT32_MAX;
    p->preprocess = 1;
    p->max_iterations = 10;
    p->characterize = 0;
    p->filter_close = 1;
    p->cpu_processed = 0;
    return GT_OK;
}

gt_status gt_params_rust_default(gt_params* p) {
    if (!p) return fail(GT_E_INVALID, "params is NULL");
    std::memset(p, 0, sizeof(*p));
    // execute_gpu.rs:40-65
    p->diameter = 9;
    p->minmass = 0.f;
    p->maxsize = 0.f;
    p->separation = 11;
    p->noise_size = 1.f;
    p->smoothing_size = 9;
    p->threshold = 0.f;
    p->invert = 0;
    p->percentile = 0.f;
    p->topn = 0;
    p->preprocess = 1;
    p->max_iterations = 10;
    p->characterize = 0;
    p->filter_close = 1;
    p->cpu_processed = 1;
    return GT_OK;
}

gt_status gt_params_validate(const gt_params* p, uint32_t h, uint32_t w) {
    return check_params(p, h, w);
}

gt_status gt_column_names(const gt_params* p, const char** names, const char** dtypes,
                          uint32_t cap, uint32_t* ncols) {
    if (!p || !ncols) return fail(GT_E_INVALID, "NULL argument");
    static const char* const kInt = "int";
    static const char* const kFloat = "float";
    const char* n[16];
    const char* d[16];
    uint32_t c = 0;
    auto push = [&](const char* name, const char* dt) { n[c] = name; d[c] = dt; ++c; };
    // execute_gpu.rs:197-226
    push(kColFrame, kInt);
    push("mass", kFloat);
    push("y", kFloat);
    push("x", kFloat);
    if (p->has_search_range) push("particle", kInt);
    if (p->cpu_processed) {
        push("Rg", kFloat);
        push("ecc", kFloat);
        push("signal", kFloat);
    }
    if (p->has_sig_radius) push("raw_mass", kFloat);
    if (p->has_bg_radius) {
        push("raw_bg_median", kFloat);
        push("raw_mass_corrected", kFloat);
    }
    if (p->characterize) {
        push("Rg_gpu", kFloat);
        push("raw_mass_gpu", kFloat);
        push("signal_gpu", kFloat);
        push("ecc_gpu", kFloat);
    }
    *ncols = c;
    for (uint32_t i = 0; i < c && i < cap; ++i) {
        if (names) names[i] = n[i];
        if (dtypes) dtypes[i] = d[i];
    }
    return c > cap ? fail(GT_E_INVALID, "capacity %u < %u columns", cap, c) : GT_OK;
}

gt_status gt_context_create(int device, gt_context** out) {
    if (!out) return fail(GT_E_INVALID, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(GT_E_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(GT_E_INVALID, "device %d out of range (%d)", device, n);
    HIP_TRY(hipSetDevice(device));
    gt_context* c = new gt_context();
    c->device = device;
    bool ok = true;
    for (int i = 0; i < kStreams; ++i)
        ok = ok && hipStreamCreateWithFlags(&c->streams[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&c->ev_scan[i], hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess &&
         hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&c->ev_na, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        delete c;
        return fail(GT_E_HIP, "stream/event creation failed");
    }
    c->stream = c->streams[0];
    // gt_locate_noalloc's scratch on this device (the device is current)
    void* sp = nullptr;
    const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (se == hipSuccess && sp) {
        c->scratch = sp;
    } else {
        (void)hipGetLastError();
        c->scratch = nullptr;   // gt_locate_noalloc reports the failure
    }
    c->na_ftab.resize(gt_context::kNaFrames);
    c->na_bases.resize(gt_context::kNaBases);
    *out = c;
    return GT_OK;
}

gt_status gt_context_destroy(gt_context* ctx) {
    if (!ctx) return GT_OK;
    (void)hipSetDevice(ctx->device);
    for (int i = 0; i < kStreams; ++i) (void)hipStreamSynchronize(ctx->streams[i]);
    ctx->drop_graphs();
    DevBuf* bufs[] = {&ctx->geom, &ctx->frames_tab, &ctx->bases, &ctx->rows, &ctx->rows_cand,
                      &ctx->counters};
    for (DevBuf* b : bufs) b->release();
    for (int i = 0; i < kStreams; ++i) ctx->ws[i].release();
    for (void* b : ctx->pinned)
        if (b) (void)hipHostFree(b);
    if (ctx->link_cache) link_cache_destroy(ctx->link_cache);
    collect_stage_times(ctx);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipEventDestroy(ctx->ev_join);
    if (ctx->ev_na) (void)hipEventDestroy(ctx->ev_na);
    for (int i = 0; i < kStreams; ++i) {
        (void)hipEventDestroy(ctx->ev_scan[i]);
        (void)hipStreamDestroy(ctx->streams[i]);
    }
    delete ctx;
    return GT_OK;
}

gt_status gt_context_timing(const gt_context* ctx, gt_timing* out) {
    if (!ctx || !out) return fail(GT_E_INVALID, "NULL argument");
    *out = ctx->last;
    return GT_OK;
}

gt_status gt_context_set_profiling(gt_context* ctx, int enable) {
    if (!ctx) return fail(GT_E_INVALID, "context is NULL");
    ctx->profiling = enable ? 1 : 0;
    for (int i = 0; i < GT_NSTAGES; ++i) {
        ctx->stage_ms[i] = 0.0;
        ctx->stage_launches[i] = 0;
    }
    return GT_OK;
}

gt_status gt_context_stage_times(const gt_context* ctx, double* ms, uint64_t* launches) {
    if (!ctx) return fail(GT_E_INVALID, "context is NULL");
    for (int i = 0; i < GT_NSTAGES; ++i) {
        if (ms) ms[i] = ctx->stage_ms[i];
        if (launches) launches[i] = ctx->stage_launches[i];
    }
    return GT_OK;
}

gt_status gt_context_set_launch_shape(gt_context* ctx, uint32_t seg_rows) {
    if (!ctx) return fail(GT_E_INVALID, "context is NULL");
    if (seg_rows < 16 || seg_rows > 272)
        return fail(GT_E_INVALID, "rows per bandpass work item must be in [16, 272]");
    ctx->seg_rows = (int)seg_rows;
    ctx->seg_set = true;
    ctx->tuned.clear();
    return GT_OK;
}

gt_status gt_context_set_graphs(gt_context* ctx, int enable) {
    if (!ctx) return fail(GT_E_INVALID, "context is NULL");
    ctx->graphs = enable ? 1 : 0;
    if (!enable && !ctx->gcache.empty()) {
        (void)hipSetDevice(ctx->device);
        ctx->drop_graphs();
    }
    return GT_OK;
}

gt_status gt_context_autotune(gt_context* ctx, const gt_params* p, uint32_t h, uint32_t w,
                              uint32_t n_frames, gt_tuning* out) {
    return guarded([&]() -> gt_status {
    if (!ctx || !p) return fail(GT_E_INVALID, "NULL argument");
    GT_TRY(check_params(p, h, w));
    if (n_frames == 0) n_frames = 1;
    HIP_TRY(hipSetDevice(ctx->device));
    Geom g;
    GT_TRY(make_geom(p, true, g));
    // synthetic uint16 frames: background 100 + Poisson, one spot per 800 px^2
    const uint64_t npx = (uint64_t)h * w;
    struct ScopedBuf : DevBuf {
        ~ScopedBuf() { release(); }
    } movie;
    HIP_TRY(movie.ensure(npx * n_frames * 2));
    {
        uint64_t st = 0x9e3779b97f4a7c15ull ^ ((uint64_t)h << 20) ^ w;
        auto rnd = [&]() {  // splitmix64 -> [0, 1)
            uint64_t z = (st += 0x9e3779b97f4a7c15ull);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
        };
        const uint64_t per = std::max<uint64_t>(1, npx / 800);
        std::vector<float> spots;
        std::vector<uint64_t> off(n_frames + 1, 0);
        std::vector<float> bg(n_frames, 100.f);
        for (uint32_t f = 0; f < n_frames; ++f) {
            for (uint64_t i = 0; i < per; ++i) {
                spots.push_back((float)(rnd() * (h - 1)));
                spots.push_back((float)(rnd() * (w - 1)));
                spots.push_back((float)(500.0 + 1500.0 * rnd()));
                spots.push_back((float)(1.5 + rnd()));
            }
            off[f + 1] = off[f] + per;
        }
        GT_TRY(gt_synth_u16(ctx, movie.as<uint16_t>(), n_frames, h, w, spots.data(), off.data(),
                            bg.data(), 7));
    }
    const int saved_seg = ctx->seg_rows;
    const uint32_t saved_batch = ctx->batch_frames;
    const int saved_graphs = ctx->graphs;
    const int saved_fuse = ctx->fuse_walk;
    ctx->graphs = 0;
    ctx->tuned.erase(tune_key(h, w, g));
    const int segs[] = {64, 128, 256};
    const uint32_t batches[] = {0, 16, 32};
    // the walk inside detect or in its own kernel, where both are possible
    const int nfuse = (saved_fuse && walk_in_detect_supported(g)) ? 2 : 1;
    double best = 1e30;
    int best_seg = saved_seg, best_fuse = saved_fuse;
    uint32_t best_batch = 0;
    gt_status st = GT_OK;
    for (int fz = 0; fz < nfuse && st == GT_OK; ++fz) {
        ctx->fuse_walk = fz == 0 ? saved_fuse : 0;
        for (int sg : segs) {
            if (g.has_pct && (int64_t)K1_STRIP * sg >= 65536) continue;
            for (uint32_t bf : batches) {
                if (bf && bf >= n_frames) continue;
                ctx->seg_rows = sg;
                ctx->batch_frames = bf;
                double ms = 1e30;
                for (int rep = 0; rep < 3 && st == GT_OK; ++rep) {
                    gt_result* r = nullptr;
                    st = gt_locate_u16(ctx, movie.as<uint16_t>(), n_frames, h, w, p,
                                       GT_FRAMES_ON_DEVICE | GT_RESULT_ON_DEVICE, &r);
                    gt_result_free(r);
                    if (rep > 0) ms = std::min(ms, (double)ctx->last.device_ms);
                }
                if (st != GT_OK) break;
                if (ms < best) {
                    best = ms;
                    best_seg = sg;
                    best_batch = bf;
                    best_fuse = ctx->fuse_walk;
                }
            }
            if (st != GT_OK) break;
        }
    }
    ctx->seg_rows = saved_seg;
    ctx->batch_frames = saved_batch;
    ctx->graphs = saved_graphs;
    ctx->fuse_walk = saved_fuse;
    if (st != GT_OK) return st;
    ctx->tuned[tune_key(h, w, g)] = {best_seg, best_batch, best_fuse};
    if (out) {
        out->seg_rows = (uint32_t)best_seg;
        out->frames_per_launch = best_batch;
        out->device_ms_per_frame = best / n_frames;
        out->fused_walk = (uint32_t)(best_fuse && walk_in_detect_supported(g));
    }
    return GT_OK;
    });
}

gt_status gt_context_set_fused_walk(gt_context* ctx, int enable) {
    if (!ctx) return fail(GT_E_INVALID, "context is NULL");
    ctx->fuse_walk = enable ? 1 : 0;
    return GT_OK;
}

gt_status gt_context_set_batch(gt_context* ctx, uint32_t frames_per_launch) {
    if (!ctx) return fail(GT_E_INVALID, "context is NULL");
    ctx->batch_frames = frames_per_launch;
    return GT_OK;
}

// Frames located chunk by chunk (about 256 MiB of input each).  Host frames
// (fill given) go through the context's two pinned buffers: while chunk c is
// uploaded and located, up to 8 threads fill chunk c + 1 (fill(frame, dst)
// copies or decodes one frame).  Device frames (fill empty) are used in
// place.  With search_range each located chunk is linked on a worker thread
// while the next one is located.
static gt_status locate_staged(gt_context* ctx, const std::vector<FrameIn>& frames, int in_type,
                               const gt_params* p, uint32_t flags,
                               const std::function<gt_status(size_t, void*)>& fill,
                               gt_result** out) {
    const bool host = (bool)fill;
    const auto t_start = std::chrono::steady_clock::now();
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t es = elem_size(in_type);
    const size_t limit = (size_t)256 << 20;
    std::vector<size_t> starts(1, 0), offs(frames.size());
    size_t cur = 0, max_chunk = 0;
    for (size_t i = 0; i < frames.size(); ++i) {
        const size_t fb = al256((size_t)frames[i].h * frames[i].w * es);
        if (cur > 0 && cur + fb > limit) {
            starts.push_back(i);
            max_chunk = std::max(max_chunk, cur);
            cur = 0;
        }
        offs[i] = cur;
        cur += fb;
    }
    max_chunk = std::max(max_chunk, cur);
    starts.push_back(frames.size());
    const size_t nchunks = frames.empty() ? 0 : starts.size() - 1;
    if (host && nchunks && ctx->pinned_bytes < max_chunk) {
        for (auto& b : ctx->pinned) {
            if (b) (void)hipHostFree(b);
            b = nullptr;
        }
        ctx->pinned_bytes = 0;
        for (auto& b : ctx->pinned) {
            if (injected_failure(0)) return fail(GT_E_OOM, "hipHostMalloc: out of memory (injected)");
            HIP_TRY(hipHostMalloc(&b, max_chunk, hipHostMallocDefault));
        }
        ctx->pinned_bytes = max_chunk;
    }
    char* const buf[2] = {(char*)ctx->pinned[0], (char*)ctx->pinned[1]};
    // (error messages are per thread: the filler's travels back with it)
    using Status = std::pair<gt_status, std::string>;
    auto fill_chunk = [&](size_t c) -> Status {
        const size_t a = starts[c], m = starts[c + 1] - a;
        const size_t nt = std::min<size_t>(8, m), per = (m + nt - 1) / nt;
        std::vector<std::future<Status>> parts;
        for (size_t t = 0; t < nt; ++t) {
            const size_t f0 = a + t * per, f1 = std::min(a + m, f0 + per);
            if (f0 >= f1) break;
            parts.push_back(std::async(std::launch::async, [&, f0, f1, c]() -> Status {
                for (size_t fi = f0; fi < f1; ++fi) {
                    const gt_status r = fill(fi, buf[c & 1] + offs[fi]);
                    if (r != GT_OK) return {r, gt_last_error()};
                }
                return {GT_OK, {}};
            }));
        }
        Status st{GT_OK, {}};
        for (auto& q : parts) {
            Status r = q.get();
            if (st.first == GT_OK) st = std::move(r);
        }
        return st;
    };
    gt_params q = *p;
    q.has_search_range = 0;   // linked once over all frames below
    std::unique_ptr<gt_result> res(new gt_result());
    {
        const char* names[16];
        const char* dtypes[16];
        uint32_t nc = 0;
        GT_TRY(gt_column_names(&q, names, dtypes, 16, &nc));
        res->ncols = nc;
    }
    gt_timing total{};
    // with linking, the rows are stored with the particle column (index 4,
    // after x) in place, and the link worker fills it in chunk by chunk
    const uint32_t gap = p->has_search_range ? 1u : 0u;
    const uint32_t stride = res->ncols + gap;
    std::mutex data_mu;  // res->data grows here while the worker writes ids into it
    std::unique_ptr<LinkPipeline> lp;
    if (p->has_search_range)
        lp.reset(new LinkPipeline(ctx, p, frames.size(), &res->data, &data_mu, stride));
    if (host && nchunks) {
        const Status r0 = fill_chunk(0);
        if (r0.first != GT_OK) return fail(r0.first, "%s", r0.second.c_str());
    }
    for (size_t c = 0; c < nchunks; ++c) {
        std::future<Status> next;
        if (host && c + 1 < nchunks) next = std::async(std::launch::async, fill_chunk, c + 1);
        std::vector<FrameIn> fr(frames.begin() + starts[c], frames.begin() + starts[c + 1]);
        if (host)
            for (size_t i = 0; i < fr.size(); ++i) fr[i].ptr = buf[c & 1] + offs[starts[c] + i];
        gt_result* part = nullptr;
        const gt_status st = run_locate(ctx, fr, in_type, &q, flags, &part);
        const Status rs = next.valid() ? next.get() : Status{GT_OK, {}};
        if (st != GT_OK) return st;
        std::unique_ptr<gt_result> partp(part);
        if (rs.first != GT_OK) return fail(rs.first, "%s", rs.second.c_str());
        if (part->ncols + gap != stride) return fail(GT_E_INTERNAL, "locate: column count mismatch");
        const uint64_t r0 = res->rows;
        {
            std::lock_guard<std::mutex> lk(data_mu);
            if (c == 0 && nchunks > 1) {   // room for all chunks at the first one's density
                const size_t est = (size_t)((double)part->rows * frames.size() / fr.size() * 1.1);
                res->data.reserve(est * stride);
                res->cand.reserve(est);
            }
            if (!gap) {
                res->data.insert(res->data.end(), part->data.begin(), part->data.end());
            } else {
                const uint32_t nc = part->ncols;
                const size_t old = res->data.size();
                res->data.resize(old + part->rows * stride);
                float* dst = res->data.data() + old;
                const float* src = part->data.data();
                for (uint64_t i = 0; i < part->rows; ++i, dst += stride, src += nc) {
                    for (uint32_t k = 0; k < 4; ++k) dst[k] = src[k];
                    for (uint32_t k = 4; k < nc; ++k) dst[k + 1] = src[k];
                }
            }
        }
        res->cand.insert(res->cand.end(), part->cand.begin(), part->cand.end());
        res->rows += part->rows;
        if (lp) GT_TRY(lp->submit(fr, part->data.data(), part->rows, part->ncols, r0));
        total.device_ms += ctx->last.device_ms;
        total.frames += ctx->last.frames;
        total.pixels += ctx->last.pixels;
        total.candidates += ctx->last.candidates;
        total.rows += ctx->last.rows;
        total.launches += ctx->last.launches;
    }
    if (lp) GT_TRY(lp->finish());
    res->ncols = stride;
    {
        const char* names[16];
        const char* dtypes[16];
        uint32_t nc = 0;
        GT_TRY(gt_column_names(p, names, dtypes, 16, &nc));
        res->names.assign(names, names + nc);
        res->dtypes.assign(dtypes, dtypes + nc);
    }
    total.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    ctx->last = total;
    *out = res.release();
    return GT_OK;
}

// Host frames go through the pinned staging unless the call asks for the
// per-frame path or device-resident rows; device frames are chunked only to
// link beside locate.
static bool staged_path(uint32_t flags) {
    return !(flags & (GT_FRAMES_ON_DEVICE | GT_PLAIN | GT_RESULT_ON_DEVICE));
}
static bool chunked_link_path(uint32_t flags, const gt_params* p) {
    return p && p->has_search_range && (flags & GT_FRAMES_ON_DEVICE) &&
           !(flags & (GT_PLAIN | GT_RESULT_ON_DEVICE));
}

static gt_status locate_uniform(gt_context* ctx, const void* frames, uint64_t n, uint32_t h,
                                uint32_t w, int in_type, const gt_params* p, uint32_t flags,
                                gt_result** out) {
    if (!frames && n) return fail(GT_E_INVALID, "frames is NULL");
    std::vector<FrameIn> v;
    v.reserve(n);
    const size_t fb = (size_t)h * w * elem_size(in_type);
    for (uint64_t i = 0; i < n; ++i) v.push_back({(const char*)frames + i * fb, h, w, (int64_t)i});
    if (n == 0) {
        GT_TRY(check_params(p, h, w));
    }
    if (staged_path(flags) && n > 0 && ctx && out) {
        GT_TRY(check_params(p, h, w));
        return locate_staged(ctx, v, in_type, p, flags, [&](size_t fi, void* dst) {
            std::memcpy(dst, v[fi].ptr, fb);
            return GT_OK;
        }, out);
    }
    if (chunked_link_path(flags, p) && n > 0 && ctx && out) {
        GT_TRY(check_params(p, h, w));
        return locate_staged(ctx, v, in_type, p, flags, {}, out);
    }
    return run_locate(ctx, v, in_type, p, flags, out);
}

// gt_locate_noalloc: the launch sequence of run_locate's batched path on one
// stream, with every workspace carved out of the device's `scratch` array
// and the rows written into the caller's device buffers.  Nothing is
// allocated: the host frame tables live in the context (sized at creation),
// the parameter block and the counters in the scratch.
static gt_status locate_noalloc(gt_context* ctx, const void* frames, int in_type, uint64_t n,
                                uint32_t h, uint32_t w, const gt_params* p, float* d_rows,
                                int64_t* d_cand, uint64_t rows_cap, uint64_t* n_rows) {
    if (!ctx) return fail(GT_E_INVALID, "context is NULL");
    if (!n_rows) return fail(GT_E_INVALID, "n_rows is NULL");
    *n_rows = 0;
    if (n && !frames) return fail(GT_E_INVALID, "frames is NULL");
    if (rows_cap && (!d_rows || !d_cand)) return fail(GT_E_INVALID, "rows or cand is NULL");
    GT_TRY(check_params(p, h, w));
    if (p->has_search_range)
        return fail(GT_E_NOT_IMPLEMENTED, "gt_locate_noalloc does not link (search_range)");
    if (!ctx->scratch)
        return fail(GT_E_HIP, "the no-allocation scratch has no device address on device %d",
                    ctx->device);
    HIP_TRY(hipSetDevice(ctx->device));
    const bool integer_input = in_type != IN_F32;
    Geom g;
    GT_TRY(make_geom(p, integer_input, g));
    g.walk_in_detect = ctx->fuse_walk && walk_in_detect_supported(g) ? 1 : 0;
    const size_t es = elem_size(in_type);
    const bool vec = w % 4 == 0 && ((uintptr_t)frames) % (4 * es) == 0;
    bool generic;
    {
        Batch probe;
        std::memset(&probe, 0, sizeof(probe));
        probe.in_type = in_type;
        probe.vec = vec ? 1 : 0;
        generic = !bandpass_fast_path(probe, g);
    }
    const bool tmp_planes = generic || detect_needs_plane(g);
    const int seg = ctx->seg_set ? ctx->seg_rows : 128;
    const FrameSizes fs = sizes_of(h, w, seg);
    const uint64_t dblocks_pf = (uint64_t)((fs.ty + 1) / 2) * ((fs.tx + 1) / 2);  // dedup blocks

    // scratch layout for nb frames, every part 256-byte aligned
    struct Layout {
        size_t geom, counters, ftab, bases, P, tmp, list, slots, rec, keep, tile3, perframe, end;
        uint64_t tiles, k1, nbases;
    };
    auto layout = [&](uint64_t nb) {
        Layout L{};
        size_t o = 0;
        auto take = [&](size_t bytes) { const size_t at = o; o += al256(bytes); return at; };
        L.tiles = nb * fs.tiles;
        L.k1 = nb * fs.k1_items;
        L.nbases = 3 * (nb + 1) + L.k1 + L.tiles + nb * dblocks_pf;
        const uint64_t slots = L.tiles * TILE_PX;
        L.geom = take(sizeof(Geom));
        L.counters = take(16);
        L.ftab = take(nb * sizeof(Frame));
        L.bases = take(L.nbases * 4);
        L.P = take(nb * fs.px * 4);
        L.tmp = tmp_planes ? take(nb * fs.px * 8) : 0;
        L.list = g.has_pct ? take(slots * 4) : 0;
        L.slots = take(slots * 2);
        L.rec = take(slots * 4 * (uint64_t)rec_planes(g));
        L.keep = take(slots);
        L.tile3 = take(al256(L.tiles * 4) * 4 + 2 * al256(L.tiles * 4 * TILE_H));
        L.perframe = take(al256(nb * HIST_BINS * 4) + 3 * al256(nb * 4) + al256(nb * 16) +
                          2 * al256(nb * 8));
        L.end = o;
        return L;
    };
    auto fits = [&](uint64_t nb) {
        const Layout L = layout(nb);
        return L.end <= (size_t)GT_NOALLOC_SCRATCH_BYTES && nb <= gt_context::kNaFrames &&
               3 * (nb + 1) <= gt_context::kNaBases;
    };
    if (n == 0) return GT_OK;
    if (!fits(1))
        return fail(GT_E_INVALID,
                    "a %ux%u frame needs %zu bytes of scratch; gt_locate_noalloc holds %u",
                    h, w, layout(1).end, (unsigned)GT_NOALLOC_SCRATCH_BYTES);
    uint64_t nbmax = 1;
    while (nbmax < n && fits(nbmax * 2)) nbmax *= 2;
    {
        uint64_t lo = nbmax, hi = std::min<uint64_t>(n, nbmax * 2);   // fits(lo), maybe not hi
        while (lo < hi) {
            const uint64_t mid = (lo + hi + 1) / 2;
            if (fits(mid)) lo = mid; else hi = mid - 1;
        }
        nbmax = lo;
    }
    const Layout L = layout(nbmax);
    char* const base = (char*)ctx->scratch;
    Geom* d_geom = (Geom*)(base + L.geom);
    uint64_t* counters = (uint64_t*)(base + L.counters);
    hipStream_t st = ctx->streams[0];
    HIP_TRY(hipMemcpyAsync(d_geom, &g, sizeof(Geom), hipMemcpyHostToDevice, st));
    HIP_TRY(launch_fill_u64(counters, 2, 0, st));
    Frame* ftab = ctx->na_ftab.data();
    uint32_t* bases = ctx->na_bases.data();
    for (uint64_t b0 = 0; b0 < n; b0 += nbmax) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(nbmax, n - b0);
        if (b0 > 0) HIP_TRY(hipEventSynchronize(ctx->ev_na));   // host tables copied out
        // frame tables: k1 bases [nb + 1], tile bases [nb + 1] and dedup-block
        // bases [nb + 1] from the host; the per-item tables after them are
        // filled on the device (k_fill_tables)
        uint32_t* k1b = bases;
        uint32_t* tb = bases + nb + 1;
        uint32_t* dbb = tb + nb + 1;
        for (uint32_t fi = 0; fi < nb; ++fi) {
            Frame& F = ftab[fi];
            std::memset(&F, 0, sizeof(F));
            F.in = (const char*)frames + (b0 + fi) * fs.px * es;
            F.P = (float*)(base + L.P) + (uint64_t)fi * fs.px;
            F.tmp = tmp_planes ? (float*)(base + L.tmp) + 2 * (uint64_t)fi * fs.px : nullptr;
            F.H = h;
            F.W = w;
            F.k1_tx = fs.k1_tx;
            F.k1_ty = fs.k1_ty;
            F.tx = fs.tx;
            F.ty = fs.ty;
            F.tile_base = fi * fs.tiles;
            F.k1_base = fi * fs.k1_items;
            F.slot_base = (uint64_t)F.tile_base * TILE_PX;
            F.frame_index = (int64_t)(b0 + fi);
            k1b[fi] = F.k1_base;
            tb[fi] = F.tile_base;
            dbb[fi] = (uint32_t)(fi * dblocks_pf);
        }
        k1b[nb] = nb * fs.k1_items;
        tb[nb] = nb * fs.tiles;
        dbb[nb] = (uint32_t)(nb * dblocks_pf);
        Frame* d_ftab = (Frame*)(base + L.ftab);
        uint32_t* d_bases = (uint32_t*)(base + L.bases);
        HIP_TRY(hipMemcpyAsync(d_ftab, ftab, nb * sizeof(Frame), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_bases, bases, 3 * ((size_t)nb + 1) * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipEventRecord(ctx->ev_na, st));

        Batch b;
        std::memset(&b, 0, sizeof(b));
        b.frames = d_ftab;
        b.k1_bases = d_bases;
        b.tile_bases = d_bases + nb + 1;
        b.k1_frame = d_bases + 3 * ((size_t)nb + 1);
        b.tile_frame = b.k1_frame + (uint64_t)nb * fs.k1_items;
        b.dblock_tile = b.tile_frame + (uint64_t)nb * fs.tiles;
        b.dblocks = (uint32_t)(nb * dblocks_pf);
        b.nframes = nb;
        b.k1_items = nb * fs.k1_items;
        b.tiles = nb * fs.tiles;
        b.in_type = in_type;
        b.vec = vec ? 1 : 0;
        char* t3 = base + L.tile3;
        const size_t t3s = al256(L.tiles * 4);
        b.tile_count = (uint32_t*)t3;
        b.tile_keep = (uint32_t*)(t3 + t3s);
        b.tile_row_off = (uint32_t*)(t3 + 2 * t3s);
        b.tile_lcount = (uint32_t*)(t3 + 3 * t3s);
        b.rowk = (uint32_t*)(t3 + 4 * t3s);
        b.rowoff = (uint32_t*)(t3 + 4 * t3s + al256(L.tiles * 4 * TILE_H));
        char* pf = base + L.perframe;
        b.hist1 = g.has_pct ? (uint32_t*)pf : nullptr;
        pf += al256(nbmax * HIST_BINS * 4);
        b.list = g.has_pct ? (uint32_t*)(base + L.list) : nullptr;
        b.frame_rows = (uint32_t*)pf; pf += al256(nbmax * 4);
        b.frame_cands = (uint32_t*)pf; pf += al256(nbmax * 4);
        pf += al256(nbmax * 4);
        b.selstate = (int32_t*)pf; pf += al256(nbmax * 16);
        b.thr = (double*)pf; pf += al256(nbmax * 8);
        b.frame_row_base = (uint64_t*)pf;
        b.slots = (uint16_t*)(base + L.slots);
        b.rec = (float*)(base + L.rec);
        b.keep = (uint8_t*)(base + L.keep);
        b.slot_cap = L.tiles * TILE_PX;
        b.cand_total = counters;
        b.row_carry = counters + 1;
        b.rows = d_rows;
        b.rows_cand = d_cand;
        b.rows_cap = rows_cap;

        HIP_TRY(launch_fill_tables(b, d_bases + 2 * ((size_t)nb + 1),
                                   (uint32_t)(fs.k1_items + fs.tiles + dblocks_pf), st));
        HIP_TRY(launch_zero_batch(b, st));
        if (generic)
            HIP_TRY(launch_bandpass_generic(b, d_geom, h * w, g.has_pct, st));
        else
            HIP_TRY(launch_bandpass(b, g, d_geom, seg, st));
        if (g.has_pct) HIP_TRY(launch_select1(b, g, st));
        HIP_TRY(launch_detect(b, g, d_geom, st));
        if (g.has_pct) HIP_TRY(launch_select2(b, g, st));
        HIP_TRY(launch_refine(b, g, d_geom, st));
        HIP_TRY(launch_dedup(b, g, d_geom, st));
        HIP_TRY(launch_tile_scan(b, fs.ty, st));
        HIP_TRY(launch_frame_scan(b, st));
        HIP_TRY(launch_emit(b, g, d_geom, st));
        if (g.raw_cols) HIP_TRY(launch_raw(b, g, d_geom, st));
    }
    uint64_t cnt[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(cnt, counters, 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *n_rows = cnt[1];
    ctx->last = gt_timing{};
    ctx->last.frames = n;
    ctx->last.pixels = n * fs.px;
    ctx->last.candidates = cnt[0];
    ctx->last.rows = cnt[1];
    if (cnt[1] > rows_cap)
        return fail(GT_E_INVALID,
                    "%llu rows found but rows_capacity is %llu; only the first %llu were written",
                    (unsigned long long)cnt[1], (unsigned long long)rows_cap,
                    (unsigned long long)rows_cap);
    return GT_OK;
}

gt_status gt_locate_u16(gt_context* ctx, const uint16_t* frames, uint64_t n_frames, uint32_t h,
                        uint32_t w, const gt_params* p, uint32_t flags, gt_result** out) {
    return guarded([&]() -> gt_status {
    return locate_uniform(ctx, frames, n_frames, h, w, IN_U16, p, flags, out);
    });
}

gt_status gt_locate_u8(gt_context* ctx, const uint8_t* frames, uint64_t n_frames, uint32_t h,
                       uint32_t w, const gt_params* p, uint32_t flags, gt_result** out) {
    return guarded([&]() -> gt_status {
    return locate_uniform(ctx, frames, n_frames, h, w, IN_U8, p, flags, out);
    });
}

gt_status gt_locate_f32(gt_context* ctx, const float* frames, uint64_t n_frames, uint32_t h,
                        uint32_t w, const gt_params* p, uint32_t flags, gt_result** out) {
    return guarded([&]() -> gt_status {
    return locate_uniform(ctx, frames, n_frames, h, w, IN_F32, p, flags, out);
    });
}

gt_status gt_locate_noalloc(gt_context* ctx, const void* frames, int dtype, uint64_t n_frames,
                            uint32_t h, uint32_t w, const gt_params* p, float* rows,
                            int64_t* cand, uint64_t rows_capacity, uint64_t* n_rows) {
    return guarded([&]() -> gt_status {
    int in_type;
    if (dtype == 0) in_type = IN_U16;
    else if (dtype == 1) in_type = IN_F32;
    else if (dtype == 2) in_type = IN_U8;
    else return fail(GT_E_INVALID, "dtype %d (0 u16, 1 f32, 2 u8)", dtype);
    return locate_noalloc(ctx, frames, in_type, n_frames, h, w, p, rows, cand, rows_capacity,
                          n_rows);
    });
}

gt_status gt_locate_batch(gt_context* ctx, const gt_frame_desc* frames, uint64_t n_frames,
                          int dtype, const gt_params* p, uint32_t flags, gt_result** out) {
    return guarded([&]() -> gt_status {
    if (!frames && n_frames) return fail(GT_E_INVALID, "frames is NULL");
    int in_type;
    if (dtype == 0) in_type = IN_U16;
    else if (dtype == 1) in_type = IN_F32;
    else if (dtype == 2) in_type = IN_U8;
    else return fail(GT_E_INVALID, "dtype must be 0 (u16), 1 (f32) or 2 (u8)");
    std::vector<FrameIn> v;
    v.reserve(n_frames);
    for (uint64_t i = 0; i < n_frames; ++i)
        v.push_back({frames[i].data, frames[i].h, frames[i].w, frames[i].frame_index});
    if ((staged_path(flags) || chunked_link_path(flags, p)) && n_frames > 0 && ctx && out) {
        for (const auto& f : v) {
            GT_TRY(check_params(p, f.h, f.w));
            if (!f.ptr) return fail(GT_E_INVALID, "frame pointer is NULL");
        }
        if (p->has_search_range)
            for (size_t f = 1; f < v.size(); ++f)
                if ((float)v[f].fi == (float)v[f - 1].fi)
                    return fail(GT_E_INVALID, "search_range needs consecutive frames with distinct indices");
        const size_t es = elem_size(in_type);
        if (!staged_path(flags)) return locate_staged(ctx, v, in_type, p, flags, {}, out);
        return locate_staged(ctx, v, in_type, p, flags, [&](size_t fi, void* dst) {
            std::memcpy(dst, v[fi].ptr, (size_t)v[fi].h * v[fi].w * es);
            return GT_OK;
        }, out);
    }
    return run_locate(ctx, v, in_type, p, flags, out);
    });
}

gt_status gt_locate_file(gt_context* ctx, const char* path, uint32_t channel, const gt_params* p,
                         uint32_t flags, gt_result** out) {
    return guarded([&]() -> gt_status {
    if (!out) return fail(GT_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!ctx) return fail(GT_E_INVALID, "context is NULL");
    if (flags & ~(uint32_t)GT_PLAIN)
        return fail(GT_E_INVALID, "gt_locate_file takes flags 0 or GT_PLAIN");
    GT_TRY(check_params(p, 0, 0));
    gt_file* fp = nullptr;
    GT_TRY(gt_file_open(path, channel, &fp));
    std::unique_ptr<gt_file, void (*)(gt_file*)> file(fp, gt_file_close);
    int kind = 0, dtype = 0;
    uint64_t n = 0;
    uint32_t h = 0, w = 0;
    GT_TRY(gt_file_info(fp, &kind, &n, &h, &w, &dtype));
    // execute_file (execute_gpu.rs:525-531) hands ETS frames on with the
    // ETS header's (dims[0], dims[1]) as (rows, columns)
    if (kind == 1) std::swap(h, w);
    GT_TRY(check_params(p, h, w));
    std::vector<FrameIn> all;
    all.reserve(n);
    for (uint64_t i = 0; i < n; ++i) all.push_back({nullptr, h, w, (int64_t)i});
    return locate_staged(ctx, all, dtype == 0 ? IN_U16 : IN_U8, p, flags,
                         [fp](size_t fi, void* dst) { return gt_file_read(fp, fi, 1, dst); }, out);
    });
}

gt_status gt_locate_adjoint(gt_context* ctx, const void* frames, int dtype, uint64_t n_frames,
                            uint32_t h, uint32_t w, const gt_params* p, uint32_t flags,
                            gt_adjoint** out) {
    return guarded([&]() -> gt_status {
    if (!out) return fail(GT_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!ctx) return fail(GT_E_INVALID, "context is NULL");
    if (!frames && n_frames) return fail(GT_E_INVALID, "frames is NULL");
    if (dtype < 0 || dtype > 2) return fail(GT_E_INVALID, "dtype must be 0 (u16), 1 (f32) or 2 (u8)");
    if (flags & GT_PLAIN) return fail(GT_E_INVALID, "the adjoint runs on the batched path");
    const int in_type = dtype == 0 ? IN_U16 : (dtype == 1 ? IN_F32 : IN_U8);
    GT_TRY(check_params(p, h, w));
    std::vector<FrameIn> v;
    v.reserve(n_frames);
    const size_t fb = (size_t)h * w * elem_size(in_type);
    for (uint64_t i = 0; i < n_frames; ++i)
        v.push_back({(const char*)frames + i * fb, h, w, (int64_t)i});
    AdjReq req;
    req.jacobian = (flags & GT_ADJ_JACOBIAN) != 0;
    gt_result* res = nullptr;
    GT_TRY(run_locate(ctx, v, in_type, p, flags & ~(uint32_t)(GT_RESULT_ON_DEVICE | GT_ADJ_JACOBIAN),
                      &res, &req));
    gt_adjoint* a = new gt_adjoint();
    a->fwd = res;
    a->device = ctx->device;
    {
        Geom g;
        make_geom(p, in_type != IN_F32, g);
        a->S = (uint32_t)adjoint_patch(g);
    }
    a->nframes = n_frames;
    a->h = h;
    a->w = w;
    a->jacobian = req.jacobian;
    const uint64_t n = res->rows;
    a->sigma.resize(n * 3);
    a->origin.resize(n * 2);
    a->rowpos.resize(n);
    auto cleanup = [&](hipError_t e) {
        delete a;
        return fail(GT_E_HIP, "adjoint readback: %s", hipGetErrorString(e));
    };
    hipError_t e = hipSuccess;
    if (n) {
        e = hipMemcpy(a->sigma.data(), req.sigma.p, n * 12, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(a->origin.data(), req.origin.p, n * 8, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(a->rowpos.data(), req.rowpos.p, n * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return cleanup(e);
    }
    if (req.jacobian) {
        // keep the patches and origins on the device; row range of every frame
        std::swap(a->d_jac.p, req.jac.p);
        std::swap(a->d_jac.bytes, req.jac.bytes);
        std::swap(a->d_origin.p, req.origin.p);
        std::swap(a->d_origin.bytes, req.origin.bytes);
        std::vector<uint64_t> fr(n_frames + 1, 0);
        for (uint64_t r = 0; r < n; ++r) fr[(size_t)a->rowpos[r] + 1] += 1;
        for (uint64_t f = 0; f < n_frames; ++f) fr[f + 1] += fr[f];
        e = a->d_frame_rows.ensure(fr.size() * 8);
        if (e == hipSuccess) e = hipMemcpy(a->d_frame_rows.p, fr.data(), fr.size() * 8, hipMemcpyHostToDevice);
        if (e != hipSuccess) return cleanup(e);
    }
    *out = a;
    return GT_OK;
    });
}

uint64_t gt_adjoint_rows(const gt_adjoint* a) { return a ? a->fwd->rows : 0; }
uint32_t gt_adjoint_ncols(const gt_adjoint* a) { return a ? a->fwd->ncols : 0; }
uint32_t gt_adjoint_patch(const gt_adjoint* a) { return a ? a->S : 0; }
const float* gt_adjoint_data(const gt_adjoint* a) {
    return a && !a->fwd->data.empty() ? a->fwd->data.data() : nullptr;
}
const int64_t* gt_adjoint_candidates(const gt_adjoint* a) {
    return a && !a->fwd->cand.empty() ? a->fwd->cand.data() : nullptr;
}
const float* gt_adjoint_sigma(const gt_adjoint* a) {
    return a && !a->sigma.empty() ? a->sigma.data() : nullptr;
}
const int32_t* gt_adjoint_origin(const gt_adjoint* a) {
    return a && !a->origin.empty() ? a->origin.data() : nullptr;
}

gt_status gt_adjoint_jacobian(const gt_adjoint* a, float* out, uint64_t capacity_floats) {
    if (!a || !out) return fail(GT_E_INVALID, "NULL argument");
    if (!a->jacobian) return fail(GT_E_INVALID, "the adjoint was computed without GT_ADJ_JACOBIAN");
    const uint64_t n = a->fwd->rows * 3ull * a->S * a->S;
    if (capacity_floats < n) return fail(GT_E_INVALID, "capacity too small");
    if (n) HIP_TRY(hipMemcpy(out, a->d_jac.p, n * 4, hipMemcpyDeviceToHost));
    return GT_OK;
}

gt_status gt_adjoint_vjp(gt_context* ctx, const gt_adjoint* a, const float* grad_rows,
                         float* grad_frames, uint32_t flags) {
    return guarded([&]() -> gt_status {
    if (!ctx || !a) return fail(GT_E_INVALID, "NULL argument");
    if (!a->jacobian) return fail(GT_E_INVALID, "the adjoint was computed without GT_ADJ_JACOBIAN");
    if (!grad_frames && a->nframes) return fail(GT_E_INVALID, "grad_frames is NULL");
    const uint64_t n = a->fwd->rows;
    if (!grad_rows && n) return fail(GT_E_INVALID, "grad_rows is NULL");
    HIP_TRY(hipSetDevice(ctx->device));
    const bool on_device = flags & GT_FRAMES_ON_DEVICE;
    const uint64_t gbytes = a->nframes * (uint64_t)a->h * a->w * 4;
    struct ScopedBuf : DevBuf {
        ~ScopedBuf() { release(); }
    } dg, dgrad;
    const float* grows = grad_rows;
    float* gout = grad_frames;
    if (!on_device) {
        HIP_TRY(dg.ensure(n * 12));
        HIP_TRY(dgrad.ensure(gbytes));
        if (n) HIP_TRY(hipMemcpy(dg.p, grad_rows, n * 12, hipMemcpyHostToDevice));
        grows = dg.as<float>();
        gout = dgrad.as<float>();
    }
    hipStream_t s = ctx->streams[0];
    hipError_t e = launch_adjoint_vjp(a->d_jac.as<float>(), a->d_origin.as<int32_t>(),
                                      a->d_frame_rows.as<uint64_t>(), (uint32_t)a->nframes, grows,
                                      gout, (int)a->S, (int)a->h, (int)a->w, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && !on_device && gbytes)
        e = hipMemcpy(grad_frames, gout, gbytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(GT_E_HIP, "gt_adjoint_vjp: %s", hipGetErrorString(e));
    return GT_OK;
    });
}

void gt_adjoint_free(gt_adjoint* a) { delete a; }

uint64_t gt_result_rows(const gt_result* r) { return r ? r->rows : 0; }
uint32_t gt_result_ncols(const gt_result* r) { return r ? r->ncols : 0; }
const float* gt_result_data(const gt_result* r) {
    return (r && !r->ddata) ? r->data.data() : nullptr;
}
const int64_t* gt_result_candidates(const gt_result* r) {
    return (r && !r->ddata) ? r->cand.data() : nullptr;
}
const float* gt_result_device_data(const gt_result* r) { return r ? r->ddata : nullptr; }

gt_status gt_result_columns(const gt_result* r, const char** names, const char** dtypes,
                            uint32_t cap, uint32_t* ncols) {
    if (!r || !ncols) return fail(GT_E_INVALID, "NULL argument");
    *ncols = (uint32_t)r->names.size();
    for (uint32_t i = 0; i < *ncols && i < cap; ++i) {
        if (names) names[i] = r->names[i].c_str();
        if (dtypes) dtypes[i] = r->dtypes[i].c_str();
    }
    return *ncols > cap ? fail(GT_E_INVALID, "capacity too small") : GT_OK;
}

void gt_result_free(gt_result* r) { delete r; }

gt_status gt_synth_u16(gt_context* ctx, uint16_t* d_frames, uint64_t n_frames, uint32_t h,
                       uint32_t w, const float* spots, const uint64_t* spot_off, const float* bg,
                       uint64_t seed) {
    return guarded([&]() -> gt_status {
    if (!ctx || !d_frames || !spot_off || !bg) return fail(GT_E_INVALID, "NULL argument");
    if (h == 0 || w == 0) return fail(GT_E_INVALID, "empty frame");
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t cell = 32;
    const uint32_t cx = (w + cell - 1) / cell, cy = (h + cell - 1) / cell;
    const uint32_t ncell = cx * cy;
    const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(64, (1ull << 28) / ((uint64_t)h * w)));
    DevBuf d_spots, d_off, d_bg;
    for (uint64_t f0 = 0; f0 < n_frames; f0 += chunk) {
        const uint64_t nf = std::min<uint64_t>(chunk, n_frames - f0);
        // counting sort of each frame's spots into cells
        std::vector<uint32_t> off(nf * (ncell + 1), 0);
        const uint64_t s_begin = spot_off[f0], s_end = spot_off[f0 + nf];
        std::vector<float4> sorted(s_end - s_begin);
        for (uint64_t f = 0; f < nf; ++f) {
            uint32_t* co = off.data() + f * (ncell + 1);
            std::vector<uint32_t> cnt(ncell, 0);
            for (uint64_t s = spot_off[f0 + f]; s < spot_off[f0 + f + 1]; ++s) {
                const float y = spots[4 * s], x = spots[4 * s + 1];
                int cyi = (int)std::floor(y / cell), cxi = (int)std::floor(x / cell);
                cyi = std::max(0, std::min((int)cy - 1, cyi));
                cxi = std::max(0, std::min((int)cx - 1, cxi));
                cnt[(uint32_t)cyi * cx + (uint32_t)cxi]++;
            }
            uint32_t run = (uint32_t)(spot_off[f0 + f] - s_begin);
            for (uint32_t c = 0; c < ncell; ++c) {
                co[c] = run;
                run += cnt[c];
            }
            co[ncell] = run;
            std::vector<uint32_t> fill(co, co + ncell);
            for (uint64_t s = spot_off[f0 + f]; s < spot_off[f0 + f + 1]; ++s) {
                const float y = spots[4 * s], x = spots[4 * s + 1];
                int cyi = (int)std::floor(y / cell), cxi = (int)std::floor(x / cell);
                cyi = std::max(0, std::min((int)cy - 1, cyi));
                cxi = std::max(0, std::min((int)cx - 1, cxi));
                const uint32_t c = (uint32_t)cyi * cx + (uint32_t)cxi;
                sorted[fill[c]++] = make_float4(y, x, spots[4 * s + 2], spots[4 * s + 3]);
            }
        }
        HIP_TRY(d_spots.ensure(std::max<size_t>(16, sorted.size() * sizeof(float4))));
        HIP_TRY(d_off.ensure(off.size() * 4));
        HIP_TRY(d_bg.ensure(nf * 4));
        if (!sorted.empty())
            HIP_TRY(hipMemcpy(d_spots.p, sorted.data(), sorted.size() * sizeof(float4), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_off.p, off.data(), off.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_bg.p, bg + f0, nf * 4, hipMemcpyHostToDevice));
        HIP_TRY(launch_synth(d_frames + f0 * (uint64_t)h * w, nf, h, w, d_spots.as<const float4>(),
                             d_off.as<const uint32_t>(), cx, cy, cell, d_bg.as<const float>(),
                             seed, f0, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    d_spots.release();
    d_off.release();
    d_bg.release();
    return GT_OK;
    });
}

gt_status gt_device_alloc(gt_context* ctx, uint64_t bytes, void** out) {
    if (!ctx || !out) return fail(GT_E_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMalloc(out, bytes ? bytes : 1));
    return GT_OK;
}

gt_status gt_device_free(gt_context* ctx, void* p) {
    if (!ctx) return fail(GT_E_INVALID, "NULL argument");
    if (p) HIP_TRY(hipFree(p));
    return GT_OK;
}

gt_status gt_memcpy_h2d(gt_context* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx) return fail(GT_E_INVALID, "NULL argument");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return GT_OK;
}

gt_status gt_memcpy_d2h(gt_context* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx) return fail(GT_E_INVALID, "NULL argument");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GT_OK;
}

gt_status gt_link(gt_context* ctx, const float* rows, uint64_t n_rows, uint32_t ncols,
                  float search_range, uint64_t memory, uint64_t* ids) {
    return guarded([&]() -> gt_status {
    if (!ctx) return fail(GT_E_INVALID, "context is NULL");
    if (n_rows && (!rows || !ids)) return fail(GT_E_INVALID, "NULL argument");
    if (ncols < 4) return fail(GT_E_INVALID, "rows need at least 4 columns (frame, mass, y, x)");
    HIP_TRY(hipSetDevice(ctx->device));
    std::string err;
    if (!ctx->link_cache) ctx->link_cache = link_cache_create();
    const hipError_t e = link_rows(ctx->streams[0], ctx->link_cache, rows, n_rows, ncols, search_range,
                                   (uint32_t)std::min<uint64_t>(memory, UINT32_MAX), ids,
                                   &ctx->last_link, err);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? GT_E_OOM
                    : e == hipErrorInvalidValue ? GT_E_INVALID : GT_E_HIP, "%s", err.c_str());
    return GT_OK;
    });
}

gt_status gt_context_link_stats(const gt_context* ctx, gt_link_stats* out) {
    if (!ctx || !out) return fail(GT_E_INVALID, "NULL argument");
    const LinkStats& l = ctx->last_link;
    out->steps = l.steps;
    out->pairs = l.pairs;
    out->subnets = l.subnets;
    out->max_subnet = l.max_subnet;
    out->ids = l.ids;
    out->search_ms = l.search_ms;
    out->solve_ms = l.solve_ms;
    return GT_OK;
}

gt_status gt_memcpy_d2d(gt_context* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(GT_E_INVALID, "NULL argument");
    if ((((uintptr_t)dst | (uintptr_t)src) & 15) != 0)
        return fail(GT_E_INVALID, "gt_memcpy_d2d: pointers must be 16-byte aligned");
    if (bytes == 0) return GT_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(launch_copy(dst, src, bytes, ctx->streams[0]));
    HIP_TRY(hipStreamSynchronize(ctx->streams[0]));
    return GT_OK;
}

gt_status gt_synchronize(gt_context* ctx) {
    if (!ctx) return fail(GT_E_INVALID, "NULL argument");
    for (int i = 0; i < kStreams; ++i) HIP_TRY(hipStreamSynchronize(ctx->streams[i]));
    return GT_OK;
}

}  // extern "C"
