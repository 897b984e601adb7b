// gt_link.hip -- frame-to-frame linker (gpu-tracking/src/linking.rs) with
// the candidate search on the GPU.
//
// The caller hands over the positions of all steps at once (a step is one
// call of Linker::advance: a FrameSubsetter run of rows for link_rust, one
// frame for locate with search_range).  The particles of every step are
// sorted by y on the device (rocPRIM segmented radix sort on
// order-preserving keys); one thread per (particle p of step s, target step
// s + 1 + j, j <= memory) binary-searches the target step's y band and
// applies within_radius_rd's rule (linking.rs:73-105): the per-axis box
// [q - r, q + r] with binary32 bounds, then (0 + dy*dy) + dx*dx < r*r in
// binary32, dy = dest - src.  Two passes (count, exclusive scan, write) give
// every (p, j) its exact candidate list, so a memory entry of any age finds
// its candidates in the same table.
//
// The subnet solver (gt_link_solve.h) then runs on the host in the order of
// Linker::advance / link (linking.rs:234-492).  A source's candidates are
// taken in ascending destination order (each thread sorts its own list);
// the reference takes its kd-tree's traversal order, which only decides
// between assignments of equal binary32 cost.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "gt_internal.h"

#pragma clang fp contract(off)

namespace gt {

namespace {

// Order-preserving key of a float (+0 and -0 share a key).
__device__ __forceinline__ uint32_t ykey(float y) {
    const uint32_t b = __float_as_uint(y + 0.0f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// step_of[i] = the step row i belongs to: one workgroup per step fills its
// rows (steps may be empty).
__global__ void k_link_stepof(const uint32_t* __restrict__ step_off, uint32_t* __restrict__ step_of) {
    const uint32_t s = blockIdx.x;
    const uint32_t a = step_off[s], e = step_off[s + 1];
    for (uint32_t i = a + threadIdx.x; i < e; i += blockDim.x) step_of[i] = s;
}

__global__ void k_link_keys(const float* __restrict__ ys, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ vals, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        keys[i] = ykey(ys[i]);
        vals[i] = i;
    }
}

// One thread per (source particle p, age j).  Without `off` it counts the
// candidates into counts[t]; with `off` it writes them from off[t] on, as
// (destination index within the target step, squared distance), in
// ascending destination order (the search finds them in y order; the thread
// sorts its own short list).
__global__ void k_link_pairs(const float* __restrict__ ys, const float* __restrict__ xs,
                             const uint32_t* __restrict__ step_of,
                             const uint32_t* __restrict__ step_off, uint32_t nsteps,
                             const uint32_t* __restrict__ skeys,
                             const uint32_t* __restrict__ sidx, uint32_t nm, float r,
                             uint32_t n, uint32_t* __restrict__ counts,
                             const uint32_t* __restrict__ off, uint32_t* __restrict__ out_d,
                             float* __restrict__ out_dist) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)n * nm) return;
    const uint32_t p = (uint32_t)(t / nm), j = (uint32_t)(t % nm);
    const uint32_t ts = step_of[p] + 1 + j;
    uint32_t c = 0;
    if (ts < nsteps) {
        const float sy = ys[p], sx = xs[p];
        const float lo0 = sy - r, hi0 = sy + r, lo1 = sx - r, hi1 = sx + r;
        const float r2 = r * r;
        const uint32_t klo = ykey(lo0), khi = ykey(hi0);
        const uint32_t a = step_off[ts], e = step_off[ts + 1];
        uint32_t lo = a, hi = e;   // first key >= klo
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (skeys[mid] < klo) lo = mid + 1;
            else hi = mid;
        }
        const uint32_t base = off ? off[t] : 0;
        for (uint32_t k = lo; k < e && skeys[k] <= khi; ++k) {
            const uint32_t d = sidx[k];
            const float dy = ys[d], dx = xs[d];
            if (dy < lo0 || dy > hi0 || dx < lo1 || dx > hi1) continue;
            const float a0 = dy - sy, a1 = dx - sx;
            float dist = 0.0f + a0 * a0;
            dist = dist + a1 * a1;
            if (dist < r2) {
                if (off) {
                    out_d[base + c] = d - a;
                    out_dist[base + c] = dist;
                }
                ++c;
            }
        }
        if (off) {
            for (uint32_t q = base + 1; q < base + c; ++q) {  // insertion sort by destination
                const uint32_t dq = out_d[q];
                const float vq = out_dist[q];
                uint32_t k = q;
                for (; k > base && out_d[k - 1] > dq; --k) {
                    out_d[k] = out_d[k - 1];
                    out_dist[k] = out_dist[k - 1];
                }
                out_d[k] = dq;
                out_dist[k] = vq;
            }
        }
    }
    if (!off) counts[t] = c;
}


template <bool PINNED>
struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        n = std::max<size_t>(n, 256);
        if (p && n <= bytes) return hipSuccess;
        release();
        n = (n + 4095) & ~(size_t)4095;
        const hipError_t e = PINNED ? hipHostMalloc(&p, n, hipHostMallocDefault) : hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        else p = nullptr;
        return e;
    }
    void release() {
        if (p) (void)(PINNED ? hipHostFree(p) : hipFree(p));
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return (T*)p; }
    ~Buf() { release(); }
};

}  // namespace

// Device buffers, pinned host tables and host scratch of the linker, kept by
// the context between calls (grown, never shrunk), so repeated calls touch
// no fresh pages.
struct LinkCache {
    Buf<false> y, x, stepof, stepoff, k0, v0, k1, v1, cnt, off, cd, dist, sort_tmp, scan_tmp;
    Buf<true> off_h, cd_h, dist_h;
    std::vector<float> ry, rx;      // link_rows: the y and x columns
    std::vector<float> cy, cx;      // link_chunk: carried entries, then the chunk
    std::vector<uint32_t> stamp;    // link_steps: the per-id stamps of a call
    hipStream_t stream = nullptr;   // for linking beside the locate streams
    ~LinkCache() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

LinkCache* link_cache_create() { return new LinkCache(); }
void link_cache_destroy(LinkCache* c) { delete c; }
hipError_t link_cache_stream(LinkCache* c, hipStream_t* out) {
    if (!c->stream) {
        const hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) return e;
    }
    *out = c->stream;
    return hipSuccess;
}

#define LK_TRY(x)                                                 \
    do {                                                          \
        const hipError_t e_ = (x);                                \
        if (e_ != hipSuccess) {                                   \
            err = std::string(#x) + ": " + hipGetErrorString(e_); \
            return e_;                                            \
        }                                                         \
    } while (0)

LinkState* link_state_create(uint32_t memory) {
    LinkState* s = new LinkState();
    s->memory = memory;
    s->ring.resize(memory);
    return s;
}
void link_state_destroy(LinkState* s) { delete s; }

hipError_t link_chunk(hipStream_t st, LinkCache* cache, LinkState* state, const float* ys_c,
                      const float* xs_c, const std::vector<uint32_t>& soff_c, float search_range,
                      uint64_t* ids, LinkStats* stats, std::string& err) {
    if (stats) *stats = LinkStats{};
    const uint32_t memory = state->memory;
    const uint32_t nsc = (uint32_t)soff_c.size() - 1, nc = soff_c.back();
    // Table of positions: steps 0..memory-1 hold the memory entries carried
    // in (oldest first), step memory the previous step's particles, then
    // the chunk's steps; a carried particle's step keeps its age relative to
    // the chunk, so (row, age) candidate lists mean the same as in one call.
    std::vector<uint32_t> step_off(1, 0);
    std::vector<uint64_t> carry_id;
    uint32_t ncar = 0;
    for (uint32_t k = 0; k < memory; ++k) ncar += (uint32_t)state->ring[k].size();
    ncar += (uint32_t)state->prev.size();
    // with nothing carried in (a first chunk) the chunk's own columns serve
    const float* ys = ys_c;
    const float* xs = xs_c;
    {
        std::vector<float>& ysv = cache->cy;
        std::vector<float>& xsv = cache->cx;
        if (ncar) {
            ysv.resize((size_t)ncar + nc);
            xsv.resize((size_t)ncar + nc);
        }
        uint32_t w = 0;
        auto carry = [&](const std::vector<LinkState::Carry>& v) {
            for (const auto& e : v) {
                ysv[w] = e.y;
                xsv[w] = e.x;
                carry_id.push_back(e.id);
                ++w;
            }
            step_off.push_back(w);
        };
        for (uint32_t k = 0; k < memory; ++k) carry(state->ring[k]);
        carry(state->prev);
        if (ncar) {
            std::copy(ys_c, ys_c + nc, ysv.begin() + ncar);
            std::copy(xs_c, xs_c + nc, xsv.begin() + ncar);
            ys = ysv.data();
            xs = xsv.data();
        }
    }
    for (uint32_t k = 1; k <= nsc; ++k) step_off.push_back(ncar + soff_c[k]);
    const uint32_t nsteps = (uint32_t)step_off.size() - 1;
    const uint32_t first = memory + 1;   // first step of the chunk
    const uint32_t n = step_off.back();
    const uint32_t nm = memory + 1;
    const size_t nt = (size_t)n * nm;
    if (nt >= (1ull << 32)) {
        err = "link: rows * (memory + 1) must be below 2^32";
        return hipErrorInvalidValue;
    }

    if (stats) stats->steps = nsc;
    const auto t0 = std::chrono::steady_clock::now();

    // ---- candidates on the device --------------------------------------------
    LinkCache& C = *cache;
    LK_TRY(C.off_h.ensure((nt + 1) * 4));
    uint32_t* off_h = C.off_h.as<uint32_t>();
    uint32_t* cd_h = nullptr;
    float* cdist_h = nullptr;
    if (n == 0 || nsteps <= 1) std::fill(off_h, off_h + nt + 1, 0u);   // no pairs
    if (n > 0 && nsteps > 1) {
        LK_TRY(C.y.ensure((size_t)n * 4));
        LK_TRY(C.x.ensure((size_t)n * 4));
        LK_TRY(C.stepof.ensure((size_t)n * 4));
        LK_TRY(C.stepoff.ensure((size_t)(nsteps + 1) * 4));
        LK_TRY(C.k0.ensure((size_t)n * 4));
        LK_TRY(C.v0.ensure((size_t)n * 4));
        LK_TRY(C.k1.ensure((size_t)n * 4));
        LK_TRY(C.v1.ensure((size_t)n * 4));
        LK_TRY(C.cnt.ensure(nt * 4));
        LK_TRY(C.off.ensure(nt * 4));
        float *d_y = C.y.as<float>(), *d_x = C.x.as<float>();
        uint32_t *d_stepof = C.stepof.as<uint32_t>(), *d_stepoff = C.stepoff.as<uint32_t>();
        uint32_t *d_k0 = C.k0.as<uint32_t>(), *d_v0 = C.v0.as<uint32_t>();
        uint32_t *d_k1 = C.k1.as<uint32_t>(), *d_v1 = C.v1.as<uint32_t>();
        uint32_t *d_cnt = C.cnt.as<uint32_t>(), *d_off = C.off.as<uint32_t>();
        LK_TRY(hipMemcpyAsync(d_y, ys, (size_t)n * 4, hipMemcpyHostToDevice, st));
        LK_TRY(hipMemcpyAsync(d_x, xs, (size_t)n * 4, hipMemcpyHostToDevice, st));
        LK_TRY(hipMemcpyAsync(d_stepoff, step_off.data(), (size_t)(nsteps + 1) * 4,
                              hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_link_stepof, dim3(nsteps), dim3(256), 0, st, d_stepoff, d_stepof);
        LK_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_link_keys, dim3((n + 255) / 256), dim3(256), 0, st, d_y, d_k0, d_v0,
                           n);
        LK_TRY(hipGetLastError());
        size_t sort_bytes = 0, scan_bytes = 0;
        LK_TRY(rocprim::segmented_radix_sort_pairs(nullptr, sort_bytes, d_k0, d_k1, d_v0, d_v1, n,
                                                   nsteps, d_stepoff, d_stepoff + 1, 0, 32, st));
        LK_TRY(C.sort_tmp.ensure(sort_bytes));
        LK_TRY(rocprim::segmented_radix_sort_pairs(C.sort_tmp.p, sort_bytes, d_k0, d_k1, d_v0,
                                                   d_v1, n, nsteps, d_stepoff, d_stepoff + 1, 0,
                                                   32, st));
        const uint32_t blocks = (uint32_t)((nt + 255) / 256);
        hipLaunchKernelGGL(k_link_pairs, dim3(blocks), dim3(256), 0, st, d_y, d_x, d_stepof,
                           d_stepoff, nsteps, d_k1, d_v1, nm, search_range, n, d_cnt,
                           (const uint32_t*)nullptr, (uint32_t*)nullptr, (float*)nullptr);
        LK_TRY(hipGetLastError());
        LK_TRY(rocprim::exclusive_scan(nullptr, scan_bytes, d_cnt, d_off, 0u, nt,
                                       rocprim::plus<uint32_t>(), st));
        LK_TRY(C.scan_tmp.ensure(scan_bytes));
        LK_TRY(rocprim::exclusive_scan(C.scan_tmp.p, scan_bytes, d_cnt, d_off, 0u, nt,
                                       rocprim::plus<uint32_t>(), st));
        LK_TRY(hipMemcpyAsync(off_h, d_off, nt * 4, hipMemcpyDeviceToHost, st));
        LK_TRY(hipMemcpyAsync(off_h + nt, d_cnt + nt - 1, 4, hipMemcpyDeviceToHost, st));
        LK_TRY(hipStreamSynchronize(st));
        const uint64_t total = (uint64_t)off_h[nt - 1] + off_h[nt];
        if (total >= (1ull << 32)) {
            err = "link: more than 2^32 - 1 candidate pairs";
            return hipErrorInvalidValue;
        }
        off_h[nt] = (uint32_t)total;
        if (total) {
            LK_TRY(C.cd.ensure(total * 4));
            LK_TRY(C.dist.ensure(total * 4));
            LK_TRY(C.cd_h.ensure(total * 4));
            LK_TRY(C.dist_h.ensure(total * 4));
            cd_h = C.cd_h.as<uint32_t>();
            cdist_h = C.dist_h.as<float>();
            hipLaunchKernelGGL(k_link_pairs, dim3(blocks), dim3(256), 0, st, d_y, d_x, d_stepof,
                               d_stepoff, nsteps, d_k1, d_v1, nm, search_range, n, d_cnt,
                               (const uint32_t*)d_off, C.cd.as<uint32_t>(), C.dist.as<float>());
            LK_TRY(hipGetLastError());
            LK_TRY(hipMemcpyAsync(cd_h, C.cd.p, total * 4, hipMemcpyDeviceToHost, st));
            LK_TRY(hipMemcpyAsync(cdist_h, C.dist.p, total * 4, hipMemcpyDeviceToHost, st));
            LK_TRY(hipStreamSynchronize(st));
        }
        if (stats) stats->pairs = total;
    }

    const auto t1 = std::chrono::steady_clock::now();

    // ---- Linker::advance over the steps (gt_link_solve.h) ---------------------
    LinkSolveIn in;
    in.memory = memory;
    in.nm = nm;
    in.first = first;
    in.nsteps = nsteps;
    in.nsc = nsc;
    in.ncar = ncar;
    in.step_off = step_off.data();
    in.carry_id = carry_id.data();
    in.off = off_h;
    in.cd = cd_h;
    in.cdist = cdist_h;
    in.ys = ys;
    in.xs = xs;
    in.search_range = search_range;
    LinkSolveOut so;
    link_solve(in, state, ids, &so);
    if (stats) {
        using ms = std::chrono::duration<double, std::milli>;
        stats->search_ms = ms(t1 - t0).count();
        stats->solve_ms = ms(std::chrono::steady_clock::now() - t1).count();
        stats->subnets = so.subnets;
        stats->max_subnet = so.max_subnet;
        stats->ids = state->counter;
    }
    return hipSuccess;
}

hipError_t link_steps(hipStream_t st, LinkCache* cache, const float* ys, const float* xs,
                      const std::vector<uint32_t>& step_off, float search_range,
                      uint32_t memory, uint64_t* ids, LinkStats* stats, std::string& err) {
    // memory beyond the number of steps behaves as memory == steps: entries
    // never expire within the call
    const uint32_t nsteps = (uint32_t)step_off.size() - 1;
    std::unique_ptr<LinkState> state(link_state_create(std::min(memory, nsteps)));
    // the stamps of a fresh state start at zero: the cache's vector, cleared
    // (no fresh pages), lent to the state for the call
    state->stamp.swap(cache->stamp);
    state->stamp.assign((size_t)step_off.back() + 1, 0u);
    const hipError_t e =
        link_chunk(st, cache, state.get(), ys, xs, step_off, search_range, ids, stats, err);
    state->stamp.swap(cache->stamp);
    return e;
}

hipError_t link_rows(hipStream_t st, LinkCache* cache, const float* rows, uint64_t n_rows,
                     uint32_t ncols,
                     float search_range, uint32_t memory, uint64_t* ids, LinkStats* stats,
                     std::string& err) {
    if (n_rows >= (1ull << 32)) {
        err = "link: at most 2^32 - 1 rows";
        return hipErrorInvalidValue;
    }
    // FrameSubsetter (linking.rs:34-58): a new step starts where the frame
    // value differs from the previous one, the first comparison being
    // against 0.0 (a first frame other than 0 yields an empty first step);
    // no rows, no steps.
    std::vector<uint32_t> step_off(1, 0);
    std::vector<float>& ys = cache->ry;
    std::vector<float>& xs = cache->rx;
    ys.resize(n_rows);
    xs.resize(n_rows);
    float cur = 0.0f;
    for (uint64_t i = 0; i < n_rows; ++i) {
        const float fr = rows[i * ncols + 0];
        if (fr != cur) {
            cur = fr;
            step_off.push_back((uint32_t)i);
        }
        ys[i] = rows[i * ncols + 2];
        xs[i] = rows[i * ncols + 3];
    }
    if (n_rows > 0) step_off.push_back((uint32_t)n_rows);
    return link_steps(st, cache, ys.data(), xs.data(), step_off, search_range, memory, ids, stats,
                      err);
}

#undef LK_TRY

}  // namespace gt
