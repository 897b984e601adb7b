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
// The subnet solver then runs on the host in the order of Linker::advance /
// link (linking.rs:234-492): memory entries appended oldest first, depth-
// first subnets from each destination in order, one-to-one pairs taken
// directly, exhaustive minimum-cost assignment (null cost r*r, first
// strictly better score kept, binary32 running score), new ids in
// destination order, unmatched sources of the previous step kept as memory.
// A source's candidates are taken in ascending destination order; the
// reference takes its kd-tree's traversal order, which only decides between
// assignments of equal binary32 cost.
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

// A source of link(): the particle it was (index into the positions), the
// step that particle belongs to, and its id.
struct Src {
    uint32_t p, s;
    uint64_t id;
};

// recurse2 (linking.rs:391-452) over one subnet: every destination in turn
// takes one of its sources (ascending source order) or none (cost r*r); the
// running binary32 score is compared with the best one before each choice
// and the first assignment with a strictly smaller total is kept.
struct Assign {
    const uint32_t* dests = nullptr;   // destinations of the subnet, DFS order
    size_t ndests = 0;
    const uint32_t* d2s_off = nullptr; // CSR of destination -> (source, distance)
    const uint32_t* d2s_src = nullptr;
    const float* d2s_dist = nullptr;
    const Src* srcs = nullptr;
    int64_t* out = nullptr;            // -1: no source
    std::vector<uint64_t> used_id;
    std::vector<uint32_t> used_dest;
    std::vector<uint32_t> nulls;
    float score = 0.f, best = 0.f, null_cost = 0.f;

    bool in_use(uint64_t id) const {
        return std::find(used_id.begin(), used_id.end(), id) != used_id.end();
    }

    void run(size_t progress) {
        const uint32_t cur = dests[progress];
        const uint32_t b = d2s_off[cur], e = d2s_off[cur + 1];
        for (uint32_t q = b; q <= e; ++q) {
            const bool is_null = q == e;
            const float dist = is_null ? null_cost : d2s_dist[q];
            if (score > best) return;
            score = score + dist;
            if (!is_null) {
                const uint64_t sid = srcs[d2s_src[q]].id;
                if (in_use(sid)) {
                    score = score - dist;
                    continue;
                }
                used_id.push_back(sid);
                used_dest.push_back(cur);
            } else {
                nulls.push_back(cur);
            }
            if (progress + 1 == ndests) {
                if (score < best) {
                    best = score;
                    for (size_t u = 0; u < used_id.size(); ++u) out[used_dest[u]] = (int64_t)used_id[u];
                    for (uint32_t v : nulls) out[v] = -1;
                }
            } else {
                run(progress + 1);
            }
            if (!is_null) {
                used_id.pop_back();
                used_dest.pop_back();
            } else {
                nulls.pop_back();
            }
            score = score - dist;
        }
    }
};

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

    // ---- Linker::advance over the steps --------------------------------------
    // Ids are dense (< n); a per-id stamp answers "is this id among the
    // sources" (2t + 1) and "was it handed to a destination" (2t + 2).
    uint64_t counter = state->counter;
    std::vector<uint32_t>& stamp = state->stamp;
    if (stamp.size() < counter + nc + 1) stamp.resize(counter + nc + 1, 0);
    std::vector<Src> prev, next;
    std::vector<std::vector<Src>> ring(memory);   // memory steps, ring[head] oldest
    uint32_t head = 0;
    for (uint32_t k = 0; k < memory; ++k)
        for (uint32_t i = step_off[k]; i < step_off[k + 1]; ++i) ring[k].push_back({i, k, carry_id[i]});
    for (uint32_t i = step_off[memory]; i < step_off[memory + 1]; ++i)
        prev.push_back({i, memory, carry_id[i]});
    uint32_t prev_n = (uint32_t)prev.size();
    std::vector<uint32_t> d2s_off, d2s_fill, d2s_src;
    std::vector<float> d2s_dist;
    std::vector<uint8_t> vis_s, vis_d;
    std::vector<std::pair<uint32_t, uint8_t>> stack;
    std::vector<uint32_t> path_d, sub_dests, sub_off;
    std::vector<int64_t> out;
    std::vector<std::pair<uint32_t, uint32_t>> lists;
    Assign A;
    A.null_cost = search_range * search_range;
    uint64_t n_subnets = 0, max_subnet = 0;
    auto list_of = [&](const Src& s, uint32_t t, uint32_t& b, uint32_t& e) {
        const uint32_t j = t - s.s - 1;
        if (j >= nm) {
            b = e = 0;
            return;
        }
        const size_t key = (size_t)s.p * nm + j;
        b = off_h[key];
        e = off_h[key + 1];
    };
    for (uint32_t t = first; t < nsteps; ++t) {
        const uint32_t a = step_off[t], nd = step_off[t + 1] - step_off[t];
        // stamps are unique over every step the state has seen
        const uint64_t gs = state->steps_done + (t - first);
        const uint32_t st_src = (uint32_t)(2 * gs + 1), st_out = (uint32_t)(2 * gs + 2);
        // memory entries, oldest first, ids not already among the sources:
        // the previous step's ids carry that step's output stamp (2 gs),
        // memory entries taken in this step get st_src
        if (memory > 0) {
            const uint32_t st_prev = (uint32_t)(2 * gs);
            for (uint32_t k = 0; k < memory; ++k)
                for (const Src& e : ring[(head + k) % memory]) {
                    if (stamp[e.id] == st_src || (gs > 0 && stamp[e.id] == st_prev)) continue;
                    stamp[e.id] = st_src;
                    prev.push_back(e);
                }
        }
        const uint32_t ns = (uint32_t)prev.size();
        // each source's candidate list [b, e) for this step
        lists.resize(ns);
        for (uint32_t si = 0; si < ns; ++si) list_of(prev[si], t, lists[si].first, lists[si].second);
        // every source's list is in ascending destination order (sorted by
        // k_link_pairs); the destination -> source lists in ascending source
        // order
        d2s_off.assign((size_t)nd + 1, 0);
        for (uint32_t si = 0; si < ns; ++si) {
            const uint32_t b = lists[si].first, e = lists[si].second;
            for (uint32_t q = b; q < e; ++q) ++d2s_off[cd_h[q] + 1];
        }
        for (uint32_t d = 0; d < nd; ++d) d2s_off[d + 1] += d2s_off[d];
        d2s_fill.assign(d2s_off.begin(), d2s_off.end() - 1);
        d2s_src.resize(d2s_off[nd]);
        d2s_dist.resize(d2s_off[nd]);
        for (uint32_t si = 0; si < ns; ++si) {
            const uint32_t b = lists[si].first, e = lists[si].second;
            for (uint32_t q = b; q < e; ++q) {
                const uint32_t w = d2s_fill[cd_h[q]]++;
                d2s_src[w] = si;
                d2s_dist[w] = cdist_h[q];
            }
        }
        // connected subnets, depth first from each destination in order
        // (the recursion of linking.rs:317-341 with an explicit stack)
        vis_s.assign(ns, 0);
        vis_d.assign(nd, 0);
        out.assign(nd, -1);
        sub_dests.clear();
        sub_off.assign(1, 0);
        for (uint32_t i = 0; i < nd; ++i) {
            if (vis_d[i]) continue;
            // the common subnets resolved without the stack: a destination
            // without candidates, or one whose only source has only it
            // (what the search would find: path [i] and at most one source)
            const uint32_t nsi = d2s_off[i + 1] - d2s_off[i];
            if (nsi == 0) {
                vis_d[i] = 1;
                continue;
            }
            if (nsi == 1) {
                const uint32_t si = d2s_src[d2s_off[i]];
                if (lists[si].second - lists[si].first == 1) {
                    vis_d[i] = 1;
                    vis_s[si] = 1;
                    out[i] = (int64_t)prev[si].id;
                    continue;
                }
            }
            path_d.clear();
            uint32_t n_src = 0, first_src = 0;
            stack.assign(1, {i, 1});
            while (!stack.empty()) {
                const auto [idx, kind] = stack.back();
                stack.pop_back();
                if (kind == 0) {
                    if (vis_s[idx]) continue;
                    vis_s[idx] = 1;
                    if (n_src++ == 0) first_src = idx;
                    const uint32_t b = lists[idx].first, e = lists[idx].second;
                    for (uint32_t q = e; q-- > b;) stack.push_back({cd_h[q], 1});
                } else {
                    if (vis_d[idx]) continue;
                    vis_d[idx] = 1;
                    path_d.push_back(idx);
                    for (uint32_t q = d2s_off[idx + 1]; q-- > d2s_off[idx];)
                        stack.push_back({d2s_src[q], 0});
                }
            }
            if (path_d.size() == 1 && n_src == 1) {
                out[i] = (int64_t)prev[first_src].id;
            } else if (path_d.size() > 1 || n_src > 1) {
                sub_dests.insert(sub_dests.end(), path_d.begin(), path_d.end());
                sub_off.push_back((uint32_t)sub_dests.size());
            }
        }
        A.d2s_off = d2s_off.data();
        A.d2s_src = d2s_src.data();
        A.d2s_dist = d2s_dist.data();
        A.srcs = prev.data();
        A.out = out.data();
        for (size_t k = 0; k + 1 < sub_off.size(); ++k) {
            A.dests = sub_dests.data() + sub_off[k];
            A.ndests = sub_off[k + 1] - sub_off[k];
            A.used_id.clear();
            A.used_dest.clear();
            A.nulls.clear();
            A.score = 0.f;
            A.best = __builtin_inff();
            A.run(0);
            ++n_subnets;
            max_subnet = std::max<uint64_t>(max_subnet, A.ndests);
        }
        // ids; destinations without a source get new ones in order
        next.resize(nd);
        for (uint32_t i = 0; i < nd; ++i) {
            const uint64_t id = out[i] >= 0 ? (uint64_t)out[i] : counter++;
            ids[a + i - ncar] = id;
            stamp[id] = st_out;
            next[i] = {a + i, t, id};
        }
        if (memory > 0) {
            // unmatched particles of the previous step replace the oldest
            // memory step
            std::vector<Src>& slot = ring[head];
            slot.clear();
            for (uint32_t si = 0; si < prev_n; ++si)
                if (stamp[prev[si].id] != st_out) slot.push_back(prev[si]);
            head = (head + 1) % memory;
        }
        std::swap(prev, next);
        prev_n = nd;
    }
    // the state carried to the next chunk
    auto to_carry = [&](const std::vector<Src>& v, std::vector<LinkState::Carry>& out_v) {
        out_v.clear();
        for (const Src& e : v) out_v.push_back({ys[e.p], xs[e.p], e.id});
    };
    to_carry(prev, state->prev);
    for (uint32_t k = 0; k < memory; ++k) to_carry(ring[(head + k) % memory], state->ring[k]);
    state->counter = counter;
    state->steps_done += nsc;
    if (stats) {
        using ms = std::chrono::duration<double, std::milli>;
        stats->search_ms = ms(t1 - t0).count();
        stats->solve_ms = ms(std::chrono::steady_clock::now() - t1).count();
        stats->subnets = n_subnets;
        stats->max_subnet = max_subnet;
        stats->ids = counter;
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
