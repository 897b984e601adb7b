// gt_link_solve.h -- the subnet solver of the linker (linking.rs:234-492):
// host code only, shared by gt_link.hip (after the GPU candidate search).
// Steps run in the order of Linker::advance / link: memory entries appended
// oldest first, depth-first subnets from each destination in order,
// one-to-one pairs taken directly, exhaustive minimum-cost assignment (null
// cost r*r, first strictly better score kept, binary32 running score), new
// ids in destination order, unmatched sources of the previous step kept as
// memory.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <utility>
#include <vector>

namespace gt {

// Linker state between chunks of steps: the last step's particles and the
// memory entries (oldest step first), the next id, the steps seen so far.
struct LinkState {
    struct Carry {
        float y, x;
        uint64_t id;
    };
    uint32_t memory = 0;
    uint64_t counter = 0, steps_done = 0;
    std::vector<Carry> prev;
    std::vector<std::vector<Carry>> ring;   // `memory` steps, oldest first
    std::vector<uint32_t> stamp;            // per id, solver scratch
};

namespace linksolve {

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

}  // namespace linksolve

// The position table of a chunk and its candidate lists.  Steps 0 ..
// memory - 1 hold the memory entries carried in (oldest first), step memory
// the previous step's particles (their ids in carry_id, ncar entries in
// all), then the chunk's nsc steps.  The candidates of (position p, age j)
// -- destinations in step s(p) + 1 + j, as indices within that step, in
// ascending order, with binary32 squared distances -- are cd / cdist
// [off[p * nm + j], off[p * nm + j + 1]).
struct LinkSolveIn {
    uint32_t memory = 0, nm = 1, first = 1, nsteps = 0, nsc = 0, ncar = 0;
    const uint32_t* step_off = nullptr;   // nsteps + 1
    const uint64_t* carry_id = nullptr;   // ncar
    const uint32_t* off = nullptr;        // n * nm + 1
    const uint32_t* cd = nullptr;
    const float* cdist = nullptr;
    const float* ys = nullptr;            // the table's positions
    const float* xs = nullptr;
    float search_range = 0.f;
};
struct LinkSolveOut {
    uint64_t subnets = 0, max_subnet = 0;
};

// Ids of the chunk's positions (ids[i] for table position ncar + i); the
// state advances past the chunk.
inline void link_solve(const LinkSolveIn& in, LinkState* state, uint64_t* ids, LinkSolveOut* so) {
    using linksolve::Src;
    using linksolve::Assign;
    const uint32_t memory = in.memory, nm = in.nm, first = in.first, nsteps = in.nsteps;
    const uint32_t nsc = in.nsc, ncar = in.ncar;
    const uint32_t* step_off = in.step_off;
    const uint32_t nc = step_off[nsteps] - ncar;
    const uint64_t* carry_id = in.carry_id;
    const uint32_t* off_h = in.off;
    const uint32_t* cd_h = in.cd;
    const float* cdist_h = in.cdist;
    const float* ys = in.ys;
    const float* xs = in.xs;
    const float search_range = in.search_range;
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
    if (so) {
        so->subnets = n_subnets;
        so->max_subnet = max_subnet;
    }
}

}  // namespace gt
