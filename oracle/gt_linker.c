/* gt_linker.c -- CPU oracle of the frame-to-frame linker, in C.
 *
 * TEST INFRASTRUCTURE ONLY: the second restatement of gpu-tracking's
 * linking.rs (next to oracle/linker.py) and the compiled CPU baseline of the
 * locate + link benchmark.  Only tests/ and bench.py's CPU-baseline leg use
 * it.  One thread, in the reference's order:
 *
 *   FrameSubsetter (linking.rs:34-58): a step is a run of equal frame values,
 *     the first comparison against 0.0 (a first frame other than 0 gives an
 *     empty first step);
 *   Linker::advance (linking.rs:234-283): memory entries (unmatched sources
 *     of the last `memory` steps, oldest first, ids not already present)
 *     appended after the previous step's particles;
 *   within_radius_rd (linking.rs:73-105): per-axis box against binary32
 *     bounds q - r, q + r, then (0 + dy*dy) + dx*dx < r*r in binary32; the
 *     candidates of a source are taken in ascending destination order (the
 *     kd-tree crate's order decides only between equal-cost assignments);
 *     a uniform grid of cell r over the destinations narrows the pairs
 *     before the exact test;
 *   link (linking.rs:286-492): subnets by depth-first search from each
 *     destination in order, one-to-one pairs taken directly, recurse2's
 *     exhaustive search (null cost r*r, binary32 running score, first
 *     strictly better total kept), new ids in destination order, unmatched
 *     sources of the previous step kept as memory.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "gt_oracle.h"

typedef struct {
    float y, x;
    uint64_t id;
} Src;

typedef struct {
    Src* v;
    size_t n, cap;
} SrcVec;

static int sv_push(SrcVec* s, Src e) {
    if (s->n == s->cap) {
        size_t c = s->cap ? 2 * s->cap : 64;
        Src* p = (Src*)realloc(s->v, c * sizeof(Src));
        if (!p) return -1;
        s->v = p;
        s->cap = c;
    }
    s->v[s->n++] = e;
    return 0;
}

typedef struct {
    uint32_t* v;
    size_t n, cap;
} U32Vec;

static int uv_push(U32Vec* s, uint32_t e) {
    if (s->n == s->cap) {
        size_t c = s->cap ? 2 * s->cap : 64;
        uint32_t* p = (uint32_t*)realloc(s->v, c * sizeof(uint32_t));
        if (!p) return -1;
        s->v = p;
        s->cap = c;
    }
    s->v[s->n++] = e;
    return 0;
}

/* recurse2 (linking.rs:391-452) */
typedef struct {
    const uint32_t* dests;
    size_t nd;
    const uint32_t *d2s_off, *d2s_src;
    const float* d2s_dist;
    const Src* src;
    int64_t* out;
    uint64_t* used_id;
    uint32_t* used_dest;
    size_t n_used;
    uint32_t* nulls;
    size_t n_nulls;
    float score, best, null_cost;
} Assign;

static void recurse2(Assign* A, size_t progress) {
    const uint32_t cur = A->dests[progress];
    const uint32_t b = A->d2s_off[cur], e = A->d2s_off[cur + 1];
    for (uint32_t q = b; q <= e; ++q) {
        const int is_null = q == e;
        const float dist = is_null ? A->null_cost : A->d2s_dist[q];
        if (A->score > A->best) return;
        A->score = A->score + dist;
        if (!is_null) {
            const uint64_t sid = A->src[A->d2s_src[q]].id;
            int used = 0;
            for (size_t u = 0; u < A->n_used; ++u)
                if (A->used_id[u] == sid) used = 1;
            if (used) {
                A->score = A->score - dist;
                continue;
            }
            A->used_id[A->n_used] = sid;
            A->used_dest[A->n_used++] = cur;
        } else {
            A->nulls[A->n_nulls++] = cur;
        }
        if (progress + 1 == A->nd) {
            if (A->score < A->best) {
                A->best = A->score;
                for (size_t u = 0; u < A->n_used; ++u) A->out[A->used_dest[u]] = (int64_t)A->used_id[u];
                for (size_t u = 0; u < A->n_nulls; ++u) A->out[A->nulls[u]] = -1;
            }
        } else {
            recurse2(A, progress + 1);
        }
        if (!is_null) --A->n_used;
        else --A->n_nulls;
        A->score = A->score - dist;
    }
}

/* Candidate lists of every source (ascending destination index). */
typedef struct {
    uint32_t* off;   /* [ns + 1] */
    uint32_t* dst;
    float* dist;
    size_t n, cap;
} Lists;

static int lists_push(Lists* L, uint32_t d, float dist) {
    if (L->n == L->cap) {
        size_t c = L->cap ? 2 * L->cap : 1024;
        uint32_t* a = (uint32_t*)realloc(L->dst, c * sizeof(uint32_t));
        if (!a) return -1;
        L->dst = a;
        float* b = (float*)realloc(L->dist, c * sizeof(float));
        if (!b) return -1;
        L->dist = b;
        L->cap = c;
    }
    L->dst[L->n] = d;
    L->dist[L->n++] = dist;
    return 0;
}

static int search(const Src* src, size_t ns, const float* dy, const float* dx, size_t nd, float r,
                  Lists* L) {
    L->n = 0;
    uint32_t* off = (uint32_t*)realloc(L->off, (ns + 1) * sizeof(uint32_t));
    if (!off) return -1;
    L->off = off;
    off[0] = 0;
    const float r2 = r * r;
    if (!(r > 0.0f) || nd == 0) {
        for (size_t s = 0; s < ns; ++s) off[s + 1] = 0;
        return 0;
    }
    /* grid over the finite destinations (non-finite ones never link) */
    float ymin = INFINITY, ymax = -INFINITY, xmin = INFINITY, xmax = -INFINITY;
    for (size_t d = 0; d < nd; ++d)
        if (isfinite(dy[d]) && isfinite(dx[d])) {
            ymin = fminf(ymin, dy[d]); ymax = fmaxf(ymax, dy[d]);
            xmin = fminf(xmin, dx[d]); xmax = fmaxf(xmax, dx[d]);
        }
    double cell = r;
    if (!(ymin <= ymax)) {  /* no finite destination */
        for (size_t s = 0; s < ns; ++s) off[s + 1] = 0;
        return 0;
    }
    if ((ymax - ymin) / cell > 4096.0) cell = (ymax - ymin) / 4096.0;
    if ((xmax - xmin) / cell > 4096.0) cell = (xmax - xmin) / 4096.0;
    const long gy = (long)((ymax - ymin) / cell) + 1, gx = (long)((xmax - xmin) / cell) + 1;
    uint32_t* head = (uint32_t*)calloc((size_t)(gy * gx + 1), sizeof(uint32_t));
    uint32_t* order = (uint32_t*)malloc((nd + 1) * sizeof(uint32_t));
    long* cellof = (long*)malloc((nd + 1) * sizeof(long));
    if (!head || !order || !cellof) {
        free(head); free(order); free(cellof);
        return -1;
    }
    for (size_t d = 0; d < nd; ++d) {
        if (isfinite(dy[d]) && isfinite(dx[d])) {
            long cy = (long)((dy[d] - ymin) / cell), cx = (long)((dx[d] - xmin) / cell);
            if (cy >= gy) cy = gy - 1;
            if (cx >= gx) cx = gx - 1;
            cellof[d] = cy * gx + cx;
            head[cellof[d] + 1]++;
        } else {
            cellof[d] = -1;
        }
    }
    for (long c = 0; c < gy * gx; ++c) head[c + 1] += head[c];
    {
        uint32_t* fill = (uint32_t*)malloc((size_t)(gy * gx) * sizeof(uint32_t));
        if (!fill) {
            free(head); free(order); free(cellof);
            return -1;
        }
        memcpy(fill, head, (size_t)(gy * gx) * sizeof(uint32_t));
        for (size_t d = 0; d < nd; ++d)   /* ascending d within a cell */
            if (cellof[d] >= 0) order[fill[cellof[d]]++] = (uint32_t)d;
        free(fill);
    }
    int rc = 0;
    for (size_t s = 0; s < ns && rc == 0; ++s) {
        const float qy = src[s].y, qx = src[s].x;
        const size_t first = L->n;
        if (isfinite(qy) && isfinite(qx)) {
            const float lo0 = qy - r, hi0 = qy + r, lo1 = qx - r, hi1 = qx + r;
            long cy0 = (long)floor((lo0 - ymin) / cell) - 1, cy1 = (long)floor((hi0 - ymin) / cell) + 1;
            long cx0 = (long)floor((lo1 - xmin) / cell) - 1, cx1 = (long)floor((hi1 - xmin) / cell) + 1;
            if (cy0 < 0) cy0 = 0;
            if (cx0 < 0) cx0 = 0;
            if (cy1 > gy - 1) cy1 = gy - 1;
            if (cx1 > gx - 1) cx1 = gx - 1;
            for (long cy = cy0; cy <= cy1 && rc == 0; ++cy)
                for (long cx = cx0; cx <= cx1 && rc == 0; ++cx) {
                    const long c = cy * gx + cx;
                    for (uint32_t k = head[c]; k < head[c + 1]; ++k) {
                        const uint32_t d = order[k];
                        if (dy[d] < lo0 || dy[d] > hi0 || dx[d] < lo1 || dx[d] > hi1) continue;
                        const float a = dy[d] - qy, b = dx[d] - qx;
                        float dist = 0.0f + a * a;
                        dist = dist + b * b;
                        if (dist < r2 && lists_push(L, d, dist)) {
                            rc = -1;
                            break;
                        }
                    }
                }
            /* ascending destination index */
            for (size_t i = first + 1; i < L->n; ++i) {
                const uint32_t d = L->dst[i];
                const float v = L->dist[i];
                size_t k = i;
                for (; k > first && L->dst[k - 1] > d; --k) {
                    L->dst[k] = L->dst[k - 1];
                    L->dist[k] = L->dist[k - 1];
                }
                L->dst[k] = d;
                L->dist[k] = v;
            }
        }
        off[s + 1] = (uint32_t)L->n;
    }
    free(head);
    free(order);
    free(cellof);
    return rc;
}

int orc_link_rows(const float* rows, int64_t n, int32_t ncols, float r, int64_t memory,
                  uint64_t* ids) {
    /* steps (FrameSubsetter) */
    size_t nsteps = 0;
    size_t* so = (size_t*)malloc((size_t)(n + 2) * sizeof(size_t));
    if (!so) return -1;
    so[0] = 0;
    {
        float cur = 0.0f;
        for (int64_t i = 0; i < n; ++i) {
            const float fr = rows[i * ncols];
            if (fr != cur) {
                cur = fr;
                so[++nsteps] = (size_t)i;
            }
        }
        if (n > 0) so[++nsteps] = (size_t)n;
    }
    if ((int64_t)nsteps < memory) memory = (int64_t)nsteps;
    const size_t M = (size_t)(memory > 0 ? memory : 0);
    int rc = 0;
    SrcVec prev = {0}, srcs = {0};
    SrcVec* ring = (SrcVec*)calloc(M + 1, sizeof(SrcVec));
    uint32_t* stamp = (uint32_t*)calloc((size_t)n + 2, sizeof(uint32_t));
    Lists L = {0};
    float *dy = NULL, *dx = NULL;
    uint32_t *d2s_off = NULL, *d2s_src = NULL, *fillp = NULL;
    float* d2s_dist = NULL;
    uint8_t *vis_s = NULL, *vis_d = NULL;
    int64_t* out = NULL;
    uint32_t* path = NULL;
    U32Vec stack = {0}, kinds = {0};
    uint64_t* used_id = NULL;
    uint32_t *used_dest = NULL, *nulls = NULL;
    if (!ring || !stamp) rc = -1;
    size_t head = 0;
    uint64_t counter = 0;
    for (size_t t = 0; t < nsteps && rc == 0; ++t) {
        const size_t a = so[t], nd = so[t + 1] - so[t];
        const uint32_t st = (uint32_t)(t + 1);
        /* sources: the previous step's particles, then memory entries */
        srcs.n = 0;
        for (size_t i = 0; i < prev.n && rc == 0; ++i) {
            rc = sv_push(&srcs, prev.v[i]);
            stamp[prev.v[i].id] = st;
        }
        for (size_t k = 0; k < M && rc == 0; ++k) {
            const SrcVec* m = &ring[(head + k) % M];
            for (size_t i = 0; i < m->n && rc == 0; ++i) {
                if (stamp[m->v[i].id] == st) continue;
                stamp[m->v[i].id] = st;
                rc = sv_push(&srcs, m->v[i]);
            }
        }
        if (rc) break;
        const size_t ns = srcs.n;
        dy = (float*)realloc(dy, (nd + 1) * sizeof(float));
        dx = (float*)realloc(dx, (nd + 1) * sizeof(float));
        if (!dy || !dx) { rc = -1; break; }
        for (size_t i = 0; i < nd; ++i) {
            dy[i] = rows[(a + i) * (size_t)ncols + 2];
            dx[i] = rows[(a + i) * (size_t)ncols + 3];
        }
        if (search(srcs.v, ns, dy, dx, nd, r, &L)) { rc = -1; break; }
        /* destination -> sources, ascending source order */
        d2s_off = (uint32_t*)realloc(d2s_off, (nd + 2) * sizeof(uint32_t));
        fillp = (uint32_t*)realloc(fillp, (nd + 2) * sizeof(uint32_t));
        d2s_src = (uint32_t*)realloc(d2s_src, (L.n + 1) * sizeof(uint32_t));
        d2s_dist = (float*)realloc(d2s_dist, (L.n + 1) * sizeof(float));
        vis_s = (uint8_t*)realloc(vis_s, ns + 1);
        vis_d = (uint8_t*)realloc(vis_d, nd + 1);
        out = (int64_t*)realloc(out, (nd + 1) * sizeof(int64_t));
        path = (uint32_t*)realloc(path, (nd + 1) * sizeof(uint32_t));
        used_id = (uint64_t*)realloc(used_id, (nd + 1) * sizeof(uint64_t));
        used_dest = (uint32_t*)realloc(used_dest, (nd + 1) * sizeof(uint32_t));
        nulls = (uint32_t*)realloc(nulls, (nd + 1) * sizeof(uint32_t));
        if (!d2s_off || !fillp || !d2s_src || !d2s_dist || !vis_s || !vis_d || !out || !path ||
            !used_id || !used_dest || !nulls) { rc = -1; break; }
        memset(d2s_off, 0, (nd + 2) * sizeof(uint32_t));
        for (size_t q = 0; q < L.n; ++q) d2s_off[L.dst[q] + 1]++;
        for (size_t d = 0; d < nd; ++d) d2s_off[d + 1] += d2s_off[d];
        memcpy(fillp, d2s_off, (nd + 1) * sizeof(uint32_t));
        for (size_t s = 0; s < ns; ++s)
            for (uint32_t q = L.off[s]; q < L.off[s + 1]; ++q) {
                const uint32_t w = fillp[L.dst[q]]++;
                d2s_src[w] = (uint32_t)s;
                d2s_dist[w] = L.dist[q];
            }
        memset(vis_s, 0, ns + 1);
        memset(vis_d, 0, nd + 1);
        for (size_t i = 0; i < nd; ++i) out[i] = -1;
        /* subnets: depth first from each destination (linking.rs:316-341) */
        for (size_t i = 0; i < nd && rc == 0; ++i) {
            size_t np = 0, nsrc = 0, first_src = 0;
            stack.n = kinds.n = 0;
            rc = uv_push(&stack, (uint32_t)i) | uv_push(&kinds, 1);
            while (stack.n && rc == 0) {
                const uint32_t idx = stack.v[--stack.n];
                const uint32_t kind = kinds.v[--kinds.n];
                if (kind == 0) {
                    if (vis_s[idx]) continue;
                    vis_s[idx] = 1;
                    if (nsrc++ == 0) first_src = idx;
                    for (uint32_t q = L.off[idx + 1]; q-- > L.off[idx] && rc == 0;)
                        rc = uv_push(&stack, L.dst[q]) | uv_push(&kinds, 1);
                } else {
                    if (vis_d[idx]) continue;
                    vis_d[idx] = 1;
                    path[np++] = idx;
                    for (uint32_t q = d2s_off[idx + 1]; q-- > d2s_off[idx] && rc == 0;)
                        rc = uv_push(&stack, d2s_src[q]) | uv_push(&kinds, 0);
                }
            }
            if (rc) break;
            if (np == 1 && nsrc == 1) {
                out[i] = (int64_t)srcs.v[first_src].id;
            } else if (np > 1 || (np == 1 && nsrc > 1)) {
                Assign A = {path, np, d2s_off, d2s_src, d2s_dist, srcs.v, out,
                            used_id, used_dest, 0, nulls, 0, 0.0f, INFINITY, r * r};
                recurse2(&A, 0);
            }
        }
        if (rc) break;
        /* ids, new ones in destination order; the next sources */
        const size_t prev_n = prev.n;
        SrcVec unused = {0};
        for (size_t i = 0; i < nd; ++i) {
            const uint64_t id = out[i] >= 0 ? (uint64_t)out[i] : counter++;
            ids[a + i] = id;
        }
        if (M > 0) {
            for (size_t i = 0; i < nd; ++i) stamp[ids[a + i]] = st + 0x80000000u;
            for (size_t s = 0; s < prev_n && rc == 0; ++s)
                if (stamp[prev.v[s].id] != st + 0x80000000u) rc = sv_push(&unused, prev.v[s]);
            free(ring[head].v);
            ring[head] = unused;
            head = (head + 1) % M;
        }
        prev.n = 0;
        for (size_t i = 0; i < nd && rc == 0; ++i) {
            Src e = {dy[i], dx[i], ids[a + i]};
            rc = sv_push(&prev, e);
        }
    }
    for (size_t k = 0; ring && k < M; ++k) free(ring[k].v);
    free(ring); free(stamp); free(so); free(prev.v); free(srcs.v);
    free(L.off); free(L.dst); free(L.dist); free(dy); free(dx);
    free(d2s_off); free(fillp); free(d2s_src); free(d2s_dist); free(vis_s); free(vis_d);
    free(out); free(path); free(stack.v); free(kinds.v); free(used_id); free(used_dest);
    free(nulls);
    return rc;
}
