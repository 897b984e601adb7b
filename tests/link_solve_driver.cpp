// Test driver for the linker's host subnet solver (gt_link_solve.h): reads a
// position table and its candidate lists from a binary file, runs
// link_solve, writes the chunk's ids.  Built and run by
// tests/test_link_solve_cpu.py; no GPU involved.
//
// Input (little-endian): u32 memory, nsteps, n, npairs; f32 search_range;
// u32 step_off[nsteps + 1]; f32 ys[n], xs[n]; u32 off[n * (memory + 1) + 1];
// u32 cd[npairs]; f32 cdist[npairs].  Output: u64 ids[n].
#include <cstdio>
#include <cstdint>
#include <vector>

#include "gt_link_solve.h"

template <class T>
static bool rd(FILE* f, T* p, size_t n) { return fread(p, sizeof(T), n, f) == n; }

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 3;
    uint32_t hdr[4];
    float r;
    if (!rd(f, hdr, 4) || !rd(f, &r, 1)) return 4;
    const uint32_t memory = hdr[0], nsteps = hdr[1], n = hdr[2], npairs = hdr[3];
    const uint32_t nm = memory + 1;
    std::vector<uint32_t> step_off(nsteps + 1), off((size_t)n * nm + 1), cd(npairs);
    std::vector<float> ys(n), xs(n), cdist(npairs);
    if (!rd(f, step_off.data(), step_off.size()) || !rd(f, ys.data(), n) || !rd(f, xs.data(), n) ||
        !rd(f, off.data(), off.size()) || !rd(f, cd.data(), npairs) ||
        !rd(f, cdist.data(), npairs))
        return 5;
    fclose(f);
    // a fresh state: memory empty carried steps and an empty previous step
    // open the table, as in a first chunk
    gt::LinkState st;
    st.memory = memory;
    st.ring.resize(memory);
    st.stamp.assign((size_t)n + 1, 0);
    gt::LinkSolveIn in;
    in.memory = memory;
    in.nm = nm;
    in.first = memory + 1;
    in.nsteps = nsteps;
    in.nsc = nsteps - (memory + 1);
    in.ncar = 0;
    in.step_off = step_off.data();
    in.off = off.data();
    in.cd = cd.data();
    in.cdist = cdist.data();
    in.ys = ys.data();
    in.xs = xs.data();
    in.search_range = r;
    std::vector<uint64_t> ids(n);
    gt::link_solve(in, &st, ids.data(), nullptr);
    FILE* o = fopen(argv[2], "wb");
    if (!o) return 6;
    fwrite(ids.data(), sizeof(uint64_t), n, o);
    fclose(o);
    return 0;
}
