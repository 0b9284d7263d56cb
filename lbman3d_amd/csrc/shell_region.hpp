// shell_region.hpp — the wall-shell cells of a domain with open faces, as a
// flat index space for one-thread-per-cell passes (host and device).
#pragma once

#include <algorithm>

#include "common.hpp"

namespace lbm {

// The cells within w of an open face, each once and in the class order of
// the shells: the x band, then the y band outside the x band, then the z band
// outside both. Along axis a the band is [0, lo) and [hi, n); its complement
// [lo, hi) is one interval. A face is open unless both its Q and velocity
// tags are periodic.
struct ShellRegion {
    int lo[3], hi[3];
    long part[3];   // cells of the x, y and z parts
};

inline ShellRegion make_shell_region(const BCDesc<float>& bc, int w) {
    ShellRegion r{};
    int band[3];
    for (int a = 0; a < 3; ++a) {
        const int n = bc.n[a];
        const bool open_lo = !(bc.qkind[2 * a] == QB_PERIODIC && bc.ukind[2 * a] == UB_PERIODIC);
        const bool open_hi = !(bc.qkind[2 * a + 1] == QB_PERIODIC && bc.ukind[2 * a + 1] == UB_PERIODIC);
        r.lo[a] = open_lo ? std::min(w, n) : 0;
        r.hi[a] = open_hi ? std::max(r.lo[a], n - w) : n;
        band[a] = r.lo[a] + (n - r.hi[a]);
    }
    const long cx = bc.n[0] - band[0], cy = bc.n[1] - band[1];
    r.part[0] = (long)band[0] * bc.n[1] * bc.n[2];
    r.part[1] = cx * band[1] * bc.n[2];
    r.part[2] = cx * cy * band[2];
    return r;
}

__host__ __device__ inline long region_cells(const ShellRegion& r) { return r.part[0] + r.part[1] + r.part[2]; }

// Cell (p[0], p[1], p[2]) of item i (0 <= i < region_cells(r)); x runs
// fastest in every part.
__host__ __device__ inline void region_cell(const ShellRegion& r, const int n[3], long i, int p[3]) {
    auto pick = [&](int a, int j) { return j < r.lo[a] ? j : r.hi[a] + (j - r.lo[a]); };
    const int bx = r.lo[0] + (n[0] - r.hi[0]), by = r.lo[1] + (n[1] - r.hi[1]);
    const int cx = r.hi[0] - r.lo[0], cy = r.hi[1] - r.lo[1];
    if (i < r.part[0]) {
        const long rest = i / bx;
        p[0] = pick(0, (int)(i - rest * bx));
        p[1] = (int)(rest % n[1]);
        p[2] = (int)(rest / n[1]);
        return;
    }
    i -= r.part[0];
    if (i < r.part[1]) {
        const long rest = i / cx;
        p[0] = r.lo[0] + (int)(i - rest * cx);
        p[1] = pick(1, (int)(rest % by));
        p[2] = (int)(rest / by);
        return;
    }
    i -= r.part[1];
    const long rest = i / cx;
    p[0] = r.lo[0] + (int)(i - rest * cx);
    p[1] = r.lo[1] + (int)(rest % cy);
    p[2] = pick(2, (int)(rest / cy));
}

}  // namespace lbm
