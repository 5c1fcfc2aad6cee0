// xcd_sched.hpp -- host-side per-XCD schedules of the marching kernels (FD
// step, separable SEM step).
//
// The hardware deals block b to XCD b % 8, and each XCD starts its blocks in
// order as slots on its CUs free up.  Split m (1..4) cuts the march axis into
// R = 8 >> (m - 1) plane ranges (between plane groups of `group` planes) and
// the tile columns into 8 / R bands; XCD x takes range x % R of band x / R,
// so the tiles that share one L2 march the same planes at the same time and
// every range is as long as the split allows (less ring priming per march).
// The columns that would leave a last round of an XCD's slots part-empty are
// cut into shorter pieces that fill one round instead.  The lists are padded
// to one length with empty ranges; table entry b belongs to hardware block b
// and holds (column, first plane, end plane, 0).  Coverage does not depend on
// where a block runs: every (column, plane) of [plane0, plane0 + planes) is
// in exactly one entry.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

namespace phc {

inline void build_xcd_schedule(int planes, int plane0, int ncol, int slots_per_xcd, int m, int group,
                               std::vector<int4> &tab)
{
    const int R = 8 >> (m - 1), B = 8 / R;
    const int cpx = std::max(slots_per_xcd, 1);
    std::vector<std::vector<int4>> lists(8);
    size_t longest = 0;
    const int G = (planes + group - 1) / group;  // plane groups, dealt evenly over the ranges
    for (int x = 0; x < 8; ++x) {
        const int r = x % R, band = x / R;
        const int a = group * (G * r / R), b = std::min(group * (G * (r + 1) / R), planes);
        const int c0 = ncol * band / B, c1 = ncol * (band + 1) / B;
        auto &li = lists[x];
        if (a < b && c0 < c1) {
            const int rem = (c1 - c0) % cpx, full = c1 - rem;
            for (int w = c0; w < full; ++w) li.push_back(make_int4(w, plane0 + a, plane0 + b, 0));
            const int g = (b - a + group - 1) / group;  // plane groups of the range
            const int pieces = rem ? std::max(1, std::min(cpx / rem, g)) : 0;
            for (int w = full; w < c1; ++w)
                for (int q = 0; q < pieces; ++q) {
                    const int g0 = g * q / pieces, g1 = g * (q + 1) / pieces;
                    li.push_back(make_int4(w, plane0 + a + group * g0, plane0 + std::min(a + group * g1, b), 0));
                }
        }
        longest = std::max(longest, li.size());
    }
    tab.assign(8 * longest, make_int4(0, 0, 0, 0));
    for (int x = 0; x < 8; ++x)
        for (size_t l = 0; l < lists[x].size(); ++l) tab[8 * l + x] = lists[x][l];
}

} // namespace phc
