// plan.cpp — tile planner; see plan.h.  Replaces the reference's cone bucket
// builders (src/algorithm.jl:1921-2210, size classes 3/4/5/8, <65, 65-1023,
// >=1024) with one nnz- and size-balanced tiling shared by every kernel.
#include "plan.h"

#include <algorithm>

namespace hpr {

static int32_t pow2ceil(int64_t v) {
    int32_t p = 1;
    while (p < v && p < (1 << 30)) p <<= 1;
    return p;
}

static int32_t choose_g(int64_t nnz, int64_t rows, const PlanConfig &cfg) {
    if (rows <= 0) return 1;
    if (cfg.stream && nnz <= kStreamNnz) return 0;  // CSR-stream through LDS
    const int64_t avg = (nnz + rows - 1) / rows;
    if (avg >= 1024) return kBlock;  // whole workgroup per row
    int32_t g = pow2ceil((avg + cfg.g_divisor - 1) / cfg.g_divisor);
    return std::min<int32_t>(std::max<int32_t>(g, 1), 64);
}

HostPlan build_plan(const int32_t *rowptr, int32_t nrows, int32_t lin_end, const int32_t *cone_idx,
                    int32_t ncones, const PlanConfig &cfg, const int32_t *row_inst) {
    HostPlan P;
    auto nnz_of = [&](int32_t a, int32_t b) -> int64_t {
        return rowptr ? (int64_t)rowptr[b] - rowptr[a] : (int64_t)(b - a);
    };
    auto inst_of = [&](int32_t row) -> int32_t {
        return (row_inst && row >= 0 && row < nrows) ? row_inst[row] : 0;
    };
    auto cone_inst = [&](int32_t c) -> int32_t { return inst_of(cone_idx[c]); };
    const int64_t target = rowptr ? cfg.nnz_target : kTileRows;
    // linear rows
    int32_t r = 0;
    while (r < lin_end) {
        int32_t e = r + 1;
        int64_t nz = nnz_of(r, e);
        while (e < lin_end && e - r < kTileRows && nz + nnz_of(e, e + 1) <= target && inst_of(e) == inst_of(r)) {
            nz += nnz_of(e, e + 1);
            ++e;
        }
        Tile t{r, e, TILE_LINEAR, choose_g(nz, e - r, cfg), -1, -1, 1, inst_of(r)};
        P.tiles.push_back(t);
        r = e;
    }
    // cones
    std::vector<Tile> huge;
    const int64_t huge_nnz = (int64_t)cfg.huge_nnz_factor * target;
    int32_t c = 0;
    while (c < ncones) {
        const int32_t s = cone_idx[c], e = cone_idx[c + 1];
        const int32_t sz = e - s;
        if (sz > kTileRows || nnz_of(s, e) > huge_nnz) {
            const int32_t h = (int32_t)P.hstart.size();
            P.hstart.push_back(s);
            P.hend.push_back(e);
            P.hslot0.push_back(P.nslots);
            int32_t a = s, k = 0;
            while (a < e) {
                int32_t b = a + 1;
                int64_t nz = nnz_of(a, b);
                const int64_t htarget = rowptr ? std::max<int64_t>(target, cfg.huge_target) : target;
                while (b < e && b - a < kTileRows && nz + nnz_of(b, b + 1) <= htarget) {
                    nz += nnz_of(b, b + 1);
                    ++b;
                }
                int32_t g = choose_g(nz, b - a, cfg);
                if (g == kBlock && nz < (int64_t)cfg.huge_whole * (b - a)) g = 64;
                Tile t{a, b, TILE_HUGE, g, h, P.nslots + k, 1, inst_of(s)};
                huge.push_back(t);
                ++k;
                a = b;
            }
            P.hnslot.push_back(k);
            P.nslots += k;
            ++c;
            continue;
        }
        // pack whole small cones
        int32_t c1 = c;
        int32_t rows = 0, maxsz = 0;
        int64_t nz = 0;
        while (c1 < ncones) {
            const int32_t s1 = cone_idx[c1], e1 = cone_idx[c1 + 1];
            const int32_t sz1 = e1 - s1;
            if (sz1 > kTileRows || nnz_of(s1, e1) > huge_nnz) break;
            if (c1 > c && (rows + sz1 > kTileRows || nz + nnz_of(s1, e1) > target)) break;
            if (c1 > c && cone_inst(c1) != cone_inst(c)) break;
            rows += sz1;
            nz += nnz_of(s1, e1);
            maxsz = std::max(maxsz, sz1);
            ++c1;
        }
        const int32_t rb = cone_idx[c], re = cone_idx[c1];
        const int32_t g2 = std::min<int32_t>(std::max<int32_t>(pow2ceil(maxsz), 1), 64);
        Tile t{rb, re, TILE_CONES, choose_g(nz, std::max(rows, 1), cfg), c, c1, g2, cone_inst(c)};
        P.tiles.push_back(t);
        c = c1;
    }
    P.first_huge = (int32_t)P.tiles.size();
    P.tiles.insert(P.tiles.end(), huge.begin(), huge.end());
    // rows after the last cone (none in the canonical layout) are covered by lin_end checks upstream
    (void)nrows;
    return P;
}

}  // namespace hpr
