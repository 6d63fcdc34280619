// plan.h — host-side tile planner (cone-size-histogram load balancing).
//
// Rows of one operator (A or A^T) are cut into workgroup tiles:
//   * linear rows are packed greedily up to kTileRows rows / nnz_target nonzeros;
//   * cones of at most kTileRows rows (and a bounded nonzero count) are packed
//     whole, so their ||xi|| reduction never leaves the workgroup;
//   * longer cones are split into HUGE segments, reduced in two phases.
// The SpMV lane-group width g is chosen per tile from its mean row length and
// the projection group width g2 from the largest cone of the tile.
#pragma once
#include <stdint.h>

#include <vector>

#include "kernels.h"

namespace hpr {

struct HostPlan {
    std::vector<Tile> tiles;
    int32_t first_huge = 0;
    std::vector<int32_t> hstart, hend, hslot0, hnslot;
    int32_t nslots = 0;
};

struct PlanConfig {
    int32_t nnz_target = 2048;     // nonzeros per tile (SpMV work balance)
    int32_t huge_nnz_factor = 16;  // cones above factor*nnz_target nonzeros are split
    // HUGE segments (rows of a split cone): up to huge_target nonzeros per tile;
    // segments whose rows average between 1024 and huge_whole nonzeros take one
    // wavefront per row (several rows per tile) instead of the whole workgroup
    // per row, which leaves a one-row tile latency-bound
    int32_t huge_target = 16384;
    int32_t huge_whole = 8192;
    int32_t g_divisor = 2;         // g = pow2ceil(mean_row_nnz / g_divisor), clamped to 1..64
    bool stream = true;            // CSR-stream tiles when the tile fits kStreamNnz
};

// rowptr may be null (no SpMV: rows weigh one unit each).  row_inst (null for
// a single problem) gives each row's instance; tiles never straddle two
// instances and carry the instance in Tile::inst.
HostPlan build_plan(const int32_t *rowptr, int32_t nrows, int32_t lin_end, const int32_t *cone_idx,
                    int32_t ncones, const PlanConfig &cfg, const int32_t *row_inst = nullptr);

}  // namespace hpr
