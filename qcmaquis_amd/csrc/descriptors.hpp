// Conversion of C-ABI descriptors into the host data model (shared by the C ABI and the host-side
// table emulator of the tests).
#pragma once

#include <vector>

#include "core.hpp"

namespace mq {

// boundary index with its arena layout (non-transposed cohorts concatenated in cohort order)
BoundaryIndex boundary_index_from(const mq_bindex_desc& d);
// all-ones edge boundary index (reference: Boundary ctor, boundary.h:257-283)
BoundaryIndex initial_boundary_index(Index const& I, uint32_t aux_dim);

// right-paired <-> left-paired segments (one per (left sector, sigma) with a right sector), each
// checked to stay inside the tensor
struct ReshapeSegHost {
    int64_t rp, lp;
    int32_t ls, rs, d, lp_ld;
};
std::vector<ReshapeSegHost> reshape_segments(MPSLayout const& L);

}  // namespace mq
