// Conversion of C-ABI descriptors into the host data model.
#include "descriptors.hpp"

namespace mq {

BoundaryIndex boundary_index_from(const mq_bindex_desc& d) {
    require(d.n_cohorts >= 0 && d.aux_dim >= 1, "boundary descriptor: bad sizes");
    require(d.n_cohorts == 0 || (d.cohort_lr && d.offsets), "boundary descriptor: null arrays");
    BoundaryIndex B(Index(d.bra), Index(d.ket), d.aux_dim);
    for (int ci = 0; ci < d.n_cohorts; ++ci) {
        uint32_t lb = d.cohort_lr[2 * ci], rb = d.cohort_lr[2 * ci + 1];
        require(lb < B.bra.size() && rb < B.ket.size(), "boundary descriptor: cohort sector out of range");
        std::vector<int64_t> off(d.offsets + int64_t(ci) * d.aux_dim, d.offsets + int64_t(ci + 1) * d.aux_dim);
        uint32_t c = B.add_cohort(lb, rb, off);
        if (d.conj_scales)
            std::copy(d.conj_scales + int64_t(ci) * d.aux_dim, d.conj_scales + int64_t(ci + 1) * d.aux_dim,
                      B.conj_scales[c].begin());
        if (d.transposes)
            std::copy(d.transposes + int64_t(ci) * d.aux_dim, d.transposes + int64_t(ci + 1) * d.aux_dim,
                      B.transposes[c].begin());
        if (d.tr) B.tr[c] = d.tr[ci];
    }
    // a transposed cohort carries no blocks of its own (reference: complement_transpose)
    for (uint32_t c = 0; c < B.n_cohorts(); ++c)
        if (B.tr[c]) {
            B.nblk[c] = 0;
            require(B.partner(c) < B.n_cohorts() && !B.tr[B.partner(c)],
                    "boundary descriptor: transposed cohort without stored partner");
        }
    B.layout();
    return B;
}

BoundaryIndex initial_boundary_index(Index const& I, uint32_t aux_dim) {
    BoundaryIndex B(I, I, aux_dim);
    for (size_t i = 0; i < I.size(); ++i) {
        uint64_t bs = uint64_t(I.sizes[i]) * I.sizes[i];
        std::vector<int64_t> off(aux_dim);
        for (uint32_t k = 0; k < aux_dim; ++k) off[k] = int64_t(k * bs);
        B.add_cohort(i, i, off);
    }
    B.layout();
    return B;
}

std::vector<ReshapeSegHost> reshape_segments(MPSLayout const& L) {
    ProductBasis rpb = ProductBasis::right_paired(L.phys, L.right);
    ProductBasis lpb = ProductBasis::left_paired(L.phys, L.left);
    std::vector<ReshapeSegHost> segs;
    const uint64_t n = L.elements();
    for (size_t lb = 0; lb < L.left.size(); ++lb)
        for (size_t s = 0; s < L.phys.size(); ++s) {
            Charge rc = L.left.charges[lb] + L.phys.charges[s];
            size_t rb = L.right.position(rc);
            if (rb == L.right.size()) continue;
            ReshapeSegHost g{};
            g.ls = int32_t(L.left.sizes[lb]);
            g.rs = int32_t(L.right.sizes[rb]);
            g.d = int32_t(L.phys.sizes[s]);
            g.lp_ld = int32_t(L.lp_rows[rb]);
            g.rp = int64_t(L.rp_off[lb] + rpb(L.phys.charges[s], rc) * L.left.sizes[lb]);
            g.lp = int64_t(L.lp_off[rb] + lpb(L.phys.charges[s], L.left.charges[lb]));
            if (uint64_t(g.ls) * g.rs * g.d == 0) continue;
            // last element touched on each side stays inside the tensor
            uint64_t rp_last = uint64_t(g.rp) + (g.ls - 1) + uint64_t(g.d * g.rs - 1) * g.ls;
            uint64_t lp_last = uint64_t(g.lp) + uint64_t(g.d * g.ls - 1) + uint64_t(g.rs - 1) * g.lp_ld;
            require(rp_last < n && lp_last < n && uint64_t(g.d) * g.ls <= uint64_t(g.lp_ld),
                    "MPS reshape segment outside the tensor");
            segs.push_back(g);
        }
    return segs;
}

}  // namespace mq
