// Quantum-chemistry Hamiltonian MPOs (the input producer of the contraction path): the flattened
// site tensor shared by the TwoU1 and SU2U1 builders.
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "qcmaquis_amd.h"

namespace mq {

// flattened MPO tensor of one site (the arrays of an mq_mpo_desc)
struct FlatMPO {
    uint32_t row_dim = 1, col_dim = 1;
    std::vector<int32_t> op_spin;
    std::vector<int64_t> op_block_ptr{0};
    std::vector<int32_t> block_charges;
    std::vector<int64_t> block_entry_ptr{0};
    std::vector<uint32_t> entry_rc;
    std::vector<int32_t> entry_spins;  // empty: abelian, or single-site SU2 operators (|2S| of the sectors)
    std::vector<double> entry_coeff;
    std::vector<uint32_t> term_b, term_op;
    std::vector<double> term_scale;
    std::vector<int32_t> lspin, rspin, lphase, rphase;
    std::vector<uint32_t> lherm, rherm;

    void fill_desc(mq_mpo_desc& d) const {
        std::memset(&d, 0, sizeof(d));
        d.n_ops = int32_t(op_spin.size());
        d.op_spin = op_spin.data();
        d.op_block_ptr = op_block_ptr.data();
        d.block_charges = block_charges.data();
        d.block_entry_ptr = block_entry_ptr.data();
        d.entry_rc = entry_rc.data();
        d.entry_spins = entry_spins.empty() ? nullptr : entry_spins.data();
        d.entry_coeff = entry_coeff.data();
        d.row_dim = row_dim;
        d.col_dim = col_dim;
        d.n_terms = int64_t(term_op.size());
        d.term_b = term_b.data();
        d.term_op = term_op.data();
        d.term_scale = term_scale.data();
        d.left_spin = lspin.data();
        d.right_spin = rspin.data();
        d.left_herm = lherm.data();
        d.left_phase = lphase.data();
        d.right_herm = rherm.data();
        d.right_phase = rphase.data();
    }
};

// SU2U1 quantum-chemistry MPO chain (qc_su2.cpp)
struct SU2QC;
struct SU2QCDeleter {
    void operator()(SU2QC* p) const;
};
// gradient_layout: keep every canonical integral's terms (zeros included), single-site terms one
// per operator, and the map from integrals to MPO terms (su2qc_integral_gradient)
std::unique_ptr<SU2QC, SU2QCDeleter> su2qc_build(int L, const double* h1, const double* eri, double cutoff,
                                                 bool gradient_layout = false);
// dE/dh and dE/d(pq|rs) at the canonical integral positions (zero elsewhere) from dE/d(term scale)
// of every site MPO (term_grads[p]: one value per term of site p's MPO)
void su2qc_integral_gradient(SU2QC const& q, std::vector<const double*> const& term_grads, double* dh, double* deri);
int su2qc_length(SU2QC const& q);
std::vector<uint32_t> const& su2qc_bond_dims(SU2QC const& q);
FlatMPO const& su2qc_site(SU2QC& q, int p);
FlatMPO const& su2qc_twosite(SU2QC& q, int p);
// physical index (charges n x 2, sizes) of one site or of two sites
void su2qc_phys(int twosite, std::vector<int32_t>& charges, std::vector<uint32_t>& sizes);

}  // namespace mq
