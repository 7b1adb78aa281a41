// Host-side data model of the contraction path: charges, sector index, product basis, MPO tensor
// and boundary index.  Everything here is integer bookkeeping built once per site; the device
// work lists in plans.hpp are derived from it.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "qcmaquis_amd.h"

namespace mq {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
// the message is made into a string only when the check fails (checks sit in loops that run tens
// of millions of times per plan)
inline void require(bool ok, const char* what) {
    if (__builtin_expect(!ok, 0)) throw Error(MQ_ERR_INVALID, what);
}
inline void require(bool ok, const std::string& what) {
    if (!ok) throw Error(MQ_ERR_INVALID, what);
}

enum class Sym : int { TwoU1 = MQ_SYM_TWOU1, SU2U1 = MQ_SYM_SU2U1 };

// Two-component abelian charge, lexicographic order (reference: NU1Charge<2>, nu1_tpl.h:42-296).
struct Charge {
    int32_t q[2] = {0, 0};
    Charge() = default;
    Charge(int32_t a, int32_t b) { q[0] = a; q[1] = b; }
    int32_t operator[](int i) const { return q[i]; }
};
inline bool operator<(Charge a, Charge b) { return a.q[0] < b.q[0] || (a.q[0] == b.q[0] && a.q[1] < b.q[1]); }
inline bool operator>(Charge a, Charge b) { return b < a; }
inline bool operator==(Charge a, Charge b) { return a.q[0] == b.q[0] && a.q[1] == b.q[1]; }
inline bool operator!=(Charge a, Charge b) { return !(a == b); }
inline Charge operator+(Charge a, Charge b) { return Charge(a.q[0] + b.q[0], a.q[1] + b.q[1]); }
inline Charge operator-(Charge a) { return Charge(-a.q[0], -a.q[1]); }
inline Charge operator-(Charge a, Charge b) { return a + (-b); }

// particle number: TwoU1 sums both components (nu1_tpl.h:277), SU2U1 takes N (su2u1.h:49)
inline int particle_number(Sym s, Charge c) { return s == Sym::TwoU1 ? c.q[0] + c.q[1] : c.q[0]; }
// 2S of an SU2U1 charge (su2u1.h:51-52)
inline int spin_of(Charge c) { return c.q[1]; }

struct ChargeHash {
    size_t operator()(Charge c) const {
        return std::hash<int64_t>()((int64_t(c.q[0]) << 32) ^ uint32_t(c.q[1]));
    }
};
struct ChargePairHash {
    size_t operator()(std::pair<Charge, Charge> const& p) const {
        size_t h = ChargeHash()(p.first);
        return h ^ (ChargeHash()(p.second) + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2));
    }
};

// Sector index sorted strictly descending (reference: Index, indexing_stable.hpp:97-305).
class Index {
public:
    std::vector<Charge> charges;
    std::vector<uint32_t> sizes;

    Index() = default;
    explicit Index(const mq_index_desc& d) {
        require(d.n >= 0, "index: negative sector count");
        require(d.n == 0 || (d.charges && d.sizes), "index: null arrays");
        charges.resize(d.n);
        sizes.resize(d.n);
        for (int i = 0; i < d.n; ++i) {
            charges[i] = Charge(d.charges[2 * i], d.charges[2 * i + 1]);
            sizes[i] = d.sizes[i];
        }
        for (int i = 1; i < d.n; ++i)
            require(charges[i - 1] > charges[i], "index: charges must be sorted strictly descending");
        rebuild();
    }
    size_t size() const { return charges.size(); }
    // position of charge c, size() if absent (binary search on the descending order,
    // indexing_stable.hpp:129-140)
    size_t position(Charge c) const {
        auto it = pos_.find(c);
        return it == pos_.end() ? charges.size() : it->second;
    }
    bool has(Charge c) const { return pos_.count(c) > 0; }
    uint64_t sum_of_sizes() const {
        uint64_t s = 0;
        for (auto v : sizes) s += v;
        return s;
    }
    bool operator==(Index const& o) const { return charges == o.charges && sizes == o.sizes; }
    void push_sorted(Charge c, uint32_t s) {
        charges.push_back(c);
        sizes.push_back(s);
    }
    void rebuild() {
        pos_.clear();
        for (size_t i = 0; i < charges.size(); ++i) pos_[charges[i]] = i;
    }

private:
    std::unordered_map<Charge, size_t, ChargeHash> pos_;
};

inline bool operator==(std::vector<Charge> const& a, std::vector<Charge> const& b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i] != b[i]) return false;
    return true;
}

// Offsets of (a, b) sector pairs inside the fused sector f(a, b), a-outer / b-inner insertion
// order (reference: ProductBasis::init, indexing_stable.hpp:331-345).
class ProductBasis {
public:
    // right-paired MPS columns: fuse(-sigma, rc) (site_hamil_schedule.hpp:61-63)
    static ProductBasis right_paired(Index const& phys, Index const& right) {
        ProductBasis pb;
        for (size_t i = 0; i < phys.size(); ++i)
            for (size_t j = 0; j < right.size(); ++j) {
                Charge pc = right.charges[j] - phys.charges[i];
                uint64_t& sz = pb.size_[pc];
                pb.off_[{phys.charges[i], right.charges[j]}] = sz;
                sz += uint64_t(phys.sizes[i]) * right.sizes[j];
            }
        return pb;
    }
    // left-paired MPS rows: fuse(sigma, lc) (move_boundary.hpp:174)
    static ProductBasis left_paired(Index const& phys, Index const& left) {
        ProductBasis pb;
        for (size_t i = 0; i < phys.size(); ++i)
            for (size_t j = 0; j < left.size(); ++j) {
                Charge pc = phys.charges[i] + left.charges[j];
                uint64_t& sz = pb.size_[pc];
                pb.off_[{phys.charges[i], left.charges[j]}] = sz;
                sz += uint64_t(phys.sizes[i]) * left.sizes[j];
            }
        return pb;
    }
    uint64_t operator()(Charge a, Charge b) const {
        auto it = off_.find({a, b});
        require(it != off_.end(), "product basis: missing charge pair");
        return it->second;
    }
    uint64_t size(Charge fused) const {
        auto it = size_.find(fused);
        return it == size_.end() ? 0 : it->second;
    }

private:
    std::unordered_map<Charge, uint64_t, ChargeHash> size_;
    std::unordered_map<std::pair<Charge, Charge>, uint64_t, ChargePairHash> off_;
};

// Block layout of one MPS site tensor in either pairing.  Right-paired block k belongs to left
// sector k (ls x stripe_k); left-paired block k belongs to right sector k (rows_k x rs).
struct MPSLayout {
    Index left, phys, right;
    std::vector<uint64_t> rp_cols;   // right-paired stripe per left sector
    std::vector<uint64_t> rp_off;    // element offset of right-paired block k (size left+1)
    std::vector<uint64_t> lp_rows;   // left-paired rows per right sector
    std::vector<uint64_t> lp_off;    // element offset of left-paired block k (size right+1)

    MPSLayout() = default;
    MPSLayout(Index l, Index p, Index r) : left(std::move(l)), phys(std::move(p)), right(std::move(r)) {
        ProductBasis rpb = ProductBasis::right_paired(phys, right);
        ProductBasis lpb = ProductBasis::left_paired(phys, left);
        rp_cols.resize(left.size());
        rp_off.assign(left.size() + 1, 0);
        for (size_t k = 0; k < left.size(); ++k) {
            rp_cols[k] = rpb.size(left.charges[k]);
            rp_off[k + 1] = rp_off[k] + uint64_t(left.sizes[k]) * rp_cols[k];
        }
        lp_rows.resize(right.size());
        lp_off.assign(right.size() + 1, 0);
        for (size_t k = 0; k < right.size(); ++k) {
            lp_rows[k] = lpb.size(right.charges[k]);
            lp_off[k + 1] = lp_off[k] + lp_rows[k] * uint64_t(right.sizes[k]);
        }
    }
    uint64_t elements() const { return rp_off.back(); }
};

// ---- MPO ------------------------------------------------------------------------------------

struct SparseEntry {
    uint32_t row, col;
    int32_t row_spin, col_spin;
    double coeff;
};
struct OpBlock {
    Charge lc, rc;  // W.basis().left_charge / right_charge
    std::vector<SparseEntry> entries;
};
struct SiteOp {
    int spin = 0;
    std::vector<OpBlock> blocks;
};
struct Term {
    uint32_t op;
    double scale;
};
// Hermitian partners of MPO bond indices (reference: MPOTensor_detail::Hermitian,
// mpotensor_detail.h:139-180)
struct Hermitian {
    std::vector<uint32_t> herm;
    std::vector<int32_t> phase;
    // "b-index complete cohorts": skip b for l < r whenever b has a registered partner
    bool skip(uint32_t b, Charge l, Charge r) const { return l < r && herm[b] != MQ_NO_PARTNER; }
    uint32_t conj(uint32_t b) const { return herm[b]; }
    int32_t ph(uint32_t b) const { return phase[b]; }
    size_t size() const { return herm.size(); }
};

struct MPO {
    uint32_t row_dim = 0, col_dim = 0;
    std::vector<SiteOp> ops;
    // rows[b1] = ascending list of (b2, term range); cols[b2] = ascending list of (b1, term range)
    std::vector<Term> terms;
    struct Cell { uint32_t other; uint32_t t0, t1; };
    std::vector<std::vector<Cell>> rows, cols;
    std::vector<int32_t> left_spin, right_spin;
    Hermitian left_herm, right_herm;

    explicit MPO(const mq_mpo_desc& d) {
        require(d.n_ops >= 0 && d.n_terms >= 0, "mpo: negative sizes");
        ops.resize(d.n_ops);
        for (int o = 0; o < d.n_ops; ++o) {
            ops[o].spin = d.op_spin ? d.op_spin[o] : 0;
            for (int64_t bl = d.op_block_ptr[o]; bl < d.op_block_ptr[o + 1]; ++bl) {
                OpBlock B;
                B.lc = Charge(d.block_charges[4 * bl], d.block_charges[4 * bl + 1]);
                B.rc = Charge(d.block_charges[4 * bl + 2], d.block_charges[4 * bl + 3]);
                for (int64_t e = d.block_entry_ptr[bl]; e < d.block_entry_ptr[bl + 1]; ++e) {
                    SparseEntry E;
                    E.row = d.entry_rc[2 * e];
                    E.col = d.entry_rc[2 * e + 1];
                    E.row_spin = d.entry_spins ? d.entry_spins[2 * e] : 0;
                    E.col_spin = d.entry_spins ? d.entry_spins[2 * e + 1] : 0;
                    E.coeff = d.entry_coeff[e];
                    B.entries.push_back(E);
                }
                ops[o].blocks.push_back(std::move(B));
            }
        }
        row_dim = d.row_dim;
        col_dim = d.col_dim;
        rows.assign(row_dim, {});
        cols.assign(col_dim, {});
        terms.resize(d.n_terms);
        int64_t t = 0;
        int64_t prev_key = -1;
        while (t < d.n_terms) {
            uint32_t b1 = d.term_b[2 * t], b2 = d.term_b[2 * t + 1];
            require(b1 < row_dim && b2 < col_dim, "mpo: term index out of range");
            int64_t key = int64_t(b1) * col_dim + b2;
            require(key > prev_key, "mpo: terms must be sorted by (b1, b2) with each cell contiguous");
            prev_key = key;
            uint32_t t0 = uint32_t(t);
            while (t < d.n_terms && d.term_b[2 * t] == b1 && d.term_b[2 * t + 1] == b2) {
                require(d.term_op[t] < uint32_t(d.n_ops), "mpo: op index out of range");
                terms[t] = Term{d.term_op[t], d.term_scale[t]};
                ++t;
            }
            rows[b1].push_back(Cell{b2, t0, uint32_t(t)});
            cols[b2].push_back(Cell{b1, t0, uint32_t(t)});
        }
        for (auto& c : cols)
            std::sort(c.begin(), c.end(), [](Cell const& a, Cell const& b) { return a.other < b.other; });
        left_spin.assign(row_dim, 0);
        right_spin.assign(col_dim, 0);
        if (d.left_spin) std::copy(d.left_spin, d.left_spin + row_dim, left_spin.begin());
        if (d.right_spin) std::copy(d.right_spin, d.right_spin + col_dim, right_spin.begin());
        left_herm.herm.assign(row_dim, MQ_NO_PARTNER);
        left_herm.phase.assign(row_dim, 1);
        right_herm.herm.assign(col_dim, MQ_NO_PARTNER);
        right_herm.phase.assign(col_dim, 1);
        if (d.left_herm) std::copy(d.left_herm, d.left_herm + row_dim, left_herm.herm.begin());
        if (d.left_phase) std::copy(d.left_phase, d.left_phase + row_dim, left_herm.phase.begin());
        if (d.right_herm) std::copy(d.right_herm, d.right_herm + col_dim, right_herm.herm.begin());
        if (d.right_phase) std::copy(d.right_phase, d.right_phase + col_dim, right_herm.phase.begin());
        for (auto h : left_herm.herm) require(h == MQ_NO_PARTNER || h < row_dim, "mpo: left partner out of range");
        for (auto h : right_herm.herm) require(h == MQ_NO_PARTNER || h < col_dim, "mpo: right partner out of range");
    }
};

// ---- boundary index -------------------------------------------------------------------------

// SU2 factor of a transposed reduced block (reference: detail::conjugate_correction,
// boundary.h:43-69); 1 for abelian groups.
double conjugate_correction(Sym sym, Charge lc, Charge rc);

// Cohort bookkeeping of a boundary (reference: BoundaryIndex, boundary.h:76-234).
struct BoundaryIndex {
    Index bra, ket;
    uint32_t aux_dim = 0;
    std::vector<uint32_t> lbrb;  // bra.size() x ket.size(), row-major, UINT32_MAX = none
    std::vector<std::vector<int64_t>> offsets;
    std::vector<std::vector<double>> conj_scales;
    std::vector<std::vector<uint8_t>> transposes;
    std::vector<uint8_t> tr;
    std::vector<uint32_t> lsz, rsz, nblk;
    std::vector<uint64_t> data_off;  // element offset of each cohort in the boundary's arena
    std::vector<std::pair<uint32_t, uint32_t>> clr;  // (bra position, ket position) per cohort

    BoundaryIndex() = default;
    BoundaryIndex(Index b, Index k, uint32_t aux) : bra(std::move(b)), ket(std::move(k)), aux_dim(aux) {
        lbrb.assign(bra.size() * ket.size(), UINT32_MAX);
    }
    uint32_t n_cohorts() const { return uint32_t(offsets.size()); }
    uint32_t cohort_index(size_t lb, size_t rb) const {
        if (lb < bra.size() && rb < ket.size() && lbrb[lb * ket.size() + rb] != UINT32_MAX)
            return lbrb[lb * ket.size() + rb];
        return n_cohorts();
    }
    uint32_t cohort_index(Charge lc, Charge rc) const { return cohort_index(bra.position(lc), ket.position(rc)); }
    bool has_block(uint32_t ci, uint32_t b) const { return ci < n_cohorts() && offsets[ci][b] > -1; }
    int64_t offset(uint32_t ci, uint32_t b) const { return offsets[ci][b]; }
    double conjugate_scale(uint32_t ci, uint32_t b) const { return conj_scales[ci][b]; }
    uint64_t block_size(uint32_t ci) const { return uint64_t(lsz[ci]) * rsz[ci]; }
    uint64_t cohort_size(uint32_t ci) const { return block_size(ci) * nblk[ci]; }

    uint32_t add_cohort(size_t lb, size_t rb, std::vector<int64_t> const& off) {
        require(cohort_index(lb, rb) == n_cohorts(), "boundary index: duplicate cohort");
        require(off.size() == aux_dim, "boundary index: offsets size != aux_dim");
        uint32_t ci = n_cohorts();
        lbrb[lb * ket.size() + rb] = ci;
        offsets.push_back(off);
        conj_scales.push_back(std::vector<double>(off.size(), 1.0));
        transposes.push_back(std::vector<uint8_t>(off.size(), 0));
        tr.push_back(0);
        clr.push_back({uint32_t(lb), uint32_t(rb)});
        lsz.push_back(bra.sizes[lb]);
        rsz.push_back(ket.sizes[rb]);
        uint32_t nb = 0;
        for (auto o : off) nb += (o > -1);
        nblk.push_back(nb);
        return ci;
    }
    // Fill Hermitian-partner blocks by reference to the transposed partner cohort
    // (reference: BoundaryIndex::complement_transpose, boundary.h:161-191).  The reference writes
    // the partner's offsets into the cohort's own offset table, which is only consistent when the
    // completed cohort has no blocks of its own ("b-index complete cohorts"); that is asserted.
    void complement_transpose(Sym sym, Hermitian const& herm, bool forward) {
        // pass 1 (serial, cohort numbering as in the reference's loop): partner cohorts, new ones
        // appended as transposed aliases with their per-index tables still empty
        require(herm.size() == aux_dim, "complement_transpose: Hermitian table size != aux_dim");
        struct Pair { uint32_t ciA, ciB; uint32_t lb, rb; bool fresh; };
        std::vector<Pair> pairs;
        for (size_t rb = 0; rb < ket.size(); ++rb)
            for (size_t lb = 0; lb < bra.size(); ++lb) {
                uint32_t ciA = lbrb[lb * ket.size() + rb];
                if (ciA == UINT32_MAX) continue;
                uint32_t ciB = cohort_index(ket.charges[rb], bra.charges[lb]);
                bool fresh = false;
                if (ciB == n_cohorts()) {
                    size_t pl = bra.position(ket.charges[rb]), pr = ket.position(bra.charges[lb]);
                    require(pl < bra.size() && pr < ket.size(),
                            "complement_transpose: partner charges missing from the index");
                    ciB = n_cohorts();
                    lbrb[pl * ket.size() + pr] = ciB;
                    offsets.emplace_back();
                    conj_scales.emplace_back();
                    transposes.emplace_back();
                    tr.push_back(1);
                    clr.push_back({uint32_t(pl), uint32_t(pr)});
                    lsz.push_back(bra.sizes[pl]);
                    rsz.push_back(ket.sizes[pr]);
                    nblk.push_back(0);
                    fresh = true;
                }
                pairs.push_back(Pair{ciA, ciB, uint32_t(lb), uint32_t(rb), fresh});
            }
        // pass 2 (in parallel; every pair writes only its own partner cohort): tables of the new
        // cohorts, then the Hermitian-partner blocks by reference.  herm.skip(b, l, r) = l < r
        // and b has a partner, so nothing is filled unless l < r
        std::string err;
#pragma omp parallel for schedule(dynamic, 8)
        for (int64_t q = 0; q < int64_t(pairs.size()); ++q) {
            Pair const& pr = pairs[size_t(q)];
            if (pr.fresh) {
                offsets[pr.ciB].assign(aux_dim, -1);
                conj_scales[pr.ciB].assign(aux_dim, 1.0);
                transposes[pr.ciB].assign(aux_dim, 0);
            }
            const Charge lc = ket.charges[pr.rb], rc = bra.charges[pr.lb];
            if (!(lc < rc)) continue;
            const double cc = conjugate_correction(sym, lc, rc);
            const bool target_ok = pr.fresh || tr[pr.ciB];
            const int64_t* offA = offsets[pr.ciA].data();
            int64_t* offB = offsets[pr.ciB].data();
            double* csB = conj_scales[pr.ciB].data();
            uint8_t* tpB = transposes[pr.ciB].data();
            const char* bad = nullptr;
            for (uint32_t b = 0; b < herm.size() && !bad; ++b) {
                const uint32_t pb = herm.herm[b];
                if (pb == MQ_NO_PARTNER) continue;
                if (offB[b] != -1) bad = "complement_transpose: block already present";
                else if (!target_ok) bad = "complement_transpose: cohort is not b-index complete";
                else {
                    offB[b] = offA[pb];
                    csB[b] = cc * double(herm.phase[forward ? b : pb]);
                    tpB[b] = 1;
                }
            }
            if (bad) {
#pragma omp critical
                err = bad;
            }
        }
        if (!err.empty()) throw Error(MQ_ERR_INVALID, err);
        // as in the reference, n_blocks of a transposed cohort stays 0: consumers always size
        // work from the partner cohort ci_eff
    }
    // Arena layout: non-transposed cohorts own storage, tr cohorts alias their partner's.
    uint64_t layout() {
        data_off.assign(n_cohorts(), 0);
        uint64_t off = 0;
        for (uint32_t ci = 0; ci < n_cohorts(); ++ci) {
            if (tr[ci]) continue;
            data_off[ci] = off;
            off += cohort_size(ci);
        }
        for (uint32_t ci = 0; ci < n_cohorts(); ++ci)
            if (tr[ci]) data_off[ci] = data_off[partner(ci)];
        return off;
    }
    // the cohort with swapped charges
    uint32_t partner(uint32_t ci) const {
        return cohort_index(ket.charges[clr[ci].second], bra.charges[clr[ci].first]);
    }
};

}  // namespace mq
