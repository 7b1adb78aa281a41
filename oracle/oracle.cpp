// CPU FP64 oracle of the symmetry-blocked DMRG contraction path.
//
// TEST INFRASTRUCTURE ONLY: this file is the checker that tests/, the package smoke test and the
// CPU baseline of bench.py compare against.  The product library never links or calls it.
//
// It restates the reference's CPU path independently of the product code:
//   * sector index / product basis     indexing_stable.hpp:97-378
//   * BoundaryIndex, complement_transpose, conjugate_correction   mp_tensors/boundary.h:43-234
//   * schedule builders rshtm_t_tasks / shtm_tasks / rshtm_tasks / lshtm_t_tasks / lshtm_tasks
//                                      contractions/create_schedule/{right,site_hamil,left}_tasks.hpp
//   * Cohort / MPSBlock kernels create_T, create_T_left, create_s, create_s_r, contract, prop_l,
//     prop_r, compute_mpo_offsets     solver/tasks.cpp:198-313, 425-526, 545-705
//   * super_hamil_mv CPU loop (OpenMP dynamic over MPS blocks, per-output mutex)
//                                      solver/super_hamil_mv.hpp:100-155
//   * overlap_mpo_left_step / overlap_mpo_right_step CPU branches
//                                      contractions/boundary_ops/move_boundary.hpp:147-404
//   * SU(2) couplings: GSL-style 9j as a sum over three 6j (coupling.c:427-534), Racah 6j,
//     set_coupling and Wigner9jCache (gsl_coupling.h:38-239)
// BLAS dgemm comes from a dlopen'ed OpenBLAS when oracle_set_blas() is given a library, otherwise
// from a plain triple loop.
#include <dlfcn.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "qcmaquis_amd.h"

namespace orc {

// ---------------------------------------------------------------------------------------------
// BLAS
typedef void (*dgemm_t)(const char*, const char*, const int*, const int*, const int*, const double*, const double*,
                        const int*, const double*, const int*, const double*, double*, const int*);
static dgemm_t g_dgemm = nullptr;
static void (*g_set_threads)(int) = nullptr;

static void naive_gemm(char ta, char tb, int M, int N, int K, double alpha, const double* A, int lda, const double* B,
                       int ldb, double beta, double* C, int ldc) {
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            double s = 0;
            for (int k = 0; k < K; ++k) {
                double a = (ta == 'N') ? A[i + size_t(k) * lda] : A[k + size_t(i) * lda];
                double b = (tb == 'N') ? B[k + size_t(j) * ldb] : B[j + size_t(k) * ldb];
                s += a * b;
            }
            C[i + size_t(j) * ldc] = alpha * s + (beta == 0 ? 0.0 : beta * C[i + size_t(j) * ldc]);
        }
}

static void gemm(char ta, char tb, int M, int N, int K, double alpha, const double* A, int lda, const double* B,
                 int ldb, double beta, double* C, int ldc) {
    if (M <= 0 || N <= 0) return;
    if (K <= 0) {
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < M; ++i) C[i + size_t(j) * ldc] = beta == 0 ? 0.0 : beta * C[i + size_t(j) * ldc];
        return;
    }
    if (g_dgemm) {
        lda = std::max(lda, 1); ldb = std::max(ldb, 1); ldc = std::max(ldc, 1);
        g_dgemm(&ta, &tb, &M, &N, &K, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
    } else {
        naive_gemm(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    }
}

static inline void axpy(const double* first, const double* last, double* out, double a) {
    for (; first != last; ++first, ++out) *out += a * *first;
}

// ---------------------------------------------------------------------------------------------
// charges, index, product basis
struct Q {
    int a = 0, b = 0;
};
static bool operator<(Q x, Q y) { return x.a < y.a || (x.a == y.a && x.b < y.b); }
static bool operator>(Q x, Q y) { return y < x; }
static bool operator==(Q x, Q y) { return x.a == y.a && x.b == y.b; }
static bool operator!=(Q x, Q y) { return !(x == y); }
static Q operator+(Q x, Q y) { return Q{x.a + y.a, x.b + y.b}; }
static Q operator-(Q x) { return Q{-x.a, -x.b}; }

static int g_sym = MQ_SYM_TWOU1;
static bool su2() { return g_sym == MQ_SYM_SU2U1; }
static int spin(Q q) { return q.b; }
static int particle_number(Q q) { return su2() ? q.a : q.a + q.b; }

struct Idx {
    std::vector<std::pair<Q, size_t>> d;
    size_t size() const { return d.size(); }
    std::pair<Q, size_t> const& operator[](size_t i) const { return d[i]; }
    // lower_bound on the descending order
    size_t position(Q c) const {
        auto it = std::lower_bound(d.begin(), d.end(), std::make_pair(c, size_t(0)),
                                   [](std::pair<Q, size_t> const& x, std::pair<Q, size_t> const& y) { return x.first > y.first; });
        if (it != d.end() && it->first != c) it = d.end();
        return size_t(it - d.begin());
    }
    std::vector<size_t> sizes() const {
        std::vector<size_t> r;
        for (auto& p : d) r.push_back(p.second);
        return r;
    }
};
static Idx make_idx(mq_index_desc const& x) {
    Idx I;
    for (int i = 0; i < x.n; ++i) I.d.push_back({Q{x.charges[2 * i], x.charges[2 * i + 1]}, x.sizes[i]});
    return I;
}

struct PB {
    std::map<Q, size_t> size_;
    std::map<std::pair<Q, Q>, size_t> kv;
    // f(a, b) = sign_a * a + b
    PB(Idx const& a, Idx const& b, int sign_a) {
        for (auto& x : a.d)
            for (auto& y : b.d) {
                Q pc = (sign_a > 0 ? x.first : -x.first) + y.first;
                kv[{x.first, y.first}] = size_[pc];
                size_[pc] += x.second * y.second;
            }
    }
    size_t operator()(Q a, Q b) const { return kv.at({a, b}); }
    size_t size(Q c) const {
        auto it = size_.find(c);
        return it == size_.end() ? 0 : it->second;
    }
};

// ---------------------------------------------------------------------------------------------
// SU(2) couplings
static double fact(int n) {
    static std::vector<double> t = [] {
        std::vector<double> v(171, 1.0);
        for (int i = 1; i <= 170; ++i) v[i] = v[i - 1] * i;
        return v;
    }();
    return t.at(n);
}
static bool tri_fail(int a, int b, int c) { return b < std::abs(a - c) || b > a + c || ((a + b + c) % 2); }
static double sixj(int ja, int jb, int jc, int jd, int je, int jf) {
    if (ja < 0 || jb < 0 || jc < 0 || jd < 0 || je < 0 || jf < 0) return std::nan("");
    if (tri_fail(ja, jb, jc) || tri_fail(ja, je, jf) || tri_fail(jb, jd, jf) || tri_fail(je, jd, jc)) return 0.0;
    auto delta = [](int a, int b, int c) {
        return std::sqrt(fact((a + b - c) / 2) * fact((a - b + c) / 2) * fact((-a + b + c) / 2) / fact((a + b + c) / 2 + 1));
    };
    int a1 = (ja + jb + jc) / 2, a2 = (ja + je + jf) / 2, a3 = (jd + jb + jf) / 2, a4 = (jd + je + jc) / 2;
    int b1 = (ja + jb + jd + je) / 2, b2 = (jb + jc + je + jf) / 2, b3 = (jc + ja + jf + jd) / 2;
    double s = 0;
    for (int z = std::max(std::max(a1, a2), std::max(a3, a4)); z <= std::min(b1, std::min(b2, b3)); ++z)
        s += ((z % 2) ? -1.0 : 1.0) * fact(z + 1) /
             (fact(z - a1) * fact(z - a2) * fact(z - a3) * fact(z - a4) * fact(b1 - z) * fact(b2 - z) * fact(b3 - z));
    return delta(ja, jb, jc) * delta(ja, je, jf) * delta(jd, jb, jf) * delta(jd, je, jc) * s;
}
// the three-6j expansion in the argument order of the GSL extract (coupling.c:457-479)
static double ninej(int ja, int jb, int jc, int jd, int je, int jf, int jg, int jh, int ji) {
    if (ja < 0 || jb < 0 || jc < 0 || jd < 0 || je < 0 || jf < 0 || jg < 0 || jh < 0 || ji < 0) return std::nan("");
    if (tri_fail(ja, jb, jc) || tri_fail(jd, je, jf) || tri_fail(jg, jh, ji) || tri_fail(ja, jd, jg) ||
        tri_fail(jb, je, jh) || tri_fail(jc, jf, ji))
        return 0.0;
    int tkmin = std::max(std::abs(ja - ji), std::max(std::abs(jh - jd), std::abs(jb - jf)));
    int tkmax = std::min(ja + ji, std::min(jh + jd, jb + jf));
    double sum = 0;
    for (int tk = tkmin; tk <= tkmax; tk += 2)
        sum += (tk + 1) * sixj(ja, ji, tk, jh, jd, jg) * sixj(jb, jf, tk, jd, jh, je) * sixj(ja, ji, tk, jf, jb, jc);
    return ((tkmin % 2) ? -1.0 : 1.0) * sum;
}
static double mod_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    return std::sqrt((g + 1.) * (h + 1.) * (c + 1.) * (f + 1.)) * ninej(a, b, c, d, e, f, g, h, i);
}
static bool triangle(int a, int b, int c) { return std::abs(a - b) <= c && c <= a + b; }
static void set_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double cp[4]) {
    cp[0] = cp[1] = cp[2] = cp[3] = 0.0;
    double pre = std::sqrt((i + 1.) * (a + 1.) / ((g + 1.) * (c + 1.))) * init;
    if (triangle(a, b, c)) {
        cp[0] = pre * mod_coupling(a, b, c, d, e, f, g, h, i);
        cp[2] = pre * mod_coupling(a, b, c, d, e, f, g, 2, i);
    }
    if (triangle(a, 2, c)) {
        cp[1] = pre * mod_coupling(a, 2, c, d, e, f, g, h, i);
        cp[3] = pre * mod_coupling(a, 2, c, d, e, f, g, 2, i);
    }
}
struct W9j {
    int Jpmin = 0, Jpmax = 0;
    std::vector<double> coeff;
    bool active = false;
    W9j(Q lc, Q mc, Q rc) {
        if (!su2()) return;
        active = true;
        int J = spin(mc), I = spin(lc), Ip = spin(rc);
        int two_sp = std::abs(I - Ip);
        Jpmax = J + 2;
        Jpmin = std::max(J - 2, 0);
        coeff.assign(48 * (Jpmax - Jpmin + 1), 0.0);
        int K[12] = {0, 1, 2, 0, 0, 1, 2, 1, 0, 1, 2, 0};
        int Ap[12] = {0, 1, 2, 0, 1, 0, 1, 2, 2, 1, 0, 0};
        for (int Jp = Jpmin; Jp <= Jpmax; ++Jp)
            for (int i = 0; i < 12; ++i)
                set_coupling(J, std::abs(J - Jp), Jp, i / 4, K[i], Ap[i], I, two_sp, Ip, 1.0,
                             &coeff[4 * i + 48 * (Jp - Jpmin)]);
    }
    void set_scale(int A, int K, int Ap, Q jp, double scale, double cp[4]) const {
        if (!active) { cp[0] = cp[1] = cp[2] = cp[3] = scale; return; }
        int h = 4 * A + K + ((A == 1 && K == 1) ? Ap : 0);
        int Jp = spin(jp);
        if (Jp < Jpmin || Jp > Jpmax || h < 0 || h >= 12) throw std::runtime_error("oracle: 9j cache lookup out of range");
        for (int k = 0; k < 4; ++k) cp[k] = coeff[4 * h + 48 * (Jp - Jpmin) + k] * scale;
    }
};
static double conjugate_correction(Q lc, Q rc) {
    if (!su2()) return 1.0;
    int S = std::min(spin(rc), spin(lc));
    switch (spin(rc) - spin(lc)) {
        case 0: return 1.0;
        case 1: return -std::sqrt((S + 1.) / (S + 2.));
        case -1: return std::sqrt((S + 2.) / (S + 1.));
        case 2: return -std::sqrt((S + 1.) / (S + 3.));
        case -2: return -std::sqrt((S + 3.) / (S + 1.));
    }
    throw std::runtime_error("oracle: conjugate correction beyond rank 1");
}

// ---------------------------------------------------------------------------------------------
// MPO
struct Entry { unsigned row, col; int rspin, cspin; double coeff; };
struct Op { int spin; std::vector<std::pair<Q, Q>> basis; std::vector<std::vector<Entry>> blocks; };
struct Herm {
    std::vector<unsigned> h;
    std::vector<int> ph;
    bool skip(unsigned b, Q l, Q r) const { return l < r && h[b] != std::numeric_limits<unsigned>::max(); }
};
struct Mpo {
    unsigned rows, cols;
    std::vector<Op> ops;
    std::map<std::pair<unsigned, unsigned>, std::vector<std::pair<unsigned, double>>> cells;
    std::vector<std::vector<unsigned>> row_idx, col_idx;  // ascending
    std::vector<int> lspin, rspin;
    Herm lh, rh;
};
static Mpo make_mpo(mq_mpo_desc const& d) {
    Mpo m;
    m.rows = d.row_dim;
    m.cols = d.col_dim;
    for (int o = 0; o < d.n_ops; ++o) {
        Op op;
        op.spin = d.op_spin ? d.op_spin[o] : 0;
        for (int64_t bl = d.op_block_ptr[o]; bl < d.op_block_ptr[o + 1]; ++bl) {
            op.basis.push_back({Q{d.block_charges[4 * bl], d.block_charges[4 * bl + 1]},
                                Q{d.block_charges[4 * bl + 2], d.block_charges[4 * bl + 3]}});
            std::vector<Entry> es;
            for (int64_t e = d.block_entry_ptr[bl]; e < d.block_entry_ptr[bl + 1]; ++e)
                es.push_back(Entry{d.entry_rc[2 * e], d.entry_rc[2 * e + 1], d.entry_spins ? d.entry_spins[2 * e] : 0,
                                   d.entry_spins ? d.entry_spins[2 * e + 1] : 0, d.entry_coeff[e]});
            op.blocks.push_back(es);
        }
        m.ops.push_back(op);
    }
    m.row_idx.resize(m.rows);
    m.col_idx.resize(m.cols);
    for (int64_t t = 0; t < d.n_terms; ++t) {
        unsigned b1 = d.term_b[2 * t], b2 = d.term_b[2 * t + 1];
        auto& cell = m.cells[{b1, b2}];
        if (cell.empty()) { m.row_idx[b1].push_back(b2); m.col_idx[b2].push_back(b1); }
        cell.push_back({d.term_op[t], d.term_scale[t]});
    }
    for (auto& r : m.row_idx) std::sort(r.begin(), r.end());
    for (auto& c : m.col_idx) std::sort(c.begin(), c.end());
    m.lspin.assign(m.rows, 0);
    m.rspin.assign(m.cols, 0);
    if (d.left_spin) std::copy(d.left_spin, d.left_spin + m.rows, m.lspin.begin());
    if (d.right_spin) std::copy(d.right_spin, d.right_spin + m.cols, m.rspin.begin());
    m.lh.h.assign(m.rows, std::numeric_limits<unsigned>::max());
    m.lh.ph.assign(m.rows, 1);
    m.rh.h.assign(m.cols, std::numeric_limits<unsigned>::max());
    m.rh.ph.assign(m.cols, 1);
    if (d.left_herm) std::copy(d.left_herm, d.left_herm + m.rows, m.lh.h.begin());
    if (d.left_phase) std::copy(d.left_phase, d.left_phase + m.rows, m.lh.ph.begin());
    if (d.right_herm) std::copy(d.right_herm, d.right_herm + m.cols, m.rh.h.begin());
    if (d.right_phase) std::copy(d.right_phase, d.right_phase + m.cols, m.rh.ph.begin());
    return m;
}

// ---------------------------------------------------------------------------------------------
// boundary
struct BIdx {
    Idx bra, ket;
    std::vector<unsigned> lbrb;
    std::vector<std::vector<long>> offsets;
    std::vector<std::vector<double>> cscale;
    std::vector<std::vector<char>> transp;
    std::vector<char> tr;
    std::vector<size_t> lsz, rsz;
    std::vector<unsigned> nb;
    std::vector<std::pair<unsigned, unsigned>> lr;
    unsigned aux = 0;
    static constexpr unsigned NONE = std::numeric_limits<unsigned>::max();
    BIdx() {}
    BIdx(Idx const& b, Idx const& k, unsigned a) : bra(b), ket(k), lbrb(b.size() * k.size(), NONE), aux(a) {}
    unsigned n_cohorts() const { return unsigned(offsets.size()); }
    unsigned cohort_index(size_t l, size_t r) const {
        if (l < bra.size() && r < ket.size() && lbrb[l * ket.size() + r] != NONE) return lbrb[l * ket.size() + r];
        return n_cohorts();
    }
    unsigned ci(Q l, Q r) const { return cohort_index(bra.position(l), ket.position(r)); }
    bool has_block(unsigned c, unsigned b) const { return c < n_cohorts() && offsets[c][b] > -1; }
    size_t block_size(unsigned c) const { return lsz[c] * rsz[c]; }
    size_t cohort_size(unsigned c) const { return nb[c] * block_size(c); }
    unsigned add_cohort(size_t l, size_t r, std::vector<long> const& off) {
        unsigned c = n_cohorts();
        lbrb[l * ket.size() + r] = c;
        offsets.push_back(off);
        cscale.push_back(std::vector<double>(off.size(), 1.0));
        transp.push_back(std::vector<char>(off.size(), 0));
        tr.push_back(0);
        lsz.push_back(bra[l].second);
        rsz.push_back(ket[r].second);
        nb.push_back(unsigned(std::count_if(off.begin(), off.end(), [](long o) { return o > -1; })));
        lr.push_back({unsigned(l), unsigned(r)});
        return c;
    }
    void complement_transpose(Herm const& herm, bool forward) {
        for (size_t rb = 0; rb < ket.size(); ++rb)
            for (size_t lb = 0; lb < bra.size(); ++lb) {
                if (lbrb[lb * ket.size() + rb] == NONE) continue;
                unsigned A = lbrb[lb * ket.size() + rb];
                unsigned B = ci(ket[rb].first, bra[lb].first);
                if (B == n_cohorts()) {
                    B = add_cohort(bra.position(ket[rb].first), ket.position(bra[lb].first), std::vector<long>(herm.h.size(), -1));
                    tr[B] = 1;
                }
                for (unsigned b = 0; b < herm.h.size(); ++b)
                    if (herm.skip(b, ket[rb].first, bra[lb].first)) {
                        if (offsets[B][b] != -1) throw std::runtime_error("oracle: complement_transpose overwrite");
                        offsets[B][b] = offsets[A][herm.h[b]];
                        cscale[B][b] = conjugate_correction(ket[rb].first, bra[lb].first) *
                                       double(herm.ph[forward ? b : herm.h[b]]);
                        transp[B][b] = 1;
                    }
            }
    }
};
struct Bnd {
    BIdx idx;
    std::vector<std::vector<double>> data;  // per cohort (empty for transposed cohorts)
    std::vector<const double*> view() const {
        std::vector<const double*> v;
        for (auto& d : data) v.push_back(d.data());
        return v;
    }
};
static BIdx make_bidx(mq_bindex_desc const& d) {
    BIdx B(make_idx(d.bra), make_idx(d.ket), d.aux_dim);
    for (int c = 0; c < d.n_cohorts; ++c) {
        std::vector<long> off(d.offsets + size_t(c) * d.aux_dim, d.offsets + size_t(c + 1) * d.aux_dim);
        unsigned k = B.add_cohort(d.cohort_lr[2 * c], d.cohort_lr[2 * c + 1], off);
        for (unsigned b = 0; b < d.aux_dim; ++b) {
            if (d.conj_scales) B.cscale[k][b] = d.conj_scales[size_t(c) * d.aux_dim + b];
            if (d.transposes) B.transp[k][b] = d.transposes[size_t(c) * d.aux_dim + b];
        }
        if (d.tr) B.tr[k] = d.tr[c];
        if (B.tr[k]) B.nb[k] = 0;
    }
    return B;
}
static Bnd make_bnd(mq_bindex_desc const& d, const double* data) {
    Bnd b;
    b.idx = make_bidx(d);
    size_t off = 0;
    b.data.resize(b.idx.n_cohorts());
    for (unsigned c = 0; c < b.idx.n_cohorts(); ++c) {
        if (b.idx.tr[c]) continue;
        size_t n = b.idx.cohort_size(c);
        b.data[c].assign(data + off, data + off + n);
        off += n;
    }
    return b;
}

// ---------------------------------------------------------------------------------------------
// Cohort / MPSBlock
struct SUnit {
    unsigned offset = 0, ms = 0;
    std::vector<unsigned> tidx;
    std::vector<double> alpha;
    std::vector<unsigned> b2s, b1;
    unsigned b2count = 0;
    void push_back(double s, unsigned ti, unsigned col) { alpha.push_back(s); ++b2count; tidx.push_back(ti); tidx.push_back(col); }
    unsigned add_line(unsigned b) {
        unsigned r = b2count;
        if (b2count) { b2s.push_back(b2count); b1.push_back(b); b2count = 0; }
        return r;
    }
};
struct Cohort {
    unsigned lb, rb, ls, rs, ci, ci_eff, sblock, nSrows, stripe = 0;
    std::vector<long> mpo_offsets;
    std::vector<unsigned> sfold;
    std::vector<SUnit> suv;
    Cohort(std::vector<size_t> const& phys, unsigned l, unsigned r, unsigned lsize, unsigned rsize, unsigned c,
           unsigned ce, unsigned mpodim, bool left = false)
        : lb(l), rb(r), ls(lsize), rs(rsize), ci(c), ci_eff(ce), nSrows(mpodim), mpo_offsets(mpodim, 0),
          sfold(phys.size()) {
        unsigned s = 0;
        for (size_t i = 0; i < phys.size(); ++i) { sfold[i] = s; s += unsigned(phys[i]); }
        suv.resize(s);
        sblock = left ? rs : ls;
    }
    void push_back(unsigned s, unsigned ss2, double sc, unsigned ti, unsigned col) { suv[sfold[s] + ss2].push_back(sc, ti, col); }
    void add_line(unsigned b1) { for (auto& u : suv) mpo_offsets[b1] += u.add_line(b1); }
    void add_unit(unsigned s, unsigned ss, unsigned msz, unsigned off) {
        for (unsigned i = 0; i < ss; ++i) { suv[sfold[s] + i].offset = i * msz + off; suv[sfold[s] + i].ms = msz; stripe += msz; }
    }
    size_t n_tasks() const { size_t n = 0; for (auto& u : suv) n += u.alpha.size(); return n; }
    size_t n_flops() const {
        size_t r = 2ull * rs * ls * nSrows * stripe;
        for (auto& x : suv) r += 2ull * ls * x.ms * x.alpha.size();
        return r;
    }
    void finalize() {
        std::vector<long> cpy = mpo_offsets;
        unsigned n = 0;
        for (auto& b : cpy) if (b) b = n++;
        nSrows = n;
        for (auto& x : suv) for (auto& b : x.b1) b = unsigned(cpy[b]);
        size_t bs = size_t(ls) * rs;
        unsigned cnt = 0;
        for (auto& b : mpo_offsets) b = b ? long(bs * cnt++) : -1;
    }
    size_t S_size() const { return size_t(nSrows) * stripe * sblock; }
    // left S: stripe x nSrows*rs   (tasks.cpp:425-454)
    std::vector<double> create_s(std::vector<std::vector<double>> const& T) const {
        std::vector<double> ret(S_size(), 0.0);
        for (auto const& x : suv) {
            if (x.alpha.empty()) continue;
            std::vector<double> buf(size_t(x.ms) * rs);
            unsigned seek = 0;
            for (size_t b = 0; b < x.b2s.size(); ++b) {
                std::fill(buf.begin(), buf.end(), 0.0);
                for (unsigned ia = seek; ia < seek + x.b2s[b]; ++ia) {
                    const double* src = &T[x.tidx[2 * ia]][size_t(x.tidx[2 * ia + 1]) * rs];
                    axpy(src, src + size_t(x.ms) * rs, buf.data(), x.alpha[ia]);
                }
                unsigned bb = x.b1[b];
                for (unsigned c = 0; c < rs; ++c)
                    std::copy(&buf[size_t(c) * x.ms], &buf[size_t(c) * x.ms] + x.ms, ret.data() + size_t(stripe) * (size_t(bb) * rs + c) + x.offset);
                seek += x.b2s[b];
            }
        }
        return ret;
    }
    // right S: nSrows*ls x stripe   (tasks.cpp:466-494)
    std::vector<double> create_s_r(std::vector<std::vector<double>> const& T) const {
        std::vector<double> ret(S_size(), 0.0);
        for (auto const& x : suv) {
            if (x.alpha.empty()) continue;
            std::vector<double> buf(size_t(ls) * x.ms);
            unsigned seek = 0;
            for (size_t b = 0; b < x.b2s.size(); ++b) {
                std::fill(buf.begin(), buf.end(), 0.0);
                for (unsigned ia = seek; ia < seek + x.b2s[b]; ++ia) {
                    const double* src = &T[x.tidx[2 * ia]][size_t(x.tidx[2 * ia + 1]) * ls];
                    axpy(src, src + size_t(x.ms) * ls, buf.data(), x.alpha[ia]);
                }
                unsigned bb = x.b1[b];
                for (unsigned c = 0; c < x.ms; ++c)
                    std::copy(&buf[size_t(c) * ls], &buf[size_t(c) * ls] + ls, ret.data() + size_t(nSrows) * ls * (x.offset + c) + size_t(bb) * ls);
                seek += x.b2s[b];
            }
        }
        return ret;
    }
    void prop_l(const double* bra, std::vector<std::vector<double>> const& T, double* out) const {
        auto s = create_s(T);
        gemm('T', 'N', int(ls), int(nSrows * rs), int(stripe), 1.0, bra, int(stripe), s.data(), int(stripe), 0.0, out, int(ls));
    }
    void prop_r(const double* bra, std::vector<std::vector<double>> const& T, double* out) const {
        auto s = create_s_r(T);
        int M = int(nSrows * ls), N = int(rs);
        std::vector<double> buf(size_t(M) * N);
        gemm('N', 'T', M, N, int(stripe), 1.0, s.data(), M, bra, int(rs), 0.0, buf.data(), M);
        for (unsigned b = 0; b < nSrows; ++b)
            for (unsigned col = 0; col < rs; ++col)
                std::copy(&buf[size_t(ls) * b + size_t(col) * M], &buf[size_t(ls) * b + size_t(col) * M] + ls,
                          out + (size_t(b) * rs + col) * ls);
    }
    void contract(std::vector<const double*> const& left, std::vector<std::vector<double>> const& T, double* out,
                  std::mutex& mtx) const {
        auto s = create_s_r(T);
        int M = int(rs), N = int(stripe), K = int(nSrows * ls);
        const double* luse = left[ci_eff];
        std::vector<double> lbuf;
        if (ci != ci_eff) {  // tr_tile_v (tasks.cpp:49-62)
            lbuf.resize(size_t(M) * K);
            for (size_t b = 0; b < nSrows; ++b) {
                size_t o = b * size_t(ls) * rs;
                for (unsigned i = 0; i < ls; ++i)
                    for (unsigned j = 0; j < rs; ++j) lbuf[o + size_t(rs) * i + j] = left[ci_eff][o + size_t(ls) * j + i];
            }
            luse = lbuf.data();
        }
        std::vector<double> buf(size_t(M) * N);
        gemm('N', 'N', M, N, K, 1.0, luse, M, s.data(), K, 0.0, buf.data(), M);
        std::lock_guard<std::mutex> lk(mtx);
        axpy(buf.data(), buf.data() + buf.size(), out, 1.0);
    }
};

struct TEntry { unsigned mps_offset, ci, ci_eff, lb_ket; size_t sza; };
struct MPSBlock {
    std::vector<Cohort> data;
    std::vector<TEntry> t;
    unsigned rb_ket = 0;
    unsigned get_ti(unsigned off, unsigned ci) const {
        for (unsigned i = 0; i < t.size(); ++i)
            if (t[i].mps_offset == off && t[i].ci == ci) return i;
        return std::numeric_limits<unsigned>::max();
    }
    // T = psi slice * R  (tasks.cpp:650-705)
    std::vector<std::vector<double>> create_T(BIdx const& R, std::vector<const double*> const& right,
                                              std::vector<const double*> const& mps, std::vector<size_t> const& lr_sizes) const {
        std::vector<std::vector<double>> ret(t.size());
        for (size_t ti = 0; ti < t.size(); ++ti) {
            auto& e = t[ti];
            int bls = int(R.lsz[e.ci]), brs = int(R.rsz[e.ci]);
            int M = int(lr_sizes[e.lb_ket]), K = bls;
            const double* mpsdata = mps[e.lb_ket] + size_t(M) * e.mps_offset;
            ret[ti].assign(size_t(M) * R.nb[e.ci_eff] * brs, 0.0);
            for (unsigned b = 0; b < R.nb[e.ci_eff]; ++b) {
                int N = brs;
                size_t roff = size_t(b) * K * N, ooff = size_t(b) * M * N;
                if (e.ci != e.ci_eff)
                    gemm('N', 'T', M, N, K, 1.0, mpsdata, M, right[e.ci_eff] + roff, N, 0.0, ret[ti].data() + ooff, M);
                else
                    gemm('N', 'N', M, N, K, 1.0, mpsdata, M, right[e.ci_eff] + roff, K, 0.0, ret[ti].data() + ooff, M);
            }
        }
        return ret;
    }
    // T^b = L_b * psi slice  (tasks.cpp:545-601)
    std::vector<std::vector<double>> create_T_left(BIdx const& L, std::vector<const double*> const& left,
                                                   std::vector<const double*> const& mps, std::vector<size_t> const& lr_sizes) const {
        std::vector<std::vector<double>> ret(t.size());
        for (size_t ti = 0; ti < t.size(); ++ti) {
            auto& e = t[ti];
            int M = int(L.lsz[e.ci]), K = int(L.rsz[e.ci]), N = int(lr_sizes[rb_ket]);
            unsigned nb = L.nb[e.ci_eff];
            const double* mpsdata = mps[e.lb_ket] + size_t(K) * e.mps_offset;
            ret[ti].assign(size_t(M) * N * nb, 0.0);
            for (unsigned b = 0; b < nb; ++b) {
                size_t loff = size_t(b) * M * K, ooff = size_t(b) * M * N;
                if (e.ci != e.ci_eff)
                    gemm('T', 'N', M, N, K, 1.0, left[e.ci_eff] + loff, K, mpsdata, K, 0.0, ret[ti].data() + ooff, M);
                else
                    gemm('N', 'N', M, N, K, 1.0, left[e.ci_eff] + loff, M, mpsdata, K, 0.0, ret[ti].data() + ooff, M);
            }
        }
        return ret;
    }
};

static bool tri(Q a, int b, Q c) { return !su2() || triangle(spin(a), b, spin(c)); }

static void op_iterate(Op const& W, size_t wb, double cp[4], Cohort& cg, unsigned s, unsigned m2, MPSBlock const& mb,
                       unsigned mps_offset, unsigned ci, unsigned col) {
    for (auto const& e : W.blocks[wb]) {
        int cn = 0;
        if (e.rspin == 2 && e.cspin == 2) cn = 3;
        else if (e.rspin == 2) cn = 1;
        else if (e.cspin == 2) cn = 2;
        unsigned ti = mb.get_ti(mps_offset + e.row * m2, ci);
        if (ti < std::numeric_limits<unsigned>::max()) cg.push_back(s, e.col, e.coeff * cp[cn], ti, col);
    }
}

// rshtm_t_tasks (right_tasks.hpp:39-78)
static void rshtm_t(BIdx const& R, Idx const& left_i, Idx const& right_i, Idx const& phys, PB const& pb, unsigned lb_ket,
                    MPSBlock& mb, Idx const& bra_right) {
    Q lc = left_i[lb_ket].first;
    for (unsigned s = 0; s < phys.size(); ++s) {
        Q rc = lc + phys[s].first;
        size_t rb = right_i.position(rc);
        if (rb == right_i.size()) continue;
        unsigned rs = unsigned(right_i[rb].second);
        unsigned off = unsigned(pb(phys[s].first, rc));
        for (unsigned rb_bra = 0; rb_bra < bra_right.size(); ++rb_bra) {
            unsigned ci = R.ci(rc, bra_right[rb_bra].first);
            if (ci == R.n_cohorts()) continue;
            unsigned ce = R.tr[ci] ? R.ci(bra_right[rb_bra].first, rc) : ci;
            for (unsigned ss = 0; ss < phys[s].second; ++ss) mb.t.push_back(TEntry{off + ss * rs, ci, ce, lb_ket, 0});
        }
    }
}

}  // namespace orc

using namespace orc;

struct oracle_boundary {
    Bnd b;
};

extern "C" {

int oracle_set_blas(const char* path, int nthreads) {
    if (!path) { g_dgemm = nullptr; return 0; }
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    g_dgemm = (dgemm_t)dlsym(h, "scipy_dgemm_");
    if (!g_dgemm) g_dgemm = (dgemm_t)dlsym(h, "dgemm_");
    g_set_threads = (void (*)(int))dlsym(h, "scipy_openblas_set_num_threads");
    if (!g_set_threads) g_set_threads = (void (*)(int))dlsym(h, "openblas_set_num_threads");
    if (g_set_threads && nthreads > 0) g_set_threads(nthreads);
    return g_dgemm ? 0 : -2;
}

int oracle_blas_threads(int n) {
    if (g_set_threads) { g_set_threads(n); return 0; }
    return -1;
}

// Wigner 9j / set_coupling exposed for the coupling tests
double oracle_ninej(int a, int b, int c, int d, int e, int f, int g, int h, int i) { return ninej(a, b, c, d, e, f, g, h, i); }
double oracle_sixj(int a, int b, int c, int d, int e, int f) { return sixj(a, b, c, d, e, f); }

// H*psi on the CPU: schedule (create_contraction_schedule) + super_hamil_mv's OpenMP branch.
// psi / out: right-paired blocks of left_i concatenated without padding.
// lb_mask (may be NULL): only the input blocks lb_in with lb_mask[lb_in] != 0 are scheduled and
// contracted (a bounded sample of the H*psi); flops_out then counts those blocks' flops and
// seconds_out (may be NULL) receives the time of the contraction loop alone.
int oracle_hpsi_sample(int sym, const mq_mpo_desc* mpod, const mq_bindex_desc* Ld, const double* Ldata,
                       const mq_bindex_desc* Rd, const double* Rdata, const mq_index_desc* li, const mq_index_desc* pi,
                       const mq_index_desc* ri, const double* psi, double* out, int nthreads, double* flops_out,
                       const unsigned char* lb_mask, double* seconds_out) {
    try {
        g_sym = sym;
        Mpo mpo = make_mpo(*mpod);
        Bnd L = make_bnd(*Ld, Ldata), R = make_bnd(*Rd, Rdata);
        Idx left_i = make_idx(*li), phys = make_idx(*pi), right_i = make_idx(*ri);
        PB pb(phys, right_i, -1);
        PB pbsz = pb;
        std::vector<size_t> phys_s = phys.sizes();
        size_t nl = left_i.size();
        std::vector<MPSBlock> sched(nl);
        std::vector<size_t> bsz(nl), boff(nl + 1, 0);
        for (size_t b = 0; b < nl; ++b) {
            bsz[b] = left_i[b].second * pbsz.size(left_i[b].first);
            boff[b + 1] = boff[b] + bsz[b];
        }
        std::vector<size_t> lrs = left_i.sizes();
        BIdx const& LI = L.idx;
        BIdx const& RI = R.idx;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        for (long lb_in = 0; lb_in < long(nl); ++lb_in) {
            if (lb_mask && !lb_mask[lb_in]) continue;
            MPSBlock& mb = sched[lb_in];
            rshtm_t(RI, left_i, right_i, phys, pb, unsigned(lb_in), mb, right_i);
            // shtm_tasks (site_hamil_tasks.hpp:41-137)
            Q lc_in = left_i[lb_in].first;
            unsigned ls_in = unsigned(left_i[lb_in].second);
            for (unsigned lb_out = 0; lb_out < nl; ++lb_out) {
                Q lc_out = left_i[lb_out].first;
                unsigned ls_out = unsigned(left_i[lb_out].second);
                unsigned ci = LI.ci(lc_out, lc_in);
                if (ci == LI.n_cohorts()) continue;
                unsigned ce = LI.tr[ci] ? LI.ci(lc_in, lc_out) : ci;
                Cohort coh(phys_s, unsigned(lb_in), lb_out, ls_in, ls_out, ci, ce, LI.nb[ce]);
                for (unsigned s = 0; s < phys.size(); ++s) {
                    Q po = phys[s].first;
                    Q rc_out = lc_out + po;
                    size_t rb_out = right_i.position(rc_out);
                    if (rb_out == right_i.size()) continue;
                    unsigned rs_out = unsigned(right_i[rb_out].second);
                    coh.add_unit(s, unsigned(phys[s].second), rs_out, unsigned(pb(po, rc_out)));
                    W9j w9j(lc_out, lc_in, rc_out);
                    for (unsigned b1 = 0; b1 < mpo.rows; ++b1) {
                        if (!LI.has_block(ci, b1)) continue;
                        unsigned left_idx = unsigned(LI.offsets[ci][b1] / (long(ls_in) * ls_out));
                        int A = mpo.lspin[b1];
                        if (!tri(lc_in, A, lc_out)) continue;
                        for (unsigned b2 : mpo.row_idx[b1]) {
                            auto const& cell = mpo.cells.at({b1, b2});
                            for (auto const& term : cell) {
                                Op const& W = mpo.ops[term.first];
                                int K = W.spin, Ap = mpo.rspin[b2];
                                for (size_t wb = 0; wb < W.basis.size(); ++wb) {
                                    if (po != W.basis[wb].second) continue;
                                    Q pin = W.basis[wb].first;
                                    Q rc_in = lc_in + pin;
                                    unsigned cr = RI.ci(rc_in, rc_out);
                                    if (!RI.has_block(cr, b2)) continue;
                                    size_t rb_in = right_i.position(rc_in);
                                    if (rb_in == right_i.size()) continue;
                                    unsigned rs_in = unsigned(right_i[rb_in].second);
                                    double cp[4];
                                    double scale = RI.cscale[cr][b2] * term.second * LI.cscale[ci][b1];
                                    w9j.set_scale(A, K, Ap, rc_in, scale, cp);
                                    op_iterate(W, wb, cp, coh, s, rs_in, mb, unsigned(pb(pin, rc_in)), cr,
                                               unsigned(RI.offsets[cr][b2] / rs_in));
                                }
                            }
                        }
                        coh.add_line(left_idx);
                    }
                }
                if (coh.n_tasks()) mb.data.push_back(std::move(coh));
            }
        }
        // flops (ScheduleNew::total_flops)
        double fl = 0;
        for (size_t b = 0; b < nl; ++b) {
            for (auto& e : sched[b].t) fl += 2.0 * lrs[b] * RI.cohort_size(e.ci_eff);
            for (auto& c : sched[b].data) fl += double(c.n_flops());
        }
        if (flops_out) *flops_out = fl;
        // super_hamil_mv: OpenMP dynamic over MPS blocks, per-output-block mutex
        std::vector<const double*> psiv(nl), lview = L.view(), rview = R.view();
        for (size_t b = 0; b < nl; ++b) psiv[b] = psi + boff[b];
        std::fill(out, out + boff[nl], 0.0);
        std::vector<std::mutex> mtx(nl);
        double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        for (long lb_in = 0; lb_in < long(nl); ++lb_in) {
            if (lb_mask && !lb_mask[lb_in]) continue;
            auto T = sched[lb_in].create_T(RI, rview, psiv, lrs);
            for (auto const& c : sched[lb_in].data) c.contract(lview, T, out + boff[c.rb], mtx[c.rb]);
        }
        if (seconds_out) *seconds_out = omp_get_wtime() - t0;
        return 0;
    } catch (std::exception const& e) {
        fprintf(stderr, "oracle_hpsi: %s\n", e.what());
        return -1;
    }
}

int oracle_hpsi(int sym, const mq_mpo_desc* mpod, const mq_bindex_desc* Ld, const double* Ldata, const mq_bindex_desc* Rd,
                const double* Rdata, const mq_index_desc* li, const mq_index_desc* pi, const mq_index_desc* ri,
                const double* psi, double* out, int nthreads, double* flops_out) {
    return oracle_hpsi_sample(sym, mpod, Ld, Ldata, Rd, Rdata, li, pi, ri, psi, out, nthreads, flops_out, nullptr,
                              nullptr);
}

// overlap_mpo_left_step CPU branch.  bra: left-paired data, ket: right-paired data.
oracle_boundary* oracle_lstep(int sym, const mq_mpo_desc* mpod, const mq_bindex_desc* Ld, const double* Ldata,
                              const mq_index_desc* bra_l, const mq_index_desc* bra_p, const mq_index_desc* bra_r,
                              const double* bra_lp, const mq_index_desc* ket_l, const mq_index_desc* ket_p,
                              const mq_index_desc* ket_r, const double* ket_rp, int symmetric, int nthreads) {
    try {
        g_sym = sym;
        Mpo mpo = make_mpo(*mpod);
        Bnd L = make_bnd(*Ld, Ldata);
        Idx bl = make_idx(*bra_l), phys = make_idx(*bra_p), br = make_idx(*bra_r);
        Idx kl = make_idx(*ket_l), kr = make_idx(*ket_r);
        (void)ket_p;
        PB bra_left_pb(phys, bl, +1);
        PB ket_right_pb(phys, kr, -1);
        std::vector<size_t> phys_s = phys.sizes();
        // data views
        std::vector<const double*> ketv(kl.size()), brav(br.size());
        {
            size_t o = 0;
            for (size_t b = 0; b < kl.size(); ++b) { ketv[b] = ket_rp + o; o += kl[b].second * ket_right_pb.size(kl[b].first); }
            o = 0;
            for (size_t b = 0; b < br.size(); ++b) { brav[b] = bra_lp + o; o += bra_left_pb.size(br[b].first) * br[b].second; }
        }
        BIdx const& LI = L.idx;
        size_t loop_max = kr.size();
        std::vector<MPSBlock> sched(loop_max);
        std::vector<size_t> lr_ket = kr.sizes();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        for (long rb_ket = 0; rb_ket < long(loop_max); ++rb_ket) {
            MPSBlock& mb = sched[rb_ket];
            mb.rb_ket = unsigned(rb_ket);
            Q rc_ket = kr[rb_ket].first;
            unsigned rs_ket = unsigned(kr[rb_ket].second);
            // lshtm_t_tasks (left_tasks.hpp:38-72)
            for (unsigned s = 0; s < phys.size(); ++s) {
                Q lc_ket = rc_ket + (-phys[s].first);
                size_t lb_ket = kl.position(lc_ket);
                if (lb_ket == kl.size()) continue;
                unsigned off = unsigned(ket_right_pb(phys[s].first, rc_ket));
                for (unsigned lb_bra = 0; lb_bra < kl.size(); ++lb_bra) {
                    unsigned ci = LI.ci(kl[lb_bra].first, lc_ket);
                    if (ci == LI.n_cohorts()) continue;
                    unsigned ce = LI.tr[ci] ? LI.ci(lc_ket, kl[lb_bra].first) : ci;
                    for (unsigned ss = 0; ss < phys[s].second; ++ss)
                        mb.t.push_back(TEntry{off + ss * rs_ket, ci, ce, unsigned(lb_ket), 0});
                }
            }
            // lshtm_tasks (left_tasks.hpp:74-172)
            for (unsigned rb_bra = 0; rb_bra < br.size(); ++rb_bra) {
                Q rc_bra = br[rb_bra].first;
                if (std::abs(particle_number(rc_bra) - particle_number(rc_ket)) > 2) continue;
                unsigned rs_bra = unsigned(br[rb_bra].second);
                Cohort coh(phys_s, rb_bra, unsigned(rb_ket), rs_bra, rs_ket, 0, 0, mpo.cols, true);
                for (unsigned s = 0; s < phys.size(); ++s) {
                    Q po = phys[s].first;
                    Q lc_bra = rc_bra + (-po);
                    size_t lb_bra = bl.position(lc_bra);
                    if (lb_bra == bl.size()) continue;
                    unsigned ls_bra = unsigned(bl[lb_bra].second);
                    coh.add_unit(s, unsigned(phys[s].second), ls_bra, unsigned(bra_left_pb(po, lc_bra)));
                    for (unsigned b2 = 0; b2 < mpo.cols; ++b2) {
                        if (mpo.rh.skip(b2, rc_bra, rc_ket) && symmetric) continue;
                        int Ap = mpo.rspin[b2];
                        if (!tri(rc_ket, Ap, rc_bra)) continue;
                        for (unsigned b1 : mpo.col_idx[b2]) {
                            for (auto const& term : mpo.cells.at({b1, b2})) {
                                Op const& W = mpo.ops[term.first];
                                int K = W.spin, A = mpo.lspin[b1];
                                for (size_t wb = 0; wb < W.basis.size(); ++wb) {
                                    if (po != W.basis[wb].second) continue;
                                    Q pin = W.basis[wb].first;
                                    Q lc_ket = rc_ket + (-pin);
                                    size_t lb_ket = kl.position(lc_ket);
                                    if (lb_ket == kl.size()) continue;
                                    unsigned ci = LI.ci(lc_bra, lc_ket);
                                    if (!LI.has_block(ci, b1)) continue;
                                    unsigned ls_ket = unsigned(kl[lb_ket].second);
                                    double cp[4];
                                    double scale = LI.cscale[ci][b1] * term.second;
                                    if (su2())
                                        set_coupling(spin(lc_ket), std::abs(spin(rc_ket + (-lc_ket))), spin(rc_ket), A, K, Ap,
                                                     spin(lc_bra), std::abs(spin(rc_bra + (-lc_bra))), spin(rc_bra), scale, cp);
                                    else
                                        cp[0] = cp[1] = cp[2] = cp[3] = scale;
                                    op_iterate(W, wb, cp, coh, s, rs_ket, mb, unsigned(ket_right_pb(pin, rc_ket)), ci,
                                               unsigned(LI.offsets[ci][b1] / ls_ket));
                                }
                            }
                        }
                        coh.add_line(b2);
                    }
                }
                coh.finalize();
                if (coh.n_tasks()) mb.data.push_back(std::move(coh));
            }
        }
        auto* res = new oracle_boundary();
        BIdx bi(br, kr, mpo.cols);
        for (unsigned rb_ket = 0; rb_ket < loop_max; ++rb_ket)
            for (auto& e : sched[rb_ket].data) bi.add_cohort(e.lb, rb_ket, e.mpo_offsets);
        if (symmetric) bi.complement_transpose(mpo.rh, true);
        res->b.idx = bi;
        res->b.data.resize(bi.n_cohorts());
        for (unsigned c = 0; c < bi.n_cohorts(); ++c)
            if (!bi.tr[c]) res->b.data[c].assign(bi.cohort_size(c), 0.0);
        std::vector<const double*> lview = L.view();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        for (long rb_ket = 0; rb_ket < long(loop_max); ++rb_ket) {
            auto T = sched[rb_ket].create_T_left(LI, lview, ketv, lr_ket);
            for (auto const& c : sched[rb_ket].data) {
                unsigned ci = bi.cohort_index(c.lb, unsigned(rb_ket));
                c.prop_l(brav[c.lb], T, res->b.data[ci].data());
            }
        }
        return res;
    } catch (std::exception const& e) {
        fprintf(stderr, "oracle_lstep: %s\n", e.what());
        return nullptr;
    }
}

// overlap_mpo_right_step CPU branch.  bra and ket: right-paired data.
oracle_boundary* oracle_rstep(int sym, const mq_mpo_desc* mpod, const mq_bindex_desc* Rd, const double* Rdata,
                              const mq_index_desc* bra_l, const mq_index_desc* bra_p, const mq_index_desc* bra_r,
                              const double* bra_rp, const mq_index_desc* ket_l, const mq_index_desc* ket_p,
                              const mq_index_desc* ket_r, const double* ket_rp, int symmetric, int nthreads) {
    try {
        g_sym = sym;
        Mpo mpo = make_mpo(*mpod);
        Bnd R = make_bnd(*Rd, Rdata);
        Idx bl = make_idx(*bra_l), phys = make_idx(*bra_p), br = make_idx(*bra_r);
        Idx kl = make_idx(*ket_l), kr = make_idx(*ket_r);
        (void)ket_p;
        PB bra_right_pb(phys, br, -1), ket_right_pb(phys, kr, -1);
        std::vector<size_t> phys_s = phys.sizes();
        std::vector<const double*> ketv(kl.size()), brav(bl.size());
        {
            size_t o = 0;
            for (size_t b = 0; b < kl.size(); ++b) { ketv[b] = ket_rp + o; o += kl[b].second * ket_right_pb.size(kl[b].first); }
            o = 0;
            for (size_t b = 0; b < bl.size(); ++b) { brav[b] = bra_rp + o; o += bl[b].second * bra_right_pb.size(bl[b].first); }
        }
        BIdx const& RI = R.idx;
        size_t loop_max = kl.size();
        std::vector<MPSBlock> sched(loop_max);
        std::vector<size_t> lr_ket = kl.sizes();
        const int site_basis_max_diff = 2;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        for (long lb_ket = 0; lb_ket < long(loop_max); ++lb_ket) {
            MPSBlock& mb = sched[lb_ket];
            // T tasks over the ket's right index, bra charges from the bra right index
            rshtm_t(RI, kl, kr, phys, ket_right_pb, unsigned(lb_ket), mb, br);
            // rshtm_tasks (right_tasks.hpp:80-167)
            Q lc_ket = kl[lb_ket].first;
            unsigned ls_ket = unsigned(kl[lb_ket].second);
            for (unsigned lb_bra = 0; lb_bra < bl.size(); ++lb_bra) {
                Q lc_bra = bl[lb_bra].first;
                if (std::abs(particle_number(lc_ket) - particle_number(lc_bra)) > site_basis_max_diff) continue;
                unsigned ls_bra = unsigned(bl[lb_bra].second);
                Cohort coh(phys_s, unsigned(lb_ket), lb_bra, ls_ket, ls_bra, 0, 0, mpo.rows);
                for (unsigned s = 0; s < phys.size(); ++s) {
                    Q po = phys[s].first;
                    Q rc_bra = lc_bra + po;
                    size_t rb_bra = br.position(rc_bra);
                    if (rb_bra == br.size()) continue;
                    unsigned rs_bra = unsigned(br[rb_bra].second);
                    coh.add_unit(s, unsigned(phys[s].second), rs_bra, unsigned(bra_right_pb(po, rc_bra)));
                    W9j w9j(lc_bra, lc_ket, rc_bra);
                    for (unsigned b1 = 0; b1 < mpo.rows; ++b1) {
                        if (mpo.lh.skip(b1, lc_ket, lc_bra) && symmetric) continue;
                        int A = mpo.lspin[b1];
                        if (!tri(lc_ket, A, lc_bra)) continue;
                        for (unsigned b2 : mpo.row_idx[b1]) {
                            for (auto const& term : mpo.cells.at({b1, b2})) {
                                Op const& W = mpo.ops[term.first];
                                int K = W.spin, Ap = mpo.rspin[b2];
                                for (size_t wb = 0; wb < W.basis.size(); ++wb) {
                                    if (po != W.basis[wb].second) continue;
                                    Q pin = W.basis[wb].first;
                                    Q rc_ket = lc_ket + pin;
                                    unsigned ci = RI.ci(rc_ket, rc_bra);
                                    if (!RI.has_block(ci, b2)) continue;
                                    size_t rb_ket = kr.position(rc_ket);
                                    if (rb_ket == kr.size()) continue;
                                    unsigned rs_ket = unsigned(kr[rb_ket].second);
                                    double cp[4];
                                    double scale = RI.cscale[ci][b2] * term.second;
                                    w9j.set_scale(A, K, Ap, rc_ket, scale, cp);
                                    op_iterate(W, wb, cp, coh, s, rs_ket, mb, unsigned(ket_right_pb(pin, rc_ket)), ci,
                                               unsigned(RI.offsets[ci][b2] / rs_ket));
                                }
                            }
                        }
                        coh.add_line(b1);
                    }
                }
                coh.finalize();
                if (coh.n_tasks()) mb.data.push_back(std::move(coh));
            }
        }
        auto* res = new oracle_boundary();
        BIdx bi(kl, bl, mpo.rows);
        for (unsigned lb_ket = 0; lb_ket < loop_max; ++lb_ket)
            for (auto& e : sched[lb_ket].data) bi.add_cohort(lb_ket, e.rb, e.mpo_offsets);
        bi.complement_transpose(mpo.lh, false);
        res->b.idx = bi;
        res->b.data.resize(bi.n_cohorts());
        for (unsigned c = 0; c < bi.n_cohorts(); ++c)
            if (!bi.tr[c]) res->b.data[c].assign(bi.cohort_size(c), 0.0);
        std::vector<const double*> rview = R.view();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        for (long lb_ket = 0; lb_ket < long(loop_max); ++lb_ket) {
            auto T = sched[lb_ket].create_T(RI, rview, ketv, lr_ket);
            for (auto const& c : sched[lb_ket].data) {
                unsigned ci = bi.cohort_index(unsigned(lb_ket), c.rb);
                c.prop_r(brav[c.rb], T, res->b.data[ci].data());
            }
        }
        return res;
    } catch (std::exception const& e) {
        fprintf(stderr, "oracle_rstep: %s\n", e.what());
        return nullptr;
    }
}

int oracle_boundary_info(oracle_boundary* h, int* n_cohorts, unsigned* aux, int* nbra, int* nket, uint64_t* n_elems) {
    if (!h) return -1;
    BIdx const& B = h->b.idx;
    *n_cohorts = int(B.n_cohorts());
    *aux = B.aux;
    *nbra = int(B.bra.size());
    *nket = int(B.ket.size());
    uint64_t n = 0;
    for (unsigned c = 0; c < B.n_cohorts(); ++c)
        if (!B.tr[c]) n += B.cohort_size(c);
    *n_elems = n;
    return 0;
}

int oracle_boundary_export(oracle_boundary* h, uint32_t* cohort_lr, int64_t* offsets, double* cscale, uint8_t* transp,
                           uint8_t* tr, uint32_t* nblocks, double* data) {
    if (!h) return -1;
    BIdx const& B = h->b.idx;
    size_t o = 0;
    for (unsigned c = 0; c < B.n_cohorts(); ++c) {
        cohort_lr[2 * c] = B.lr[c].first;
        cohort_lr[2 * c + 1] = B.lr[c].second;
        for (unsigned b = 0; b < B.aux; ++b) {
            offsets[size_t(c) * B.aux + b] = B.offsets[c][b];
            cscale[size_t(c) * B.aux + b] = B.cscale[c][b];
            transp[size_t(c) * B.aux + b] = uint8_t(B.transp[c][b]);
        }
        tr[c] = uint8_t(B.tr[c]);
        nblocks[c] = B.nb[c];
        if (!B.tr[c]) {
            std::copy(h->b.data[c].begin(), h->b.data[c].end(), data + o);
            o += h->b.data[c].size();
        }
    }
    return 0;
}

void oracle_boundary_free(oracle_boundary* h) { delete h; }

}  // extern "C"
