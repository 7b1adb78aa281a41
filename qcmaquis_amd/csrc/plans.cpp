#include "plans.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <memory>
#include <unordered_map>
#include <unordered_set>

#include "coupling.hpp"

namespace mq {

namespace {

// Output-tile accumulator of one cohort: units (sigma', ss2) with their term lines, in the
// reference's insertion order (Cohort / SUnit, solver/tasks.h:50-190, tasks.cpp:72-196).
struct CohortAcc {
    struct Term {
        double alpha;
        uint32_t ti, col;  // T task and its column block
    };
    struct Unit {
        uint32_t offset = 0, ms = 0;
        uint32_t s = UINT32_MAX;  // physical sector
        std::vector<Term> terms;
        std::vector<uint32_t> line_count, line_b;
        uint32_t pending = 0;
        // with term sources: per term, the MPO term it comes from and its alpha at unit term scale
        std::vector<std::pair<uint32_t, double>> src;
    };
    std::vector<uint32_t> sfold;
    std::vector<Unit> units;
    uint64_t stripe = 0;
    std::vector<uint32_t> bcount;  // terms per MPO index b (growth steps)
    uint64_t n_terms = 0;
    std::vector<uint32_t> dirty;   // units with pending terms, in order of their first term

    CohortAcc(Index const& phys, uint32_t mpodim) : sfold(phys.size()), bcount(mpodim, 0) {
        uint32_t acc = 0;
        for (size_t s = 0; s < phys.size(); ++s) {
            sfold[s] = acc;
            acc += phys.sizes[s];
        }
        units.resize(acc);
    }
    void add_unit(size_t s, uint32_t d, uint32_t ms, uint64_t offset) {
        for (uint32_t i = 0; i < d; ++i) {
            units[sfold[s] + i].offset = uint32_t(i * ms + offset);
            units[sfold[s] + i].ms = ms;
            units[sfold[s] + i].s = uint32_t(s);
            stripe += ms;
        }
    }
    void push(size_t s, uint32_t ss2, double alpha, uint32_t ti, uint32_t col) {
        const uint32_t ui = sfold[s] + ss2;
        Unit& u = units[ui];
        u.terms.push_back(Term{alpha, ti, col});
        if (!u.pending++) dirty.push_back(ui);
        ++n_terms;
    }
    void push_src(size_t s, uint32_t ss2, double alpha, uint32_t ti, uint32_t col, uint32_t mpo_term, double beta) {
        units[sfold[s] + ss2].src.push_back({mpo_term, beta});
        push(s, ss2, alpha, ti, col);
    }
    // closes the line of MPO index b in every unit that received terms since the last line
    void add_line(uint32_t b) {
        for (uint32_t ui : dirty) {
            Unit& u = units[ui];
            u.line_count.push_back(u.pending);
            u.line_b.push_back(b);
            if (b < bcount.size()) bcount[b] += u.pending;
            u.pending = 0;
        }
        dirty.clear();
    }
};

// T work list of one MPS block with the reference's (mps_offset, ci) -> ti lookup
// (MPSBlock::get_ti, tasks.cpp:766-775).
struct TTask {
    uint64_t psi_col;  // column offset in the ket block
    uint32_t ci, ci_eff;
    uint32_t lb_ket;
    bool used = false;
    uint64_t t_off = 0;
    uint64_t size = 0;   // elements of the T intermediate
    uint32_t rows = 0;   // its leading dimension
};
// The (column, cohort) -> task lookup is an open-addressing table (linear probing, at most half
// full): the schedule builders look up one task per operator entry, tens of millions per site
// problem.  A list is filled by one thread and then only read.
struct TList {
    std::vector<TTask> tasks;
    std::vector<uint64_t> keys;   // key + 1 per slot (0 = empty)
    std::vector<uint32_t> vals;
    static uint64_t key_of(uint64_t col, uint32_t ci) { return (col << 24) ^ ci; }
    size_t slot(uint64_t k) const {
        return size_t((k * 0x9E3779B97F4A7C15ull) >> 17) & (keys.size() - 1);
    }
    void insert(uint64_t k, uint32_t v) {
        size_t h = slot(k);
        while (keys[h] != 0) {
            if (keys[h] == k + 1) return;  // the first task of a key is kept
            h = (h + 1) & (keys.size() - 1);
        }
        keys[h] = k + 1;
        vals[h] = v;
    }
    void add(uint64_t col, uint32_t ci, uint32_t ci_eff, uint32_t lb_ket) {
        require(col < (uint64_t(1) << 40) && ci < (1u << 24), "T list: key overflow");
        if (2 * (tasks.size() + 1) > keys.size()) {  // grow and re-insert in task order
            std::vector<uint64_t> old_keys;
            old_keys.swap(keys);
            keys.assign(std::max<size_t>(16, 2 * old_keys.size()), 0);
            vals.assign(keys.size(), 0);
            for (size_t i = 0; i < tasks.size(); ++i) insert(key_of(tasks[i].psi_col, tasks[i].ci), uint32_t(i));
        }
        insert(key_of(col, ci), uint32_t(tasks.size()));
        tasks.push_back(TTask{col, ci, ci_eff, lb_ket});
    }
    uint32_t get(uint64_t col, uint32_t ci) const {
        if (keys.empty()) return UINT32_MAX;
        const uint64_t k = key_of(col, ci);
        for (size_t h = slot(k);; h = (h + 1) & (keys.size() - 1)) {
            if (keys[h] == k + 1) return vals[h];
            if (keys[h] == 0) return UINT32_MAX;
        }
    }
};

inline int casenr(int32_t rspin, int32_t cspin) {
    if (rspin == 2 && cspin == 2) return 3;
    if (rspin == 2) return 1;
    if (cspin == 2) return 2;
    return 0;
}

// iterate the sparse entries of one W block (reference: detail::op_iterate, op_iterate.hpp:37-69);
// Src: also record each term's MPO term and its alpha at unit term scale (coupling cpl1)
template <bool Src = false>
inline void op_iterate(OpBlock const& blk, const double cpl[4], CohortAcc& coh, size_t s, uint32_t m2_size,
                       TList& tl, uint64_t mps_offset, uint32_t ci, uint32_t boundary_col,
                       uint32_t mpo_term = 0, const double* cpl1 = nullptr) {
    for (auto const& e : blk.entries) {
        const int cn = casenr(e.row_spin, e.col_spin);
        double alpha = e.coeff * cpl[cn];
        uint32_t ti = tl.get(mps_offset + uint64_t(e.row) * m2_size, ci);
        if (ti != UINT32_MAX) {
            // cohorts of one T list may be built concurrently: every writer stores the same flag
            // (read first, so the threads do not keep invalidating each other's copy of the line)
            if (!__atomic_load_n(&tl.tasks[ti].used, __ATOMIC_RELAXED))
                __atomic_store_n(&tl.tasks[ti].used, true, __ATOMIC_RELAXED);
            if (Src) coh.push_src(s, e.col, alpha, ti, boundary_col, mpo_term, e.coeff * cpl1[cn]);
            else coh.push(s, e.col, alpha, ti, boundary_col);
        }
    }
}

// MPO rows of one unit that received no terms: emitted as lines without terms, so the term sum
// writes their S tiles as zeros and S needs no separate clear
template <class Out, class MakeLine>
void add_empty_lines(CohortAcc::Unit const& u, uint32_t nrows, int32_t term_pos, Out& out, MakeLine make_line) {
    if (!u.ms) return;
    auto empty = [&](uint32_t b) {
        SLine line = make_line(b);
        line.t0 = line.t1 = term_pos;
        out.push_back(line);
    };
    if (std::is_sorted(u.line_b.begin(), u.line_b.end(), std::less_equal<uint32_t>())) {
        // rows with terms strictly increasing (the builders add lines in MPO-row order)
        require(u.line_b.empty() || u.line_b.back() < nrows, "S lines: MPO row outside its cohort");
        size_t i = 0;
        for (uint32_t b = 0; b < nrows; ++b) {
            if (i < u.line_b.size() && u.line_b[i] == b) ++i;
            else empty(b);
        }
        return;
    }
    std::vector<char> has(nrows, 0);
    for (uint32_t b : u.line_b) {
        require(b < nrows, "S lines: MPO row outside its cohort");
        has[b] = 1;
    }
    for (uint32_t b = 0; b < nrows; ++b)
        if (!has[b]) empty(b);
}

// lines a unit emits: one per MPO row with terms, plus the empty lines of add_empty_lines
size_t unit_lines(CohortAcc::Unit const& u, uint32_t nrows) {
    if (!u.ms) return u.line_count.size();
    if (std::is_sorted(u.line_b.begin(), u.line_b.end(), std::less_equal<uint32_t>()))
        return u.line_count.size() + (nrows - std::min<size_t>(nrows, u.line_b.size()));
    std::vector<char> has(nrows, 0);
    size_t n = 0;
    for (uint32_t b : u.line_b)
        if (b < nrows && !has[b]) has[b] = 1, ++n;
    return u.line_count.size() + (nrows - n);
}

// fixed-size output window of one cohort in the plan's line or term table; size() is the absolute
// index of the next entry
template <class T>
struct TableSink {
    T* p;
    size_t n, cap, base;
    void push_back(T const& v) {
        if (n == cap) throw Error(MQ_ERR_INVALID, "S table: cohort emitted more entries than counted");
        p[n++] = v;
    }
    size_t size() const { return base + n; }
};

// S lines and terms of all cohorts, written in parallel straight into the tables in cohort order:
// the lines and terms of every cohort are counted first (units_of(c): its units, nrows_of(c): its
// MPO rows, keep(c, unit): the units taken), then emit(c, lines, terms) appends cohort c's lines
// (term ranges absolute) and terms of the units taken
struct KeepAll {
    bool operator()(size_t, CohortAcc::Unit const&) const { return true; }
};
template <class Units, class Rows, class Emit, class Keep = KeepAll>
void emit_lines(size_t n_cohorts, HostTable<SLine>& lines, HostTable<STerm>& terms, Units units_of, Rows nrows_of,
                Emit emit, Keep keep = Keep()) {
    std::vector<size_t> nl(n_cohorts), nt(n_cohorts);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t c = 0; c < int64_t(n_cohorts); ++c) {
        size_t a = 0, b = 0;
        for (auto const& u : units_of(size_t(c))) {
            if (!keep(size_t(c), u)) continue;
            a += unit_lines(u, nrows_of(size_t(c)));
            b += u.terms.size();
        }
        nl[size_t(c)] = a;
        nt[size_t(c)] = b;
    }
    std::vector<size_t> loff(n_cohorts + 1, lines.size()), toff(n_cohorts + 1, terms.size());
    for (size_t c = 0; c < n_cohorts; ++c) {
        loff[c + 1] = loff[c] + nl[c];
        toff[c + 1] = toff[c] + nt[c];
    }
    require(toff[n_cohorts] < size_t(INT32_MAX), "S table: more than 2^31 terms in one step");
    lines.resize(loff[n_cohorts]);
    terms.resize(toff[n_cohorts]);
    std::string err;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < int64_t(n_cohorts); ++c) {
        try {
            TableSink<SLine> ls{lines.data() + loff[size_t(c)], 0, nl[size_t(c)], loff[size_t(c)]};
            TableSink<STerm> ts{terms.data() + toff[size_t(c)], 0, nt[size_t(c)], toff[size_t(c)]};
            emit(size_t(c), ls, ts);
            if (ls.n != ls.cap || ts.n != ts.cap)
                throw Error(MQ_ERR_INVALID, "S table: cohort emitted fewer entries than counted");
        } catch (std::exception const& e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) throw Error(MQ_ERR_INVALID, err);
}

inline uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// a term reads a contiguous rows x cols tile at element offset `off` of its T intermediate;
// refuse anything that would leave the intermediate
void check_tile(TTask const& t, uint64_t off, uint32_t rows, uint32_t cols) {
    require(t.used && t.rows == rows && off + uint64_t(rows) * cols <= t.size,
            "S-assembly term reads outside its T intermediate (boundary / MPS index mismatch)");
}

void check_cohort_shape(BoundaryIndex const& B, uint32_t ci, uint32_t rows, uint32_t cols, const char* what) {
    if (B.lsz[ci] != rows || B.rsz[ci] != cols)
        throw Error(MQ_ERR_INVALID, std::string(what) + ": boundary cohort shape does not match the MPS sectors");
}

GemmDesc gemm(int32_t M, int32_t N, int32_t K, int base_a, int64_t a, int32_t lda, int base_b, int64_t b,
              int32_t ldb, int base_c, int64_t c, int32_t ldc, int32_t flags = 0, int32_t batch = 1,
              int64_t sa = 0, int64_t sb = 0, int64_t sc = 0) {
    GemmDesc g{};
    g.a = a; g.b = b; g.c = c;
    g.sa = sa; g.sb = sb; g.sc = sc;
    g.M = M; g.N = N; g.K = K; g.batch = batch;
    g.lda = std::max(lda, 1); g.ldb = std::max(ldb, 1); g.ldc = std::max(ldc, 1);
    g.flags = flags;
    g.base_a = base_a; g.base_b = base_b; g.base_c = base_c;
    return g;
}

// Right-boundary T-GEMMs shared by H*psi and the right step (reference: MPSBlock::create_T,
// tasks.cpp:650-752, one dgemm per task): T(ls_ket x nb*rs_bra) = psi_ket(:, col : col+rs_ket) *
// R[ci_eff], per block transposed when the cohort is stored as its Hermitian partner.  The tasks
// that read one cohort in one orientation are stacked into one problem (rows from a row table),
// so each cohort is read once per row tile instead of once per task; stacks in order of first use.
struct RightTStacks {
    struct Ref {
        TTask const* t;
        uint32_t ls_ket, rs_ket, rs_bra;
        int64_t psi_block_off;
    };
    std::vector<std::vector<Ref>> stacks;
    std::unordered_map<uint64_t, size_t> pos;
    void add(TTask const& t, uint32_t ls_ket, uint32_t rs_ket, uint32_t rs_bra, int64_t psi_block_off) {
        uint64_t key = uint64_t(t.ci_eff) * 2 + (t.ci != t.ci_eff);
        auto it = pos.find(key);
        if (it == pos.end()) {
            it = pos.emplace(key, stacks.size()).first;
            stacks.emplace_back();
        }
        stacks[it->second].push_back(Ref{&t, ls_ket, rs_ket, rs_bra, psi_block_off});
    }
    // per_task (optional): the same products as one problem per task, as the reference issues them
    void emit(BoundaryIndex const& R, int psi_base, std::vector<GemmDesc>& out, std::vector<GemmRow>& rows,
              std::vector<GemmDesc>* per_task = nullptr) const {
        for (auto const& st : stacks) {
            Ref const& f = st.front();
            const uint32_t ci_eff = f.t->ci_eff, nb = R.nblk[ci_eff], rs_ket = f.rs_ket, rs_bra = f.rs_bra;
            const bool trans = f.t->ci != ci_eff;
            if (!trans) check_cohort_shape(R, ci_eff, rs_ket, rs_bra, "T-GEMM");
            else check_cohort_shape(R, ci_eff, rs_bra, rs_ket, "T-GEMM");
            const int64_t row0 = int64_t(rows.size());
            int64_t M = 0;
            for (Ref const& r : st) {
                require(r.rs_ket == rs_ket && r.rs_bra == rs_bra, "T-GEMM: stacked tasks disagree on the cohort shape");
                int64_t a = r.psi_block_off + int64_t(r.t->psi_col) * r.ls_ket;
                for (uint32_t i = 0; i < r.ls_ket; ++i)
                    rows.push_back(GemmRow{a + i, int64_t(r.t->t_off) + i, int32_t(r.ls_ket), 0});
                M += r.ls_ket;
                if (!per_task) continue;
                if (!trans)
                    per_task->push_back(gemm(int32_t(r.ls_ket), int32_t(nb * rs_bra), rs_ket, psi_base, a,
                                             int32_t(r.ls_ket), BASE_BOUNDARY, int64_t(R.data_off[ci_eff]), rs_ket,
                                             BASE_T, int64_t(r.t->t_off), int32_t(r.ls_ket)));
                else
                    per_task->push_back(gemm(int32_t(r.ls_ket), rs_bra, rs_ket, psi_base, a, int32_t(r.ls_ket),
                                             BASE_BOUNDARY, int64_t(R.data_off[ci_eff]), rs_bra, BASE_T,
                                             int64_t(r.t->t_off), int32_t(r.ls_ket), GEMM_TRANS_B, int32_t(nb), 0,
                                             int64_t(rs_bra) * rs_ket, int64_t(r.ls_ket) * rs_bra));
            }
            require(M < INT32_MAX, "T-GEMM: stacked problem too tall");
            if (!trans)
                out.push_back(gemm(int32_t(M), int32_t(nb * rs_bra), rs_ket, psi_base, row0, 1, BASE_BOUNDARY,
                                   int64_t(R.data_off[ci_eff]), rs_ket, BASE_T, 0, 1, GEMM_ROWMAP));
            else
                out.push_back(gemm(int32_t(M), rs_bra, rs_ket, psi_base, row0, 1, BASE_BOUNDARY,
                                   int64_t(R.data_off[ci_eff]), rs_bra, BASE_T, 0, 1, GEMM_ROWMAP | GEMM_TRANS_B,
                                   int32_t(nb), 0, int64_t(rs_bra) * rs_ket, int64_t(rs_bra)));
        }
    }
};

void validate_boundary_offsets(BoundaryIndex const& B) {
    for (uint32_t ci = 0; ci < B.n_cohorts(); ++ci) {
        uint64_t bs = B.block_size(ci);
        for (uint32_t b = 0; b < B.aux_dim; ++b) {
            int64_t o = B.offsets[ci][b];
            if (o < 0) continue;
            uint32_t owner = B.tr[ci] ? B.partner(ci) : ci;
            require(owner < B.n_cohorts(), "boundary: transposed cohort without partner");
            uint64_t obs = B.block_size(owner);
            require(obs > 0 && uint64_t(o) % obs == 0 && uint64_t(o) < B.cohort_size(owner),
                    "boundary: block offset outside its cohort");
            (void)bs;
        }
    }
}

}  // namespace

// ----------------------------------------------------------------------------------------------
// H*psi  (reference: create_contraction_schedule -> rshtm_t_tasks + shtm_tasks,
//         site_hamil_schedule.hpp:40-93, right_tasks.hpp:39-78, site_hamil_tasks.hpp:41-137)
// ----------------------------------------------------------------------------------------------
HpsiPlanHost build_hpsi_plan(Sym sym, MPO const& mpo, BoundaryIndex const& L, BoundaryIndex const& R,
                             Index const& left_i, Index const& phys_i, Index const& right_i, uint64_t chunk_elems,
                             bool term_sources) {
    HpsiPlanHost P;
    P.mps = MPSLayout(left_i, phys_i, right_i);
    require(L.aux_dim == mpo.row_dim, "hpsi: left boundary aux dim != MPO row dim");
    require(R.aux_dim == mpo.col_dim, "hpsi: right boundary aux dim != MPO col dim");
    validate_boundary_offsets(L);
    validate_boundary_offsets(R);
    ProductBasis rpb = ProductBasis::right_paired(phys_i, right_i);
    const bool su2 = sym == Sym::SU2U1;
    const size_t NL = left_i.size(), NR = right_i.size(), NP = phys_i.size();

    // lookup tables built once per plan
    std::vector<uint32_t> pos_lp(NL * NP);
    std::vector<uint64_t> pb_off(NP * NR, 0);
    for (size_t lb = 0; lb < NL; ++lb)
        for (size_t s = 0; s < NP; ++s)
            pos_lp[lb * NP + s] = uint32_t(right_i.position(left_i.charges[lb] + phys_i.charges[s]));
    for (size_t s = 0; s < NP; ++s)
        for (size_t rb = 0; rb < NR; ++rb) pb_off[s * NR + rb] = rpb(phys_i.charges[s], right_i.charges[rb]);
    std::vector<uint32_t> r_ci(NR * NR);
    for (size_t a = 0; a < NR; ++a)
        for (size_t b = 0; b < NR; ++b) r_ci[a * NR + b] = R.cohort_index(right_i.charges[a], right_i.charges[b]);
    // operator blocks grouped by the physical sector of their output charge
    struct BlkRef { uint32_t blk; uint32_t s_in; };
    std::vector<std::vector<BlkRef>> blk_by_out(mpo.ops.size() * NP);
    for (size_t o = 0; o < mpo.ops.size(); ++o)
        for (size_t bi = 0; bi < mpo.ops[o].blocks.size(); ++bi) {
            OpBlock const& blk = mpo.ops[o].blocks[bi];
            size_t so = phys_i.position(blk.rc), si = phys_i.position(blk.lc);
            if (so == NP || si == NP) continue;  // a block outside the physical index never matches
            blk_by_out[o * NP + so].push_back(BlkRef{uint32_t(bi), uint32_t(si)});
        }
    // present MPO rows of every left cohort
    std::vector<std::vector<uint32_t>> l_rows(L.n_cohorts());
    for (uint32_t ci = 0; ci < L.n_cohorts(); ++ci)
        for (uint32_t b1 = 0; b1 < mpo.row_dim; ++b1)
            if (L.has_block(ci, b1)) l_rows[ci].push_back(b1);

    struct CohRec {
        uint32_t lb_in, lb_out, ci, ci_eff, nb;
        CohortAcc acc;
    };
    std::vector<TList> tlists(NL);
    std::unordered_set<uint32_t> l_used, r_used;
    std::string err;

    // T work lists (rshtm_t_tasks), one per input block
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t lb_in = 0; lb_in < NL; ++lb_in) {
        try {
            TList& tl = tlists[lb_in];
            for (size_t s = 0; s < NP; ++s) {
                size_t rb_ket = pos_lp[lb_in * NP + s];
                if (rb_ket == NR) continue;
                uint32_t rs_ket = right_i.sizes[rb_ket];
                uint64_t ket_offset = pb_off[s * NR + rb_ket];
                for (size_t rb_bra = 0; rb_bra < NR; ++rb_bra) {
                    uint32_t ci = r_ci[rb_ket * NR + rb_bra];
                    if (ci == R.n_cohorts()) continue;
                    uint32_t ci_eff = R.tr[ci] ? r_ci[rb_bra * NR + rb_ket] : ci;
                    require(ci_eff < R.n_cohorts(), "hpsi: transposed right cohort without partner");
                    for (uint32_t ss = 0; ss < phys_i.sizes[s]; ++ss) tl.add(ket_offset + ss * rs_ket, ci, ci_eff, lb_in);
                }
            }
        } catch (std::exception const& e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) throw Error(MQ_ERR_INVALID, err);

    // cohorts (shtm_tasks), one per (lb_in, lb_out) pair with a left cohort, built in parallel; the
    // T lists are only read (a task's used flag is set atomically)
    std::vector<std::unique_ptr<CohRec>> recs(size_t(NL) * NL);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t pr = 0; pr < int64_t(NL) * int64_t(NL); ++pr) {
        const uint32_t lb_in = uint32_t(pr / int64_t(NL)), lb_out = uint32_t(pr % int64_t(NL));
        try {
            Charge lc_in = left_i.charges[lb_in];
            uint32_t ls_in = left_i.sizes[lb_in];
            TList& tl = tlists[lb_in];
            Charge lc_out = left_i.charges[lb_out];
            uint32_t ls_out = left_i.sizes[lb_out];
            uint32_t ci = L.cohort_index(lc_out, lc_in);
            if (ci == L.n_cohorts()) continue;  // no left cohort for this pair
            uint32_t ci_eff = L.tr[ci] ? L.cohort_index(lc_in, lc_out) : ci;
            require(ci_eff < L.n_cohorts(), "hpsi: transposed left cohort without partner");
            CohRec rec{lb_in, lb_out, ci, ci_eff, L.nblk[ci_eff], CohortAcc(phys_i, 0)};
            CohortAcc& coh = rec.acc;
            for (size_t s = 0; s < NP; ++s) {
                size_t rb_out = pos_lp[lb_out * NP + s];
                if (rb_out == NR) continue;
                Charge rc_out = right_i.charges[rb_out];
                uint32_t rs_out = right_i.sizes[rb_out];
                coh.add_unit(s, phys_i.sizes[s], rs_out, pb_off[s * NR + rb_out]);
                std::shared_ptr<const Wigner9jCache> w9j;
                if (su2) w9j = wigner9j_cache(spin_of(lc_in), spin_of(lc_out), spin_of(rc_out));

                for (uint32_t b1 : l_rows[ci]) {
                    uint32_t left_idx = uint32_t(L.offset(ci, b1) / (int64_t(ls_in) * ls_out));
                    int A = mpo.left_spin[b1];
                    if (su2 && !triangle(spin_of(lc_in), A, spin_of(lc_out))) continue;
                    for (auto const& cell : mpo.rows[b1]) {
                        uint32_t b2 = cell.other;
                        for (uint32_t t = cell.t0; t < cell.t1; ++t) {
                            uint32_t op = mpo.terms[t].op;
                            SiteOp const& W = mpo.ops[op];
                            int K = W.spin, Ap = mpo.right_spin[b2];
                            for (BlkRef const& br : blk_by_out[size_t(op) * NP + s]) {
                                OpBlock const& blk = W.blocks[br.blk];
                                size_t rb_in = pos_lp[lb_in * NP + br.s_in];
                                if (rb_in == NR) continue;
                                uint32_t ci_right = r_ci[rb_in * NR + rb_out];
                                if (!R.has_block(ci_right, b2)) continue;
                                Charge rc_in = right_i.charges[rb_in];
                                uint32_t rs_in = right_i.sizes[rb_in];
                                int64_t right_offset = R.offset(ci_right, b2);
                                double cpl[4];
                                double scale = R.conjugate_scale(ci_right, b2) * mpo.terms[t].scale *
                                               L.conjugate_scale(ci, b1);
                                if (su2)
                                    w9j->set_scale(A, K, Ap, spin_of(rc_in), scale, cpl);
                                else
                                    cpl[0] = cpl[1] = cpl[2] = cpl[3] = scale;
                                if (term_sources) {
                                    // the same coupling at unit term scale: d alpha / d(term scale)
                                    double cpl1[4];
                                    const double scale1 = R.conjugate_scale(ci_right, b2) * L.conjugate_scale(ci, b1);
                                    if (su2)
                                        w9j->set_scale(A, K, Ap, spin_of(rc_in), scale1, cpl1);
                                    else
                                        cpl1[0] = cpl1[1] = cpl1[2] = cpl1[3] = scale1;
                                    op_iterate<true>(blk, cpl, coh, s, rs_in, tl, pb_off[size_t(br.s_in) * NR + rb_in],
                                                     ci_right, uint32_t(right_offset / rs_in), t, cpl1);
                                } else {
                                    op_iterate(blk, cpl, coh, s, rs_in, tl, pb_off[size_t(br.s_in) * NR + rb_in],
                                               ci_right, uint32_t(right_offset / rs_in));
                                }
                            }
                        }
                    }
                    coh.add_line(left_idx);
                }
            }
            if (coh.n_terms) recs[size_t(pr)].reset(new CohRec(std::move(rec)));
        } catch (std::exception const& e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) throw Error(MQ_ERR_INVALID, err);
    std::vector<CohRec> cohorts;
    for (auto& r : recs)
        if (r) cohorts.push_back(std::move(*r));

    // --- S panels: per output block the cohorts stacked along K (the same in every chunk) ---
    std::vector<uint64_t> ktot(NL, 0);
    std::vector<uint64_t> koff(cohorts.size());
    for (size_t c = 0; c < cohorts.size(); ++c) {
        koff[c] = ktot[cohorts[c].lb_out];
        ktot[cohorts[c].lb_out] += uint64_t(cohorts[c].nb) * left_i.sizes[cohorts[c].lb_in];
    }
    // column stripes of every output block: one per physical sector with a right sector, at the
    // stripe's column offset of the psi block; stripes of one block in column order
    struct Stripe { uint32_t s, rb; uint64_t col, width; };
    std::vector<std::vector<Stripe>> stripes(NL);
    for (size_t lb = 0; lb < NL; ++lb) {
        for (size_t s = 0; s < NP; ++s) {
            size_t rb = pos_lp[lb * NP + s];
            if (rb == NR) continue;
            stripes[lb].push_back(Stripe{uint32_t(s), uint32_t(rb), pb_off[s * NR + rb],
                                         uint64_t(phys_i.sizes[s]) * right_i.sizes[rb]});
        }
        std::sort(stripes[lb].begin(), stripes[lb].end(), [](Stripe const& a, Stripe const& b) { return a.col < b.col; });
    }
    // used T tasks: their output right sector (the ket sector of their right cohort)
    auto task_rb = [&](TTask const& t) {
        auto lr = R.clr[t.ci];
        size_t rb = right_i.position(R.ket.charges[lr.second]);
        require(rb < NR, "hpsi: right boundary index does not match the MPS right index");
        return rb;
    };
    // --- chunks: right sectors gathered in index order up to chunk_elems of T + S ---
    std::vector<uint64_t> t_rb(NR, 0), s_rb(NR, 0);
    for (uint32_t lb_in = 0; lb_in < NL; ++lb_in)
        for (auto const& t : tlists[lb_in].tasks)
            if (t.used) t_rb[task_rb(t)] += round_up(uint64_t(left_i.sizes[lb_in]) * R.nblk[t.ci_eff] *
                                                         R.ket.sizes[R.clr[t.ci].second], 32);
    for (size_t lb = 0; lb < NL; ++lb)
        for (auto const& st : stripes[lb]) s_rb[st.rb] += ktot[lb] * st.width;
    std::vector<uint32_t> chunk_of(NR, 0);
    uint32_t n_chunks = 1;
    if (chunk_elems) {
        uint64_t acc = 0;
        bool any = false;
        for (size_t rb = 0; rb < NR; ++rb) {
            const uint64_t need = t_rb[rb] + s_rb[rb];
            if (any && acc + need > chunk_elems) {
                ++n_chunks;
                acc = 0;
            }
            chunk_of[rb] = n_chunks - 1;
            acc += need;
            any = true;
        }
    }
    P.chunks.resize(n_chunks);
    // chunk-local S panel of every (chunk, output block): the block's stripes of the chunk side by
    // side; loc[lb][i] = local column of stripe i
    std::vector<std::vector<uint64_t>> s_off(n_chunks, std::vector<uint64_t>(NL, 0));
    std::vector<std::vector<uint64_t>> loc(NL);
    std::vector<uint64_t> l_off(NL);
    for (size_t lb = 0; lb < NL; ++lb) {
        l_off[lb] = P.lpanel_elems;
        P.lpanel_elems += round_up(uint64_t(left_i.sizes[lb]) * ktot[lb], 32);
        loc[lb].resize(stripes[lb].size());
    }
    for (uint32_t ch = 0; ch < n_chunks; ++ch) {
        uint64_t& se = P.chunks[ch].s_elems;
        for (size_t lb = 0; lb < NL; ++lb) {
            s_off[ch][lb] = se;
            uint64_t w = 0;
            for (size_t i = 0; i < stripes[lb].size(); ++i)
                if (chunk_of[stripes[lb][i].rb] == ch) {
                    loc[lb][i] = w;
                    w += stripes[lb][i].width;
                }
            se += round_up(ktot[lb] * w, 32);
        }
    }
    // local column of a unit's stripe in its chunk's S panel
    auto unit_col = [&](uint32_t lb_out, CohortAcc::Unit const& u, uint32_t& ch) -> uint64_t {
        const size_t rb = pos_lp[size_t(lb_out) * NP + u.s];
        auto const& v = stripes[lb_out];
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i].s == u.s) {
                ch = chunk_of[rb];
                return loc[lb_out][i] + (u.offset - v[i].col);
            }
        throw Error(MQ_ERR_INVALID, "hpsi: unit outside the stripes of its output block");
    };

    // --- T buffer layout per chunk (only tasks referenced by some alpha are executed) ---
    std::vector<RightTStacks> tstacks(n_chunks);
    for (uint32_t lb_in = 0; lb_in < NL; ++lb_in) {
        uint32_t ls_in = left_i.sizes[lb_in];
        for (auto& t : tlists[lb_in].tasks) {
            uint64_t csz = R.cohort_size(t.ci_eff);
            // the reference's flop model counts every T task (tasks.cpp:778-791)
            P.flops.flops_total += 2ull * ls_in * csz;
            P.flops.flops_gemm += 2ull * ls_in * csz;
            ++P.n_t_tasks;
            if (!t.used) continue;
            auto lr = R.clr[t.ci];
            uint32_t rs_ket = R.bra.sizes[lr.first], rs_bra = R.ket.sizes[lr.second];
            size_t rb = right_i.position(R.bra.charges[lr.first]);
            require(rb < right_i.size() && right_i.sizes[rb] == rs_ket,
                    "hpsi: right boundary index does not match the MPS right index");
            const uint32_t ch = chunk_of[task_rb(t)];
            t.rows = ls_in;
            t.size = uint64_t(ls_in) * R.nblk[t.ci_eff] * rs_bra;
            t.t_off = P.chunks[ch].t_elems;
            P.chunks[ch].t_elems += round_up(t.size, 32);
            r_used.insert(t.ci_eff);
            tstacks[ch].add(t, ls_in, rs_ket, rs_bra, int64_t(P.mps.rp_off[lb_in]));
            P.flops.flops_gemm_exec += 2ull * ls_in * csz;
        }
    }
    for (uint32_t ch = 0; ch < n_chunks; ++ch) {
        P.chunks[ch].t_gemm0 = P.t_gemms.size();
        P.chunks[ch].t_row0 = P.t_rows.size();
        P.chunks[ch].t_task0 = P.t_gemms_task.size();
        tstacks[ch].emit(R, BASE_PSI_IN, P.t_gemms, P.t_rows, &P.t_gemms_task);
        P.chunks[ch].t_gemm1 = P.t_gemms.size();
        P.chunks[ch].t_row1 = P.t_rows.size();
        P.chunks[ch].t_task1 = P.t_gemms_task.size();
    }

    // --- L panels and flop counts per cohort ---
    for (size_t c = 0; c < cohorts.size(); ++c) {
        auto& rec = cohorts[c];
        uint32_t ls_in = left_i.sizes[rec.lb_in], ls_out = left_i.sizes[rec.lb_out];
        require(rec.acc.stripe == P.mps.rp_cols[rec.lb_out], "hpsi: cohort stripe != psi block width");
        // L panel copy job
        CopyJob job{};
        job.dst = int64_t(l_off[rec.lb_out] + uint64_t(ls_out) * koff[c]);
        job.src = int64_t(L.data_off[rec.ci_eff]);
        if (rec.ci == rec.ci_eff) {
            check_cohort_shape(L, rec.ci_eff, ls_out, ls_in, "hpsi L panel");
            job.rows = int32_t(ls_out);
            job.cols = int32_t(uint64_t(rec.nb) * ls_in);
            job.nblocks = 1;
            job.trans = 0;
        } else {
            check_cohort_shape(L, rec.ci_eff, ls_in, ls_out, "hpsi L panel");
            job.rows = int32_t(ls_out);
            job.cols = int32_t(ls_in);
            job.nblocks = int32_t(rec.nb);
            job.trans = 1;
        }
        if (uint64_t(rec.nb) * ls_in * ls_out != 0) P.lpanel_jobs.push_back(job);
        l_used.insert(rec.ci_eff);
        for (auto& u : rec.acc.units) P.flops.flops_total += 2ull * ls_in * u.ms * u.terms.size();
        uint64_t g = 2ull * ls_out * ls_in * rec.nb * rec.acc.stripe;
        P.flops.flops_total += g;
        P.flops.flops_gemm += g;
        P.flops.flops_gemm_exec += g;
        P.flops.n_terms += rec.acc.n_terms;
    }
    // --- S lines, chunk by chunk ---
    uint64_t n_all_terms = 0;
    if (term_sources) {  // every term is emitted once, in some chunk
        for (auto const& rec : cohorts) n_all_terms += rec.acc.n_terms;
        P.term_src.resize(n_all_terms);
        P.term_beta.resize(n_all_terms);
    }
    for (uint32_t ch = 0; ch < n_chunks; ++ch) {
        P.chunks[ch].line0 = P.lines.size();
        emit_lines(cohorts.size(), P.lines, P.terms, [&](size_t c) -> auto const& { return cohorts[c].acc.units; },
                   [&](size_t c) { return cohorts[c].nb; }, [&](size_t c, auto& lines, auto& terms) {
            auto const& rec = cohorts[c];
            uint32_t ls_in = left_i.sizes[rec.lb_in];
            uint64_t ldS = ktot[rec.lb_out];
            for (auto const& u : rec.acc.units) {
                if (!u.ms) continue;
                uint32_t uch = 0;
                const uint64_t col = unit_col(rec.lb_out, u, uch);
                if (uch != ch) continue;
                const uint64_t base = s_off[ch][rec.lb_out] + koff[c] + col * ldS;
                uint32_t seek = 0;
                for (size_t l = 0; l < u.line_count.size(); ++l) {
                    SLine line{};
                    line.s_off = int64_t(base + uint64_t(u.line_b[l]) * ls_in);
                    line.lds = int32_t(ldS);
                    line.rows = int32_t(ls_in);
                    line.cols = int32_t(u.ms);
                    line.t0 = int32_t(terms.size());
                    for (uint32_t k = seek; k < seek + u.line_count[l]; ++k) {
                        TTask const& t = tlists[rec.lb_in].tasks[u.terms[k].ti];
                        check_tile(t, uint64_t(u.terms[k].col) * ls_in, ls_in, u.ms);
                        if (term_sources) {
                            P.term_src[terms.size()] = u.src[k].first;
                            P.term_beta[terms.size()] = u.src[k].second;
                        }
                        terms.push_back(STerm{u.terms[k].alpha, int64_t(t.t_off + uint64_t(u.terms[k].col) * ls_in)});
                    }
                    line.t1 = int32_t(terms.size());
                    seek += u.line_count[l];
                    lines.push_back(line);
                }
                add_empty_lines(u, rec.nb, int32_t(terms.size()), lines, [&](uint32_t b) {
                    SLine line{};
                    line.s_off = int64_t(base + uint64_t(b) * ls_in);
                    line.lds = int32_t(ldS);
                    line.rows = int32_t(ls_in);
                    line.cols = int32_t(u.ms);
                    return line;
                });
            }
        }, [&](size_t c, CohortAcc::Unit const& u) {
            // the units of the chunk's right sectors (a unit without columns has no terms)
            if (!u.ms) return false;
            uint32_t uch = 0;
            unit_col(cohorts[c].lb_out, u, uch);
            return uch == ch;
        });
        P.chunks[ch].line1 = P.lines.size();
    }
    require(!term_sources || P.terms.size() == n_all_terms, "hpsi: term sources do not cover the term table");
    P.n_cohorts = cohorts.size();
    // L*S products: per chunk and output block, one per run of the chunk's stripes that are
    // adjacent both in the psi block and in the chunk's S panel (one per block with one chunk);
    // and the reference's products, one per cohort (Cohort::contract, tasks.cpp:286-344), each
    // added into its output block: the K range of the cohort in its block's stacked L panel and
    // S panel, per run of the chunk's columns when the plan has several chunks
    for (uint32_t ch = 0; ch < n_chunks; ++ch) {
        // runs of the chunk's stripes of every output block: (first stripe, width)
        std::vector<std::vector<std::pair<size_t, uint64_t>>> runs(NL);
        for (uint32_t lb = 0; lb < NL; ++lb) {
            auto const& v = stripes[lb];
            for (size_t i = 0; i < v.size();) {
                if (chunk_of[v[i].rb] != ch) {
                    ++i;
                    continue;
                }
                size_t j = i + 1;
                uint64_t w = v[i].width;
                while (j < v.size() && chunk_of[v[j].rb] == ch && v[j].col == v[i].col + w && loc[lb][j] == loc[lb][i] + w) {
                    w += v[j].width;
                    ++j;
                }
                runs[lb].push_back({i, w});
                i = j;
            }
        }
        P.chunks[ch].c_gemm0 = P.c_gemms.size();
        for (uint32_t lb = 0; lb < NL; ++lb) {
            uint32_t ls = left_i.sizes[lb];
            for (auto const& r : runs[lb])
                P.c_gemms.push_back(gemm(int32_t(ls), int32_t(r.second), int32_t(ktot[lb]), BASE_LPANEL, int64_t(l_off[lb]),
                                         int32_t(ls), BASE_S, int64_t(s_off[ch][lb] + loc[lb][r.first] * ktot[lb]),
                                         int32_t(ktot[lb]), BASE_PSI_OUT,
                                         int64_t(P.mps.rp_off[lb] + stripes[lb][r.first].col * ls), int32_t(ls)));
        }
        P.chunks[ch].c_gemm1 = P.c_gemms.size();
        P.chunks[ch].c_coh0 = P.c_gemms_cohort.size();
        for (size_t c = 0; c < cohorts.size(); ++c) {
            auto const& rec = cohorts[c];
            const uint32_t lb = rec.lb_out, ls_out = left_i.sizes[lb];
            const int32_t kc = int32_t(uint64_t(rec.nb) * left_i.sizes[rec.lb_in]);
            for (auto const& r : runs[lb])
                P.c_gemms_cohort.push_back(gemm(int32_t(ls_out), int32_t(r.second), kc, BASE_LPANEL,
                                                int64_t(l_off[lb] + uint64_t(ls_out) * koff[c]), int32_t(ls_out), BASE_S,
                                                int64_t(s_off[ch][lb] + koff[c] + loc[lb][r.first] * ktot[lb]),
                                                int32_t(ktot[lb]), BASE_PSI_OUT,
                                                int64_t(P.mps.rp_off[lb] + stripes[lb][r.first].col * ls_out),
                                                int32_t(ls_out)));
        }
        P.chunks[ch].c_coh1 = P.c_gemms_cohort.size();
    }
    if (n_chunks == 1)  // one chunk: the S panel columns are the psi block's columns
        for (size_t lb = 0; lb < NL; ++lb)
            for (size_t i = 0; i < stripes[lb].size(); ++i)
                require(loc[lb][i] == stripes[lb][i].col, "hpsi: psi block columns are not its stripes side by side");
    for (auto const& ch : P.chunks) {
        P.t_elems += ch.t_elems;
        P.s_elems += ch.s_elems;
        P.t_buf = std::max(P.t_buf, ch.t_elems);
        P.s_buf = std::max(P.s_buf, ch.s_elems);
    }
    uint64_t bytes = 0;
    for (auto ci : l_used) bytes += L.cohort_size(ci);
    for (auto ci : r_used) bytes += R.cohort_size(ci);
    bytes += 2 * P.mps.elements();
    P.bytes_compulsory = 8 * bytes;
    return P;
}

// ----------------------------------------------------------------------------------------------
// left growth (reference: overlap_mpo_left_step, move_boundary.hpp:147-279, with
//               lshtm_t_tasks / lshtm_tasks, left_tasks.hpp:38-172)
// ----------------------------------------------------------------------------------------------
GrowPlanHost build_lstep_plan(Sym sym, MPO const& mpo, BoundaryIndex const& L, MPSLayout const& bra,
                              MPSLayout const& ket, bool symmetric) {
    GrowPlanHost P;
    require(bra.phys == ket.phys, "lstep: bra and ket physical indices differ");
    require(L.aux_dim == mpo.row_dim, "lstep: boundary aux dim != MPO row dim");
    validate_boundary_offsets(L);
    Index const& phys = ket.phys;
    Index const& left_i = ket.left;
    ProductBasis ket_rpb = ProductBasis::right_paired(phys, ket.right);
    ProductBasis bra_lpb = ProductBasis::left_paired(phys, bra.left);
    const bool su2 = sym == Sym::SU2U1;

    const size_t NP = phys.size(), NKL = left_i.size(), NKR = ket.right.size(), NBL = bra.left.size(),
                 NBR = bra.right.size();
    // lookup tables built once per step
    std::vector<uint32_t> kl_of(NKR * NP), bl_of(NBR * NP);  // left sector of (right sector - sigma)
    std::vector<uint64_t> kpb(NP * NKR), bpb(NP * NBL);
    for (size_t r = 0; r < NKR; ++r)
        for (size_t x = 0; x < NP; ++x) {
            kl_of[r * NP + x] = uint32_t(left_i.position(ket.right.charges[r] - phys.charges[x]));
            kpb[x * NKR + r] = ket_rpb(phys.charges[x], ket.right.charges[r]);
        }
    for (size_t r = 0; r < NBR; ++r)
        for (size_t x = 0; x < NP; ++x) bl_of[r * NP + x] = uint32_t(bra.left.position(bra.right.charges[r] - phys.charges[x]));
    for (size_t l = 0; l < NBL; ++l)
        for (size_t x = 0; x < NP; ++x) bpb[x * NBL + l] = bra_lpb(phys.charges[x], bra.left.charges[l]);
    std::vector<uint32_t> l_ci_kk(NKL * NKL), l_ci_bk(NBL * NKL);
    for (size_t x = 0; x < NKL; ++x)
        for (size_t y = 0; y < NKL; ++y) l_ci_kk[x * NKL + y] = L.cohort_index(left_i.charges[x], left_i.charges[y]);
    for (size_t x = 0; x < NBL; ++x)
        for (size_t y = 0; y < NKL; ++y) l_ci_bk[x * NKL + y] = L.cohort_index(bra.left.charges[x], left_i.charges[y]);
    struct BlkRef { uint32_t blk; uint32_t s_in; };
    std::vector<std::vector<BlkRef>> blk_by_out(mpo.ops.size() * NP);
    for (size_t o = 0; o < mpo.ops.size(); ++o)
        for (size_t bi = 0; bi < mpo.ops[o].blocks.size(); ++bi) {
            OpBlock const& blk = mpo.ops[o].blocks[bi];
            size_t so = phys.position(blk.rc), si = phys.position(blk.lc);
            if (so == NP || si == NP) continue;
            blk_by_out[o * NP + so].push_back(BlkRef{uint32_t(bi), uint32_t(si)});
        }

    struct CohRec {
        uint32_t rb_bra, rb_ket;
        CohortAcc acc;
        std::vector<int64_t> offs;
        uint32_t nrows;
    };
    std::vector<TList> tlists(NKR);
    std::string err;

    // MPO rows with a block in each left cohort
    std::vector<std::vector<uint32_t>> lblk(L.n_cohorts());
    for (uint32_t ci = 0; ci < L.n_cohorts(); ++ci)
        for (uint32_t b = 0; b < L.aux_dim; ++b)
            if (L.has_block(ci, b)) lblk[ci].push_back(b);

    // T work lists, one per right ket sector (lshtm_t_tasks)
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t rb_ket = 0; rb_ket < NKR; ++rb_ket) {
        try {
            uint32_t rs_ket = ket.right.sizes[rb_ket];
            TList& tl = tlists[rb_ket];
            for (size_t s = 0; s < NP; ++s) {
                size_t lb_ket = kl_of[rb_ket * NP + s];
                if (lb_ket == NKL) continue;
                uint64_t ket_offset = kpb[s * NKR + rb_ket];
                for (size_t lb_bra = 0; lb_bra < NKL; ++lb_bra) {
                    uint32_t ci = l_ci_kk[lb_bra * NKL + lb_ket];
                    if (ci == L.n_cohorts()) continue;
                    uint32_t ci_eff = L.tr[ci] ? l_ci_kk[lb_ket * NKL + lb_bra] : ci;
                    require(ci_eff < L.n_cohorts(), "lstep: transposed cohort without partner");
                    for (uint32_t ss = 0; ss < phys.sizes[s]; ++ss)
                        tl.add(ket_offset + ss * rs_ket, ci, ci_eff, uint32_t(lb_ket));
                }
            }
        } catch (std::exception const& e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) throw Error(MQ_ERR_INVALID, err);

    // cohorts, one per (right ket, right bra) sector pair (lshtm_tasks), in parallel over the pairs
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (uint32_t rb_ket = 0; rb_ket < NKR; ++rb_ket)
        for (uint32_t rb_bra = 0; rb_bra < NBR; ++rb_bra)
            if (std::abs(particle_number(sym, bra.right.charges[rb_bra]) -
                         particle_number(sym, ket.right.charges[rb_ket])) <= 2)
                pairs.emplace_back(rb_ket, rb_bra);
    std::vector<std::unique_ptr<CohRec>> recs(pairs.size());
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t pi = 0; pi < int64_t(pairs.size()); ++pi) {
        try {
            const uint32_t rb_ket = pairs[size_t(pi)].first, rb_bra = pairs[size_t(pi)].second;
            Charge rc_ket = ket.right.charges[rb_ket];
            uint32_t rs_ket = ket.right.sizes[rb_ket];
            TList& tl = tlists[rb_ket];
            Charge rc_bra = bra.right.charges[rb_bra];
            std::unique_ptr<CohRec> rp(new CohRec{rb_bra, rb_ket, CohortAcc(phys, mpo.col_dim), {}, 0});
            CohRec& rec = *rp;
            CohortAcc& coh = rec.acc;
            for (size_t s = 0; s < NP; ++s) {
                size_t lb_bra = bl_of[rb_bra * NP + s];
                if (lb_bra == NBL) continue;
                Charge lc_bra = bra.left.charges[lb_bra];
                uint32_t ls_bra = bra.left.sizes[lb_bra];
                coh.add_unit(s, phys.sizes[s], ls_bra, bpb[s * NBL + lb_bra]);
                // left cohorts reachable from this unit, one per input sector (or none)
                uint32_t ci_of[16];
                require(NP <= 16, "lstep: more than 16 physical sectors");
                std::vector<uint32_t> cand;
                for (size_t si = 0; si < NP; ++si) {
                    size_t lk = kl_of[rb_ket * NP + si];
                    ci_of[si] = lk == NKL ? L.n_cohorts() : l_ci_bk[lb_bra * NKL + lk];
                    if (ci_of[si] < L.n_cohorts()) cand.insert(cand.end(), lblk[ci_of[si]].begin(), lblk[ci_of[si]].end());
                }
                if (cand.empty()) continue;
                std::sort(cand.begin(), cand.end());
                cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
                // the terms of this unit, enumerated from the MPO rows b1 that have a left block and
                // then put in the reference's order: MPO column b2, then b1, term, operator block
                // (the b2-outer loop of lshtm_tasks, left_tasks.hpp:38-172)
                struct Item { uint32_t b2, b1, t, k; };
                std::vector<Item> items;
                for (uint32_t b1 : cand)
                    for (auto const& cell : mpo.rows[b1]) {
                        uint32_t b2 = cell.other;
                        if (symmetric && mpo.right_herm.skip(b2, rc_bra, rc_ket)) continue;
                        if (su2 && !triangle(spin_of(rc_ket), mpo.right_spin[b2], spin_of(rc_bra))) continue;
                        for (uint32_t t = cell.t0; t < cell.t1; ++t) {
                            auto const& brs = blk_by_out[size_t(mpo.terms[t].op) * NP + s];
                            for (uint32_t k = 0; k < brs.size(); ++k) {
                                uint32_t ci = ci_of[brs[k].s_in];
                                if (ci < L.n_cohorts() && L.has_block(ci, b1)) items.push_back(Item{b2, b1, t, k});
                            }
                        }
                    }
                std::stable_sort(items.begin(), items.end(), [](Item const& x, Item const& y) { return x.b2 < y.b2; });
                for (size_t q = 0; q < items.size(); ++q) {
                    Item const& it = items[q];
                    uint32_t op = mpo.terms[it.t].op;
                    SiteOp const& W = mpo.ops[op];
                    BlkRef const& br = blk_by_out[size_t(op) * NP + s][it.k];
                    OpBlock const& blk = W.blocks[br.blk];
                    size_t lb_ket = kl_of[rb_ket * NP + br.s_in];
                    uint32_t ci = ci_of[br.s_in];
                    Charge lc_ket = left_i.charges[lb_ket];
                    uint32_t ls_ket = left_i.sizes[lb_ket];
                    uint64_t ket_offset = kpb[size_t(br.s_in) * NKR + rb_ket];
                    int64_t left_offset = L.offset(ci, it.b1);
                    double cpl[4];
                    double scale = L.conjugate_scale(ci, it.b1) * mpo.terms[it.t].scale;
                    if (su2)
                        set_coupling(spin_of(lc_ket), std::abs(spin_of(rc_ket - lc_ket)), spin_of(rc_ket),
                                     mpo.left_spin[it.b1], W.spin, mpo.right_spin[it.b2], spin_of(lc_bra),
                                     std::abs(spin_of(rc_bra - lc_bra)), spin_of(rc_bra), scale, cpl);
                    else
                        cpl[0] = cpl[1] = cpl[2] = cpl[3] = scale;
                    op_iterate(blk, cpl, coh, s, rs_ket, tl, ket_offset, ci, uint32_t(left_offset / ls_ket));
                    if (q + 1 == items.size() || items[q + 1].b2 != it.b2) coh.add_line(it.b2);
                }
            }
            if (!coh.n_terms) continue;
            // finalize: compact the MPO indices with terms (Cohort::compute_mpo_offsets, tasks.cpp:507-526)
            uint64_t bs = uint64_t(bra.right.sizes[rb_bra]) * rs_ket;
            rec.offs.assign(mpo.col_dim, -1);
            std::vector<uint32_t> pos(mpo.col_dim, 0);
            uint32_t cnt = 0;
            for (uint32_t b = 0; b < mpo.col_dim; ++b)
                if (coh.bcount[b]) { pos[b] = cnt; rec.offs[b] = int64_t(bs * cnt); ++cnt; }
            rec.nrows = cnt;
            for (auto& u : coh.units)
                for (auto& b : u.line_b) b = pos[b];
            recs[size_t(pi)] = std::move(rp);
        } catch (std::exception const& e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) throw Error(MQ_ERR_INVALID, err);
    std::vector<CohRec> cohorts;
    for (auto& r : recs)
        if (r) cohorts.push_back(std::move(*r));

    // new boundary index (move_boundary.hpp:202-208)
    P.out = BoundaryIndex(bra.right, ket.right, mpo.col_dim);
    std::vector<uint32_t> new_ci(cohorts.size());
    for (size_t c = 0; c < cohorts.size(); ++c)
        new_ci[c] = P.out.add_cohort(cohorts[c].rb_bra, cohorts[c].rb_ket, cohorts[c].offs);
    if (symmetric) P.out.complement_transpose(sym, mpo.right_herm, true);
    P.out_elems = P.out.layout();

    // T buffer + T-GEMMs: T^b (ls_bra x rs_ket) = L_b (or L_b^T) * psi_ket slice (create_T_left, tasks.cpp:545-648)
    for (uint32_t rb_ket = 0; rb_ket < ket.right.size(); ++rb_ket) {
        uint32_t rs_ket = ket.right.sizes[rb_ket];
        for (auto& t : tlists[rb_ket].tasks) {
            auto lr = L.clr[t.ci];
            uint32_t bls = L.bra.sizes[lr.first], brs = L.ket.sizes[lr.second];
            uint32_t nb = L.nblk[t.ci_eff];
            uint64_t flops = 2ull * bls * brs * rs_ket * nb;
            P.flops.flops_total += flops;
            P.flops.flops_gemm += flops;
            if (!t.used) continue;
            require(brs == left_i.sizes[t.lb_ket], "lstep: boundary ket sector size != MPS left sector size");
            t.rows = bls;
            t.size = uint64_t(bls) * rs_ket * nb;
            t.t_off = P.t_elems;
            P.t_elems += round_up(t.size, 32);
            int64_t b = int64_t(ket.rp_off[t.lb_ket] + t.psi_col * brs);
            if (t.ci == t.ci_eff) {
                check_cohort_shape(L, t.ci_eff, bls, brs, "lstep T");
                P.t_gemms.push_back(gemm(bls, rs_ket, brs, BASE_BOUNDARY, int64_t(L.data_off[t.ci_eff]), bls, BASE_KET, b,
                                         brs, BASE_T, int64_t(t.t_off), bls, 0, int32_t(nb), int64_t(bls) * brs, 0,
                                         int64_t(bls) * rs_ket));
            } else {
                check_cohort_shape(L, t.ci_eff, brs, bls, "lstep T");
                P.t_gemms.push_back(gemm(bls, rs_ket, brs, BASE_BOUNDARY, int64_t(L.data_off[t.ci_eff]), brs, BASE_KET, b,
                                         brs, BASE_T, int64_t(t.t_off), bls, GEMM_TRANS_A, int32_t(nb),
                                         int64_t(bls) * brs, 0, int64_t(bls) * rs_ket));
            }
            P.flops.flops_gemm_exec += flops;
        }
    }

    // S buffers and final GEMMs: L'(rs_bra x nS*rs_ket) = A_bra^T (stripe x rs_bra)^T * S (prop_l, tasks.cpp:198-241)
    std::vector<uint64_t> sbase(cohorts.size());
    for (size_t c = 0; c < cohorts.size(); ++c) {
        auto& rec = cohorts[c];
        uint32_t rs_ket = ket.right.sizes[rec.rb_ket], rs_bra = bra.right.sizes[rec.rb_bra];
        uint64_t stripe = rec.acc.stripe;
        require(stripe == bra.lp_rows[rec.rb_bra], "lstep: cohort stripe != bra left-paired rows");
        sbase[c] = P.s_elems;
        P.s_elems += round_up(stripe * rec.nrows * rs_ket, 32);
        for (auto& u : rec.acc.units) P.flops.flops_total += 2ull * rs_ket * u.ms * u.terms.size();
        uint32_t ci_new = new_ci[c];
        P.c_gemms.push_back(gemm(int32_t(rs_bra), int32_t(rec.nrows * rs_ket), int32_t(stripe), BASE_BRA,
                                 int64_t(bra.lp_off[rec.rb_bra]), int32_t(stripe), BASE_S, int64_t(sbase[c]),
                                 int32_t(stripe), BASE_NEW, int64_t(P.out.data_off[ci_new]), int32_t(rs_bra),
                                 GEMM_TRANS_A));
        uint64_t g = 2ull * rs_bra * rec.nrows * rs_ket * stripe;
        P.flops.flops_total += g;
        P.flops.flops_gemm += g;
        P.flops.flops_gemm_exec += g;
        P.flops.n_terms += rec.acc.n_terms;
    }
    emit_lines(cohorts.size(), P.lines, P.terms, [&](size_t c) -> auto const& { return cohorts[c].acc.units; },
               [&](size_t c) { return cohorts[c].nrows; }, [&](size_t c, auto& lines, auto& terms) {
        auto const& rec = cohorts[c];
        uint32_t rs_ket = ket.right.sizes[rec.rb_ket];
        uint64_t stripe = rec.acc.stripe;
        for (auto const& u : rec.acc.units) {
            uint32_t seek = 0;
            for (size_t l = 0; l < u.line_count.size(); ++l) {
                SLine line{};
                line.s_off = int64_t(sbase[c] + u.offset + uint64_t(u.line_b[l]) * rs_ket * stripe);
                line.lds = int32_t(stripe);
                line.rows = int32_t(u.ms);
                line.cols = int32_t(rs_ket);
                line.t0 = int32_t(terms.size());
                for (uint32_t k = seek; k < seek + u.line_count[l]; ++k) {
                    TTask const& t = tlists[rec.rb_ket].tasks[u.terms[k].ti];
                    check_tile(t, uint64_t(u.terms[k].col) * rs_ket, u.ms, rs_ket);
                    terms.push_back(STerm{u.terms[k].alpha, int64_t(t.t_off + uint64_t(u.terms[k].col) * rs_ket)});
                }
                line.t1 = int32_t(terms.size());
                seek += u.line_count[l];
                lines.push_back(line);
            }
            add_empty_lines(u, rec.nrows, int32_t(terms.size()), lines, [&](uint32_t b) {
                SLine line{};
                line.s_off = int64_t(sbase[c] + u.offset + uint64_t(b) * rs_ket * stripe);
                line.lds = int32_t(stripe);
                line.rows = int32_t(u.ms);
                line.cols = int32_t(rs_ket);
                return line;
            });
        }
    });
    return P;
}

// ----------------------------------------------------------------------------------------------
// right growth (reference: overlap_mpo_right_step, move_boundary.hpp:281-404, with
//                rshtm_t_tasks / rshtm_tasks, right_tasks.hpp:39-167)
// ----------------------------------------------------------------------------------------------
GrowPlanHost build_rstep_plan(Sym sym, MPO const& mpo, BoundaryIndex const& R, MPSLayout const& bra,
                              MPSLayout const& ket, bool symmetric) {
    GrowPlanHost P;
    require(bra.phys == ket.phys, "rstep: bra and ket physical indices differ");
    require(R.aux_dim == mpo.col_dim, "rstep: boundary aux dim != MPO col dim");
    validate_boundary_offsets(R);
    Index const& phys = ket.phys;
    ProductBasis ket_rpb = ProductBasis::right_paired(phys, ket.right);
    ProductBasis bra_rpb = ProductBasis::right_paired(phys, bra.right);
    const bool su2 = sym == Sym::SU2U1;

    const size_t NP = phys.size(), NKL = ket.left.size(), NKR = ket.right.size(), NBL = bra.left.size(),
                 NBR = bra.right.size();
    std::vector<uint32_t> kr_of(NKL * NP), br_of(NBL * NP);  // right sector of (left sector + sigma)
    std::vector<uint64_t> kpb(NKL * NP), bpb(NBL * NP);
    for (size_t l = 0; l < NKL; ++l)
        for (size_t x = 0; x < NP; ++x) {
            Charge rc = ket.left.charges[l] + phys.charges[x];
            size_t r = ket.right.position(rc);
            kr_of[l * NP + x] = uint32_t(r);
            kpb[l * NP + x] = r < NKR ? ket_rpb(phys.charges[x], rc) : 0;
        }
    for (size_t l = 0; l < NBL; ++l)
        for (size_t x = 0; x < NP; ++x) {
            Charge rc = bra.left.charges[l] + phys.charges[x];
            size_t r = bra.right.position(rc);
            br_of[l * NP + x] = uint32_t(r);
            bpb[l * NP + x] = r < NBR ? bra_rpb(phys.charges[x], rc) : 0;
        }
    std::vector<uint32_t> r_ci_kb(NKR * NBR), r_ci_bk(NBR * NKR);
    for (size_t x = 0; x < NKR; ++x)
        for (size_t y = 0; y < NBR; ++y) {
            r_ci_kb[x * NBR + y] = R.cohort_index(ket.right.charges[x], bra.right.charges[y]);
            r_ci_bk[y * NKR + x] = R.cohort_index(bra.right.charges[y], ket.right.charges[x]);
        }
    struct BlkRef { uint32_t blk; uint32_t s_in; };
    std::vector<std::vector<BlkRef>> blk_by_out(mpo.ops.size() * NP);
    for (size_t o = 0; o < mpo.ops.size(); ++o)
        for (size_t bi = 0; bi < mpo.ops[o].blocks.size(); ++bi) {
            OpBlock const& blk = mpo.ops[o].blocks[bi];
            size_t so = phys.position(blk.rc), si = phys.position(blk.lc);
            if (so == NP || si == NP) continue;
            blk_by_out[o * NP + so].push_back(BlkRef{uint32_t(bi), uint32_t(si)});
        }

    struct CohRec {
        uint32_t lb_ket, lb_bra;
        CohortAcc acc;
        std::vector<int64_t> offs;
        uint32_t nrows;
    };
    std::vector<TList> tlists(NKL);
    std::string err;

    // MPO columns with a block in each right cohort
    std::vector<std::vector<uint32_t>> rblk(R.n_cohorts());
    for (uint32_t ci = 0; ci < R.n_cohorts(); ++ci)
        for (uint32_t b = 0; b < R.aux_dim; ++b)
            if (R.has_block(ci, b)) rblk[ci].push_back(b);

    // T work lists, one per left ket sector (rshtm_t_tasks)
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t lb_ket = 0; lb_ket < NKL; ++lb_ket) {
        try {
            TList& tl = tlists[lb_ket];
            for (size_t s = 0; s < NP; ++s) {
                size_t rb_ket = kr_of[lb_ket * NP + s];
                if (rb_ket == NKR) continue;
                uint32_t rs_ket = ket.right.sizes[rb_ket];
                uint64_t ket_offset = kpb[lb_ket * NP + s];
                for (size_t rb_bra = 0; rb_bra < NBR; ++rb_bra) {
                    uint32_t ci = r_ci_kb[rb_ket * NBR + rb_bra];
                    if (ci == R.n_cohorts()) continue;
                    uint32_t ci_eff = R.tr[ci] ? r_ci_bk[rb_bra * NKR + rb_ket] : ci;
                    require(ci_eff < R.n_cohorts(), "rstep: transposed cohort without partner");
                    for (uint32_t ss = 0; ss < phys.sizes[s]; ++ss) tl.add(ket_offset + ss * rs_ket, ci, ci_eff, lb_ket);
                }
            }
        } catch (std::exception const& e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) throw Error(MQ_ERR_INVALID, err);

    // cohorts, one per (left ket, left bra) sector pair (rshtm_tasks), in parallel over the pairs
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (uint32_t lb_ket = 0; lb_ket < NKL; ++lb_ket)
        for (uint32_t lb_bra = 0; lb_bra < NBL; ++lb_bra)
            if (std::abs(particle_number(sym, ket.left.charges[lb_ket]) -
                         particle_number(sym, bra.left.charges[lb_bra])) <= 2)
                pairs.emplace_back(lb_ket, lb_bra);
    std::vector<std::unique_ptr<CohRec>> recs(pairs.size());
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t pi = 0; pi < int64_t(pairs.size()); ++pi) {
        try {
            const uint32_t lb_ket = pairs[size_t(pi)].first, lb_bra = pairs[size_t(pi)].second;
            Charge lc_ket = ket.left.charges[lb_ket];
            uint32_t ls_ket = ket.left.sizes[lb_ket];
            TList& tl = tlists[lb_ket];
            Charge lc_bra = bra.left.charges[lb_bra];
            std::unique_ptr<CohRec> rp(new CohRec{lb_ket, lb_bra, CohortAcc(phys, mpo.row_dim), {}, 0});
            CohRec& rec = *rp;
            CohortAcc& coh = rec.acc;
            for (size_t s = 0; s < NP; ++s) {
                size_t rb_bra = br_of[lb_bra * NP + s];
                if (rb_bra == NBR) continue;
                Charge rc_bra = bra.right.charges[rb_bra];
                uint32_t rs_bra = bra.right.sizes[rb_bra];
                coh.add_unit(s, phys.sizes[s], rs_bra, bpb[lb_bra * NP + s]);
                // right cohorts reachable from this unit, one per input sector (or none)
                uint32_t ci_of[16];
                require(NP <= 16, "rstep: more than 16 physical sectors");
                std::vector<uint32_t> cand;
                for (size_t si = 0; si < NP; ++si) {
                    size_t rk = kr_of[lb_ket * NP + si];
                    ci_of[si] = rk == NKR ? R.n_cohorts() : r_ci_kb[rk * NBR + rb_bra];
                    if (ci_of[si] < R.n_cohorts()) cand.insert(cand.end(), rblk[ci_of[si]].begin(), rblk[ci_of[si]].end());
                }
                if (cand.empty()) continue;
                std::sort(cand.begin(), cand.end());
                cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
                std::shared_ptr<const Wigner9jCache> w9j;
                if (su2) w9j = wigner9j_cache(spin_of(lc_ket), spin_of(lc_bra), spin_of(rc_bra));
                // the terms of this unit, enumerated from the MPO columns b2 that have a right block
                // and then put in the reference's order: MPO row b1, then b2, term, operator block
                // (the b1-outer loop of rshtm_tasks, right_tasks.hpp:79-167)
                struct Item { uint32_t b1, b2, t, k; };
                std::vector<Item> items;
                for (uint32_t b2 : cand)
                    for (auto const& cell : mpo.cols[b2]) {
                        uint32_t b1 = cell.other;
                        if (symmetric && mpo.left_herm.skip(b1, lc_ket, lc_bra)) continue;
                        if (su2 && !triangle(spin_of(lc_ket), mpo.left_spin[b1], spin_of(lc_bra))) continue;
                        for (uint32_t t = cell.t0; t < cell.t1; ++t) {
                            auto const& brs = blk_by_out[size_t(mpo.terms[t].op) * NP + s];
                            for (uint32_t k = 0; k < brs.size(); ++k) {
                                uint32_t ci = ci_of[brs[k].s_in];
                                if (ci < R.n_cohorts() && R.has_block(ci, b2)) items.push_back(Item{b1, b2, t, k});
                            }
                        }
                    }
                std::stable_sort(items.begin(), items.end(), [](Item const& x, Item const& y) { return x.b1 < y.b1; });
                for (size_t q = 0; q < items.size(); ++q) {
                    Item const& it = items[q];
                    uint32_t op = mpo.terms[it.t].op;
                    SiteOp const& W = mpo.ops[op];
                    BlkRef const& br = blk_by_out[size_t(op) * NP + s][it.k];
                    OpBlock const& blk = W.blocks[br.blk];
                    size_t rb_ket = kr_of[lb_ket * NP + br.s_in];
                    uint32_t ci = ci_of[br.s_in];
                    Charge rc_ket = ket.right.charges[rb_ket];
                    uint32_t rs_ket = ket.right.sizes[rb_ket];
                    uint64_t ket_offset = kpb[lb_ket * NP + br.s_in];
                    int64_t right_offset = R.offset(ci, it.b2);
                    double cpl[4];
                    double scale = R.conjugate_scale(ci, it.b2) * mpo.terms[it.t].scale;
                    if (su2)
                        w9j->set_scale(mpo.left_spin[it.b1], W.spin, mpo.right_spin[it.b2], spin_of(rc_ket), scale, cpl);
                    else
                        cpl[0] = cpl[1] = cpl[2] = cpl[3] = scale;
                    op_iterate(blk, cpl, coh, s, rs_ket, tl, ket_offset, ci, uint32_t(right_offset / rs_ket));
                    if (q + 1 == items.size() || items[q + 1].b1 != it.b1) coh.add_line(it.b1);
                }
            }
            if (!coh.n_terms) continue;
            uint64_t bs = uint64_t(ls_ket) * bra.left.sizes[lb_bra];
            rec.offs.assign(mpo.row_dim, -1);
            std::vector<uint32_t> pos(mpo.row_dim, 0);
            uint32_t cnt = 0;
            for (uint32_t b = 0; b < mpo.row_dim; ++b)
                if (coh.bcount[b]) { pos[b] = cnt; rec.offs[b] = int64_t(bs * cnt); ++cnt; }
            rec.nrows = cnt;
            for (auto& u : coh.units)
                for (auto& b : u.line_b) b = pos[b];
            recs[size_t(pi)] = std::move(rp);
        } catch (std::exception const& e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) throw Error(MQ_ERR_INVALID, err);
    std::vector<CohRec> cohorts;
    for (auto& r : recs)
        if (r) cohorts.push_back(std::move(*r));

    // new boundary index (move_boundary.hpp:332-338): complement_transpose is always applied
    P.out = BoundaryIndex(ket.left, bra.left, mpo.row_dim);
    std::vector<uint32_t> new_ci(cohorts.size());
    for (size_t c = 0; c < cohorts.size(); ++c)
        new_ci[c] = P.out.add_cohort(cohorts[c].lb_ket, cohorts[c].lb_bra, cohorts[c].offs);
    P.out.complement_transpose(sym, mpo.left_herm, false);
    P.out_elems = P.out.layout();

    RightTStacks tstacks;
    for (uint32_t lb_ket = 0; lb_ket < ket.left.size(); ++lb_ket) {
        uint32_t ls_ket = ket.left.sizes[lb_ket];
        for (auto& t : tlists[lb_ket].tasks) {
            uint64_t flops = 2ull * ls_ket * R.cohort_size(t.ci_eff);
            P.flops.flops_total += flops;
            P.flops.flops_gemm += flops;
            if (!t.used) continue;
            auto lr = R.clr[t.ci];
            uint32_t rs_ket = R.bra.sizes[lr.first], rs_bra = R.ket.sizes[lr.second];
            t.rows = ls_ket;
            t.size = uint64_t(ls_ket) * R.nblk[t.ci_eff] * rs_bra;
            t.t_off = P.t_elems;
            P.t_elems += round_up(t.size, 32);
            require(ket.right.position(R.bra.charges[lr.first]) < ket.right.size() &&
                        ket.right.sizes[ket.right.position(R.bra.charges[lr.first])] == rs_ket,
                    "rstep: right boundary index does not match the ket right index");
            tstacks.add(t, ls_ket, rs_ket, rs_bra, int64_t(ket.rp_off[lb_ket]));
            P.flops.flops_gemm_exec += flops;
        }
    }
    tstacks.emit(R, BASE_KET, P.t_gemms, P.t_rows);

    // S (nS*ls_ket x stripe) and R'_b = S_b * B_bra^T per MPO row (prop_r, tasks.cpp:243-284)
    std::vector<uint64_t> sbase(cohorts.size());
    for (size_t c = 0; c < cohorts.size(); ++c) {
        auto& rec = cohorts[c];
        uint32_t ls_ket = ket.left.sizes[rec.lb_ket], ls_bra = bra.left.sizes[rec.lb_bra];
        uint64_t stripe = rec.acc.stripe;
        require(stripe == bra.rp_cols[rec.lb_bra], "rstep: cohort stripe != bra right-paired columns");
        uint64_t lds = uint64_t(rec.nrows) * ls_ket;
        sbase[c] = P.s_elems;
        P.s_elems += round_up(lds * stripe, 32);
        for (auto& u : rec.acc.units) P.flops.flops_total += 2ull * ls_ket * u.ms * u.terms.size();
        uint32_t ci_new = new_ci[c];
        P.c_gemms.push_back(gemm(int32_t(ls_ket), int32_t(ls_bra), int32_t(stripe), BASE_S, int64_t(sbase[c]),
                                 int32_t(lds), BASE_BRA, int64_t(bra.rp_off[rec.lb_bra]), int32_t(ls_bra), BASE_NEW,
                                 int64_t(P.out.data_off[ci_new]), int32_t(ls_ket), GEMM_TRANS_B, int32_t(rec.nrows),
                                 int64_t(ls_ket), 0, int64_t(ls_ket) * ls_bra));
        uint64_t g = 2ull * rec.nrows * ls_ket * stripe * ls_bra;
        P.flops.flops_total += g;
        P.flops.flops_gemm += g;
        P.flops.flops_gemm_exec += g;
        P.flops.n_terms += rec.acc.n_terms;
    }
    emit_lines(cohorts.size(), P.lines, P.terms, [&](size_t c) -> auto const& { return cohorts[c].acc.units; },
               [&](size_t c) { return cohorts[c].nrows; }, [&](size_t c, auto& lines, auto& terms) {
        auto const& rec = cohorts[c];
        uint32_t ls_ket = ket.left.sizes[rec.lb_ket];
        uint64_t lds = uint64_t(rec.nrows) * ls_ket;
        for (auto const& u : rec.acc.units) {
            uint32_t seek = 0;
            for (size_t l = 0; l < u.line_count.size(); ++l) {
                SLine line{};
                line.s_off = int64_t(sbase[c] + uint64_t(u.line_b[l]) * ls_ket + uint64_t(u.offset) * lds);
                line.lds = int32_t(lds);
                line.rows = int32_t(ls_ket);
                line.cols = int32_t(u.ms);
                line.t0 = int32_t(terms.size());
                for (uint32_t k = seek; k < seek + u.line_count[l]; ++k) {
                    TTask const& t = tlists[rec.lb_ket].tasks[u.terms[k].ti];
                    check_tile(t, uint64_t(u.terms[k].col) * ls_ket, ls_ket, u.ms);
                    terms.push_back(STerm{u.terms[k].alpha, int64_t(t.t_off + uint64_t(u.terms[k].col) * ls_ket)});
                }
                line.t1 = int32_t(terms.size());
                seek += u.line_count[l];
                lines.push_back(line);
            }
            add_empty_lines(u, rec.nrows, int32_t(terms.size()), lines, [&](uint32_t b) {
                SLine line{};
                line.s_off = int64_t(sbase[c] + uint64_t(b) * ls_ket + uint64_t(u.offset) * lds);
                line.lds = int32_t(lds);
                line.rows = int32_t(ls_ket);
                line.cols = int32_t(u.ms);
                return line;
            });
        }
    });
    return P;
}

// ---- block form of the MPO-term sum ------------------------------------------------------------
// small-tile lines: ordered by lanes per line (the power of two >= rows * cols), then by term
// count (longest first), with their terms copied in that order
template <class Pred>
static void build_small_lines(SBlockTable& B, HostTable<SLine> const& lines, HostTable<STerm> const& terms,
                              Pred is_small) {
    struct Key { int32_t lg, nt; uint32_t i; };
    std::vector<Key> keys;
    for (uint32_t i = 0; i < lines.size(); ++i) {
        SLine const& l = lines[i];
        if (!is_small(l)) continue;
        require(int64_t(l.rows) * l.cols <= SMALL_LINE_LANES, "S blocks: small-tile line wider than 256 elements");
        int32_t lg = 0;
        while ((1 << lg) < l.rows * l.cols) ++lg;
        keys.push_back(Key{lg, l.t1 - l.t0, i});
    }
    std::stable_sort(keys.begin(), keys.end(), [](Key const& a, Key const& b) {
        return a.lg != b.lg ? a.lg < b.lg : a.nt > b.nt;
    });
    std::vector<int64_t> t_start(keys.size() + 1, 0);
    for (size_t q = 0; q < keys.size(); ++q) t_start[q + 1] = t_start[q] + keys[q].nt;
    require(t_start.back() < int64_t(INT32_MAX), "S blocks: more than 2^31 small-tile terms");
    B.small_lines.resize(keys.size());
    B.small_terms.resize(size_t(t_start.back()));
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t q = 0; q < int64_t(keys.size()); ++q) {
        SLine l = lines[keys[size_t(q)].i];
        std::copy(terms.begin() + l.t0, terms.begin() + l.t1, B.small_terms.begin() + t_start[size_t(q)]);
        l.t0 = int32_t(t_start[size_t(q)]);
        l.t1 = int32_t(t_start[size_t(q) + 1]);
        B.small_lines[size_t(q)] = l;
    }
    for (size_t q = 0; q < keys.size();) {
        const int32_t lg = keys[q].lg, per = SMALL_LINE_LANES >> lg;
        int32_t n = 0;
        while (q + size_t(n) < keys.size() && keys[q + size_t(n)].lg == lg && n < per) ++n;
        B.small_work.push_back(SSmallWork{int32_t(q), n, lg, 0});
        q += size_t(n);
    }
    for (SLine const& l : B.small_lines) B.small_gathered += uint64_t(l.t1 - l.t0) * uint64_t(l.rows) * l.cols;
}

int32_t auto_small_tile_max(HostTable<SLine> const& lines, size_t l0, size_t l1) {
    uint64_t e = 0;
    for (size_t i = l0; i < l1; ++i) e += uint64_t(lines[i].rows) * uint64_t(lines[i].cols);
    return l1 > l0 && e <= 16 * uint64_t(l1 - l0) ? 64 : 0;
}

void mark_small_lines(HostTable<SLine>& lines, size_t l0, size_t l1, int32_t limit) {
    require(limit >= 0 && limit <= SMALL_LINE_LANES, "S blocks: small-tile limit outside 0..256");
    for (size_t i = l0; i < l1; ++i) {
        SLine& l = lines[i];
        const int64_t E = int64_t(l.rows) * l.cols;
        l.flags = (l.flags & ~SLINE_SMALL) | (E > 0 && E <= limit ? SLINE_SMALL : 0);
    }
}

SBlockTable build_sblocks(HostTable<SLine> const& lines, HostTable<STerm> const& terms, int32_t panel) {
    require(panel >= 0, "S blocks: negative panel");
    SBlockTable B;
    auto is_small = [&](SLine const& l) {
        return (l.flags & SLINE_SMALL) && l.rows > 0 && l.cols > 0;
    };
    bool any_small = false;
    for (SLine const& l : lines) any_small |= is_small(l);
    const int64_t nl = int64_t(lines.size());
    // first T tile of every line (lines without terms, i.e. zero tiles, sort last)
    std::vector<int64_t> tmin(lines.size(), INT64_MAX);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nl; ++i) {
        SLine const& l = lines[size_t(i)];
        int64_t mn = INT64_MAX;
        for (int32_t k = l.t0; k < l.t1; ++k) mn = std::min(mn, terms[size_t(k)].t_off);
        tmin[size_t(i)] = mn;
    }
    // lines by tile shape; inside a shape by their first T tile (lines reading the same T region
    // share most of their tiles), longer lines first
    std::map<std::pair<int32_t, int32_t>, std::vector<uint32_t>> by_shape;
    for (uint32_t i = 0; i < lines.size(); ++i)
        if (lines[i].rows > 0 && lines[i].cols > 0 && !is_small(lines[i]))
            by_shape[{lines[i].rows, lines[i].cols}].push_back(i);
    std::vector<std::vector<uint32_t>*> shapes;
    for (auto& kv : by_shape) shapes.push_back(&kv.second);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < int64_t(shapes.size()); ++q) {
        std::vector<uint32_t>& v = *shapes[size_t(q)];
        std::stable_sort(v.begin(), v.end(), [&](uint32_t x, uint32_t y) {
            if (tmin[x] != tmin[y]) return tmin[x] < tmin[y];
            return lines[x].t1 - lines[x].t0 > lines[y].t1 - lines[y].t0;
        });
    }
    // blocks: up to SBLOCK_LINES consecutive lines of one shape
    struct Span { std::vector<uint32_t> const* v; uint32_t s0, n; };
    std::vector<Span> spans;
    for (auto* v : shapes)
        for (uint32_t s0 = 0; s0 < v->size(); s0 += SBLOCK_LINES)
            spans.push_back(Span{v, s0, std::min<uint32_t>(SBLOCK_LINES, uint32_t(v->size()) - s0)});
    const int64_t nb = int64_t(spans.size());
    struct BlockTabs {
        std::vector<int64_t> slots;   // 4 per k-step
        std::vector<uint32_t> words;  // per k-step
        std::vector<double> alpha;    // sub-blocks
        uint32_t nu = 0;
    };
    std::vector<BlockTabs> tabs(spans.size());
#pragma omp parallel
    {
        // union of one block's T tiles, found with an open-addressing table in order of first use
        // (lines in block order, each line's terms in term order) with the groups reading each tile
        // and each line's coefficient (summed in the line's term order); then ordered by
        // (groups, T offset) and cut into k-steps
        std::vector<int64_t> key;
        std::vector<int32_t> slot;
        std::vector<int64_t> u;
        std::vector<uint32_t> pat;
        std::vector<double> al;
        std::vector<std::pair<std::pair<uint32_t, int64_t>, int32_t>> order;
#pragma omp for schedule(dynamic, 16)
        for (int64_t bi = 0; bi < nb; ++bi) {
            Span const& sp = spans[size_t(bi)];
            size_t nref = 0;
            for (uint32_t q = 0; q < sp.n; ++q) {
                SLine const& l = lines[(*sp.v)[sp.s0 + q]];
                nref += size_t(l.t1 - l.t0);
            }
            size_t cap = 16;
            while (cap < 2 * nref) cap *= 2;
            key.assign(cap, -1);
            slot.resize(cap);
            u.clear();
            pat.clear();
            al.clear();
            for (uint32_t q = 0; q < sp.n; ++q) {
                SLine const& l = lines[(*sp.v)[sp.s0 + q]];
                for (int32_t k = l.t0; k < l.t1; ++k) {
                    const int64_t t = terms[size_t(k)].t_off;
                    size_t h = size_t((uint64_t(t) * 0x9E3779B97F4A7C15ull) >> 20) & (cap - 1);
                    while (key[h] != -1 && key[h] != t) h = (h + 1) & (cap - 1);
                    if (key[h] == -1) {
                        key[h] = t;
                        slot[h] = int32_t(u.size());
                        u.push_back(t);
                        pat.push_back(0);
                        al.resize(al.size() + SBLOCK_LINES, 0.0);
                    }
                    pat[size_t(slot[h])] |= 1u << (q / SBLOCK_GROUP);
                    al[size_t(slot[h]) * SBLOCK_LINES + q] += terms[size_t(k)].alpha;
                }
            }
            order.resize(u.size());
            for (size_t i = 0; i < u.size(); ++i) order[i] = std::make_pair(std::make_pair(pat[i], u[i]), int32_t(i));
            std::sort(order.begin(), order.end());
            BlockTabs& bt = tabs[size_t(bi)];
            bt.nu = uint32_t(u.size());
            const size_t nk = (u.size() + 3) / 4;
            bt.slots.assign(4 * nk, -1);
            bt.words.resize(nk);
            uint32_t na = 0;
            for (size_t k = 0; k < nk; ++k) {
                uint32_t m = 0;
                for (size_t t = 4 * k; t < std::min(u.size(), 4 * k + 4); ++t) {
                    bt.slots[t] = order[t].first.second;
                    m |= order[t].first.first;
                }
                require(na < (1u << 24), "S blocks: too many coefficient sub-blocks in one block");
                bt.words[k] = m | (na << 8);
                for (int g = 0; g < SBLOCK_GROUPS; ++g) {
                    if (!((m >> g) & 1)) continue;
                    const size_t a0 = bt.alpha.size();
                    bt.alpha.resize(a0 + 64, 0.0);
                    for (size_t t = 4 * k; t < std::min(u.size(), 4 * k + 4); ++t) {
                        const int32_t src = order[t].second;
                        for (int i = 0; i < SBLOCK_GROUP; ++i)
                            bt.alpha[a0 + (t - 4 * k) * SBLOCK_GROUP + size_t(i)] =
                                al[size_t(src) * SBLOCK_LINES + size_t(g * SBLOCK_GROUP + i)];
                    }
                    ++na;
                }
            }
        }
    }
    // concatenate in block order
    std::vector<int64_t> k_off(spans.size() + 1, 0), a_off(spans.size() + 1, 0);
    for (size_t bi = 0; bi < spans.size(); ++bi) {
        k_off[bi + 1] = k_off[bi] + int64_t(tabs[bi].words.size());
        a_off[bi + 1] = a_off[bi] + int64_t(tabs[bi].alpha.size() / 64);
    }
    require(k_off.back() < int64_t(INT32_MAX), "S blocks: more than 2^31 k-steps");
    B.utile.resize(size_t(4 * k_off.back()));
    B.kstep.resize(size_t(k_off.back()));
    B.ualpha.resize(size_t(64 * a_off.back()));
    B.lsoff.assign(spans.size() * SBLOCK_LINES, 0);
    B.llds.assign(spans.size() * SBLOCK_LINES, 1);
    B.blocks.resize(spans.size());
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t bi = 0; bi < nb; ++bi) {
        Span const& sp = spans[size_t(bi)];
        SLine const& first = lines[(*sp.v)[sp.s0]];
        BlockTabs& bt = tabs[size_t(bi)];
        std::copy(bt.slots.begin(), bt.slots.end(), B.utile.begin() + 4 * k_off[size_t(bi)]);
        std::copy(bt.words.begin(), bt.words.end(), B.kstep.begin() + k_off[size_t(bi)]);
        std::copy(bt.alpha.begin(), bt.alpha.end(), B.ualpha.begin() + 64 * a_off[size_t(bi)]);
        for (uint32_t q = 0; q < sp.n; ++q) {
            SLine const& l = lines[(*sp.v)[sp.s0 + q]];
            B.lsoff[size_t(bi) * SBLOCK_LINES + q] = l.s_off;
            B.llds[size_t(bi) * SBLOCK_LINES + q] = l.lds;
        }
        SBlock& blk = B.blocks[size_t(bi)];
        blk.a_off = a_off[size_t(bi)];
        blk.k_off = int32_t(k_off[size_t(bi)]);
        blk.nk = int32_t(bt.words.size());
        blk.na = int32_t(bt.alpha.size() / 64);
        blk.rows = first.rows;
        blk.E = first.rows * first.cols;
        blk.nlines = int32_t(sp.n);
        blk.line0 = int32_t(bi * SBLOCK_LINES);
        blk.pad = 0;
    }
    // work items: block by block, or (panel > 0) per panel of up to `panel` consecutive blocks of
    // one tile shape, element range by element range, so that neighbouring blocks -- which share
    // most of their union tiles -- read the same slice of those tiles in consecutive items
    for (int32_t p0 = 0; p0 < int32_t(nb);) {
        int32_t p1 = p0 + 1;
        if (panel > 0)
            while (p1 < int32_t(nb) && p1 - p0 < panel && B.blocks[size_t(p1)].E == B.blocks[size_t(p0)].E &&
                   B.blocks[size_t(p1)].rows == B.blocks[size_t(p0)].rows)
                ++p1;
        for (int32_t e = 0; e < B.blocks[size_t(p0)].E; e += SBLOCK_ELEMS)
            for (int32_t bi = p0; bi < p1; ++bi) B.work.push_back(SWork{bi, e});
        p0 = p1;
    }
    for (int32_t bi = 0; bi < int32_t(nb); ++bi) {
        SBlock const& blk = B.blocks[size_t(bi)];
        B.gathered += uint64_t(tabs[size_t(bi)].nu) * uint64_t(blk.E);
        B.max_union = std::max<uint64_t>(B.max_union, tabs[size_t(bi)].nu);
        B.mac += uint64_t(blk.na) * 64 * uint64_t(blk.E);
    }
    if (any_small) build_small_lines(B, lines, terms, is_small);
    return B;
}

std::vector<GemmDesc> adjoint_y_gemms(std::vector<GemmDesc> const& c_gemms) {
    std::vector<GemmDesc> ys;
    ys.reserve(c_gemms.size());
    for (GemmDesc const& g : c_gemms) {
        require(g.flags == 0 && g.batch == 1 && g.base_a == BASE_LPANEL && g.base_b == BASE_S && g.base_c == BASE_PSI_OUT,
                "MPO-term adjoint: unexpected L*S product");
        // Y (K x N, where the product reads S) = L_panel^T (K x M) * psi (M x N, where it writes psi')
        ys.push_back(gemm(g.K, g.N, g.M, BASE_LPANEL, g.a, g.lda, BASE_PSI_IN, g.c, g.ldc, BASE_S, g.b, g.ldb,
                          GEMM_TRANS_A));
    }
    return ys;
}

std::vector<TermGradWork> term_grad_work(HostTable<SLine> const& lines, size_t l0, size_t l1) {
    std::vector<TermGradWork> w;
    for (size_t li = l0; li < l1; ++li) {
        SLine const& l = lines[li];
        const int64_t E = int64_t(l.rows) * l.cols;
        int lg = 0;
        while (lg < 6 && (int64_t(1) << lg) < E) ++lg;
        require(li < size_t(INT32_MAX), "MPO-term adjoint: line index overflow");
        for (int32_t t = l.t0; t < l.t1; t += TERM_GRAD_RUN)
            w.push_back(TermGradWork{int32_t(li), t, l.t1 - t > TERM_GRAD_RUN ? t + TERM_GRAD_RUN : l.t1, lg});
    }
    return w;
}

void check_sblocks(SBlockTable const& B, uint64_t t_size, uint64_t s_size) {
    require(B.lsoff.size() == B.blocks.size() * SBLOCK_LINES && B.llds.size() == B.lsoff.size() &&
                B.utile.size() == 4 * B.kstep.size() && B.ualpha.size() % 64 == 0, "S blocks: malformed tables");
    const uint64_t n_sub = B.ualpha.size() / 64;
    int bad = 0;  // 1 malformed block, 2 tile outside T, 3 line outside S, 4 malformed work item
#pragma omp parallel for schedule(dynamic, 256) reduction(max : bad)
    for (int64_t bi = 0; bi < int64_t(B.blocks.size()); ++bi) {
        SBlock const& b = B.blocks[size_t(bi)];
        if (!(b.nk >= 0 && b.k_off >= 0 && uint64_t(b.k_off) + uint64_t(b.nk) <= B.kstep.size() && b.na >= 0 &&
              b.a_off >= 0 && uint64_t(b.a_off) + uint64_t(b.na) <= n_sub && b.E > 0 && b.rows > 0 &&
              b.E % b.rows == 0 && b.nlines > 0 && b.nlines <= SBLOCK_LINES && b.line0 >= 0 &&
              uint64_t(b.line0) + SBLOCK_LINES <= B.lsoff.size())) {
            bad = std::max(bad, 1);
            continue;
        }
        // k-step words: group masks within the block's groups, sub-blocks consecutive and inside
        // the block's range
        uint32_t next = 0;
        const uint32_t gmask = (1u << ((b.nlines + SBLOCK_GROUP - 1) / SBLOCK_GROUP)) - 1;
        for (int32_t k = 0; k < b.nk; ++k) {
            const uint32_t w = B.kstep[size_t(b.k_off) + k], m = w & 0xffu;
            if ((w >> 8) != next || (m & ~gmask)) bad = std::max(bad, 1);
            next += uint32_t(__builtin_popcount(m));
        }
        if (next != uint32_t(b.na)) bad = std::max(bad, 1);
        for (int64_t k = 4 * int64_t(b.k_off); k < 4 * (int64_t(b.k_off) + b.nk); ++k) {
            int64_t o = B.utile[size_t(k)];
            if (!(o == -1 || (o >= 0 && uint64_t(o) + uint64_t(b.E) <= t_size))) bad = std::max(bad, 2);
        }
        int32_t cols = b.E / b.rows;
        for (int32_t i = 0; i < b.nlines; ++i) {
            int64_t so = B.lsoff[size_t(b.line0) + i];
            int32_t ld = B.llds[size_t(b.line0) + i];
            if (!(so >= 0 && ld >= b.rows && uint64_t(so + (b.rows - 1) + int64_t(cols - 1) * ld) < s_size))
                bad = std::max(bad, 3);
        }
    }
    // every element range of every block in exactly one work item (in any order)
    if (bad == 0) {
        std::vector<uint64_t> first(B.blocks.size() + 1, 0);
        for (size_t bi = 0; bi < B.blocks.size(); ++bi)
            first[bi + 1] = first[bi] + uint64_t((B.blocks[bi].E + SBLOCK_ELEMS - 1) / SBLOCK_ELEMS);
        std::vector<char> seen(first.back(), 0);
        if (B.work.size() != first.back()) bad = std::max(bad, 4);
        for (SWork const& w : B.work) {
            if (!(w.blk >= 0 && size_t(w.blk) < B.blocks.size() && w.n0 >= 0 && w.n0 < B.blocks[size_t(w.blk)].E &&
                  w.n0 % SBLOCK_ELEMS == 0)) {
                bad = std::max(bad, 4);
                continue;
            }
            char& c = seen[first[size_t(w.blk)] + uint64_t(w.n0 / SBLOCK_ELEMS)];
            if (c) bad = std::max(bad, 4);
            c = 1;
        }
    }
    // small-tile tables: 5 malformed line, 6 term tile outside T, 7 malformed work item
    const int64_t nsl = int64_t(B.small_lines.size());
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : bad)
    for (int64_t i = 0; i < nsl; ++i) {
        SLine const& l = B.small_lines[size_t(i)];
        const int64_t E = int64_t(l.rows) * l.cols;
        if (!(l.rows > 0 && l.cols > 0 && E <= SMALL_LINE_LANES && l.lds >= l.rows && l.s_off >= 0 &&
              uint64_t(l.s_off + (l.rows - 1) + int64_t(l.cols - 1) * l.lds) < s_size && l.t0 >= 0 && l.t0 <= l.t1 &&
              uint64_t(l.t1) <= B.small_terms.size())) {
            bad = std::max(bad, 5);
            continue;
        }
        for (int32_t k = l.t0; k < l.t1; ++k) {
            int64_t o = B.small_terms[size_t(k)].t_off;
            if (!(o >= 0 && uint64_t(o) + uint64_t(E) <= t_size)) bad = std::max(bad, 6);
        }
    }
    for (SSmallWork const& w : B.small_work) {
        bool ok = w.lg >= 0 && (1 << w.lg) <= SMALL_LINE_LANES && w.n > 0 && w.n <= (SMALL_LINE_LANES >> w.lg) &&
                  w.line0 >= 0 && int64_t(w.line0) + w.n <= nsl;
        for (int32_t q = 0; ok && q < w.n; ++q) {
            SLine const& l = B.small_lines[size_t(w.line0 + q)];
            ok = int64_t(l.rows) * l.cols <= (int64_t(1) << w.lg);
        }
        if (!ok) bad = std::max(bad, 7);
    }
    require(bad != 1, "S blocks: malformed block");
    require(bad != 2, "S blocks: tile outside the T buffer");
    require(bad != 3, "S blocks: line outside the S buffer");
    require(bad != 4, "S blocks: malformed work item");
    require(bad != 5, "S blocks: malformed small-tile line");
    require(bad != 6, "S blocks: small-tile term outside the T buffer");
    require(bad != 7, "S blocks: malformed small-tile work item");
}

// ---- grouped GEMM schedule ---------------------------------------------------------------------
GemmPolicy gemm_policy(int i) {
    require(i >= 0 && i < GEMM_N_POLICIES, "gemm policy: index out of range");
    GemmPolicy p;
    switch (i) {
        case 1: p.wide = 1; break;              // 64- and 128-column tiles only by M
        case 2: p.wide = 2; break;              // 128-column tiles whenever N > 64
        case 3: p.tiles96 = 0; break;           // 128-row tiles from 65 rows
        case 4: p.target_items = 8192; break;   // split along K more readily
        case 5: p.target_items = 2048; break;   // split along K less
        case 6: p.min_chunk = 512; break;       // longer K ranges per split tile
        default: break;
    }
    return p;
}

uint64_t gemm_stage_key(std::vector<GemmDesc> const& probs) {
    double w = 0, lm = 0, ln = 0, lk = 0, f32 = 0, f64 = 0, f96 = 0, tiles = 0;
    bool rowmap = false;
    for (auto const& g : probs) {
        if (g.M <= 0 || g.N <= 0 || g.K <= 0) continue;
        const double f = 2.0 * g.M * g.N * g.K * g.batch;
        w += f;
        lm += f * std::log2(double(g.M));
        ln += f * std::log2(double(g.N));
        lk += f * std::log2(double(g.K));
        if (g.M <= 32) f32 += f;
        else if (g.M <= 64) f64 += f;
        else if (g.M <= 96) f96 += f;
        tiles += std::ceil(g.M / 128.0) * std::ceil(g.N / 64.0) * g.batch;
        rowmap |= (g.flags & GEMM_ROWMAP) != 0;
    }
    if (w == 0) return 0;
    auto q = [](double x, double scale, uint64_t bits) {
        return std::min<uint64_t>((uint64_t(1) << bits) - 1, uint64_t(std::max(0.0, std::floor(x * scale + 0.5))));
    };
    uint64_t k = 0;
    k = (k << 7) | q(lm / w, 4, 7);
    k = (k << 7) | q(ln / w, 4, 7);
    k = (k << 7) | q(lk / w, 4, 7);
    k = (k << 4) | q(f32 / w, 8, 4);
    k = (k << 4) | q(f64 / w, 8, 4);
    k = (k << 4) | q(f96 / w, 8, 4);
    k = (k << 6) | q(std::log2(std::max(1.0, tiles)), 1, 6);
    k = (k << 1) | uint64_t(rowmap);
    return k + 1;
}

GemmSchedule build_gemm_schedule(std::vector<GemmDesc> const& probs, GemmPolicy const& pol) {
    const int64_t target_items = pol.target_items;
    const int32_t min_chunk = pol.min_chunk;
    require(min_chunk >= 16 && min_chunk % 16 == 0, "gemm schedule: chunk must be a multiple of 16");
    require(pol.wide >= 0 && pol.wide <= 2, "gemm schedule: wide mode 0, 1 or 2");
    GemmSchedule S;
    // 96-row tiles (for 64 < M <= 96) only for problems with a plain A operand: a stacked problem's
    // threads each load one row, which needs a tile height that divides 256.  64x128 tiles (for
    // 32 < M <= 64, N > 64) only when the stage has too few tiles to fill the device and is split
    // along K anyway: then the wider tile halves the A re-reads at no cost in parallelism
    bool wide = pol.wide == 2;
    auto shape_of = [&](GemmDesc const& g) {
        const bool w = wide && g.N > 64 && !(g.flags & GEMM_ROWMAP);
        return g.M <= 32 ? GEMM_32x128
             : g.M <= 64 ? ((wide && g.N > 64) ? GEMM_64x128 : GEMM_64x64)
             : (pol.tiles96 && g.M <= 96 && !(g.flags & GEMM_ROWMAP)) ? GEMM_96x64
             : w ? GEMM_128x128 : GEMM_128x64;
    };
    auto tiles_of = [&](GemmDesc const& g) {
        int sh = shape_of(g);
        return int64_t((g.M + gemm_bm(sh) - 1) / gemm_bm(sh)) * ((g.N + gemm_bn(sh) - 1) / gemm_bn(sh)) * g.batch;
    };
    auto count = [&](int64_t& n_tiles, double& work) {
        n_tiles = 0;
        work = 0;
        for (auto const& g : probs) {
            if (g.M <= 0 || g.N <= 0) continue;
            n_tiles += tiles_of(g);
            work += double(tiles_of(g)) * g.K;
        }
    };
    int64_t n_tiles = 0;
    double work = 0;
    count(n_tiles, work);
    if (n_tiles < target_items && target_items > 0 && pol.wide == 0) {
        wide = true;
        count(n_tiles, work);
    }
    // K range per workgroup when the stage alone cannot fill the device
    int64_t chunk = INT64_MAX;
    if (n_tiles < target_items && target_items > 0) {
        int64_t c = int64_t(work / double(target_items));
        chunk = std::max<int64_t>(min_chunk, (c + 15) / 16 * 16);
    }
    for (int32_t p = 0; p < int32_t(probs.size()); ++p) {
        GemmDesc const& g = probs[p];
        if (g.M <= 0 || g.N <= 0) continue;
        int sh = shape_of(g);
        int bm = gemm_bm(sh), bn = gemm_bn(sh);
        int32_t tm = (g.M + bm - 1) / bm, tn = (g.N + bn - 1) / bn;
        int64_t splits = 1, step = g.K;
        if (g.K > chunk && !(g.flags & GEMM_ROWMAP)) {  // stacked problems have no split path
            splits = (g.K + chunk - 1) / chunk;
            step = ((g.K + splits - 1) / splits + 15) / 16 * 16;
            splits = (g.K + step - 1) / step;
        }
        for (int32_t b = 0; b < g.batch; ++b)
            for (int32_t j = 0; j < tn; ++j)
                for (int32_t i = 0; i < tm; ++i) {
                    if (splits == 1) {
                        S.tiles[sh].push_back(GemmTile{p, b, i, j, 0, g.K, -1});
                        continue;
                    }
                    int64_t base = int64_t(S.ws_elems);
                    for (int64_t s = 0; s < splits; ++s)
                        S.tiles[sh].push_back(GemmTile{p, b, i, j, int32_t(s * step),
                                                       int32_t(std::min<int64_t>(g.K, (s + 1) * step)),
                                                       base + s * int64_t(bm) * bn});
                    S.reduces.push_back(GemmReduce{p, b, i, j, int32_t(splits), sh, base});
                    S.ws_elems += uint64_t(splits) * bm * bn;
                }
    }
    // longest reductions first for load balance
    for (auto& v : S.tiles)
        std::stable_sort(v.begin(), v.end(), [](GemmTile const& a, GemmTile const& b) { return a.k1 - a.k0 > b.k1 - b.k0; });
    return S;
}

// ---- host-side bounds validation of every device table (a table entry that would leave its
//      buffer is rejected before anything is launched) -----------------------------------------
void check_gemms(std::vector<GemmDesc> const& v, const uint64_t* sizes, std::vector<GemmRow> const& rows) {
    auto span = [](int64_t off, int64_t stride, int32_t batch, int64_t rows, int64_t cols, int64_t ld) {
        // [lo, hi] accessed element offsets of a rows x cols column-major operand
        int64_t last = off + int64_t(batch - 1) * stride + (rows - 1) + (cols - 1) * ld;
        return std::make_pair(off, last);
    };
    for (auto const& g : v) {
        require(g.M >= 0 && g.N >= 0 && g.K >= 0 && g.batch >= 1, "gemm table: negative dimension");
        require(g.sa >= 0 && g.sb >= 0 && g.sc >= 0, "gemm table: negative stride");
        for (int b : {g.base_a, g.base_b, g.base_c}) require(b >= 0 && b < BASE_COUNT, "gemm table: bad base");
        if (g.M == 0 || g.N == 0) continue;
        bool ta = g.flags & GEMM_TRANS_A, tb = g.flags & GEMM_TRANS_B;
        if (g.flags & GEMM_ROWMAP) {
            require(!ta && g.sa == 0 && g.a >= 0 && uint64_t(g.a) + uint64_t(g.M) <= rows.size(),
                    "gemm table: stacked problem outside its row table");
            for (int64_t m = g.a; m < g.a + g.M; ++m) {
                GemmRow const& r = rows[size_t(m)];
                require(r.ld >= 1, "gemm table: row with ld < 1");
                if (g.K > 0)
                    require(r.a >= 0 && uint64_t(r.a + int64_t(g.K - 1) * r.ld) < sizes[g.base_a],
                            "gemm table: A row outside its buffer");
                require(r.c >= 0 && uint64_t(r.c + (int64_t(g.batch - 1) * g.sc + g.N - 1) * r.ld) < sizes[g.base_c],
                        "gemm table: C row outside its buffer");
            }
            if (g.K > 0) {
                int64_t br = tb ? g.N : g.K, bc = tb ? g.K : g.N;
                require(g.ldb >= br, "gemm table: ldb too small");
                auto sb = span(g.b, g.sb, g.batch, br, bc, g.ldb);
                require(sb.first >= 0 && uint64_t(sb.second) < sizes[g.base_b], "gemm table: B outside its buffer");
            }
            continue;
        }
        if (g.K > 0) {
            int64_t ar = ta ? g.K : g.M, ac = ta ? g.M : g.K;
            require(g.lda >= ar, "gemm table: lda too small");
            auto sa = span(g.a, g.sa, g.batch, ar, ac, g.lda);
            require(sa.first >= 0 && uint64_t(sa.second) < sizes[g.base_a], "gemm table: A outside its buffer");
            int64_t br = tb ? g.N : g.K, bc = tb ? g.K : g.N;
            require(g.ldb >= br, "gemm table: ldb too small");
            auto sb = span(g.b, g.sb, g.batch, br, bc, g.ldb);
            require(sb.first >= 0 && uint64_t(sb.second) < sizes[g.base_b], "gemm table: B outside its buffer");
        }
        require(g.ldc >= g.M, "gemm table: ldc too small");
        auto sc = span(g.c, g.sc, g.batch, g.M, g.N, g.ldc);
        require(sc.first >= 0 && uint64_t(sc.second) < sizes[g.base_c], "gemm table: C outside its buffer");
    }
}

void check_lines(HostTable<SLine> const& lines, HostTable<STerm> const& terms, uint64_t t_size, uint64_t s_size) {
    int bad = 0;  // 1 malformed line, 2 line outside S, 3 term outside T
#pragma omp parallel for schedule(static) reduction(max : bad)
    for (int64_t i = 0; i < int64_t(lines.size()); ++i) {
        SLine const& l = lines[size_t(i)];
        if (!(l.rows >= 0 && l.cols >= 0 && l.lds >= l.rows && l.t0 >= 0 && l.t1 >= l.t0 &&
              uint64_t(l.t1) <= terms.size())) {
            bad = std::max(bad, 1);
            continue;
        }
        if (!l.rows || !l.cols) continue;
        int64_t last = l.s_off + (l.rows - 1) + int64_t(l.cols - 1) * l.lds;
        if (!(l.s_off >= 0 && uint64_t(last) < s_size)) bad = std::max(bad, 2);
        for (int t = l.t0; t < l.t1; ++t)
            if (!(terms[size_t(t)].t_off >= 0 &&
                  uint64_t(terms[size_t(t)].t_off) + uint64_t(l.rows) * l.cols <= t_size))
                bad = std::max(bad, 3);
    }
    require(bad != 1, "S table: malformed line");
    require(bad != 2, "S table: line outside the S buffer");
    require(bad != 3, "S table: term outside the T buffer");
}


}  // namespace mq
