// Host-side executor of the product's device tables -- test code.
//
// Builds the product's plans (qcmaquis_amd/csrc/plans.cpp) and runs their GEMM problems, S lines
// and copy / reshape jobs on the CPU with the same index arithmetic as the gfx950 kernels, so the
// plan builder and its bounds checks are tested without a GPU.  Not linked into the product.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

#include "descriptors.hpp"
#include "plans.hpp"

using namespace mq;

namespace {

// executes the stage as the device does: the tiles of the GEMM schedule, each over its K range,
// writing C or a workspace partial, then the split-K reduce jobs.  The schedule is built with a
// K chunk of 16 so that every reduction longer than one K slice goes through the split path.
// GEMM schedule: < 0 every reduction longer than one K slice split (below), else the autotuner's
// candidate of that index
int g_policy = -1;

void run_gemms(std::vector<GemmDesc> const& G, std::vector<GemmRow> const& rows, double* const* base) {
    GemmSchedule sch = g_policy < 0 ? build_gemm_schedule(G, int64_t(1) << 40, 16)
                                    : build_gemm_schedule(G, gemm_policy(g_policy));
    std::vector<double> ws(sch.ws_elems + 1, std::nan(""));
    std::vector<char> seen(sch.ws_elems + 1, 0);
    for (int sh = 0; sh < GEMM_N_SHAPES; ++sh) {
        const int bm = gemm_bm(sh), bn = gemm_bn(sh);
        for (GemmTile const& t : sch.tiles[sh]) {
            GemmDesc const& g = G[t.prob];
            bool ta = g.flags & GEMM_TRANS_A, tb = g.flags & GEMM_TRANS_B, rmap = g.flags & GEMM_ROWMAP;
            require(!rmap || t.ws < 0, "emulated gemm: stacked problem split along K");
            const double* A = base[g.base_a] + (rmap ? 0 : g.a + int64_t(t.batch) * g.sa);
            const double* B = base[g.base_b] + g.b + int64_t(t.batch) * g.sb;
            double* Cm = base[g.base_c] + (rmap ? 0 : g.c + int64_t(t.batch) * g.sc);
            for (int nn = 0; nn < bn; ++nn)
                for (int mm = 0; mm < bm; ++mm) {
                    int m = t.tm * bm + mm, n = t.tn * bn + nn;
                    if (m >= g.M || n >= g.N) continue;
                    GemmRow row{};
                    if (rmap) row = rows.at(size_t(g.a + m));
                    double s = 0.0;
                    for (int k = t.k0; k < t.k1; ++k) {
                        double a = rmap ? A[row.a + int64_t(k) * row.ld]
                                        : ta ? A[int64_t(m) * g.lda + k] : A[m + int64_t(k) * g.lda];
                        double v = tb ? B[n + int64_t(k) * g.ldb] : B[k + int64_t(n) * g.ldb];
                        s += a * v;
                    }
                    if (rmap) {
                        Cm[row.c + (n + int64_t(t.batch) * g.sc) * row.ld] = s;
                    } else if (t.ws < 0) {
                        Cm[m + int64_t(n) * g.ldc] = s;
                    } else {
                        uint64_t w = uint64_t(t.ws) + mm + uint64_t(nn) * bm;
                        require(w < sch.ws_elems && !seen[w], "emulated gemm: workspace slot out of range or reused");
                        seen[w] = 1;
                        ws[w] = s;
                    }
                }
        }
    }
    for (GemmReduce const& r : sch.reduces) {
        GemmDesc const& g = G[r.prob];
        const int bm = gemm_bm(r.shape), bn = gemm_bn(r.shape);
        double* Cm = base[g.base_c] + g.c + int64_t(r.batch) * g.sc;
        for (int nn = 0; nn < bn; ++nn)
            for (int mm = 0; mm < bm; ++mm) {
                int m = r.tm * bm + mm, n = r.tn * bn + nn;
                if (m >= g.M || n >= g.N) continue;
                double s = 0.0;
                for (int q = 0; q < r.nsplit; ++q) {
                    uint64_t w = uint64_t(r.ws) + uint64_t(q) * bm * bn + mm + uint64_t(nn) * bm;
                    require(w < sch.ws_elems && seen[w], "emulated gemm: reduce reads an unwritten partial");
                    s += ws[w];
                }
                Cm[m + int64_t(n) * g.ldc] = s;
            }
    }
}

// the MPO-term sum as the device runs it: the block table built from the lines, each work item
// covering SBLOCK_ELEMS tile elements of one block, the groups of 16 lines summing the sub-blocks of
// the k-steps their masks name; every (line, element) of the table must be written exactly once
// (tiles of at most small_max elements: the small-tile work items, 2^lg lanes per line, lane e
// summing element e over the line's terms)
// Returns the number of small-tile lines.
// panel of the term-sum work order (MQ_OPT_SBLOCK_PANEL; the product's automatic value by default)
int32_t g_panel = 16;

uint64_t run_lines(HostTable<SLine>& lines, HostTable<STerm> const& terms, const double* T, double* S,
                   uint64_t t_size, uint64_t s_size, int32_t small_max = 0) {
    mark_small_lines(lines, 0, lines.size(),
                     small_max == MQ_SMALL_TILE_AUTO ? auto_small_tile_max(lines, 0, lines.size()) : small_max);
    SBlockTable B = build_sblocks(lines, terms, g_panel);
    check_sblocks(B, t_size, s_size);
    std::vector<char> sdone(B.small_lines.size(), 0);
    uint64_t small_covered = 0;
    for (SSmallWork const& w : B.small_work)
        for (int lane = 0; lane < SMALL_LINE_LANES; ++lane) {
            const int li = lane >> w.lg, e = lane & ((1 << w.lg) - 1);
            if (li >= w.n) continue;
            SLine const& l = B.small_lines[size_t(w.line0 + li)];
            if (e >= l.rows * l.cols) continue;
            double acc = 0.0;
            for (int32_t k = l.t0; k < l.t1; ++k) acc += B.small_terms[size_t(k)].alpha * T[B.small_terms[size_t(k)].t_off + e];
            S[l.s_off + e % l.rows + int64_t(e / l.rows) * l.lds] = acc;
            if (e == 0) {
                require(!sdone[size_t(w.line0 + li)], "emulated small lines: line written twice");
                sdone[size_t(w.line0 + li)] = 1;
            }
            ++small_covered;
        }
    for (char d : sdone) require(d, "emulated small lines: line not written");
    std::vector<std::vector<char>> done(B.blocks.size());
    for (size_t i = 0; i < B.blocks.size(); ++i) done[i].assign(size_t(B.blocks[i].E) * SBLOCK_LINES, 0);
    for (SWork const& w : B.work) {
        SBlock const& b = B.blocks[size_t(w.blk)];
        for (int e = w.n0; e < std::min(b.E, w.n0 + SBLOCK_ELEMS); ++e) {
            // the k-steps in order, each group with a sub-block adding alpha_g(16 x 4) * X(4 tiles)
            double acc[SBLOCK_LINES] = {0.0};
            for (int32_t k = 0; k < b.nk; ++k) {
                const uint32_t word = B.kstep[size_t(b.k_off + k)];
                int64_t ai = b.a_off + int64_t(word >> 8);
                for (int g = 0; g < SBLOCK_GROUPS; ++g) {
                    if (!((word >> g) & 1u)) continue;
                    for (int t = 0; t < 4; ++t) {
                        const int64_t to = B.utile[size_t(4 * int64_t(b.k_off + k) + t)];
                        const double x = to >= 0 ? T[to + e] : 0.0;
                        for (int i = 0; i < SBLOCK_GROUP; ++i)
                            acc[g * SBLOCK_GROUP + i] += B.ualpha[size_t(64 * ai + 16 * t + i)] * x;
                    }
                    ++ai;
                }
            }
            for (int i = 0; i < b.nlines; ++i) {
                char& d = done[size_t(w.blk)][size_t(e) * SBLOCK_LINES + i];
                require(!d, "emulated S blocks: element written twice");
                d = 1;
                S[B.lsoff[size_t(b.line0 + i)] + e % b.rows + int64_t(e / b.rows) * B.llds[size_t(b.line0 + i)]] = acc[i];
            }
        }
    }
    for (size_t i = 0; i < B.blocks.size(); ++i)
        for (int e = 0; e < B.blocks[i].E; ++e)
            for (int l = 0; l < B.blocks[i].nlines; ++l)
                require(done[i][size_t(e) * SBLOCK_LINES + l], "emulated S blocks: element not written");
    uint64_t covered = small_covered;
    for (auto const& b : B.blocks) covered += uint64_t(b.E) * b.nlines;
    uint64_t expected = 0;
    for (auto const& l : lines) expected += uint64_t(l.rows) * l.cols;
    require(covered == expected, "emulated S blocks: lines missing from the block table");
    return B.small_lines.size();
}

void run_copy(std::vector<CopyJob> const& jobs, const double* src, double* dst) {
    for (auto const& j : jobs) {
        int64_t per = int64_t(j.rows) * j.cols;
        for (int64_t i = 0; i < per * j.nblocks; ++i) {
            int64_t b = i / per, e = i % per;
            int r = int(e % j.rows), c = int(e / j.rows);
            dst[j.dst + b * per + e] = j.trans ? src[j.src + b * per + c + int64_t(r) * j.cols] : src[j.src + b * per + e];
        }
    }
}

std::vector<double> to_left(MPSLayout const& L, const double* rp) {
    std::vector<double> lp(L.elements(), 0.0);
    for (auto const& s : reshape_segments(L)) {
        int64_t total = int64_t(s.ls) * s.rs * s.d;
        for (int64_t i = 0; i < total; ++i) {
            int ll = int(i % s.ls);
            int64_t q = i / s.ls;
            int rr = int(q % s.rs), ss = int(q / s.rs);
            lp[s.lp + int64_t(ss) * s.ls + ll + int64_t(rr) * s.lp_ld] = rp[s.rp + ll + q * s.ls];
        }
    }
    return lp;
}

struct EmuBoundary {
    BoundaryIndex idx;
    std::vector<double> data;
};

}  // namespace

extern "C" {

void emu_set_sblock_panel(int panel) { g_panel = panel; }
void emu_set_gemm_policy(int policy) { g_policy = policy; }

int emu_hpsi(int sym, const mq_mpo_desc* md, const mq_bindex_desc* Ld, const double* Ldata, const mq_bindex_desc* Rd,
             const double* Rdata, const mq_index_desc* li, const mq_index_desc* pi, const mq_index_desc* ri,
             const double* psi, double* out, int small_max, uint64_t* n_small_lines, uint64_t chunk_elems,
             uint64_t* n_chunks, uint64_t* ts_elems, int plain) {
    try {
        MPO mpo(*md);
        BoundaryIndex L = boundary_index_from(*Ld), R = boundary_index_from(*Rd);
        uint64_t Ln = L.layout(), Rn = R.layout();
        HpsiPlanHost P = build_hpsi_plan(Sym(sym), mpo, L, R, Index(*li), Index(*pi), Index(*ri), chunk_elems);
        uint64_t n = P.mps.elements();
        for (auto const& j : P.lpanel_jobs)
            require(uint64_t(j.src) + uint64_t(j.rows) * j.cols * j.nblocks <= Ln &&
                    uint64_t(j.dst) + uint64_t(j.rows) * j.cols * j.nblocks <= P.lpanel_elems, "copy job out of range");
        std::vector<double> Lp(P.lpanel_elems + 1, std::nan(""));
        run_copy(P.lpanel_jobs, Ldata, Lp.data());
        // psi' starts as NaN: every element must be written by exactly the chunks' output GEMMs
        std::fill(out, out + n, std::nan(""));
        *n_small_lines = 0;
        *n_chunks = P.chunks.size();
        *ts_elems = P.t_elems + P.s_elems;
        require(!P.chunks.empty(), "plan without chunks");
        for (HpsiChunk const& ch : P.chunks) {
            require(ch.t_elems <= P.t_buf && ch.s_elems <= P.s_buf, "chunk larger than the buffers");
            uint64_t sizes[BASE_COUNT] = {0};
            sizes[BASE_PSI_IN] = sizes[BASE_PSI_OUT] = n;
            sizes[BASE_T] = ch.t_elems;
            sizes[BASE_S] = ch.s_elems;
            sizes[BASE_LPANEL] = P.lpanel_elems;
            sizes[BASE_BOUNDARY] = Rn;
            std::vector<GemmDesc> tg(P.t_gemms.begin() + int64_t(ch.t_gemm0), P.t_gemms.begin() + int64_t(ch.t_gemm1));
            std::vector<GemmRow> tr(P.t_rows.begin() + int64_t(ch.t_row0), P.t_rows.begin() + int64_t(ch.t_row1));
            for (auto& g : tg)
                if (g.flags & GEMM_ROWMAP) g.a -= int64_t(ch.t_row0);
            std::vector<GemmDesc> cg(P.c_gemms.begin() + int64_t(ch.c_gemm0), P.c_gemms.begin() + int64_t(ch.c_gemm1));
            HostTable<SLine> cl(P.lines.begin() + int64_t(ch.line0), P.lines.begin() + int64_t(ch.line1));
            check_gemms(tg, sizes, tr);
            check_gemms(cg, sizes);
            check_lines(cl, P.terms, ch.t_elems, ch.s_elems);
            // fresh NaN buffers per chunk: a chunk that read another chunk's intermediates shows
            std::vector<double> T(P.t_buf + 1, std::nan("")), S(P.s_buf + 1, std::nan(""));
            double* base[BASE_COUNT] = {nullptr};
            base[BASE_PSI_IN] = const_cast<double*>(psi);
            base[BASE_PSI_OUT] = out;
            base[BASE_T] = T.data();
            base[BASE_S] = S.data();
            base[BASE_LPANEL] = Lp.data();
            base[BASE_BOUNDARY] = const_cast<double*>(Rdata);
            if (!plain) {
                run_gemms(tg, tr, base);
                *n_small_lines += run_lines(cl, P.terms, T.data(), S.data(), ch.t_elems, ch.s_elems, small_max);
                run_gemms(cg, {}, base);
                continue;
            }
            // the reference-style tables: one product per T task, one gather per line, one product
            // per (cohort, column run) added into psi' (cleared once before the first chunk)
            if (&ch == &P.chunks.front()) std::fill(out, out + n, 0.0);
            std::vector<GemmDesc> tt(P.t_gemms_task.begin() + int64_t(ch.t_task0),
                                     P.t_gemms_task.begin() + int64_t(ch.t_task1));
            std::vector<GemmDesc> cc(P.c_gemms_cohort.begin() + int64_t(ch.c_coh0),
                                     P.c_gemms_cohort.begin() + int64_t(ch.c_coh1));
            check_gemms(tt, sizes);
            check_gemms(cc, sizes);
            run_gemms(tt, {}, base);
            for (SLine const& l : cl)
                for (int e = 0; e < l.rows * l.cols; ++e) {
                    double acc = 0.0;
                    for (int32_t k = l.t0; k < l.t1; ++k) acc += P.terms[size_t(k)].alpha * T[size_t(P.terms[size_t(k)].t_off + e)];
                    S[size_t(l.s_off + e % l.rows + int64_t(e / l.rows) * l.lds)] = acc;
                }
            for (GemmDesc const& g : cc) {
                require(g.batch == 1 && !(g.flags & (GEMM_TRANS_A | GEMM_TRANS_B | GEMM_ROWMAP)), "plain L*S problem shape");
                const double* A = base[g.base_a] + g.a;
                const double* B = base[g.base_b] + g.b;
                double* Cm = base[g.base_c] + g.c;
                for (int nn = 0; nn < g.N; ++nn)
                    for (int mm = 0; mm < g.M; ++mm) {
                        double acc = 0.0;
                        for (int k = 0; k < g.K; ++k) acc += A[mm + int64_t(k) * g.lda] * B[k + int64_t(nn) * g.ldb];
                        Cm[mm + int64_t(nn) * g.ldc] += acc;
                    }
            }
        }
        return 0;
    } catch (std::exception const& e) {
        fprintf(stderr, "emu_hpsi: %s\n", e.what());
        return -1;
    }
}

// the MPO-term adjoint through the product's tables (term sources, Y products, term work items),
// with the device kernel's arithmetic: grad[src] += beta * <Y tile, T tile> per term
int emu_term_grad(int sym, const mq_mpo_desc* md, const mq_bindex_desc* Ld, const double* Ldata,
                  const mq_bindex_desc* Rd, const double* Rdata, const mq_index_desc* li, const mq_index_desc* pi,
                  const mq_index_desc* ri, const double* psi, double* grad, uint64_t chunk_elems) {
    try {
        MPO mpo(*md);
        BoundaryIndex L = boundary_index_from(*Ld), R = boundary_index_from(*Rd);
        uint64_t Rn = R.layout();
        L.layout();
        HpsiPlanHost P = build_hpsi_plan(Sym(sym), mpo, L, R, Index(*li), Index(*pi), Index(*ri), chunk_elems, true);
        require(P.term_src.size() == P.terms.size() && P.term_beta.size() == P.terms.size(), "term sources missing");
        uint64_t n = P.mps.elements();
        std::vector<double> Lp(P.lpanel_elems + 1, std::nan(""));
        run_copy(P.lpanel_jobs, Ldata, Lp.data());
        std::fill(grad, grad + mpo.terms.size(), 0.0);
        for (HpsiChunk const& ch : P.chunks) {
            uint64_t sizes[BASE_COUNT] = {0};
            sizes[BASE_PSI_IN] = sizes[BASE_PSI_OUT] = n;
            sizes[BASE_T] = ch.t_elems;
            sizes[BASE_S] = ch.s_elems;
            sizes[BASE_LPANEL] = P.lpanel_elems;
            sizes[BASE_BOUNDARY] = Rn;
            std::vector<GemmDesc> tg(P.t_gemms.begin() + int64_t(ch.t_gemm0), P.t_gemms.begin() + int64_t(ch.t_gemm1));
            std::vector<GemmRow> tr(P.t_rows.begin() + int64_t(ch.t_row0), P.t_rows.begin() + int64_t(ch.t_row1));
            for (auto& g : tg)
                if (g.flags & GEMM_ROWMAP) g.a -= int64_t(ch.t_row0);
            std::vector<GemmDesc> cg(P.c_gemms.begin() + int64_t(ch.c_gemm0), P.c_gemms.begin() + int64_t(ch.c_gemm1));
            std::vector<GemmDesc> yg = adjoint_y_gemms(cg);
            HostTable<SLine> cl(P.lines.begin() + int64_t(ch.line0), P.lines.begin() + int64_t(ch.line1));
            check_gemms(tg, sizes, tr);
            check_gemms(yg, sizes);
            check_lines(cl, P.terms, ch.t_elems, ch.s_elems);
            std::vector<double> T(P.t_buf + 1, std::nan("")), Y(P.s_buf + 1, std::nan(""));
            double* base[BASE_COUNT] = {nullptr};
            base[BASE_PSI_IN] = const_cast<double*>(psi);
            base[BASE_T] = T.data();
            base[BASE_S] = Y.data();
            base[BASE_LPANEL] = Lp.data();
            base[BASE_BOUNDARY] = const_cast<double*>(Rdata);
            run_gemms(tg, tr, base);
            run_gemms(yg, {}, base);
            std::vector<char> covered(P.terms.size(), 0);
            for (TermGradWork const& w : term_grad_work(P.lines, ch.line0, ch.line1)) {
                SLine const& l = P.lines[size_t(w.line)];
                require(w.lg >= 0 && w.lg <= 6 && (int64_t(1) << w.lg) >= std::min<int64_t>(64, int64_t(l.rows) * l.cols),
                        "term work: lanes per term");
                require(w.t0 >= l.t0 && w.t1 <= l.t1 && w.t1 - w.t0 <= TERM_GRAD_RUN, "term work: range");
                for (int32_t k = w.t0; k < w.t1; ++k) {
                    require(!covered[size_t(k)], "term work: a term in two items");
                    covered[size_t(k)] = 1;
                    double acc = 0.0;
                    for (int e = 0; e < l.rows * l.cols; ++e) {
                        double y = Y[size_t(l.s_off + e % l.rows + int64_t(e / l.rows) * l.lds)];
                        require(!std::isnan(y), "term work: Y element not written");
                        acc += y * T[size_t(P.terms[size_t(k)].t_off + e)];
                    }
                    require(P.term_src[size_t(k)] < mpo.terms.size(), "term source out of range");
                    grad[P.term_src[size_t(k)]] += P.term_beta[size_t(k)] * acc;
                }
            }
            for (size_t li2 = ch.line0; li2 < ch.line1; ++li2)
                for (int32_t k = P.lines[li2].t0; k < P.lines[li2].t1; ++k)
                    require(covered[size_t(k)], "term work: a term in no item");
        }
        return 0;
    } catch (std::exception const& e) {
        fprintf(stderr, "emu_term_grad: %s\n", e.what());
        return -1;
    }
}

static EmuBoundary* grow(bool left, int sym, const mq_mpo_desc* md, const mq_bindex_desc* Bd, const double* Bdata,
                         const mq_index_desc* bl, const mq_index_desc* bp, const mq_index_desc* br, const double* bra_rp,
                         const mq_index_desc* kl, const mq_index_desc* kp, const mq_index_desc* kr, const double* ket_rp,
                         int symmetric) {
    MPO mpo(*md);
    BoundaryIndex B = boundary_index_from(*Bd);
    uint64_t Bn = B.layout();
    MPSLayout bra{Index(*bl), Index(*bp), Index(*br)}, ket{Index(*kl), Index(*kp), Index(*kr)};
    GrowPlanHost P = left ? build_lstep_plan(Sym(sym), mpo, B, bra, ket, symmetric != 0)
                          : build_rstep_plan(Sym(sym), mpo, B, bra, ket, symmetric != 0);
    std::vector<double> bra_data = left ? to_left(bra, bra_rp) : std::vector<double>(bra_rp, bra_rp + bra.elements());
    uint64_t sizes[BASE_COUNT] = {0};
    sizes[BASE_T] = P.t_elems;
    sizes[BASE_S] = P.s_elems;
    sizes[BASE_BOUNDARY] = Bn;
    sizes[BASE_BRA] = bra.elements();
    sizes[BASE_KET] = ket.elements();
    sizes[BASE_NEW] = P.out_elems;
    check_gemms(P.t_gemms, sizes, P.t_rows);
    check_gemms(P.c_gemms, sizes);
    check_lines(P.lines, P.terms, P.t_elems, P.s_elems);
    auto* res = new EmuBoundary{std::move(P.out), std::vector<double>(P.out_elems, 0.0)};
    std::vector<double> T(P.t_elems + 1, std::nan("")), S(P.s_elems + 1, std::nan(""));
    double* base[BASE_COUNT] = {nullptr};
    base[BASE_T] = T.data();
    base[BASE_S] = S.data();
    base[BASE_BOUNDARY] = const_cast<double*>(Bdata);
    base[BASE_BRA] = bra_data.data();
    base[BASE_KET] = const_cast<double*>(ket_rp);
    base[BASE_NEW] = res->data.data();
    run_gemms(P.t_gemms, P.t_rows, base);
    run_lines(P.lines, P.terms, T.data(), S.data(), P.t_elems, P.s_elems);
    run_gemms(P.c_gemms, {}, base);
    return res;
}

void* emu_lstep(int sym, const mq_mpo_desc* md, const mq_bindex_desc* Ld, const double* Ldata, const mq_index_desc* bl,
                const mq_index_desc* bp, const mq_index_desc* br, const double* bra_rp, const mq_index_desc* kl,
                const mq_index_desc* kp, const mq_index_desc* kr, const double* ket_rp, int symmetric) {
    try {
        return grow(true, sym, md, Ld, Ldata, bl, bp, br, bra_rp, kl, kp, kr, ket_rp, symmetric);
    } catch (std::exception const& e) {
        fprintf(stderr, "emu_lstep: %s\n", e.what());
        return nullptr;
    }
}

void* emu_rstep(int sym, const mq_mpo_desc* md, const mq_bindex_desc* Rd, const double* Rdata, const mq_index_desc* bl,
                const mq_index_desc* bp, const mq_index_desc* br, const double* bra_rp, const mq_index_desc* kl,
                const mq_index_desc* kp, const mq_index_desc* kr, const double* ket_rp, int symmetric) {
    try {
        return grow(false, sym, md, Rd, Rdata, bl, bp, br, bra_rp, kl, kp, kr, ket_rp, symmetric);
    } catch (std::exception const& e) {
        fprintf(stderr, "emu_rstep: %s\n", e.what());
        return nullptr;
    }
}

int emu_boundary_info(void* h, int* n_cohorts, unsigned* aux, uint64_t* n_elems) {
    auto* b = static_cast<EmuBoundary*>(h);
    *n_cohorts = int(b->idx.n_cohorts());
    *aux = b->idx.aux_dim;
    *n_elems = b->data.size();
    return 0;
}

int emu_boundary_export(void* h, uint32_t* lr, int64_t* offsets, double* cs, uint8_t* tp, uint8_t* tr, double* data) {
    auto* b = static_cast<EmuBoundary*>(h);
    BoundaryIndex const& B = b->idx;
    for (uint32_t c = 0; c < B.n_cohorts(); ++c) {
        lr[2 * c] = B.clr[c].first;
        lr[2 * c + 1] = B.clr[c].second;
        for (uint32_t k = 0; k < B.aux_dim; ++k) {
            offsets[uint64_t(c) * B.aux_dim + k] = B.offsets[c][k];
            cs[uint64_t(c) * B.aux_dim + k] = B.conj_scales[c][k];
            tp[uint64_t(c) * B.aux_dim + k] = B.transposes[c][k];
        }
        tr[c] = B.tr[c];
    }
    std::copy(b->data.begin(), b->data.end(), data);
    return 0;
}

void emu_boundary_free(void* h) { delete static_cast<EmuBoundary*>(h); }

}  // extern "C"
