// Host-built work lists of the three contraction kernels (H*psi, left growth, right growth).
//
// Each contraction has the reference's three stages (solver/tasks.cpp):
//   1. T-GEMMs:   T = psi_slice * R_cohort  (H*psi, right step)   or  L_b * psi_slice (left step)
//   2. S-assembly: S(tile) = sum_k alpha_k * T(tile_k)   -- the MPO-term sum
//   3. output GEMMs: psi' = L_panel * S  |  L' = A_bra^T * S  |  R'_b = S_b * B_bra^T
// Here all three stages are flattened into device tables: a list of GEMM problems per stage and a
// list of S "lines" (one output tile each) with their (alpha, T offset) terms.  The loops that
// generate alpha follow the reference's schedule builders term by term, so the S tiles are summed
// in the reference's order.
#pragma once

#include <cstdint>
#include <vector>

#include "core.hpp"

namespace mq {

// One GEMM problem: C = op(A) * op(B), column-major, optionally batched with element strides.
// Address fields hold an element offset into the role's base pointer given at launch.
enum : int32_t {
    GEMM_TRANS_A = 1,
    GEMM_TRANS_B = 2,
};
enum : int32_t {  // operand base selectors
    BASE_PSI_IN = 0, BASE_PSI_OUT, BASE_T, BASE_S, BASE_LPANEL, BASE_BOUNDARY, BASE_BRA, BASE_KET,
    BASE_NEW, BASE_COUNT
};
struct GemmDesc {
    int64_t a, b, c;        // element offsets
    int64_t sa, sb, sc;     // batch strides (elements)
    int32_t M, N, K, batch;
    int32_t lda, ldb, ldc;
    int32_t flags;          // GEMM_TRANS_*
    int32_t base_a, base_b, base_c;
    int32_t pad;
};
static_assert(sizeof(GemmDesc) == 96, "GemmDesc layout");

// S-assembly output tile (rows x cols at s_off with leading dimension lds) and its term range.
struct SLine {
    int64_t s_off;
    int32_t lds, rows, cols, t0;
    int32_t t1, pad;
};
static_assert(sizeof(SLine) == 32, "SLine layout");
struct STerm {
    double alpha;
    int64_t t_off;  // element offset of a contiguous rows x cols tile in the T buffer
};

struct FlopStats {
    uint64_t flops_total = 0, flops_gemm = 0, flops_gemm_exec = 0, n_terms = 0;
};

// One block copy job for building L panels: dst(rows x cols) <- src block, optionally transposed.
struct CopyJob {
    int64_t src, dst;       // element offsets (src in boundary arena, dst in panel buffer)
    int32_t rows, cols;     // of the destination block
    int32_t nblocks, trans; // trans: src block is cols x rows
};

struct HpsiPlanHost {
    MPSLayout mps;
    std::vector<GemmDesc> t_gemms, c_gemms;
    std::vector<SLine> lines;
    std::vector<STerm> terms;
    std::vector<CopyJob> lpanel_jobs;
    uint64_t t_elems = 0, s_elems = 0, lpanel_elems = 0;
    uint64_t n_cohorts = 0, n_t_tasks = 0;
    uint64_t bytes_compulsory = 0;
    FlopStats flops;
};

// Boundary-growth plan: new boundary index + device tables
struct GrowPlanHost {
    BoundaryIndex out;      // index of the new boundary (with data_off layout)
    uint64_t out_elems = 0;
    std::vector<GemmDesc> t_gemms, c_gemms;
    std::vector<SLine> lines;
    std::vector<STerm> terms;
    uint64_t t_elems = 0, s_elems = 0;
    FlopStats flops;
};

HpsiPlanHost build_hpsi_plan(Sym sym, MPO const& mpo, BoundaryIndex const& L, BoundaryIndex const& R,
                             Index const& left_i, Index const& phys_i, Index const& right_i);

GrowPlanHost build_lstep_plan(Sym sym, MPO const& mpo, BoundaryIndex const& L, MPSLayout const& bra,
                              MPSLayout const& ket, bool symmetric);

// Host-side bounds validation of device tables: every operand span of every GEMM problem, S line
// and T term must lie inside its buffer (sizes indexed by BASE_*); throws mq::Error otherwise.
void check_gemms(std::vector<GemmDesc> const& v, const uint64_t* sizes);
void check_lines(std::vector<SLine> const& lines, std::vector<STerm> const& terms, uint64_t t_size, uint64_t s_size);

GrowPlanHost build_rstep_plan(Sym sym, MPO const& mpo, BoundaryIndex const& R, MPSLayout const& bra,
                              MPSLayout const& ket, bool symmetric);

}  // namespace mq
