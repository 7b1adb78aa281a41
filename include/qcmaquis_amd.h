/*
 * qcmaquis_amd.h -- C ABI of the MI355X-native symmetry-blocked DMRG contraction path.
 *
 * The path: Davidson H*psi (Boundary x MPOTensor x MPS site problem) and the left / right
 * boundary-growth steps, on one GPU.  Everything that crosses this ABI is plain C: pointers,
 * sizes and small POD descriptors.  FP64 is the value type at the ABI; device-internal
 * precisions are selected with a flag.
 *
 * Reference interfaces replaced (paths relative to the reference tree dmrg/framework/dmrg/):
 *   contraction::Engine<...>::{overlap_mpo_left_step, overlap_mpo_right_step,
 *                              contraction_schedule}     mp_tensors/contractions/engine.hpp:47-140
 *   create_contraction_schedule                         contractions/create_schedule/site_hamil_schedule.hpp:40-93
 *   super_hamil_mv                                      solver/super_hamil_mv.hpp:100-155
 *   solve<double> (Jacobi-Davidson)                     solver/solver.h:64-68, solver/solver.cpp:13-28
 *   common::overlap_mpo_left_step / _right_step         contractions/boundary_ops/move_boundary.hpp:147-404
 *   Boundary / BoundaryIndex                            mp_tensors/boundary.h:76-406
 *
 * Error convention: every entry point returns MQ_OK (0) or a negative MQ_ERR_* code and never
 * exits the process or throws across the ABI (the reference calls exit() on CUDA/cuBLAS errors,
 * tasks.cpp:231-235).  mq_last_error() gives the message of the last failure on a context.
 *
 * Threading: a context is bound to one device and one host thread at a time.
 */
#ifndef QCMAQUIS_AMD_H
#define QCMAQUIS_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MQ_OK              0
#define MQ_ERR_INVALID    -1   /* bad argument / inconsistent descriptor (reference: std::runtime_error) */
#define MQ_ERR_HIP        -2   /* HIP runtime failure (reference: exit(EXIT_FAILURE)) */
#define MQ_ERR_NOMEM      -3   /* device allocation failed (reference: std::out_of_range, accelerator.cpp:129) */
#define MQ_ERR_UNSUPPORTED -4  /* feature not available in this build */

/* symmetry groups (reference: block_matrix/symmetry/2u1.h, su2u1.h) */
#define MQ_SYM_TWOU1 1         /* charge = (N_up, N_down) */
#define MQ_SYM_SU2U1 2         /* charge = (N, 2S) */

/* device arithmetic of the sector GEMMs */
#define MQ_PREC_FP64   0
#define MQ_PREC_FP32   1
#define MQ_PREC_BF16X3 2         /* H*psi (mq_hpsi, mq_hpsi_timed) as MQ_PREC_FP32, with the GEMM
                                    stages' products split into BF16 parts on the BF16 matrix cores
                                    (x = hi + lo, |x - hi - lo| <= 2^-16 |x|; hi hi + hi lo + lo hi,
                                    FP32 accumulation); eager launches only (no graph); not a
                                    Davidson precision */
/* FP64 with the reference's launch structure, for comparison with the grouped path: one library
 * (rocBLAS) dgemm per T task, one S-assembly workgroup per S tile that gathers every term, and
 * one dgemm per cohort added into its output block (reference: MPSBlock::create_T_gpu,
 * Cohort::create_s_r_gpu / compute_s_stackedv, Cohort::contract_gpu; tasks.cpp:286-344, 466-504,
 * 650-752).  Needs a plan made while the context option MQ_OPT_PLAIN_TABLES is set.  A plan with
 * several chunks (MQ_OPT_CHUNK_MB) runs them one after the other, a cohort's dgemm split by the
 * runs of its output block's columns in the chunk. */
#define MQ_PREC_FP64_PLAIN 3

/* context options (mq_ctx_set_option) */
#define MQ_OPT_PLAIN_TABLES 1  /* value 1: H*psi plans made afterwards also keep the per-task tables
                                  that MQ_PREC_FP64_PLAIN runs (default 0) */
#define MQ_OPT_SMALL_TILE_MAX 2 /* value 0..256: H*psi plans made afterwards sum the MPO terms of S
                                  tiles of at most this many elements term by term, one lane per
                                  element, instead of in 16-tile blocks on the matrix cores (0: every
                                  tile in blocks); MQ_SMALL_TILE_AUTO (default): per site problem,
                                  64 when its mean S tile has at most 16 elements, else 0 */
#define MQ_SMALL_TILE_AUTO (-1)
#define MQ_OPT_CHUNK_MB 3      /* value >= 0: H*psi plans made afterwards split their output columns
                                  by right sector into chunks of at most this many MiB of FP64 T and
                                  S intermediates (a larger right sector is a chunk of its own),
                                  run one after the other on buffers of the largest chunk's size;
                                  0: one chunk; MQ_CHUNK_AUTO (default): 49152 MiB. */
#define MQ_CHUNK_AUTO (-1)
#define MQ_OPT_SBLOCK_PANEL 5  /* value >= 0: plans made afterwards order the term sum's work items
                                  element range by element range over panels of this many
                                  neighbouring blocks (which share most of their T tiles); 0: block
                                  by block; MQ_SBLOCK_PANEL_AUTO (default): 0 */
#define MQ_SBLOCK_PANEL_AUTO (-1)
#define MQ_OPT_SBLOCK_XCD 6    /* value 1: plans made afterwards give each XCD a contiguous range of
                                  the term sum's work items instead of every eighth (default 0) */
#define MQ_OPT_AUTOTUNE 7      /* value 1: H*psi plans made afterwards choose the schedule of each
                                  GEMM stage among the candidate tile rules by timing them, once per
                                  key of the stage's (m, n, k) histogram; later stages with a known
                                  key take the recorded pick without timing (mq_ctx_tuned lists
                                  them).  Default 0: the default rule. */
#define MQ_OPT_JCD_GMRES 8     /* value 0..64: GMRES steps of mq_davidson's correction equation (the
                                  reference's ietl_jcd_gmres; jcd_gmres_solver, ietl/jacobi.h:126-164,
                                  ietl_gmres, ietl/gmres.h:89-145); 0 (default): the plain correction
                                  t = -r + (r.u / u.u) u.  mq_davidson_noalloc refuses values > 0; the
                                  batch Davidson uses the plain correction. */
#define MQ_OPT_OVERLAP 9       /* value 1 (default): H*psi plans made afterwards that run in two or more
                                  chunks overlap chunk c's output GEMMs (on a second stream of the
                                  context) with chunk c + 1's T-GEMMs and term sum, with a second S
                                  buffer (a plan without room for it runs the chunks in turn);
                                  0: one chunk after the other.  Timed H*psi (mq_hpsi_timed) runs the
                                  chunks in turn. */
#define MQ_OPT_TERM_SOURCES 4  /* value 1: H*psi plans made afterwards (single site problems) also keep,
                                  per S term, the MPO term it comes from, for
                                  mq_hpsi_term_gradient (default 0) */

/* MPS pairing (reference: MPSTensor::make_left_paired / make_right_paired, mpstensor.hpp:211-237) */
#define MQ_RIGHT_PAIRED 0      /* blocks per left charge:  ls x (sum_sigma d_sigma * rs(lc+sigma)) */
#define MQ_LEFT_PAIRED  1      /* blocks per right charge: (sum_sigma d_sigma * ls(rc-sigma)) x rs */

#define MQ_NO_PARTNER 0xFFFFFFFFu

typedef struct mq_ctx_s*        mq_ctx_t;
typedef struct mq_mpo_s*        mq_mpo_t;
typedef struct mq_boundary_s*   mq_boundary_t;
typedef struct mq_mps_s*        mq_mps_t;
typedef struct mq_hpsi_plan_s*  mq_hpsi_plan_t;

/* Sector index.  Charges must be sorted strictly descending in lexicographic order, which is the
 * invariant of the reference's Index (indexing_stable.hpp:129-176). */
typedef struct {
    int32_t         n;          /* number of sectors */
    const int32_t*  charges;    /* n x 2 */
    const uint32_t* sizes;      /* n */
} mq_index_desc;

/* MPO tensor (reference: MPOTensor, mpotensor.h:45-135; term_descriptor / Hermitian / BondProperty,
 * mpotensor_detail.h:48-239; SparseOperator, sparse_operator.h:164-233). */
typedef struct {
    /* operator table: op o has blocks [op_block_ptr[o], op_block_ptr[o+1]) */
    int32_t         n_ops;
    const int32_t*  op_spin;          /* n_ops: 2K of the operator (SiteOperator::spin().get()), 0 if abelian */
    const int64_t*  op_block_ptr;     /* n_ops + 1 */
    const int32_t*  block_charges;    /* n_blocks x 4: left charge (2 ints), right charge (2 ints) */
    const int64_t*  block_entry_ptr;  /* n_blocks + 1 (global block numbering) */
    const uint32_t* entry_rc;         /* n_entries x 2: row (ss1), col (ss2) */
    const int32_t*  entry_spins;      /* n_entries x 2: row_spin, col_spin; NULL for abelian */
    const double*   entry_coeff;      /* n_entries */
    /* tensor: terms of (b1,b2) are stored consecutively, sorted by (b1, b2) */
    uint32_t        row_dim, col_dim;
    int64_t         n_terms;
    const uint32_t* term_b;           /* n_terms x 2: (b1, b2) */
    const uint32_t* term_op;          /* n_terms */
    const double*   term_scale;       /* n_terms */
    /* bond properties; any may be NULL (spins -> 0, no Hermitian partners) */
    const int32_t*  left_spin;        /* row_dim */
    const int32_t*  right_spin;       /* col_dim */
    const uint32_t* left_herm;        /* row_dim, MQ_NO_PARTNER = none, b = self-adjoint */
    const int32_t*  left_phase;       /* row_dim */
    const uint32_t* right_herm;       /* col_dim */
    const int32_t*  right_phase;      /* col_dim */
} mq_mpo_desc;

/* Boundary index (reference: BoundaryIndex, boundary.h:76-234).  Cohort ci holds the blocks of the
 * charge pair (bra[cohort_lr[2ci]], ket[cohort_lr[2ci+1]]); block b of cohort ci starts at element
 * offsets[ci*aux_dim+b] of the cohort data (-1 = absent); blocks are ls x rs column-major. */
typedef struct {
    mq_index_desc   bra, ket;
    uint32_t        aux_dim;
    int32_t         n_cohorts;
    const uint32_t* cohort_lr;        /* n_cohorts x 2 */
    const int64_t*  offsets;          /* n_cohorts x aux_dim */
    const double*   conj_scales;      /* n_cohorts x aux_dim, NULL -> 1 */
    const uint8_t*  transposes;       /* n_cohorts x aux_dim, NULL -> 0 */
    const uint8_t*  tr;               /* n_cohorts, NULL -> 0 */
} mq_bindex_desc;

typedef struct {
    int32_t  n_cohorts;
    uint32_t aux_dim;
    int32_t  n_bra, n_ket;
    uint64_t total_elements;          /* sum over cohorts with tr == 0 of n_blocks * ls * rs */
} mq_boundary_info_t;

/* Statistics of an H*psi plan.  Flop model = ScheduleNew::total_flops (solver/tasks.cpp:389-398,
 * 778-791); bytes = compulsory FP64 traffic of one H*psi (boundary cohorts referenced + psi + psi'). */
typedef struct {
    uint64_t psi_elements;
    uint64_t n_t_tasks, n_cohorts, n_lines, n_terms;
    uint64_t flops_total;             /* reference model */
    uint64_t flops_gemm;              /* T-GEMM + L*S GEMM part of the reference model */
    uint64_t flops_gemm_exec;         /* flops of the GEMMs as executed by this plan */
    uint64_t bytes_compulsory;        /* FP64 bytes */
    uint64_t t_elements, s_elements, l_panel_elements, r_tr_elements;  /* T and S: over all chunks */
    double   build_seconds;
    uint64_t s_line_elements;         /* sum over S lines of rows * cols (S elements written) */
    uint64_t gathered_elements;       /* sum over S terms of rows * cols (T elements of the term sum) */
    uint64_t n_sblocks;               /* blocks (up to 64 lines) of the matrix-core term sum */
    uint64_t union_elements;          /* sum over blocks of union tiles * tile elements (T elements read) */
    double   seconds_tables;          /* host: schedule tables built and checked (part of build_seconds) */
    double   seconds_blocks;          /* host: term-sum block tables (part of build_seconds) */
    double   seconds_plain;           /* per-task tables of MQ_PREC_FP64_PLAIN kept (part of build_seconds) */
    uint64_t small_lines;             /* S tiles summed term by term (MQ_OPT_SMALL_TILE_MAX) */
    uint64_t small_terms;             /* their terms */
    uint64_t small_gathered;          /* their T elements read: sum over terms of the tile elements */
    uint64_t max_union_tiles;         /* largest union of T tiles of one term-sum block */
    uint64_t split_reduces;           /* split-K reduce jobs of the T-GEMM and L*S GEMM stages */
    uint64_t sblock_macs;             /* matrix-core multiply-adds of the block term sum */
    uint64_t n_chunks;                /* chunks of the plan (MQ_OPT_CHUNK_MB) */
    uint64_t ts_buffer_elements;      /* T + S buffer elements (the largest chunk's) */
    int32_t  gemm_policy_t, gemm_policy_c;  /* GEMM schedule candidates of the T and L*S stages
                                               (chunk 0; 0 = the default rule, MQ_OPT_AUTOTUNE) */
    double   seconds_tune;            /* device time spent timing schedule candidates (MQ_OPT_AUTOTUNE) */
} mq_hpsi_stats_t;

typedef struct {
    int32_t  iterations;
    double   residual_norm;
    double   seconds;
    double   hpsi_seconds;
    int32_t  refine_iterations;       /* FP64 iterations after an FP32 run (precision MQ_PREC_FP32) */
} mq_davidson_stats_t;

/* How a boundary made by mq_lstep / mq_rstep was built (all zero for uploaded and initial
 * boundaries).  Flop model as for H*psi, over the growth schedule (tasks.cpp:389-398). */
typedef struct {
    uint64_t n_lines, n_terms, n_sblocks;
    uint64_t t_elements, s_elements, out_elements;
    uint64_t flops_total, flops_gemm;
    double   seconds_tables;          /* host: schedule tables built and checked */
    double   seconds_blocks;          /* host: term-sum block tables built and checked */
    double   seconds_upload;          /* device buffers allocated, tables uploaded */
    double   seconds_device;          /* the step's kernels, to completion */
    double   seconds_total;           /* the whole call */
} mq_growth_stats_t;

/* ---- context ---------------------------------------------------------------------------- */
int         mq_ctx_create(int device, int symmetry, mq_ctx_t* out);
int         mq_ctx_destroy(mq_ctx_t ctx);
const char* mq_last_error(mq_ctx_t ctx);
int         mq_ctx_sync(mq_ctx_t ctx);
/* the autotuner's table (MQ_OPT_AUTOTUNE): for up to n recorded stage keys (key, stage kind
 * 0 = T-GEMMs / 1 = L*S, chosen candidate, its time and the default rule's time in ms); returns the
 * number of recorded keys.  No counterpart in the reference (its GEMMs are library calls). */
int         mq_ctx_tuned(mq_ctx_t ctx, int n, uint64_t* keys, int32_t* kinds, int32_t* picks, float* best_ms,
                         float* default_ms);
void*       mq_ctx_stream(mq_ctx_t ctx);      /* hipStream_t of the context */
int         mq_ctx_set_option(mq_ctx_t ctx, int option, int value);
int         mq_version(void);
/* sizes in bytes of the ABI structs, in this order: mq_index_desc, mq_mpo_desc, mq_bindex_desc,
 * mq_boundary_info_t, mq_hpsi_stats_t, mq_davidson_stats_t, mq_growth_stats_t; returns how many
 * were written (a binding checks its own struct layouts against these before the first call) */
int         mq_abi_struct_sizes(uint32_t* sizes, int n);

/* ---- MPO ---------------------------------------------------------------------------------- */
int mq_mpo_create(mq_ctx_t ctx, const mq_mpo_desc* desc, mq_mpo_t* out);
int mq_mpo_destroy(mq_mpo_t mpo);

/* ---- device-resident boundary store (reference: Boundary + storage::gpu, storage.h:547-671) -- */
/* initial boundary: one 1x1 cohort per sector, aux_dim blocks, all ones (boundary.h:257-283) */
int mq_boundary_create_initial(mq_ctx_t ctx, const mq_index_desc* idx, uint32_t aux_dim, mq_boundary_t* out);
/* data: data of the cohorts with tr == 0 concatenated in cohort order (n_blocks*ls*rs elements each),
 * host memory.  Cohorts with tr != 0 alias their partner's storage and carry no data. */
int mq_boundary_upload(mq_ctx_t ctx, const mq_bindex_desc* desc, const double* data, mq_boundary_t* out);
int mq_boundary_info(mq_boundary_t b, mq_boundary_info_t* info);
/* export the index: arrays sized from mq_boundary_info (any pointer may be NULL) */
int mq_boundary_export_index(mq_boundary_t b, int32_t* bra_charges, uint32_t* bra_sizes,
                             int32_t* ket_charges, uint32_t* ket_sizes, uint32_t* cohort_lr,
                             int64_t* offsets, double* conj_scales, uint8_t* transposes, uint8_t* tr);
int mq_boundary_download(mq_boundary_t b, double* data);
/* build statistics of a boundary returned by mq_lstep / mq_rstep */
int mq_boundary_growth_stats(mq_boundary_t b, mq_growth_stats_t* stats);
int mq_boundary_destroy(mq_boundary_t b);

/* ---- MPS site tensors in HBM ------------------------------------------------------------ */
/* data: blocks of the given pairing concatenated in block order, column-major, host memory */
int mq_mps_upload(mq_ctx_t ctx, const mq_index_desc* left, const mq_index_desc* phys,
                  const mq_index_desc* right, int pairing, const double* data, mq_mps_t* out);
int mq_mps_elements(mq_mps_t m, uint64_t* n);
int mq_mps_download(mq_mps_t m, int pairing, double* data);
int mq_mps_destroy(mq_mps_t m);

/* ---- boundary growth (reference: move_boundary.hpp:147-404) -------------------------------- */
/* new L' = bra^dagger . (L x W x ket) ; symmetric: skip Hermitian partners and complete the index
 * by complement_transpose (move_boundary.hpp:191-208) */
int mq_lstep(mq_ctx_t ctx, mq_mpo_t mpo, mq_boundary_t left, mq_mps_t bra, mq_mps_t ket,
             int symmetric, int precision, mq_boundary_t* out);
int mq_rstep(mq_ctx_t ctx, mq_mpo_t mpo, mq_boundary_t right, mq_mps_t bra, mq_mps_t ket,
             int symmetric, int precision, mq_boundary_t* out);

/* ---- H*psi (reference: create_contraction_schedule + super_hamil_mv) ----------------------- */
/* The plan copies what it needs of the left boundary and reads the right boundary's storage in
 * place: it keeps that storage alive (shared with the boundary), so the right boundary may be
 * destroyed before the plan. */
int mq_hpsi_plan_create(mq_ctx_t ctx, mq_mpo_t mpo, mq_boundary_t left, mq_boundary_t right,
                        const mq_index_desc* left_i, const mq_index_desc* phys_i,
                        const mq_index_desc* right_i, mq_hpsi_plan_t* out);
/* n independent site problems as one plan (instance i: mpos[i], lefts[i], rights[i] and the MPS
 * indices left_is[i], phys_is[i], right_is[i]); psi is the concatenation of the instances' psi, and
 * every stage of an H*psi over all instances is one grouped launch.  mq_davidson refuses batch
 * plans (their operator is block-diagonal).  Reference: no counterpart (one schedule per site
 * problem). */
int mq_hpsi_plan_create_batch(mq_ctx_t ctx, int n, const mq_mpo_t* mpos, const mq_boundary_t* lefts,
                              const mq_boundary_t* rights, const mq_index_desc* left_is,
                              const mq_index_desc* phys_is, const mq_index_desc* right_is,
                              mq_hpsi_plan_t* out);
int mq_hpsi_plan_stats(mq_hpsi_plan_t plan, mq_hpsi_stats_t* stats);
/* psi: right-paired blocks of left_i concatenated, each ls x stripe column-major, no padding */
int mq_hpsi_plan_psi_elements(mq_hpsi_plan_t plan, uint64_t* n);
/* device pointers (on the context's device), stream-ordered on the context stream */
int mq_hpsi(mq_hpsi_plan_t plan, const double* psi_in_dev, double* psi_out_dev, int precision);
/* precision: MQ_PREC_FP64, MQ_PREC_FP32 (psi, L and R rounded to FP32, T-GEMMs, term sum and
 * output GEMMs on the FP32 matrix cores with FP32 intermediates, psi' returned in FP64), or
 * MQ_PREC_FP64_PLAIN (never captured in a graph) */
/* as mq_hpsi, synchronous, with the duration of each stage measured by HIP events recorded on
 * the context stream: stage_ms[0] set-up (FP32: rounding psi), [1] T-GEMMs, [2] S-assembly
 * (MPO-term sum), [3] output GEMMs
 * (reference timing: cudaEvent around the GPU H*psi, super_hamil_mv.hpp:61-84) */
int mq_hpsi_timed(mq_hpsi_plan_t plan, const double* psi_in_dev, double* psi_out_dev, int precision,
                  float* stage_ms);
/* enable = 1: mq_hpsi, mq_hpsi_host and mq_davidson run H*psi as a HIP graph captured once per
 * precision on buffers owned by the plan (psi copied in and out on the context stream); the
 * launch sequence is the one of the default path.  Reference: no counterpart (one launch per GPU
 * task in super_hamil_mv.hpp:100-155). */
int mq_hpsi_plan_use_graphs(mq_hpsi_plan_t plan, int enable);
/* host pointers, synchronous */
int mq_hpsi_host(mq_hpsi_plan_t plan, const double* psi_in, double* psi_out, int precision);
/* MPO-term adjoint of the site problem: grad[t] = dE/d(scale of term t of the plan's MPO) for
 * E = <psi| H_eff |psi> (psi: host, the plan's layout; grad: host, n_grad = the MPO's n_terms), in
 * FP64 on the device: T = psi * R as in H*psi, Y = L_panel^T psi where the term sum writes S, then
 * per S term the dot product of its T tile with its line's Y tile, times d alpha / d(scale),
 * summed per MPO term.  sum_t scale_t grad[t] = E.  Over a chain (left boundaries grown from the
 * tensors the right ones were, every site's single-site MPO) these are the energy's derivatives
 * with respect to every MPO-term scale, which mq_qc_integral_gradient turns into dE/dh and
 * dE/d(pq|rs): the 1- and 2-RDMs by Hellmann-Feynman.  Needs a plan made with
 * MQ_OPT_TERM_SOURCES.  No counterpart in the reference (its RDMs come from one expectation value
 * per element, measurements/tagged_nrankrdm.h:413-661). */
int mq_hpsi_term_gradient(mq_hpsi_plan_t plan, const double* psi, double* grad, int64_t n_grad);
int mq_hpsi_plan_destroy(mq_hpsi_plan_t plan);

/* ---- Jacobi-Davidson (reference: solve_ietl_jcd, ietl_jacobi_davidson.h:35-76,
 *      jacobi_davidson::calculate_eigenvalue, ietl/jacobi.h:386-474) ------------------------ */
/* precision MQ_PREC_FP32: one search space is built with FP32 H*psi for up to max_iter
 * iterations, stopping at a relative residual of max(tol, 1e-5); then the Ritz vector is refined
 * with FP64 H*psi from a restart for the iterations left of max_iter, at least one (so the
 * returned eigenvalue is the FP64 Rayleigh quotient of the returned vector; max_iter + 1 H*psi
 * when the FP32 phase uses the whole budget); orthogonalisation, Rayleigh matrix and residuals
 * are FP64 throughout */
int mq_davidson(mq_hpsi_plan_t plan, const double* psi0, int n_ortho, const double* const* ortho,
                double tol, int max_iter, int precision, double* eval, double* psi_out,
                mq_davidson_stats_t* stats);
/* Jacobi-Davidson that allocates no device memory.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results as mq_davidson; its search space (2 max_iter + 2
 * + n_ortho vectors of the plan's psi length) lives in one fixed device array of
 * MQ_NOALLOC_SCRATCH_ELEMS doubles per GPU, held until the process ends.  Larger problems, other
 * precisions than MQ_PREC_FP64 and plans with HIP graphs enabled return MQ_ERR_INVALID.  Every call
 * on one GPU uses that GPU's array: calls that overlap on one GPU (from several contexts or
 * threads) race. */
#define MQ_NOALLOC_SCRATCH_ELEMS (1u << 23)
int mq_davidson_noalloc(mq_hpsi_plan_t plan, const double* psi0, int n_ortho, const double* const* ortho,
                        double tol, int max_iter, int precision, double* eval, double* psi_out,
                        mq_davidson_stats_t* stats);
/* Lowest eigenpair of every instance of a batch plan (mq_hpsi_plan_create_batch), the iteration of
 * mq_davidson run in lock step: one batched H*psi per iteration, the vector operations as
 * segmented launches over all instances.  No counterpart in the reference (which solves one site
 * problem at a time; the iteration is jacobi_davidson::calculate_eigenvalue, ietl/jacobi.h:386-474);
 * for drivers of many independent problems.  psi0 / psi_out: the instances' vectors concatenated
 * (host); evals, iterations: one entry per instance; stats->iterations counts lock-step iterations.
 * No ortho vectors.  Precision MQ_PREC_FP64, or MQ_PREC_FP32 (FP32 contraction, FP64 refinement as
 * in mq_davidson). */
int mq_davidson_batch(mq_hpsi_plan_t plan, const double* psi0, double tol, int max_iter, int precision,
                      double* evals, double* psi_out, int32_t* iterations, mq_davidson_stats_t* stats);

/* ---- two-site split (reference: TwoSiteTensor::split_mps_l2r / split_mps_r2l, twositetensor.hpp:
 *      155-200, with svd_truncate, block_matrix_algorithms.h:186-244: one SVD per symmetry block) ---- */
/* Thin SVD of n dense blocks on the device: block i is rows[i] x cols[i], column-major, the blocks
 * concatenated in a (host); per block, with k = min(rows, cols): s gets its k singular values
 * (descending), u its rows x k left singular vectors and vt the k x cols right ones (column-major),
 * each concatenated over the blocks (host arrays).  One rocSOLVER dgesvd per block on the
 * context stream; synchronous.  The caller truncates (estimate_truncation) and forms the site
 * tensors. */
int mq_block_svd(mq_ctx_t ctx, int n, const int32_t* rows, const int32_t* cols, const double* a, double* s,
                 double* u, double* vt);

/* Noise term of the density-matrix split (reference: alpha_dm_direct / alpha_dm_direct_right,
 * move_boundary.hpp:38-145, with Cohort::lbtm / rbtm, tasks.cpp:346-380).  Grow the boundary with
 * an identity bra and symmetric = 0 (no block then aliases its partner's): its blocks are the S
 * tiles of that growth step (a right step completes partner cohorts in any case: give it the MPO
 * without its Hermitian-partner tables).  side 0 (a left boundary): out_k = alpha * sum over the cohorts whose
 * row (bra) charge is charges[k] of X X^T, X = the cohort's blocks side by side; side 1 (a right
 * boundary, whose columns are the bra's): out_k = alpha * sum over the blocks X_b of the cohorts
 * whose column charge is charges[k] of X_b^T X_b.  The n charges (n x 2) and sizes must match
 * entries of the grouped index; out (host) receives the n blocks sizes[k] x sizes[k],
 * column-major, concatenated.  rocBLAS dgemm on the context stream; synchronous. */
int mq_boundary_gram(mq_boundary_t b, int side, double alpha, int n, const int32_t* charges, const uint32_t* sizes,
                     double* out);

/* ---- quantum-chemistry Hamiltonian MPO (host only; input producer of the path) -------------
 * reference: models/chem/2u1/model.hpp (create_terms), generate_mpo/tagged_mpo_maker_optim.hpp,
 * mp_tensors/ts_ops.hpp:96-185 (two-site MPO).  h1: L x L row-major h_pq; eri: L^4 row-major
 * (pq|rs) (chemists' notation, all 8 permutational symmetries expected); terms with
 * |coefficient| <= cutoff are dropped (integral_cutoff, DmrgParameters.h:76). */
typedef struct mq_qc_s* mq_qc_t;
int mq_qc_build(int symmetry, int L, const double* h1, const double* eri, double ecore, double cutoff,
                mq_qc_t* out);
/* mq_qc_build with flags.  MQ_QC_GRADIENT (SU2U1): the gradient layout -- every canonical
 * integral's terms are kept whatever its value (cutoff unused, zero scales kept), so the chain's
 * structure does not depend on the integral values; single-site terms are one MPO term per
 * operator; and the handle keeps the map from integrals to MPO terms for mq_qc_integral_gradient.
 * The energy of the chain is the one of mq_qc_build. */
#define MQ_QC_GRADIENT 1
int mq_qc_build_ex(int symmetry, int L, const double* h1, const double* eri, double ecore, double cutoff, int flags,
                   mq_qc_t* out);
/* Chain rule from MPO-term scales to integrals (the builder is linear in the integrals): given
 * term_grads[p][t] = dE/d(scale of term t of site p's single-site MPO) for p = 0 .. L-1 (e.g. from
 * mq_hpsi_term_gradient), writes dE/dh1[i*L+j] (j <= i) into dh (L*L) and dE/deri[((i*L+j)*L+k)*L+l]
 * (i >= j, k >= l, (i,j) >= (k,l)) into deri (L^4): the derivatives with respect to the canonical
 * entries that the builder reads, zero at every other position.  Needs a handle built with
 * MQ_QC_GRADIENT.  No counterpart in the reference, which measures the RDMs element by element
 * (measurements/tagged_nrankrdm.h:413-661); these are the same quantities contracted once. */
int mq_qc_integral_gradient(mq_qc_t qc, const double* const* term_grads, double* dh, double* deri);
/* L, the L+1 MPO bond dimensions, the core energy (any pointer may be NULL) */
int mq_qc_info(mq_qc_t qc, int* L, uint32_t* bond_dims, double* ecore);
/* descriptor of the site MPO (twosite = 0) or of the two-site MPO of sites (site, site+1); the
 * arrays it points to are owned by the handle and stay valid until mq_qc_destroy */
int mq_qc_site_mpo(mq_qc_t qc, int site, int twosite, mq_mpo_desc* out);
/* physical index of one site (twosite = 0) or of two fused sites */
int mq_qc_phys(mq_qc_t qc, int twosite, mq_index_desc* out);
int mq_qc_destroy(mq_qc_t qc);

/* ---- host-only utilities (no device needed) ------------------------------------------------ */
/* Wigner 6j / 9j symbols, arguments as 2j (reference: gsl_sf_coupling_6j / _9j, extern "C" in
 * block_matrix/symmetry/gsl_coupling.h:30-34; selection rules of wignerj/coupling.c) */
double mq_wigner_6j(int two_ja, int two_jb, int two_jc, int two_jd, int two_je, int two_jf);
double mq_wigner_9j(int two_ja, int two_jb, int two_jc, int two_jd, int two_je, int two_jf,
                    int two_jg, int two_jh, int two_ji);
/* position of charge (q0, q1) in an index (reference Index::position, indexing_stable.hpp:129-140):
 * n if absent, MQ_ERR_INVALID if the index is not sorted strictly descending */
int mq_index_position(const mq_index_desc* idx, int32_t q0, int32_t q1);

#ifdef __cplusplus
}
#endif

#endif /* QCMAQUIS_AMD_H */
