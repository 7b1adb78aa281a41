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
#include <memory>
#include <type_traits>
#include <utility>
#include <vector>

#include "core.hpp"

namespace mq {

// Host table whose resize leaves new elements uninitialised: the big tables are written in full
// right after they are sized, in parallel, so no serial zero-fill (and no first touch by one
// thread) is paid.
template <class T>
struct DefaultInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = DefaultInitAlloc<U>;
    };
    DefaultInitAlloc() = default;
    template <class U>
    DefaultInitAlloc(DefaultInitAlloc<U> const&) noexcept {}
    template <class U>
    void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};
template <class T>
using HostTable = std::vector<T, DefaultInitAlloc<T>>;


// One GEMM problem: C = op(A) * op(B), column-major, optionally batched with element strides.
// Address fields hold an element offset into the role's base pointer given at launch.
enum : int32_t {
    GEMM_TRANS_A = 1,
    GEMM_TRANS_B = 2,
    GEMM_ROWMAP = 4,  // stacked problem: rows from a row table (A not transposed, never split along K)
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

// Row of a stacked problem (GEMM_ROWMAP).  Several products that share their B operand (T-GEMMs
// reading one right-boundary cohort) run as one problem whose M rows are the rows of their A
// operands.  Row m of problem g is entry g.a + m of the row table: it reads A at element
// a + k * ld of base_a and writes C at element c + (n + batch * g.sc) * ld of base_c (g.sc counts
// columns for these problems; g.lda, g.ldc and g.sa are unused).
struct GemmRow {
    int64_t a, c;
    int32_t ld, pad;
};
static_assert(sizeof(GemmRow) == 24, "GemmRow layout");

// S-assembly output tile (rows x cols at s_off with leading dimension lds) and its term range;
// flags: SLINE_SMALL = summed by the small-tile term sum, not in a 16-line block.
constexpr int32_t SLINE_SMALL = 1;
struct SLine {
    int64_t s_off;
    int32_t lds, rows, cols, t0;
    int32_t t1, flags;
};
static_assert(sizeof(SLine) == 32, "SLine layout");
struct STerm {
    double alpha;
    int64_t t_off;  // element offset of a contiguous rows x cols tile in the T buffer
};

// Block form of the MPO-term sum for the matrix cores.  Lines of equal tile shape are grouped in
// blocks of up to SBLOCK_LINES, SBLOCK_GROUPS groups of 16.  A block stores the union of the T
// tiles its lines read, ordered by the set of groups that read each tile (then by T offset) and cut
// into k-steps of 4 tile slots (the last padded with empty slots, offset -1); utile holds the 4
// slots of k-step k of the block at 4 (k_off + k).  For every k-step, each group that reads one
// of its tiles has a 16 x 4 coefficient sub-block (64 values, [slot t][line i] at 16 t + i: alpha of
// line i for that tile, zero where the line does not read it, summed where it reads it twice);
// the sub-blocks of a block are stored in k-step order, groups ascending, from ualpha[64 a_off].
// kstep word of k-step k: bits 0..7 = the groups with a sub-block, bits 8..31 = its first sub-block
// relative to a_off.  Group g of the block's S tiles is then Y_g = sum over k-steps of
// alpha_g,k * X_k with X_k the 4 x E matrix of the step's tile elements: a group does matrix-core
// work only on the k-steps whose tiles it reads, while the tiles are read once per block.
constexpr int SBLOCK_GROUP = 16;
constexpr int SBLOCK_GROUPS = 4;
constexpr int SBLOCK_LINES = SBLOCK_GROUP * SBLOCK_GROUPS;
static_assert(SBLOCK_GROUPS <= 8, "k-step group masks are 8 bits");
struct SBlock {
    int64_t a_off;            // first coefficient sub-block (64 values each)
    int32_t k_off, nk, na;    // first k-step (kstep index; utile from 4 k_off), k-steps, sub-blocks
    int32_t E, rows;          // tile elements (rows x cols), tile rows
    int32_t nlines, line0;    // lines in the block, first entry of lsoff / llds
    int32_t pad;
};
static_assert(sizeof(SBlock) == 40, "SBlock layout");
struct SWork {
    int32_t blk, n0;          // block and first tile element of one workgroup
};
// Lines marked SLINE_SMALL (tiles of at most SMALL_LINE_LANES elements) are summed term by term:
// each line gets a power-of-two group of lanes (one lane per tile element) and a workgroup of 256
// lanes takes 256 / group consecutive lines of one group size (lines sorted by group size, then by
// term count, so the lanes of a wave run loops of similar length).
struct SSmallWork {
    int32_t line0, n;         // first line (small line table) and number of lines
    int32_t lg, pad;          // log2 of the lanes per line
};
struct SBlockTable {
    std::vector<SBlock> blocks;
    std::vector<SWork> work;
    HostTable<int64_t> utile;     // T element offset of each union tile slot (-1: empty slot)
    HostTable<uint32_t> kstep;    // k-step words
    HostTable<double> ualpha;     // coefficient sub-blocks
    std::vector<int64_t> lsoff;   // S offset of each line (SBLOCK_LINES entries per block)
    std::vector<int32_t> llds;    // its leading dimension
    uint64_t gathered = 0;        // sum over blocks of union tiles * E (T elements read)
    uint64_t max_union = 0;       // largest union of one block (tiles)
    uint64_t mac = 0;             // matrix-core multiply-adds: 64 * E per sub-block
    // small-tile lines: their own line and term tables (t0 / t1 index small_terms)
    std::vector<SLine> small_lines;
    HostTable<STerm> small_terms;
    std::vector<SSmallWork> small_work;
    uint64_t small_gathered = 0;  // sum over small lines of terms * tile elements
};
// Work item of the MPO-term adjoint: the terms [t0, t1) of one S line, 2^lg lanes per term (one
// dot product of the line's tile of Y = L_panel^T psi with the term's T tile each)
struct TermGradWork {
    int32_t line, t0, t1, lg;
};
constexpr int TERM_GRAD_THREADS = 256;
constexpr int32_t TERM_GRAD_RUN = 256;  // terms per work item at most
// the GEMMs Y = L_panel^T psi of the L*S products psi' = L_panel S (Y written where S is read)
std::vector<GemmDesc> adjoint_y_gemms(std::vector<GemmDesc> const& c_gemms);
// work items of the lines [l0, l1): runs of at most TERM_GRAD_RUN terms, 2^lg >= tile elements
// lanes per term (at most 64)
std::vector<TermGradWork> term_grad_work(HostTable<SLine> const& lines, size_t l0, size_t l1);
// lanes of a workgroup of the small-tile term sum; tiles up to this many elements can go there
constexpr int SMALL_LINE_LANES = 256;
// work items cover SBLOCK_ELEMS consecutive tile elements of one block
constexpr int SBLOCK_ELEMS = 256;

// Execution schedule of a grouped GEMM stage.  Output tiles come in six shapes (BM x BN):
// 32x128 for M <= 32, 64x64 for M <= 64, 96x64 for M <= 96 (problems without a row table), 128x64
// above; in a stage that is split along K (fewer output tiles than the device takes) problems
// with N > 64 take the 128-column tiles 64x128 and (M > 96, without a row table) 128x128, so the A
// panel is read once per 128 output columns (128x128 on eight waves).  So a
// problem with a short M edge is not padded to a taller tile and the B panel of a tall problem is
// read once.  When the stage has too few
// tiles to fill the device, long reductions are split along K: each split tile writes its
// partial BM x BN tile (column-major, ld BM) into a workspace and a reduce job adds the
// partials in split order into C, so the result does not depend on scheduling.
enum GemmShape {
    GEMM_32x128 = 0, GEMM_64x64 = 1, GEMM_128x64 = 2, GEMM_96x64 = 3, GEMM_64x128 = 4, GEMM_128x128 = 5,
    GEMM_N_SHAPES = 6
};
constexpr int gemm_bm(int shape) {
    return shape == GEMM_32x128 ? 32
         : (shape == GEMM_64x64 || shape == GEMM_64x128) ? 64
         : shape == GEMM_96x64 ? 96 : 128;
}
constexpr int gemm_bn(int shape) {
    return (shape == GEMM_32x128 || shape == GEMM_64x128 || shape == GEMM_128x128) ? 128 : 64;
}
struct GemmTile {
    int32_t prob, batch, tm, tn;
    int32_t k0, k1;   // reduction range of this tile
    int64_t ws;       // < 0: the tile writes C; else element offset of its partial in the workspace
};
static_assert(sizeof(GemmTile) == 32, "GemmTile layout");
struct GemmReduce {
    int32_t prob, batch, tm, tn;
    int32_t nsplit, shape;
    int64_t ws;       // first partial; partial s at ws + s * BM * BN
};
static_assert(sizeof(GemmReduce) == 32, "GemmReduce layout");
struct GemmSchedule {
    std::vector<GemmTile> tiles[GEMM_N_SHAPES];
    std::vector<GemmReduce> reduces;
    uint64_t ws_elems = 0;
    uint64_t n_tiles() const {
        uint64_t n = 0;
        for (auto const& t : tiles) n += t.size();
        return n;
    }
};
// The choices of a grouped GEMM schedule that the tile autotuner varies.  wide: 0 = 128-column
// tiles (64x128, 128x128) for N > 64 when the stage is split along K, 1 = never, 2 = always;
// tiles96: 96-row tiles for 64 < M <= 96 (problems without a row table), else 128 rows;
// target_items: workgroups a stage should provide before K is split; min_chunk: shortest K range
// of a split tile (a multiple of the K step 16)
struct GemmPolicy {
    int32_t wide = 0;
    int32_t tiles96 = 1;
    int64_t target_items = 4096;
    int32_t min_chunk = 256;
};
constexpr int GEMM_N_POLICIES = 7;
// candidate i of the autotuner (0: the default rule)
GemmPolicy gemm_policy(int i);
GemmSchedule build_gemm_schedule(std::vector<GemmDesc> const& probs, GemmPolicy const& pol);
inline GemmSchedule build_gemm_schedule(std::vector<GemmDesc> const& probs, int64_t target_items = 4096,
                                        int32_t min_chunk = 256) {
    GemmPolicy p;
    p.target_items = target_items;
    p.min_chunk = min_chunk;
    return build_gemm_schedule(probs, p);
}
// Key of a stage's (m, n, k) histogram for the autotuner: the flop-weighted means of log2 M,
// log2 N and log2 K (in quarters), the share of flops in problems of at most 32, 64 and 96 rows
// (in eighths), log2 of the output tiles at the default rule, and whether rows come from a row
// table; stages of neighbouring site problems usually share it
uint64_t gemm_stage_key(std::vector<GemmDesc> const& probs);


// block table of a list of S lines (lines without terms become zero tiles)
// lines marked SLINE_SMALL go to the small-tile tables, the others into 16-line blocks
// panel > 0: work items ordered element range by element range over panels of up to that many
// consecutive blocks of one tile shape (0: block by block)
SBlockTable build_sblocks(HostTable<SLine> const& lines, HostTable<STerm> const& terms, int32_t panel = 0);
// small-tile limit of the lines [l0, l1) of one site problem under the automatic rule: 64 when
// their mean tile has at most 16 elements (then the 16-line blocks mostly hold tiles a few
// elements wide, whose dense coefficient blocks cost more than the terms), else 0
int32_t auto_small_tile_max(HostTable<SLine> const& lines, size_t l0, size_t l1);
// marks the lines [l0, l1) whose tiles have 1 .. limit elements SLINE_SMALL (limit <= 256)
void mark_small_lines(HostTable<SLine>& lines, size_t l0, size_t l1, int32_t limit);
void check_sblocks(SBlockTable const& B, uint64_t t_size, uint64_t s_size);

struct FlopStats {
    uint64_t flops_total = 0, flops_gemm = 0, flops_gemm_exec = 0, n_terms = 0;
};

// One block copy job for building L panels: dst(rows x cols) <- src block, optionally transposed.
struct CopyJob {
    int64_t src, dst;       // element offsets (src in boundary arena, dst in panel buffer)
    int32_t rows, cols;     // of the destination block
    int32_t nblocks, trans; // trans: src block is cols x rows
};

// One chunk of an H*psi plan: the output columns of a set of right sectors of the two-site MPS.
// Its T-GEMMs, S lines and L*S GEMMs touch only the T intermediates and S panels of these
// columns, which it addresses from offset 0: the chunks run one after the other on one T and one S
// buffer of the largest chunk's size.  Ranges index the plan's tables.
struct HpsiChunk {
    size_t t_gemm0 = 0, t_gemm1 = 0;   // t_gemms
    size_t t_row0 = 0, t_row1 = 0;     // t_rows
    size_t line0 = 0, line1 = 0;       // lines
    size_t c_gemm0 = 0, c_gemm1 = 0;   // c_gemms
    size_t t_task0 = 0, t_task1 = 0;   // t_gemms_task
    size_t c_coh0 = 0, c_coh1 = 0;     // c_gemms_cohort
    uint64_t t_elems = 0, s_elems = 0;
};

struct HpsiPlanHost {
    MPSLayout mps;
    std::vector<HpsiChunk> chunks;  // empty: one chunk over all tables
    std::vector<GemmDesc> t_gemms, c_gemms;
    std::vector<GemmRow> t_rows;  // row table of the stacked T-GEMMs
    std::vector<GemmDesc> t_gemms_task;  // the same T-GEMMs, one problem per task (reference layout)
    // the L*S products per cohort (per cohort and run of the chunk's columns when the plan has
    // several chunks), to be summed into psi'
    std::vector<GemmDesc> c_gemms_cohort;
    HostTable<SLine> lines;
    HostTable<STerm> terms;
    // with term sources: per S term, the MPO term it comes from and d alpha / d(that term's scale)
    HostTable<uint32_t> term_src;
    HostTable<double> term_beta;
    std::vector<CopyJob> lpanel_jobs;
    uint64_t t_elems = 0, s_elems = 0, lpanel_elems = 0;  // T and S: sums over the chunks
    uint64_t t_buf = 0, s_buf = 0;                          // T and S buffers: the largest chunk's
    uint64_t n_cohorts = 0, n_t_tasks = 0;
    uint64_t bytes_compulsory = 0;
    FlopStats flops;
};

// Boundary-growth plan: new boundary index + device tables
struct GrowPlanHost {
    BoundaryIndex out;      // index of the new boundary (with data_off layout)
    uint64_t out_elems = 0;
    std::vector<GemmDesc> t_gemms, c_gemms;
    std::vector<GemmRow> t_rows;  // row table of stacked T-GEMMs (right step)
    HostTable<SLine> lines;
    HostTable<STerm> terms;
    uint64_t t_elems = 0, s_elems = 0;
    FlopStats flops;
};

// chunk_elems: largest T + S (elements) of one chunk; the right sectors are gathered in index
// order into chunks up to that size (a sector larger than it is a chunk of its own); 0 = one chunk.
// term_sources: also fill term_src / term_beta (the MPO-term adjoint)
HpsiPlanHost build_hpsi_plan(Sym sym, MPO const& mpo, BoundaryIndex const& L, BoundaryIndex const& R,
                             Index const& left_i, Index const& phys_i, Index const& right_i,
                             uint64_t chunk_elems = 0, bool term_sources = false);

GrowPlanHost build_lstep_plan(Sym sym, MPO const& mpo, BoundaryIndex const& L, MPSLayout const& bra,
                              MPSLayout const& ket, bool symmetric);

// Host-side bounds validation of device tables: every operand span of every GEMM problem (and
// every row of a stacked one), S line and T term must lie inside its buffer (sizes indexed by
// BASE_*); throws mq::Error otherwise.
void check_gemms(std::vector<GemmDesc> const& v, const uint64_t* sizes,
                 std::vector<GemmRow> const& rows = std::vector<GemmRow>());
void check_lines(HostTable<SLine> const& lines, HostTable<STerm> const& terms, uint64_t t_size, uint64_t s_size);

GrowPlanHost build_rstep_plan(Sym sym, MPO const& mpo, BoundaryIndex const& R, MPSLayout const& bra,
                              MPSLayout const& ket, bool symmetric);

}  // namespace mq
