// gfx950 kernels of the contraction path.
//
//  * gemm_kernel: grouped, batched FP64 / FP32 GEMM over a host-built tile table.  32x128, 64x64,
//    96x64 or 128x64 output tiles per 256-thread workgroup (four wave64s in a 2x2 grid, each a block of
//    16x16x4 matrix-core tiles), operands staged through LDS in K-slices of 8 with the next slice
//    prefetched into registers.  Each output tile is owned by one workgroup or, when K is
//    split, by a fixed set of partial tiles summed in order by gemm_reduce_kernel: no atomics.
//  * sblock_kernel: the MPO-term sum (reference create_s / create_s_r, solver/tasks.cpp:425-494)
//    as a block-sparse product on the matrix cores: up to 64 output tiles ("lines") of equal shape
//    at a time, in groups of 16, Y_g = alpha_g * X over the union of the T tiles they read, each
//    group only over the k-steps whose tiles it reads.
//  * copy / reshape / fill kernels for panel set-up and MPS pairing changes.
//  * FP64 Davidson vector kernels with device-resident coefficients.
#include "device.hpp"

#include <type_traits>

namespace mq {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int BK = 16;
constexpr int LDS_PAD = 1;

// The 16x16x4 matrix-core step of each compute precision.  Operands (measured on gfx950): lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator register r of lane l holds
// D[row(l, r)][l % 16] with row = l / 16 + 4 r for FP64 and 4 (l / 16) + r for FP32.
template <class T> struct Mma;
template <> struct Mma<double> {
    typedef double4_t acc_t;
    static __device__ inline acc_t step(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mma<float> {
    typedef float4_t acc_t;
    static __device__ inline acc_t step(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <class T> __device__ inline T ld_or_zero(const T* p, bool ok) { return ok ? *p : T(0); }

// Split-BF16 products (precision MQ_PREC_BF16X3): an FP32 operand x = hi + lo with hi its leading
// 16 bits (a BF16 value, cut) and lo the rest rounded to nearest BF16, so |x - hi - lo| <= 2^-16 |x|;
// a product is hi_a hi_b + hi_a lo_b + lo_a hi_b on the BF16 matrix cores with FP32 accumulation
// (lo_a lo_b, at most 2^-14 of it, is left out).  The 16x16x16 step: lane l supplies four
// consecutive k of A[l % 16][4 (l / 16) + q] and B[4 (l / 16) + q][l % 16]; the accumulator
// layout is the one of the FP32 16x16x4 step.
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
__device__ inline void split_bf16(float x, short& hi, short& lo) {
    const uint32_t u = __float_as_uint(x);
    hi = short(u >> 16);
    const uint32_t r = __float_as_uint(x - __uint_as_float(u & 0xffff0000u));
    lo = short((r + 0x7fffu + ((r >> 16) & 1u)) >> 16);
}

// One BM x BN output tile (BM = 32 FM, BN = 16 FN WN) per workgroup of 2 WN waves; the waves form a
// 2 x WN grid of 16 FM x 16 FN wave tiles, each FM x FN matrix-core tiles.  Operands of type TI,
// computed in TI, output converted to TO.  The tile reduces over [k0, k1) and writes C, or its
// partial tile (type TI) into the workspace (split K).
template <int FM, int FN, class TI, class TO, int KS = BK, int MINB = 1, int WN = 2, bool BF3 = false>
__global__ __launch_bounds__(128 * WN, MINB) void gemm_kernel(const GemmDesc* __restrict__ probs,
                                                              const GemmRow* __restrict__ rows,
                                                              const GemmTile* __restrict__ tiles, Bases bases,
                                                              TI* __restrict__ ws) {
    constexpr int THREADS = 128 * WN;
    constexpr int BM = 32 * FM, BN = 16 * FN * WN;
    constexpr int NA = BM * KS / THREADS, NB = BN * KS / THREADS;  // operand elements staged per thread and K slice
    static_assert(KS % 4 == 0 && NA * THREADS == BM * KS && NB * THREADS == BN * KS, "K slice geometry");
    static_assert(!BF3 || (std::is_same<TI, float>::value && KS % 16 == 0), "split-BF16 products: FP32 operands, K slices of 16");
    __shared__ TI As[KS][BM + LDS_PAD];
    __shared__ TI Bs[KS][BN + LDS_PAD];
    typedef typename Mma<TI>::acc_t acc_t;

    const GemmTile tile = tiles[blockIdx.x];
    const GemmDesc g = probs[tile.prob];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int m0 = tile.tm * BM, n0 = tile.tn * BN;
    const bool ta = g.flags & GEMM_TRANS_A, tb = g.flags & GEMM_TRANS_B, rmap = g.flags & GEMM_ROWMAP;

    const TI* A = static_cast<const TI*>(bases.p[g.base_a]) + (rmap ? 0 : g.a + int64_t(tile.batch) * g.sa);
    const TI* B = static_cast<const TI*>(bases.p[g.base_b]) + g.b + int64_t(tile.batch) * g.sb;
    // stacked problem: the tile's rows are staged in LDS once (the A loads and the C stores look them
    // up there); A is not transposed, so each thread loads from one row (tid % BM) of the tile
    // (stacked problems only get tile heights that divide the workgroup size)
    __shared__ GemmRow Rs[BM];
    const TI* a_row = A;
    int64_t a_rld = 0;
    if (rmap) {
        for (int i = tid; i < BM; i += THREADS)
            if (m0 + i < g.M) Rs[i] = rows[g.a + m0 + i];
        __syncthreads();
        if (m0 + tid % BM < g.M) {
            a_row = A + Rs[tid % BM].a;
            a_rld = Rs[tid % BM].ld;
        }
    }

    // per-thread load coordinates inside the BM x 16 A slice and the 16 x BN B slice, with the
    // contiguous dimension of the operand across consecutive threads
    int a_m[NA], a_k[NA], b_k[NB], b_n[NB];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        const int idx = tid + THREADS * i;
        if (!ta) { a_m[i] = idx % BM; a_k[i] = idx / BM; }
        else     { a_k[i] = idx % KS; a_m[i] = idx / KS; }
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int idx = tid + THREADS * i;
        if (!tb) { b_k[i] = idx % KS; b_n[i] = idx / KS; }
        else     { b_n[i] = idx % BN; b_k[i] = idx / BN; }
    }
    const int k_end = tile.k1;
    TI ra[NA], rb[NB];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            int m = m0 + a_m[i], k = k0 + a_k[i];
            bool ok = (m < g.M) && (k < k_end);
            const TI* pa = rmap ? a_row + int64_t(k) * a_rld
                                : ta ? A + int64_t(m) * g.lda + k : A + m + int64_t(k) * g.lda;
            ra[i] = ld_or_zero(pa, ok);
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            int n = n0 + b_n[i], kb = k0 + b_k[i];
            bool okb = (n < g.N) && (kb < k_end);
            const TI* pb = tb ? B + n + int64_t(kb) * g.ldb : B + kb + int64_t(n) * g.ldb;
            rb[i] = ld_or_zero(pb, okb);
        }
    };

    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    const int wm = (wave & 1) * 16 * FM, wn = (wave >> 1) * 16 * FN;
    const int fr = lane & 15, fk = lane >> 4;

    if (tile.k0 < k_end) load(tile.k0);
    for (int k0 = tile.k0; k0 < k_end; k0 += KS) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NA; ++i) As[a_k[i]][a_m[i]] = ra[i];
#pragma unroll
        for (int i = 0; i < NB; ++i) Bs[b_k[i]][b_n[i]] = rb[i];
        __syncthreads();
        if (k0 + KS < k_end) load(k0 + KS);
        if constexpr (BF3) {
#pragma unroll
            for (int kk = 0; kk < KS; kk += 16) {
                bf16x4_t ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        short h, l;
                        split_bf16(As[kk + 4 * fk + q][wm + 16 * i + fr], h, l);
                        ah[i][q] = h;
                        al[i][q] = l;
                    }
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        short h, l;
                        split_bf16(Bs[kk + 4 * fk + q][wn + 16 * j + fr], h, l);
                        bh[j][q] = h;
                        bl[j][q] = l;
                    }
                // operands swapped as below: the instruction accumulates C^T = B^T A^T
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bh[j], ah[i], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bh[j], al[i], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bl[j], ah[i], acc[i][j], 0, 0, 0);
                    }
            }
            continue;
        }
#pragma unroll
        for (int kk = 0; kk < KS; kk += 4) {
            TI af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) af[i] = As[kk + fk][wm + 16 * i + fr];
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = Bs[kk + fk][wn + 16 * j + fr];
            // operands swapped: the instruction accumulates the transposed tile C^T = B^T A^T
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = Mma<TI>::step(bf[j], af[i], acc[i][j]);
        }
    }

    // D = C^T: register r of lane l holds C[l % 16][row(l, r)] of its 16 x 16 tile, so
    // consecutive lanes store consecutive rows of a column
    TO* C = static_cast<TO*>(bases.p[g.base_c]) + (rmap ? 0 : g.c + int64_t(tile.batch) * g.sc);
    const bool split = tile.ws >= 0;  // never for stacked problems
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int mm = wm + 16 * i + fr, m = m0 + mm;
        if (m >= g.M) continue;
        // column n of this row at c_row + n * c_ld
        TO* c_row = C + m;
        int64_t c_ld = g.ldc;
        if (rmap) {
            const GemmRow r = Rs[mm];
            c_row = C + r.c + int64_t(tile.batch) * g.sc * r.ld;
            c_ld = r.ld;
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int nn = wn + 16 * j + Mma<TI>::row(lane, r), n = n0 + nn;
                if (n >= g.N) continue;
                if (split) ws[tile.ws + mm + int64_t(nn) * BM] = acc[i][j][r];
                else c_row[int64_t(n) * c_ld] = TO(acc[i][j][r]);
            }
        }
    }
}

// C tile = sum of its split partials, added in split order; blockIdx.y selects 256 elements of
// the tile, so one output tile is reduced by up to 64 workgroups
constexpr int REDUCE_CHUNKS = 64;  // 16384 (largest tile) / 256
constexpr bool reduce_covers(int sh) { return sh == GEMM_N_SHAPES || (gemm_bm(sh) * gemm_bn(sh) <= REDUCE_CHUNKS * 256 && reduce_covers(sh + 1)); }
static_assert(reduce_covers(0), "reduce grid covers every tile shape");
template <class TW, class TO>
__global__ __launch_bounds__(256) void gemm_reduce_kernel(const GemmDesc* __restrict__ probs,
                                                          const GemmReduce* __restrict__ jobs, Bases bases,
                                                          const TW* __restrict__ ws) {
    const GemmReduce job = jobs[blockIdx.x];
    const GemmDesc g = probs[job.prob];
    const int bm = gemm_bm(job.shape), bn = gemm_bn(job.shape);
    const int m0 = job.tm * bm, n0 = job.tn * bn;
    TO* C = static_cast<TO*>(bases.p[g.base_c]) + g.c + int64_t(job.batch) * g.sc;
    const int64_t tile_elems = int64_t(bm) * bn;
    const int idx = int(blockIdx.y) * 256 + int(threadIdx.x);
    if (idx < bm * bn) {
        const int mm = idx % bm, nn = idx / bm;
        const int m = m0 + mm, n = n0 + nn;
        if (m < g.M && n < g.N) {
            TW sum = 0;
            for (int s = 0; s < job.nsplit; ++s) sum += ws[job.ws + s * tile_elems + idx];
            C[m + int64_t(n) * g.ldc] = TO(sum);
        }
    }
}

// Block-form MPO-term sum on the matrix cores.  A work item is SBLOCK_ELEMS consecutive tile
// elements of one block of up to SBLOCK_GROUPS groups of 16 lines; each of the SB_WAVES waves owns
// 16 * SB_NC of them.  Per k-step (4 union tiles) every group that reads one of its tiles computes
// Y_g(16 lines x 16 elements) += alpha_g(16 x 4) * X(4 tiles x 16 elements) on the matrix cores:
// lane l loads element l % 16 of tile l / 16 (16 consecutive values per tile, shared by the
// groups) and alpha of line l % 16 for tile l / 16.  The tile offsets, k-step words and
// coefficient sub-blocks are staged in LDS, SB_KCHUNK k-steps at a time; the next chunk's tables
// are loaded into registers while the current chunk is summed, and the tiles of the next
// SB_UNROLL k-steps are loaded while the current ones go through the matrix cores.  The k-step's
// group mask is read into a scalar register, so the groups without a sub-block are skipped by the
// whole wave.
#ifndef MQ_SB_NC
#define MQ_SB_NC 2
#endif
#ifndef MQ_SB_UNROLL
#define MQ_SB_UNROLL 2
#endif
#ifndef MQ_SB_KCHUNK
#define MQ_SB_KCHUNK 8
#endif
#ifndef MQ_SB_WAVES_PER_EU
#define MQ_SB_WAVES_PER_EU 4
#endif
constexpr int SB_NC = MQ_SB_NC;                         // 16-element column tiles per wave
constexpr int SB_XCDS = 8;                              // XCDs of an MI355X
constexpr int SB_WAVES = SBLOCK_ELEMS / (16 * SB_NC);   // waves per workgroup
constexpr int SB_THREADS = 64 * SB_WAVES;
constexpr int SB_UNROLL = MQ_SB_UNROLL;                 // k-steps whose loads are issued together
constexpr int SB_KCHUNK = MQ_SB_KCHUNK;                 // k-steps staged in LDS at a time
constexpr int SB_ST_ALPHA = SB_KCHUNK * SBLOCK_GROUPS * 64 / SB_THREADS;  // staged values per thread
static_assert(SBLOCK_GROUP == 16 && SB_WAVES * 16 * SB_NC == SBLOCK_ELEMS && SB_WAVES <= 16, "S block geometry");
static_assert(SB_KCHUNK % (2 * SB_UNROLL) == 0, "staged chunks hold whole double-buffered k-step groups");
static_assert(4 * SB_KCHUNK <= SB_THREADS && SB_ST_ALPHA * SB_THREADS == SB_KCHUNK * SBLOCK_GROUPS * 64,
              "one staging pass covers a chunk's tables");

template <class T>
__global__ __launch_bounds__(SB_THREADS) __attribute__((amdgpu_waves_per_eu(MQ_SB_WAVES_PER_EU))) void sblock_kernel(const SBlock* __restrict__ blocks, const SWork* __restrict__ work,
                                                     const int64_t* __restrict__ utile, const uint32_t* __restrict__ kstep,
                                                     const T* __restrict__ ualpha, const int64_t* __restrict__ lsoff,
                                                     const int32_t* __restrict__ llds, const T* __restrict__ X,
                                                     T* __restrict__ S, int xcd_contiguous) {
    typedef typename Mma<T>::acc_t acc_t;
    __shared__ int64_t uts[4 * SB_KCHUNK];
    __shared__ uint32_t kws[SB_KCHUNK];
    __shared__ T uas[SB_KCHUNK * SBLOCK_GROUPS * 64];
    // Workgroups are dispatched to the XCDs in turn (workgroup i on XCD i mod 8); with
    // xcd_contiguous each XCD takes a contiguous range of the work list instead, so consecutive
    // work items (neighbouring blocks' slices of shared T tiles) meet in one XCD's L2.  The map is
    // a bijection of [0, gridDim.x).
    int64_t wi = blockIdx.x;
    if (xcd_contiguous) {
        const int64_t n = gridDim.x, q = n / SB_XCDS, r = n % SB_XCDS;
        const int64_t x = blockIdx.x % SB_XCDS, k = blockIdx.x / SB_XCDS;
        wi = x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
    }
    const SWork w = work[wi];
    const SBlock b = blocks[w.blk];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n0 = w.n0 + wave * 16 * SB_NC;
    // a wave past the tile end loads and stores nothing but takes part in the barriers
    const bool active = n0 < b.E;
    const int li = lane & 15, lk = lane >> 4;
    acc_t acc[SBLOCK_GROUPS][SB_NC];
#pragma unroll
    for (int g = 0; g < SBLOCK_GROUPS; ++g)
#pragma unroll
        for (int j = 0; j < SB_NC; ++j) acc[g][j] = acc_t{0, 0, 0, 0};
    const int64_t* ut = utile + 4 * int64_t(b.k_off);
    const uint32_t* kw = kstep + b.k_off;
    const T* ua = ualpha + 64 * b.a_off;

    // the tables of the chunk from k-step c0 in registers: this thread's slot offset, k-step word
    // and coefficient values; ra0 / ra1 = the chunk's sub-block range
    int64_t r_ut = -1;
    uint32_t r_kw = 0;
    T r_a[SB_ST_ALPHA];
    int ra0 = 0, ra1 = 0;
    auto fetch = [&](int c0) {
        const int cn = b.nk - c0 < SB_KCHUNK ? b.nk - c0 : SB_KCHUNK;
        ra0 = int(kw[c0] >> 8);
        ra1 = c0 + cn < b.nk ? int(kw[c0 + cn] >> 8) : b.na;
        if (tid < 4 * cn) r_ut = ut[4 * c0 + tid];
        if (tid < cn) r_kw = kw[c0 + tid];
#pragma unroll
        for (int q = 0; q < SB_ST_ALPHA; ++q) {
            const int i = tid + SB_THREADS * q;
            r_a[q] = i < 64 * (ra1 - ra0) ? ua[64 * int64_t(ra0) + i] : T(0);
        }
    };
    // X of the k-steps [k, k + SB_UNROLL) of the staged chunk (cn k-steps)
    auto load_x = [&](T (&x)[SB_UNROLL][SB_NC], int k, int cn) {
#pragma unroll
        for (int u = 0; u < SB_UNROLL; ++u) {
            const int64_t to = k + u < cn ? uts[4 * (k + u) + lk] : -1;
#pragma unroll
            for (int j = 0; j < SB_NC; ++j) {
                const int e = n0 + 16 * j + li;
                x[u][j] = (to >= 0 && e < b.E) ? X[to + e] : T(0);  // B[tile lk][element e]
            }
        }
    };
    auto mma = [&](T (&x)[SB_UNROLL][SB_NC], int k, int cn, int a0) {
#pragma unroll
        for (int u = 0; u < SB_UNROLL; ++u) {
            if (k + u >= cn) break;
            const uint32_t word = __builtin_amdgcn_readfirstlane(kws[k + u]);
            int ai = int(word >> 8) - a0;
#pragma unroll
            for (int g = 0; g < SBLOCK_GROUPS; ++g) {
                if (!((word >> g) & 1u)) continue;
                const T a = uas[64 * ai + lane];  // A[line li][tile lk]
                ++ai;
#pragma unroll
                for (int j = 0; j < SB_NC; ++j) acc[g][j] = Mma<T>::step(a, x[u][j], acc[g][j]);
            }
        }
    };

    if (b.nk > 0) fetch(0);
    for (int c0 = 0; c0 < b.nk; c0 += SB_KCHUNK) {
        const int cn = b.nk - c0 < SB_KCHUNK ? b.nk - c0 : SB_KCHUNK;
        const int a0 = ra0;
        __syncthreads();  // the previous chunk's tables are consumed
        if (tid < 4 * cn) uts[tid] = r_ut;
        if (tid < cn) kws[tid] = r_kw;
#pragma unroll
        for (int q = 0; q < SB_ST_ALPHA; ++q) {
            const int i = tid + SB_THREADS * q;
            if (i < 64 * (ra1 - ra0)) uas[i] = r_a[q];
        }
        __syncthreads();
        if (c0 + SB_KCHUNK < b.nk) fetch(c0 + SB_KCHUNK);
        if (!active) continue;
        T xa[SB_UNROLL][SB_NC], xb[SB_UNROLL][SB_NC];
        load_x(xa, 0, cn);
        for (int ks = 0; ks < cn; ks += 2 * SB_UNROLL) {
            load_x(xb, ks + SB_UNROLL, cn);
            mma(xa, ks, cn, a0);
            load_x(xa, ks + 2 * SB_UNROLL, cn);
            mma(xb, ks + SB_UNROLL, cn, a0);
        }
    }
    if (!active) return;
    // register r of lane l holds Y_g[line row(l, r)][element l % 16]
#pragma unroll
    for (int j = 0; j < SB_NC; ++j) {
        const int e = n0 + 16 * j + li;
        if (e >= b.E) continue;
        const int r_ = e % b.rows, c_ = e / b.rows;
#pragma unroll
        for (int g = 0; g < SBLOCK_GROUPS; ++g)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = SBLOCK_GROUP * g + Mma<T>::row(lane, r);
                if (i < b.nlines) S[lsoff[b.line0 + i] + r_ + int64_t(c_) * llds[b.line0 + i]] = acc[g][j][r];
            }
    }
}

// Small-tile MPO-term sum: a work item is up to 256 >> lg lines of 2^lg lanes each (tile elements
// <= 2^lg); lane e of a line sums alpha_k * T[t_off_k + e] over the line's terms in term order,
// four terms' loads ahead of their multiply-adds, and writes element e of the line's S tile.  The
// lanes of a line read the same term entries; no LDS and no barrier, so lanes past their line's
// tile simply return.
template <class T>
__global__ __launch_bounds__(SMALL_LINE_LANES) void small_lines_kernel(const SSmallWork* __restrict__ work,
                                                                       const SLine* __restrict__ lines,
                                                                       const STerm* __restrict__ terms,
                                                                       const T* __restrict__ X, T* __restrict__ S) {
    const SSmallWork w = work[blockIdx.x];
    const int li = int(threadIdx.x) >> w.lg, e = int(threadIdx.x) & ((1 << w.lg) - 1);
    if (li >= w.n) return;
    const SLine l = lines[w.line0 + li];
    if (e >= l.rows * l.cols) return;
    T acc = T(0);
    int k = l.t0;
    for (; k + 4 <= l.t1; k += 4) {
        const STerm a = terms[k], b = terms[k + 1], c = terms[k + 2], d = terms[k + 3];
        const T xa = X[a.t_off + e], xb = X[b.t_off + e], xc = X[c.t_off + e], xd = X[d.t_off + e];
        acc += T(a.alpha) * xa;
        acc += T(b.alpha) * xb;
        acc += T(c.alpha) * xc;
        acc += T(d.alpha) * xd;
    }
    for (; k < l.t1; ++k) acc += T(terms[k].alpha) * X[terms[k].t_off + e];
    S[l.s_off + e % l.rows + int64_t(e / l.rows) * l.lds] = acc;
}

// One S line (rows x cols tile at s_off, leading dimension lds) per workgroup; element e of the
// tile is the sum over the line's terms of alpha * T[t_off + e], in term order (reference:
// compute_s_stackedv, gpu_dgemm_ddot.cu:282-359, one gather per term).  Lines without terms
// write zeros.
__global__ __launch_bounds__(256) void sline_kernel(const SLine* __restrict__ lines, const STerm* __restrict__ terms,
                                                    const double* __restrict__ T, double* __restrict__ S) {
    const SLine l = lines[blockIdx.x];
    const int E = l.rows * l.cols;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        double acc = 0.0;
        for (int k = l.t0; k < l.t1; ++k) acc += terms[k].alpha * T[terms[k].t_off + e];
        S[l.s_off + e % l.rows + int64_t(e / l.rows) * l.lds] = acc;
    }
}

// MPO-term adjoint of the term sum.  E = <psi|H psi> = sum over terms k of alpha_k <Y_line(k),
// T_k> with Y = L_panel^T psi laid out like S, so dE/d(scale of MPO term m) is the sum over the
// terms k from m of beta_k <Y_line(k), T_k> (beta_k = d alpha_k / d scale_m).  A workgroup takes
// a run of one line's terms; each group of 2^lg lanes forms one term's dot product over the tile
// elements and reduces it with cross-lane exchanges inside the group.
__global__ __launch_bounds__(TERM_GRAD_THREADS) void term_grad_kernel(
    const TermGradWork* __restrict__ work, const SLine* __restrict__ lines, const STerm* __restrict__ terms,
    const uint32_t* __restrict__ src, const double* __restrict__ beta, const double* __restrict__ Y,
    const double* __restrict__ X, double* __restrict__ grad) {
    const TermGradWork w = work[blockIdx.x];
    const SLine l = lines[w.line];
    const int E = l.rows * l.cols;
    const int G = 1 << w.lg;
    const int grp = int(threadIdx.x) >> w.lg, e0 = int(threadIdx.x) & (G - 1);
    const int ngroups = TERM_GRAD_THREADS >> w.lg;
    for (int k = w.t0 + grp; k < w.t1; k += ngroups) {
        const int64_t toff = terms[k].t_off;
        double acc = 0.0;
        for (int e = e0; e < E; e += G) acc += Y[l.s_off + e % l.rows + int64_t(e / l.rows) * l.lds] * X[toff + e];
        // the lanes of one group share k, so a group is active or idle as a whole
        for (int off = G >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
        if (e0 == 0) atomicAdd(grad + src[k], beta[k] * acc);
    }
}

template <class TI, class TO>
__global__ void convert_kernel(const TI* __restrict__ src, TO* __restrict__ dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = TO(src[i]);
}

__global__ void fill_kernel(double* p, uint64_t n, double v) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// grid: x = job, y = chunk of the job's elements
__global__ void copy_jobs_kernel(const CopyJob* __restrict__ jobs, const double* __restrict__ src,
                                 double* __restrict__ dst) {
    const CopyJob j = jobs[blockIdx.x];
    const int64_t per = int64_t(j.rows) * j.cols;
    const int64_t total = per * j.nblocks;
    for (int64_t i = int64_t(blockIdx.y) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.y) * blockDim.x) {
        int64_t b = i / per, e = i % per;
        int r = int(e % j.rows), c = int(e / j.rows);
        // destination block rows x cols; transposed source block is cols x rows
        double v = j.trans ? src[j.src + b * per + c + int64_t(r) * j.cols] : src[j.src + b * per + e];
        dst[j.dst + b * per + e] = v;
    }
}

// element (ss, rr, ll) of a segment: rp[ll + (ss*rs + rr)*ls]  <->  lp[ss*ls + ll + rr*lp_ld]
__global__ void reshape_kernel(const ReshapeSeg* __restrict__ segs, const double* __restrict__ src,
                               double* __restrict__ dst, int to_left) {
    const ReshapeSeg s = segs[blockIdx.x];
    const int64_t total = int64_t(s.ls) * s.rs * s.d;
    for (int64_t i = int64_t(blockIdx.y) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.y) * blockDim.x) {
        int ll = int(i % s.ls);
        int64_t q = i / s.ls;  // = ss*rs + rr
        int rr = int(q % s.rs), ss = int(q / s.rs);
        int64_t rp = s.rp + ll + q * s.ls;
        int64_t lp = s.lp + int64_t(ss) * s.ls + ll + int64_t(rr) * s.lp_ld;
        if (to_left) dst[lp] = src[rp];
        else dst[rp] = src[lp];
    }
}

constexpr int DOT_THREADS = 256;

__device__ inline double block_sum(double v) {
    __shared__ double red[DOT_THREADS / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < DOT_THREADS / 64; ++i) r += red[i];
    return r;
}

__global__ __launch_bounds__(DOT_THREADS) void dot_partial_kernel(const double* __restrict__ x,
                                                                  const double* __restrict__ y, uint64_t n,
                                                                  double* __restrict__ work) {
    double s = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * DOT_THREADS + threadIdx.x; i < n; i += uint64_t(gridDim.x) * DOT_THREADS)
        s += x[i] * y[i];
    double r = block_sum(s);
    if (threadIdx.x == 0) work[blockIdx.x] = r;
}

__global__ __launch_bounds__(DOT_THREADS) void dot_final_kernel(const double* __restrict__ work, int nparts,
                                                                double* __restrict__ out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += DOT_THREADS) s += work[i];
    double r = block_sum(s);
    if (threadIdx.x == 0) out[0] = r;
}

__global__ void axpy_dev_kernel(double* __restrict__ y, const double* __restrict__ x, uint64_t n,
                                const double* __restrict__ num, const double* __restrict__ den, double sign) {
    const double coef = sign * num[0] / (den ? den[0] : 1.0);
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        y[i] += coef * x[i];
}

// Segmented forms for a batch of independent vectors concatenated in one buffer (segment s is
// elements off[s] .. off[s+1]-1; blockIdx.y = segment).  Dot: SEG_CHUNKS partial sums per
// segment, then one thread per segment adds them in chunk order (deterministic).
__global__ __launch_bounds__(DOT_THREADS) void seg_dot_partial_kernel(const double* __restrict__ x,
                                                                      const double* __restrict__ y,
                                                                      const int64_t* __restrict__ off,
                                                                      double* __restrict__ part) {
    const int seg = blockIdx.y;
    const int64_t a = off[seg], b = off[seg + 1];
    double s = 0.0;
    for (int64_t i = a + int64_t(blockIdx.x) * DOT_THREADS + threadIdx.x; i < b; i += int64_t(SEG_CHUNKS) * DOT_THREADS)
        s += x[i] * y[i];
    double r = block_sum(s);
    if (threadIdx.x == 0) part[int64_t(seg) * SEG_CHUNKS + blockIdx.x] = r;
}

__global__ void seg_dot_final_kernel(const double* __restrict__ part, int nseg, double* __restrict__ out) {
    const int seg = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (seg >= nseg) return;
    double s = 0.0;
    for (int c = 0; c < SEG_CHUNKS; ++c) s += part[int64_t(seg) * SEG_CHUNKS + c];
    out[seg] = s;
}

// segment s: y = a[s]*y + b[s]*x (y not read where a[s] == 0)
__global__ void seg_axpby_kernel(double* __restrict__ y, const double* __restrict__ x, const int64_t* __restrict__ off,
                                 const double* __restrict__ ca, const double* __restrict__ cb) {
    const int seg = blockIdx.y;
    const int64_t i0 = off[seg], i1 = off[seg + 1];
    const double a = ca[seg], b = cb[seg];
    for (int64_t i = i0 + int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < i1; i += int64_t(gridDim.x) * blockDim.x)
        y[i] = a == 0.0 ? b * x[i] : a * y[i] + b * x[i];
}

// y = a*y + b*x; with a == 0 y is not read (it may hold anything, NaN included: BLAS beta = 0)
__global__ void axpby_kernel(double* __restrict__ y, const double* __restrict__ x, uint64_t n, double a, double b) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        y[i] = a == 0.0 ? b * x[i] : a * y[i] + b * x[i];
}

__global__ void scale_rsqrt_kernel(double* __restrict__ x, uint64_t n, const double* __restrict__ nrm2) {
    const double s = 1.0 / sqrt(nrm2[0]);
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        x[i] *= s;
}

inline unsigned grid_for(uint64_t n, int threads, unsigned cap = 8192) {
    uint64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    return unsigned(g > cap ? cap : g);
}

}  // namespace

namespace {

// K slices of 8 with at most 128 VGPRs (two workgroups per SIMD): measured fastest for every tile
// shape on the H*psi stages of a 24-orbital, M = 512 site problem (short and long reductions alike)
constexpr int GEMM_KS = 8, GEMM_MINB = 4;

template <class TI, class TO, bool BF3 = false>
void launch_gemm_t(const GemmDesc* probs, const GemmRow* rows, const GemmTile* const* tiles, const int64_t* n_tiles,
                   const GemmReduce* reduces, int64_t n_reduces, void* ws, Bases bases, hipStream_t stream) {
    TI* w = static_cast<TI*>(ws);
    constexpr int KSL = BF3 ? 16 : GEMM_KS;  // split-BF16 steps take 16 k at a time
    for (int sh = 0; sh < GEMM_N_SHAPES; ++sh) {
        for (int64_t off = 0; off < n_tiles[sh]; off += 0x7fffffff) {
            int64_t cnt = n_tiles[sh] - off < 0x7fffffff ? n_tiles[sh] - off : 0x7fffffff;
            if (sh == GEMM_32x128)
                hipLaunchKernelGGL((gemm_kernel<1, 4, TI, TO, KSL, GEMM_MINB, 2, BF3>), dim3(unsigned(cnt)), dim3(256), 0, stream, probs,
                                   rows, tiles[sh] + off, bases, w);
            else if (sh == GEMM_64x64)
                hipLaunchKernelGGL((gemm_kernel<2, 2, TI, TO, KSL, GEMM_MINB, 2, BF3>), dim3(unsigned(cnt)), dim3(256), 0, stream, probs,
                                   rows, tiles[sh] + off, bases, w);
            else if (sh == GEMM_96x64)
                hipLaunchKernelGGL((gemm_kernel<3, 2, TI, TO, KSL, GEMM_MINB, 2, BF3>), dim3(unsigned(cnt)), dim3(256), 0, stream, probs,
                                   rows, tiles[sh] + off, bases, w);
            else if (sh == GEMM_64x128)
                hipLaunchKernelGGL((gemm_kernel<2, 4, TI, TO, KSL, GEMM_MINB, 2, BF3>), dim3(unsigned(cnt)), dim3(256), 0, stream, probs,
                                   rows, tiles[sh] + off, bases, w);
            else if (sh == GEMM_128x128)
                hipLaunchKernelGGL((gemm_kernel<4, 2, TI, TO, KSL, GEMM_MINB / 2, 4, BF3>), dim3(unsigned(cnt)), dim3(512), 0,
                                   stream, probs, rows, tiles[sh] + off, bases, w);
            else
                hipLaunchKernelGGL((gemm_kernel<4, 2, TI, TO, KSL, GEMM_MINB, 2, BF3>), dim3(unsigned(cnt)), dim3(256), 0, stream, probs,
                                   rows, tiles[sh] + off, bases, w);
        }
    }
    for (int64_t off = 0; off < n_reduces; off += 0x7fffffff) {
        int64_t cnt = n_reduces - off < 0x7fffffff ? n_reduces - off : 0x7fffffff;
        hipLaunchKernelGGL((gemm_reduce_kernel<TI, TO>), dim3(unsigned(cnt), REDUCE_CHUNKS), dim3(256), 0, stream,
                           probs, reduces + off, bases, w);
    }
}

}  // namespace

hipError_t launch_gemm(GemmPrec prec, const GemmDesc* probs, const GemmRow* rows, const GemmTile* const* tiles,
                       const int64_t* n_tiles, const GemmReduce* reduces, int64_t n_reduces, void* ws, Bases bases,
                       hipStream_t stream) {
    switch (prec) {
        case GEMM_F64:
            launch_gemm_t<double, double>(probs, rows, tiles, n_tiles, reduces, n_reduces, ws, bases, stream);
            break;
        case GEMM_F32:
            launch_gemm_t<float, float>(probs, rows, tiles, n_tiles, reduces, n_reduces, ws, bases, stream);
            break;
        case GEMM_F32_OUT_F64:
            launch_gemm_t<float, double>(probs, rows, tiles, n_tiles, reduces, n_reduces, ws, bases, stream);
            break;
        case GEMM_BF3:
            launch_gemm_t<float, float, true>(probs, rows, tiles, n_tiles, reduces, n_reduces, ws, bases, stream);
            break;
        case GEMM_BF3_OUT_F64:
            launch_gemm_t<float, double, true>(probs, rows, tiles, n_tiles, reduces, n_reduces, ws, bases, stream);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_slines(const SLine* lines, int64_t n_lines, const STerm* terms, const double* T, double* S,
                         hipStream_t stream) {
    for (int64_t off = 0; off < n_lines; off += 0x7fffffff) {
        int64_t cnt = n_lines - off < 0x7fffffff ? n_lines - off : 0x7fffffff;
        hipLaunchKernelGGL(sline_kernel, dim3(unsigned(cnt)), dim3(256), 0, stream, lines + off, terms, T, S);
    }
    return hipGetLastError();
}

template <class T>
hipError_t launch_sblocks(const SBlock* blocks, const SWork* work, int64_t n_work, const int64_t* utile,
                          const uint32_t* kstep, const T* ualpha, const int64_t* lsoff, const int32_t* llds,
                          const T* X, T* S, bool xcd_contiguous, hipStream_t stream) {
    for (int64_t off = 0; off < n_work; off += 0x7fffffff) {
        int64_t cnt = n_work - off < 0x7fffffff ? n_work - off : 0x7fffffff;
        hipLaunchKernelGGL(sblock_kernel<T>, dim3(unsigned(cnt)), dim3(SB_THREADS), 0, stream, blocks, work + off, utile,
                           kstep, ualpha, lsoff, llds, X, S, int(xcd_contiguous));
    }
    return hipGetLastError();
}
template hipError_t launch_sblocks<double>(const SBlock*, const SWork*, int64_t, const int64_t*, const uint32_t*,
                                           const double*, const int64_t*, const int32_t*, const double*, double*, bool,
                                           hipStream_t);
template hipError_t launch_sblocks<float>(const SBlock*, const SWork*, int64_t, const int64_t*, const uint32_t*,
                                          const float*, const int64_t*, const int32_t*, const float*, float*, bool,
                                          hipStream_t);

template <class T>
hipError_t launch_small_lines(const SSmallWork* work, int64_t n_work, const SLine* lines, const STerm* terms,
                              const T* X, T* S, hipStream_t stream) {
    for (int64_t off = 0; off < n_work; off += 0x7fffffff) {
        int64_t cnt = n_work - off < 0x7fffffff ? n_work - off : 0x7fffffff;
        hipLaunchKernelGGL(small_lines_kernel<T>, dim3(unsigned(cnt)), dim3(SMALL_LINE_LANES), 0, stream, work + off,
                           lines, terms, X, S);
    }
    return hipGetLastError();
}
template hipError_t launch_small_lines<double>(const SSmallWork*, int64_t, const SLine*, const STerm*, const double*,
                                               double*, hipStream_t);
template hipError_t launch_small_lines<float>(const SSmallWork*, int64_t, const SLine*, const STerm*, const float*,
                                              float*, hipStream_t);

hipError_t launch_term_grad(const TermGradWork* work, int64_t n_work, const SLine* lines, const STerm* terms,
                            const uint32_t* src, const double* beta, const double* Y, const double* T, double* grad,
                            hipStream_t stream) {
    for (int64_t off = 0; off < n_work; off += 0x7fffffff) {
        const int64_t cnt = n_work - off < 0x7fffffff ? n_work - off : 0x7fffffff;
        hipLaunchKernelGGL(term_grad_kernel, dim3(unsigned(cnt)), dim3(TERM_GRAD_THREADS), 0, stream, work + off, lines,
                           terms, src, beta, Y, T, grad);
    }
    return hipGetLastError();
}

hipError_t launch_convert(const double* src, float* dst, uint64_t n, hipStream_t stream) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(n, 256)), dim3(256), 0, stream, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_fill(double* p, uint64_t n, double v, hipStream_t stream) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

hipError_t launch_copy_jobs(const CopyJob* jobs, int32_t n_jobs, const double* src, double* dst, hipStream_t stream) {
    if (n_jobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_jobs_kernel, dim3(unsigned(n_jobs), 16), dim3(256), 0, stream, jobs, src, dst);
    return hipGetLastError();
}

hipError_t launch_reshape(const ReshapeSeg* segs, int32_t n_segs, const double* src, double* dst, int to_left,
                          hipStream_t stream) {
    if (n_segs <= 0) return hipSuccess;
    hipLaunchKernelGGL(reshape_kernel, dim3(unsigned(n_segs), 8), dim3(256), 0, stream, segs, src, dst, to_left);
    return hipGetLastError();
}

hipError_t launch_dot(const double* x, const double* y, uint64_t n, double* work, double* out, hipStream_t stream) {
    hipLaunchKernelGGL(dot_partial_kernel, dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, stream, x, y, n, work);
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(DOT_THREADS), 0, stream, work, DOT_BLOCKS, out);
    return hipGetLastError();
}

hipError_t launch_axpy_dev(double* y, const double* x, uint64_t n, const double* num, const double* den, double sign,
                           hipStream_t stream) {
    hipLaunchKernelGGL(axpy_dev_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, y, x, n, num, den, sign);
    return hipGetLastError();
}

hipError_t launch_seg_dot(const double* x, const double* y, const int64_t* off, int nseg, double* part, double* out,
                          hipStream_t stream) {
    if (nseg <= 0) return hipSuccess;
    hipLaunchKernelGGL(seg_dot_partial_kernel, dim3(SEG_CHUNKS, unsigned(nseg)), dim3(DOT_THREADS), 0, stream, x, y,
                       off, part);
    hipLaunchKernelGGL(seg_dot_final_kernel, dim3(unsigned((nseg + 255) / 256)), dim3(256), 0, stream, part, nseg, out);
    return hipGetLastError();
}

hipError_t launch_seg_axpby(double* y, const double* x, const int64_t* off, int nseg, const double* a, const double* b,
                            hipStream_t stream) {
    if (nseg <= 0) return hipSuccess;
    hipLaunchKernelGGL(seg_axpby_kernel, dim3(SEG_CHUNKS, unsigned(nseg)), dim3(256), 0, stream, y, x, off, a, b);
    return hipGetLastError();
}

// Search space of mq_davidson_noalloc: one copy per GPU, allocated with the code object and held
// until the process ends.  Host code passes its device address (noalloc_scratch_address), never
// the symbol itself.
static __device__ __align__(256) double noalloc_scratch[MQ_NOALLOC_SCRATCH_ELEMS];

hipError_t noalloc_scratch_address(void** ptr) { return hipGetSymbolAddress(ptr, HIP_SYMBOL(noalloc_scratch)); }

hipError_t launch_axpby(double* y, const double* x, uint64_t n, double a, double b, hipStream_t stream) {
    hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, y, x, n, a, b);
    return hipGetLastError();
}

hipError_t launch_scale_rsqrt(double* x, uint64_t n, const double* nrm2, hipStream_t stream) {
    hipLaunchKernelGGL(scale_rsqrt_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, x, n, nrm2);
    return hipGetLastError();
}

}  // namespace mq
