// Launchers of the gfx950 kernels (device.hip).  All launches are stream-ordered; none allocates.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "descriptors.hpp"
#include "plans.hpp"

namespace mq {

// operand base pointers of a grouped GEMM launch, indexed by BASE_* (element type given by the
// launch's precision)
struct Bases {
    void* p[BASE_COUNT];
};

// arithmetic of a grouped GEMM launch: operands, compute and workspace in FP64 or FP32; the
// output in the same type, or FP64 from FP32 operands
// GEMM_BF3*: FP32 operands, split-BF16 products (hi hi + hi lo + lo hi) on the BF16 matrix cores
enum GemmPrec { GEMM_F64 = 0, GEMM_F32 = 1, GEMM_F32_OUT_F64 = 2, GEMM_BF3 = 3, GEMM_BF3_OUT_F64 = 4 };

// grouped GEMM stage: tiles of each shape (GemmShape) then the split-K reduce jobs; ws is the
// workspace of the split partials (GemmSchedule::ws_elems elements of the compute type); rows is
// the row table of the stage's stacked (GEMM_ROWMAP) problems
hipError_t launch_gemm(GemmPrec prec, const GemmDesc* probs, const GemmRow* rows, const GemmTile* const* tiles,
                       const int64_t* n_tiles, const GemmReduce* reduces, int64_t n_reduces, void* ws, Bases bases,
                       hipStream_t stream);
// block-form MPO-term sum (SBlockTable uploaded; alpha, T and S of type T): writes every S tile
// of the table's lines; xcd_contiguous: each XCD takes a contiguous range of the work list
template <class T>
hipError_t launch_sblocks(const SBlock* blocks, const SWork* work, int64_t n_work, const int64_t* utile,
                          const uint32_t* kstep, const T* ualpha, const int64_t* lsoff, const int32_t* llds,
                          const T* X, T* S, bool xcd_contiguous, hipStream_t stream);
// small-tile MPO-term sum (SBlockTable small_* tables uploaded; T and S of type T, coefficients
// FP64, rounded to T): writes the S tiles of the small lines
template <class T>
hipError_t launch_small_lines(const SSmallWork* work, int64_t n_work, const SLine* lines, const STerm* terms,
                              const T* X, T* S, hipStream_t stream);
// reference-style MPO-term sum: one workgroup per S line, every element gathers all of the line's
// terms (the per-term formulation the block-form kernel replaces)
hipError_t launch_slines(const SLine* lines, int64_t n_lines, const STerm* terms, const double* T, double* S,
                         hipStream_t stream);
// MPO-term adjoint: for every term k of the work items' lines, grad[src[k]] += beta[k] * <Y tile of
// the line, T tile of the term> (FP64 atomics; the order of the additions is not fixed)
hipError_t launch_term_grad(const TermGradWork* work, int64_t n_work, const SLine* lines, const STerm* terms,
                            const uint32_t* src, const double* beta, const double* Y, const double* T, double* grad,
                            hipStream_t stream);
// dst = float(src)
hipError_t launch_convert(const double* src, float* dst, uint64_t n, hipStream_t stream);
hipError_t launch_fill(double* p, uint64_t n, double v, hipStream_t stream);
// device address (on the current device) of the fixed scratch array of the allocation-free entry
// points, MQ_NOALLOC_SCRATCH_ELEMS doubles
hipError_t noalloc_scratch_address(void** ptr);
hipError_t launch_copy_jobs(const CopyJob* jobs, int32_t n_jobs, const double* src, double* dst, hipStream_t stream);

// right-paired <-> left-paired MPS reshape segments (one per (left sector, sigma) with a right
// sector); rp / lp are the element offsets of the (0, col) / (row, 0) corners
using ReshapeSeg = ReshapeSegHost;
hipError_t launch_reshape(const ReshapeSeg* segs, int32_t n_segs, const double* src, double* dst, int to_left,
                          hipStream_t stream);

// ---- Davidson vector kernels (FP64) ---------------------------------------------------------
// out[0] = x . y  (deterministic two-level reduction; work holds DOT_BLOCKS partials)
constexpr int DOT_BLOCKS = 512;
hipError_t launch_dot(const double* x, const double* y, uint64_t n, double* work, double* out, hipStream_t stream);
// y += coef * x with coef = sign * num[0] / (den ? den[0] : 1)  (coefficients stay on the device)
hipError_t launch_axpy_dev(double* y, const double* x, uint64_t n, const double* num, const double* den, double sign,
                           hipStream_t stream);
// y = a*y + b*x (host scalars); a == 0 assigns b*x without reading y
hipError_t launch_axpby(double* y, const double* x, uint64_t n, double a, double b, hipStream_t stream);
// segmented forms over nseg (<= 65535) vectors concatenated in one buffer, segment s being
// elements off[s] .. off[s+1]-1 (off on the device): out[s] = x_s . y_s (part holds
// nseg * SEG_CHUNKS partials), and y_s = a[s]*y_s + b[s]*x_s (a, b on the device)
constexpr int SEG_CHUNKS = 64;
hipError_t launch_seg_dot(const double* x, const double* y, const int64_t* off, int nseg, double* part, double* out,
                          hipStream_t stream);
hipError_t launch_seg_axpby(double* y, const double* x, const int64_t* off, int nseg, const double* a, const double* b,
                            hipStream_t stream);
// x *= 1/sqrt(nrm2[0])
hipError_t launch_scale_rsqrt(double* x, uint64_t n, const double* nrm2, hipStream_t stream);

}  // namespace mq
