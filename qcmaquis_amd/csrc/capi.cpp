// C ABI of the contraction path: context, HBM block store, plan upload and execution, growth
// steps and the Jacobi-Davidson eigensolver.  No C++ type crosses the ABI; errors are returned as
// codes with the message kept on the context.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <omp.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "core.hpp"
#include "coupling.hpp"
#include "descriptors.hpp"
#include "device.hpp"
#include "plans.hpp"
#include "qcmaquis_amd.h"

using namespace mq;

struct mq_ctx_s {
    int device = 0;
    Sym sym = Sym::TwoU1;
    hipStream_t stream = nullptr;
    std::string err;
    double* dwork = nullptr;  // DOT_BLOCKS partials + 64 scalars
    bool plain_tables = false;         // MQ_OPT_PLAIN_TABLES
    int32_t small_tile_max = MQ_SMALL_TILE_AUTO;  // MQ_OPT_SMALL_TILE_MAX
    int32_t chunk_mb = MQ_CHUNK_AUTO;             // MQ_OPT_CHUNK_MB
    bool term_sources = false;         // MQ_OPT_TERM_SOURCES
    int32_t sblock_panel = MQ_SBLOCK_PANEL_AUTO;  // MQ_OPT_SBLOCK_PANEL
    bool sblock_xcd = false;                      // MQ_OPT_SBLOCK_XCD
    bool autotune = false;                        // MQ_OPT_AUTOTUNE
    int32_t jcd_gmres = 0;                        // MQ_OPT_JCD_GMRES
    bool overlap = true;                          // MQ_OPT_OVERLAP
    hipStream_t stream2 = nullptr;     // second stream of overlapped H*psi chunks, made on first use
    struct Tuned {
        int32_t pick;
        float best_ms, default_ms;
    };
    std::map<std::pair<uint64_t, int>, Tuned> tuned;  // (stage key, kind) -> pick
    rocblas_handle blas = nullptr;     // library GEMMs of MQ_PREC_FP64_PLAIN, made on first use
    double* noalloc_scratch = nullptr; // this device's scratch array of mq_davidson_noalloc
};

struct mq_mpo_s {
    mq_ctx_t ctx;
    MPO mpo;
};

struct mq_boundary_s {
    mq_ctx_t ctx;
    BoundaryIndex idx;
    double* data = nullptr;
    uint64_t n = 0;
    mq_growth_stats_t grow{};  // how mq_lstep / mq_rstep built it
    // owner of the device block `data`: H*psi plans that read the boundary in place share it, so
    // the block outlives mq_boundary_destroy while a plan still reads it
    std::shared_ptr<void> hold;
};

struct mq_mps_s {
    mq_ctx_t ctx;
    MPSLayout lay;
    double* rp = nullptr;
    double* lp = nullptr;
    bool rp_ok = false, lp_ok = false;
    ReshapeSeg* d_segs = nullptr;
    int32_t n_segs = 0;
};

namespace {

void hipcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(MQ_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Device blocks of 1 MiB and more are kept for reuse when freed: a sweep creates and destroys
// plans of similar sizes at every site, and hipMalloc / hipFree of gigabyte buffers cost
// milliseconds each (hipFree also waits for the device).  A freed block serves a later request
// of at most its size and at least 4/5 of it on the same device; at most CACHE_CAP bytes are kept,
// the blocks idle longest released first (a sweep's next sites reuse what its last sites freed,
// large plan buffers and small tables alike), and all idle blocks are released before an
// allocation that the device cannot satisfy is given up.  A block enters the cache after a device
// synchronisation (as hipFree waits for the device), so a reused block is idle; like a fresh
// hipMalloc it holds undefined contents.
struct DevCache {
    struct Idle {
        int dev;
        void* p;
        uint64_t seq;  // when it became idle
    };
    std::mutex m;
    std::unordered_map<void*, std::pair<int, uint64_t>> live;   // block -> (device, bytes)
    std::multimap<uint64_t, Idle> idle;                         // bytes -> idle block
    std::map<uint64_t, std::multimap<uint64_t, Idle>::iterator> age;  // seq -> its idle entry
    uint64_t idle_bytes = 0, next_seq = 0;
    void drop(std::multimap<uint64_t, Idle>::iterator it) {  // forget an idle entry (no hipFree)
        idle_bytes -= it->first;
        age.erase(it->second.seq);
        idle.erase(it);
    }
};
constexpr uint64_t CACHE_MIN = uint64_t(1) << 20;
constexpr uint64_t CACHE_CAP = uint64_t(96) << 30;

DevCache& devcache() {
    static DevCache* c = new DevCache;  // never destroyed: frees may come from static destructors
    return *c;
}

void free_on(int dev, void* p) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (dev != cur) (void)hipSetDevice(dev);
    (void)hipFree(p);
    if (dev != cur) (void)hipSetDevice(cur);
}

// releases the idle blocks (of one device, or all with dev < 0); caller holds the lock
void cache_release_locked(DevCache& c, int dev) {
    for (auto it = c.idle.begin(); it != c.idle.end();) {
        auto next = std::next(it);
        if (dev < 0 || it->second.dev == dev) {
            free_on(it->second.dev, it->second.p);
            c.drop(it);
        }
        it = next;
    }
}

template <class T>
T* dev_alloc(uint64_t n) {
    if (n == 0) n = 1;
    const uint64_t bytes = n * sizeof(T);
    int dev = 0;
    (void)hipGetDevice(&dev);
    DevCache& c = devcache();
    std::lock_guard<std::mutex> lock(c.m);
    if (bytes >= CACHE_MIN) {
        for (auto it = c.idle.lower_bound(bytes); it != c.idle.end() && it->first <= bytes + bytes / 4; ++it)
            if (it->second.dev == dev) {
                void* p = it->second.p;
                c.live[p] = {dev, it->first};
                c.drop(it);
                return static_cast<T*>(p);
            }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        cache_release_locked(c, dev);
        e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(MQ_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    c.live[p] = {dev, bytes};
    return static_cast<T*>(p);
}

// Host-to-device copies of large tables go through two pinned staging halves: a piece is copied
// into one half by the host threads while the previous piece moves from the other half by DMA
// (pinned transfers run at about twice the rate of pageable ones).  Process-wide, one copy at a
// time; smaller copies go directly from pageable memory.
struct Staging {
    std::mutex m;
    char* half[2] = {nullptr, nullptr};
};
constexpr size_t STAGE_HALF = size_t(64) << 20;
constexpr size_t STAGE_MIN = size_t(4) << 20;

Staging& staging() {
    static Staging* st = new Staging;  // pinned halves live until the process ends
    return *st;
}

void upload_bytes(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes < STAGE_MIN) {
        hipcheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "upload");
        hipcheck(hipStreamSynchronize(s), "upload");
        return;
    }
    Staging& st = staging();
    std::lock_guard<std::mutex> lock(st.m);
    for (auto& h : st.half)
        if (!h) hipcheck(hipHostMalloc(reinterpret_cast<void**>(&h), STAGE_HALF, hipHostMallocPortable), "pinned staging");
    const char* in = static_cast<const char*>(src);
    char* out = static_cast<char*>(dst);
    bool in_flight = false;
    int h = 0;
    try {
        for (size_t off = 0; off < bytes; off += STAGE_HALF, h ^= 1) {
            const size_t n = std::min(STAGE_HALF, bytes - off);
            // fill this half while the other half's transfer runs
            const int64_t pieces = int64_t((n + (size_t(1) << 20) - 1) >> 20);
#pragma omp parallel for schedule(static)
            for (int64_t q = 0; q < pieces; ++q) {
                const size_t a = size_t(q) << 20, len = std::min(size_t(1) << 20, n - a);
                std::memcpy(st.half[h] + a, in + off + a, len);
            }
            if (in_flight) hipcheck(hipStreamSynchronize(s), "upload");
            hipcheck(hipMemcpyAsync(out + off, st.half[h], n, hipMemcpyHostToDevice, s), "upload");
            in_flight = true;
        }
        hipcheck(hipStreamSynchronize(s), "upload");
    } catch (...) {
        (void)hipStreamSynchronize(s);  // no transfer may still read a half when the lock is released
        throw;
    }
}

// Copies a host table to the device and waits for the copy: callers pass vectors that may be
// destroyed as soon as this returns.
template <class T, class A>
T* dev_upload(std::vector<T, A> const& v, hipStream_t s) {
    T* p = dev_alloc<T>(v.size());
    if (!v.empty()) upload_bytes(p, v.data(), v.size() * sizeof(T), s);
    return p;
}

void dev_free(void* p) {
    if (!p) return;
    DevCache& c = devcache();
    std::lock_guard<std::mutex> lock(c.m);
    auto it = c.live.find(p);
    if (it == c.live.end() || it->second.second < CACHE_MIN) {
        if (it != c.live.end()) c.live.erase(it);
        (void)hipFree(p);
        return;
    }
    // work on any stream may still read the block (hipFree would wait for the device too)
    (void)hipDeviceSynchronize();
    const uint64_t seq = c.next_seq++;
    auto ins = c.idle.emplace(it->second.second, DevCache::Idle{it->second.first, p, seq});
    c.age.emplace(seq, ins);
    c.idle_bytes += it->second.second;
    c.live.erase(it);
    while (c.idle_bytes > CACHE_CAP) {
        auto oldest = c.age.begin()->second;
        free_on(oldest->second.dev, oldest->second.p);
        c.drop(oldest);
    }
}

// shared owner of a device block (freed with dev_free when the last owner lets go)
std::shared_ptr<void> own_device(void* p) {
    return std::shared_ptr<void>(p, [](void* q) { dev_free(q); });
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int d) {
        (void)hipGetDevice(&prev);
        if (prev != d) hipcheck(hipSetDevice(d), "hipSetDevice");
    }
    ~DeviceGuard() {
        int cur;
        if (hipGetDevice(&cur) == hipSuccess && prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

template <class F>
int guarded(mq_ctx_t ctx, F&& f) {
    if (!ctx) return MQ_ERR_INVALID;
    try {
        DeviceGuard g(ctx->device);
        f();
        return MQ_OK;
    } catch (Error const& e) {
        ctx->err = e.what();
        return e.code;
    } catch (std::bad_alloc const&) {
        ctx->err = "host allocation failed";
        return MQ_ERR_NOMEM;
    } catch (std::exception const& e) {
        ctx->err = e.what();
        return MQ_ERR_INVALID;
    }
}

// device copy of a grouped GEMM stage and its schedule (tiles per shape, split-K reduce jobs and
// the workspace of the partial tiles, sized for FP64 and shared by the FP32 launches)
struct GemmStage {
    GemmDesc* probs = nullptr;
    GemmRow* rows = nullptr;
    GemmTile* tiles[GEMM_N_SHAPES] = {};
    int64_t n_tiles[GEMM_N_SHAPES] = {};
    GemmReduce* reduces = nullptr;
    int64_t n_reduces = 0;
    double* ws = nullptr;
    bool own_ws = true;
    // sch: the stage's schedule (built here when null); ws_shared: a split-K workspace of at least
    // sch->ws_elems elements owned by the caller (allocated here when null)
    void upload(std::vector<GemmDesc> const& v, std::vector<GemmRow> const& r, hipStream_t s,
                GemmSchedule const* pre = nullptr, double* ws_shared = nullptr) {
        GemmSchedule local;
        if (!pre) {
            local = build_gemm_schedule(v);
            pre = &local;
        }
        GemmSchedule const& sch = *pre;
        probs = dev_upload(v, s);
        rows = dev_upload(r, s);
        for (int sh = 0; sh < GEMM_N_SHAPES; ++sh) {
            tiles[sh] = dev_upload(sch.tiles[sh], s);
            n_tiles[sh] = int64_t(sch.tiles[sh].size());
        }
        reduces = dev_upload(sch.reduces, s);
        n_reduces = int64_t(sch.reduces.size());
        own_ws = ws_shared == nullptr;
        ws = own_ws ? dev_alloc<double>(sch.ws_elems) : ws_shared;
    }
    void run(Bases const& b, hipStream_t s, GemmPrec prec) const {
        hipcheck(launch_gemm(prec, probs, rows, tiles, n_tiles, reduces, n_reduces, ws, b, s), "gemm launch");
    }
    void release() {
        dev_free(probs);
        dev_free(rows);
        rows = nullptr;
        for (auto& t : tiles) dev_free(t), t = nullptr;
        dev_free(reduces);
        if (own_ws) dev_free(ws);
        probs = nullptr;
        reduces = nullptr;
        ws = nullptr;
    }
};

// device copy of the block-form MPO-term sum; the FP32 coefficient table is made on first use
struct SAsmStage {
    SBlock* blocks = nullptr;
    SWork* work = nullptr;
    int64_t* utile = nullptr;
    uint32_t* kstep = nullptr;
    double* ualpha = nullptr;
    float* ualpha32 = nullptr;
    int64_t* lsoff = nullptr;
    int32_t* llds = nullptr;
    int64_t n_work = 0;
    uint64_t n_alpha = 0;
    // small-tile lines
    SSmallWork* swork = nullptr;
    SLine* slines = nullptr;
    STerm* sterms = nullptr;
    int64_t n_swork = 0;
    bool xcd_contiguous = false;  // the work list is panel-ordered: each XCD takes a range of it
    void upload(SBlockTable const& B, hipStream_t s, bool xcd = false) {
        xcd_contiguous = xcd;
        blocks = dev_upload(B.blocks, s);
        work = dev_upload(B.work, s);
        utile = dev_upload(B.utile, s);
        kstep = dev_upload(B.kstep, s);
        ualpha = dev_upload(B.ualpha, s);
        lsoff = dev_upload(B.lsoff, s);
        llds = dev_upload(B.llds, s);
        n_work = int64_t(B.work.size());
        n_alpha = B.ualpha.size();
        swork = dev_upload(B.small_work, s);
        slines = dev_upload(B.small_lines, s);
        sterms = dev_upload(B.small_terms, s);
        n_swork = int64_t(B.small_work.size());
    }
    void run(const double* T, double* S, hipStream_t s) const {
        if (n_work)
            hipcheck(launch_sblocks<double>(blocks, work, n_work, utile, kstep, ualpha, lsoff, llds, T, S, xcd_contiguous, s),
                     "S-assembly");
        if (n_swork) hipcheck(launch_small_lines<double>(swork, n_swork, slines, sterms, T, S, s), "S-assembly");
    }
    void ensure32(hipStream_t s) {
        if (!ualpha32) {
            ualpha32 = dev_alloc<float>(n_alpha);
            hipcheck(launch_convert(ualpha, ualpha32, n_alpha, s), "convert");
        }
    }
    void run32(const float* T, float* S, hipStream_t s) {
        ensure32(s);
        if (n_work)
            hipcheck(launch_sblocks<float>(blocks, work, n_work, utile, kstep, ualpha32, lsoff, llds, T, S, xcd_contiguous,
                                           s),
                     "S-assembly");
        if (n_swork) hipcheck(launch_small_lines<float>(swork, n_swork, slines, sterms, T, S, s), "S-assembly");
    }
    void release() {
        dev_free(blocks); dev_free(work); dev_free(utile); dev_free(kstep); dev_free(ualpha); dev_free(ualpha32);
        dev_free(lsoff); dev_free(llds); dev_free(swork); dev_free(slines); dev_free(sterms);
        blocks = nullptr; work = nullptr; utile = nullptr; kstep = nullptr; ualpha = nullptr; ualpha32 = nullptr;
        lsoff = nullptr;
        llds = nullptr; swork = nullptr; slines = nullptr; sterms = nullptr;
        n_work = 0;
        n_swork = 0;
    }
};

// blocks per panel of the term-sum work order (MQ_OPT_SBLOCK_PANEL)
int32_t sblock_panel(mq_ctx_t ctx) { return ctx->sblock_panel == MQ_SBLOCK_PANEL_AUTO ? 0 : ctx->sblock_panel; }

Index index_from(const mq_index_desc* d) {
    require(d != nullptr, "null index descriptor");
    return Index(*d);
}

void ensure_rp(mq_mps_t m) {
    if (m->rp_ok) return;
    require(m->lp_ok, "mps: no valid data");
    hipcheck(launch_reshape(m->d_segs, m->n_segs, m->lp, m->rp, 0, m->ctx->stream), "reshape");
    m->rp_ok = true;
}
void ensure_lp(mq_mps_t m) {
    if (m->lp_ok) return;
    require(m->rp_ok, "mps: no valid data");
    hipcheck(launch_reshape(m->d_segs, m->n_segs, m->rp, m->lp, 1, m->ctx->stream), "reshape");
    m->lp_ok = true;
}

}  // namespace

struct mq_hpsi_plan_s {
    mq_ctx_t ctx;
    HpsiPlanHost host;
    const double* rdata = nullptr;
    std::shared_ptr<void> rhold;  // the right boundary's storage, read in place by every H*psi
    // the stages of every chunk of the plan, run chunk after chunk on the T and S buffers
    struct Chunk {
        GemmStage tg, cg;
        SAsmStage sa;
    };
    std::vector<Chunk> chunks;
    double* gemm_ws = nullptr;   // split-K workspace shared by the chunks' GEMM stages
    uint64_t gemm_ws_elems = 0;
    // overlapped chunks (MQ_OPT_OVERLAP, plans of two or more chunks): chunk c's output GEMMs run on
    // the context's second stream from the S buffer of parity c % 2 (S, S2) and their own split-K
    // workspace, while chunk c + 1's T-GEMMs and term sum run on the context stream; per chunk one
    // event marks its term sum and one its output GEMMs
    double* S2 = nullptr;
    double* gemm_ws2 = nullptr;
    uint64_t gemm_ws2_elems = 0;
    std::vector<hipEvent_t> ev_sa, ev_cg;
    hipEvent_t ev_fork = nullptr;
    int32_t policy_t = 0, policy_c = 0;  // schedule candidates of chunk 0's stages
    double seconds_tune = 0;
    int64_t n_lines = 0, n_sblocks = 0, n_reduces = 0;
    uint64_t s_line_elements = 0, gathered_elements = 0, union_elements = 0, max_union_tiles = 0, sblock_macs = 0;
    double *T = nullptr, *S = nullptr, *Lpanel = nullptr;
    // FP32 operands of the mixed-precision H*psi, made on first use
    float *T32 = nullptr, *S32 = nullptr, *S32b = nullptr, *Lpanel32 = nullptr, *R32 = nullptr, *psi32 = nullptr;
    // captured H*psi per precision (FP64, FP32) on plan-owned psi buffers
    struct Graph {
        hipGraphExec_t exec = nullptr;
        double *in = nullptr, *out = nullptr;
    } graphs[2];
    bool use_graphs = false;
    uint64_t r_elems = 0;
    uint64_t psi_n = 0;          // psi elements (all instances of a batch plan)
    int n_instances = 1;         // independent site problems in the plan
    double* rown = nullptr;      // right boundaries copied into the plan (batch plans)
    std::vector<int64_t> inst_off;    // psi offset of each instance, and psi_n (batch plans)
    int64_t* d_inst_off = nullptr;    // the same on the device
    double build_seconds = 0, seconds_tables = 0, seconds_blocks = 0, seconds_plain = 0;
    uint64_t small_lines = 0, small_terms = 0, small_gathered = 0;
    // per-task tables of MQ_PREC_FP64_PLAIN (MQ_OPT_PLAIN_TABLES)
    struct Plain {
        bool ok = false;
        std::vector<GemmDesc> t, c;
        struct Range { size_t t0, t1, c0, c1, l0, l1; };  // per chunk: t, c and line ranges
        std::vector<Range> ranges;
        SLine* lines = nullptr;
        STerm* terms = nullptr;
        int64_t n_lines = 0;
    } plain;
    // MPO-term adjoint (plans made with MQ_OPT_TERM_SOURCES): per chunk the GEMMs of
    // Y = L_panel^T psi (written where the chunk's S lines sit) and the term work items
    struct Adjoint {
        bool ok = false;
        uint32_t n_mpo_terms = 0;
        std::vector<GemmStage> yg;
        std::vector<TermGradWork*> work;
        std::vector<int64_t> n_work;
        SLine* lines = nullptr;
        STerm* terms = nullptr;
        uint32_t* src = nullptr;
        double* beta = nullptr;
        double* grad = nullptr;
    } adj;
};

// ================================================================================================
extern "C" {

int mq_version(void) { return 1; }

int mq_abi_struct_sizes(uint32_t* sizes, int n) {
    const uint32_t v[] = {uint32_t(sizeof(mq_index_desc)), uint32_t(sizeof(mq_mpo_desc)),
                          uint32_t(sizeof(mq_bindex_desc)), uint32_t(sizeof(mq_boundary_info_t)),
                          uint32_t(sizeof(mq_hpsi_stats_t)), uint32_t(sizeof(mq_davidson_stats_t)),
                          uint32_t(sizeof(mq_growth_stats_t))};
    int k = 0;
    for (; sizes && k < n && k < int(sizeof(v) / sizeof(v[0])); ++k) sizes[k] = v[k];
    return k;
}

double mq_wigner_6j(int a, int b, int c, int d, int e, int f) { return wigner_6j(a, b, c, d, e, f); }
double mq_wigner_9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    return wigner_9j(a, b, c, d, e, f, g, h, i);
}
int mq_index_position(const mq_index_desc* idx, int32_t q0, int32_t q1) {
    if (!idx) return MQ_ERR_INVALID;
    try {
        Index I(*idx);
        return int(I.position(Charge(q0, q1)));
    } catch (...) {
        return MQ_ERR_INVALID;
    }
}

// device address of the scratch array of mq_davidson_noalloc, looked up once per device with that
// device current (each GPU holds its own copy of the array)
static double* noalloc_scratch_of(int device) {
    static std::mutex mtx;
    static std::vector<double*> addr;
    std::lock_guard<std::mutex> lock(mtx);
    if (int(addr.size()) <= device) addr.resize(size_t(device) + 1, nullptr);
    if (!addr[device]) {
        void* p = nullptr;
        hipcheck(noalloc_scratch_address(&p), "scratch symbol address");
        require(p != nullptr, "scratch symbol address: null");
        addr[device] = static_cast<double*>(p);
    }
    return addr[device];
}

int mq_ctx_create(int device, int symmetry, mq_ctx_t* out) {
    if (!out) return MQ_ERR_INVALID;
    *out = nullptr;
    if (symmetry != MQ_SYM_TWOU1 && symmetry != MQ_SYM_SU2U1) return MQ_ERR_INVALID;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return MQ_ERR_HIP;
    // The schedule builders' OpenMP regions alternate with the caller's own multithreaded host
    // work (LAPACK in a sweep's SVD split, for one).  libomp keeps idle workers spinning for 200 ms
    // after each region by default, which takes those cores from that work; here they sleep at
    // once, unless the caller has set KMP_BLOCKTIME.
    if (!std::getenv("KMP_BLOCKTIME")) kmp_set_blocktime(0);
    auto* c = new mq_ctx_s();
    c->device = device;
    c->sym = Sym(symmetry);
    int rc = guarded(c, [&] {
        hipcheck(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream create");
        c->dwork = dev_alloc<double>(DOT_BLOCKS + 64);
        c->noalloc_scratch = noalloc_scratch_of(device);
    });
    if (rc != MQ_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return MQ_OK;
}

int mq_ctx_destroy(mq_ctx_t ctx) {
    if (!ctx) return MQ_OK;
    int rc = guarded(ctx, [&] {
        (void)hipStreamSynchronize(ctx->stream);
        dev_free(ctx->dwork);
        {
            DevCache& c = devcache();
            std::lock_guard<std::mutex> lock(c.m);
            cache_release_locked(c, ctx->device);
        }
        if (ctx->blas) (void)rocblas_destroy_handle(ctx->blas);
        if (ctx->stream2) {
            (void)hipStreamSynchronize(ctx->stream2);
            (void)hipStreamDestroy(ctx->stream2);
        }
        (void)hipStreamDestroy(ctx->stream);
    });
    delete ctx;
    return rc;
}

const char* mq_last_error(mq_ctx_t ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mq_ctx_set_option(mq_ctx_t ctx, int option, int value) {
    return guarded(ctx, [&] {
        if (option == MQ_OPT_PLAIN_TABLES) {
            ctx->plain_tables = value != 0;
        } else if (option == MQ_OPT_SMALL_TILE_MAX) {
            require(value == MQ_SMALL_TILE_AUTO || (value >= 0 && value <= SMALL_LINE_LANES),
                    "mq_ctx_set_option: MQ_OPT_SMALL_TILE_MAX is MQ_SMALL_TILE_AUTO or 0..256");
            ctx->small_tile_max = value;
        } else if (option == MQ_OPT_CHUNK_MB) {
            require(value == MQ_CHUNK_AUTO || value >= 0, "mq_ctx_set_option: MQ_OPT_CHUNK_MB is MQ_CHUNK_AUTO or >= 0");
            ctx->chunk_mb = value;
        } else if (option == MQ_OPT_TERM_SOURCES) {
            ctx->term_sources = value != 0;
        } else if (option == MQ_OPT_SBLOCK_PANEL) {
            require(value == MQ_SBLOCK_PANEL_AUTO || value >= 0,
                    "mq_ctx_set_option: MQ_OPT_SBLOCK_PANEL is MQ_SBLOCK_PANEL_AUTO or >= 0");
            ctx->sblock_panel = value;
        } else if (option == MQ_OPT_SBLOCK_XCD) {
            ctx->sblock_xcd = value != 0;
        } else if (option == MQ_OPT_AUTOTUNE) {
            ctx->autotune = value != 0;
        } else if (option == MQ_OPT_OVERLAP) {
            require(value == 0 || value == 1, "mq_ctx_set_option: MQ_OPT_OVERLAP is 0 or 1");
            ctx->overlap = value == 1;
        } else if (option == MQ_OPT_JCD_GMRES) {
            require(value >= 0 && value <= 64, "mq_ctx_set_option: MQ_OPT_JCD_GMRES is 0..64");
            ctx->jcd_gmres = value;
        } else {
            throw Error(MQ_ERR_INVALID, "mq_ctx_set_option: unknown option");
        }
    });
}

int mq_ctx_tuned(mq_ctx_t ctx, int n, uint64_t* keys, int32_t* kinds, int32_t* picks, float* best_ms,
                 float* default_ms) {
    if (!ctx || n < 0) return MQ_ERR_INVALID;
    int i = 0;
    for (auto const& kv : ctx->tuned) {
        if (i < n) {
            if (keys) keys[i] = kv.first.first;
            if (kinds) kinds[i] = kv.first.second;
            if (picks) picks[i] = kv.second.pick;
            if (best_ms) best_ms[i] = kv.second.best_ms;
            if (default_ms) default_ms[i] = kv.second.default_ms;
        }
        ++i;
    }
    return i;
}

int mq_ctx_sync(mq_ctx_t ctx) {
    return guarded(ctx, [&] { hipcheck(hipStreamSynchronize(ctx->stream), "stream synchronize"); });
}

void* mq_ctx_stream(mq_ctx_t ctx) { return ctx ? (void*)ctx->stream : nullptr; }

// ---- MPO -------------------------------------------------------------------------------------
int mq_mpo_create(mq_ctx_t ctx, const mq_mpo_desc* desc, mq_mpo_t* out) {
    return guarded(ctx, [&] {
        require(desc && out, "mq_mpo_create: null argument");
        *out = new mq_mpo_s{ctx, MPO(*desc)};
    });
}
int mq_mpo_destroy(mq_mpo_t m) {
    delete m;
    return MQ_OK;
}

// ---- boundaries ------------------------------------------------------------------------------
int mq_boundary_create_initial(mq_ctx_t ctx, const mq_index_desc* idx, uint32_t aux_dim, mq_boundary_t* out) {
    return guarded(ctx, [&] {
        require(out && aux_dim >= 1, "mq_boundary_create_initial: bad argument");
        auto* b = new mq_boundary_s{ctx, initial_boundary_index(index_from(idx), aux_dim)};
        std::unique_ptr<mq_boundary_s> guard(b);
        b->n = b->idx.layout();
        b->data = dev_alloc<double>(b->n);
        b->hold = own_device(b->data);
        hipcheck(launch_fill(b->data, b->n, 1.0, ctx->stream), "fill");
        hipcheck(hipStreamSynchronize(ctx->stream), "sync");
        *out = guard.release();
    });
}

int mq_boundary_upload(mq_ctx_t ctx, const mq_bindex_desc* d, const double* data, mq_boundary_t* out) {
    return guarded(ctx, [&] {
        require(d && out, "mq_boundary_upload: bad argument");
        auto* b = new mq_boundary_s{ctx, boundary_index_from(*d)};
        std::unique_ptr<mq_boundary_s> guard(b);
        b->n = b->idx.layout();
        b->data = dev_alloc<double>(b->n);
        b->hold = own_device(b->data);
        if (b->n) {
            require(data != nullptr, "mq_boundary_upload: null data");
            hipcheck(hipMemcpyAsync(b->data, data, b->n * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "upload");
        }
        hipcheck(hipStreamSynchronize(ctx->stream), "sync");
        *out = guard.release();
    });
}

int mq_boundary_info(mq_boundary_t b, mq_boundary_info_t* info) {
    if (!b || !info) return MQ_ERR_INVALID;
    info->n_cohorts = int32_t(b->idx.n_cohorts());
    info->aux_dim = b->idx.aux_dim;
    info->n_bra = int32_t(b->idx.bra.size());
    info->n_ket = int32_t(b->idx.ket.size());
    info->total_elements = b->n;
    return MQ_OK;
}

int mq_boundary_export_index(mq_boundary_t b, int32_t* bra_charges, uint32_t* bra_sizes, int32_t* ket_charges,
                             uint32_t* ket_sizes, uint32_t* cohort_lr, int64_t* offsets, double* conj_scales,
                             uint8_t* transposes, uint8_t* tr) {
    if (!b) return MQ_ERR_INVALID;
    BoundaryIndex const& B = b->idx;
    for (size_t i = 0; i < B.bra.size(); ++i) {
        if (bra_charges) { bra_charges[2 * i] = B.bra.charges[i].q[0]; bra_charges[2 * i + 1] = B.bra.charges[i].q[1]; }
        if (bra_sizes) bra_sizes[i] = B.bra.sizes[i];
    }
    for (size_t i = 0; i < B.ket.size(); ++i) {
        if (ket_charges) { ket_charges[2 * i] = B.ket.charges[i].q[0]; ket_charges[2 * i + 1] = B.ket.charges[i].q[1]; }
        if (ket_sizes) ket_sizes[i] = B.ket.sizes[i];
    }
    for (uint32_t c = 0; c < B.n_cohorts(); ++c) {
        if (cohort_lr) { cohort_lr[2 * c] = B.clr[c].first; cohort_lr[2 * c + 1] = B.clr[c].second; }
        for (uint32_t k = 0; k < B.aux_dim; ++k) {
            if (offsets) offsets[uint64_t(c) * B.aux_dim + k] = B.offsets[c][k];
            if (conj_scales) conj_scales[uint64_t(c) * B.aux_dim + k] = B.conj_scales[c][k];
            if (transposes) transposes[uint64_t(c) * B.aux_dim + k] = B.transposes[c][k];
        }
        if (tr) tr[c] = B.tr[c];
    }
    return MQ_OK;
}

int mq_boundary_download(mq_boundary_t b, double* data) {
    if (!b) return MQ_ERR_INVALID;
    return guarded(b->ctx, [&] {
        require(data || !b->n, "mq_boundary_download: null data");
        hipcheck(hipStreamSynchronize(b->ctx->stream), "sync");
        if (b->n) hipcheck(hipMemcpy(data, b->data, b->n * sizeof(double), hipMemcpyDeviceToHost), "download");
    });
}

int mq_boundary_destroy(mq_boundary_t b) {
    if (!b) return MQ_OK;
    int rc = guarded(b->ctx, [&] {
        (void)hipStreamSynchronize(b->ctx->stream);
        b->hold.reset();
    });
    delete b;
    return rc;
}

// ---- MPS ---------------------------------------------------------------------------------------
int mq_mps_upload(mq_ctx_t ctx, const mq_index_desc* left, const mq_index_desc* phys, const mq_index_desc* right,
                  int pairing, const double* data, mq_mps_t* out) {
    return guarded(ctx, [&] {
        require(out && (pairing == MQ_RIGHT_PAIRED || pairing == MQ_LEFT_PAIRED), "mq_mps_upload: bad argument");
        auto* m = new mq_mps_s{ctx, MPSLayout(index_from(left), index_from(phys), index_from(right))};
        std::unique_ptr<mq_mps_s> guard(m);
        MPSLayout const& L = m->lay;
        require(L.elements() == L.lp_off.back(), "mq_mps_upload: inconsistent MPS sectors");
        std::vector<ReshapeSeg> segs = reshape_segments(L);
        m->n_segs = int32_t(segs.size());
        m->d_segs = dev_upload(segs, ctx->stream);
        uint64_t n = L.elements();
        m->rp = dev_alloc<double>(n);
        m->lp = dev_alloc<double>(n);
        if (n) {
            require(data != nullptr, "mq_mps_upload: null data");
            double* dst = pairing == MQ_RIGHT_PAIRED ? m->rp : m->lp;
            hipcheck(hipMemcpyAsync(dst, data, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "upload");
        }
        (pairing == MQ_RIGHT_PAIRED ? m->rp_ok : m->lp_ok) = true;
        hipcheck(hipStreamSynchronize(ctx->stream), "sync");
        *out = guard.release();
    });
}

int mq_mps_elements(mq_mps_t m, uint64_t* n) {
    if (!m || !n) return MQ_ERR_INVALID;
    *n = m->lay.elements();
    return MQ_OK;
}

int mq_mps_download(mq_mps_t m, int pairing, double* data) {
    if (!m) return MQ_ERR_INVALID;
    return guarded(m->ctx, [&] {
        require(pairing == MQ_RIGHT_PAIRED || pairing == MQ_LEFT_PAIRED, "mq_mps_download: bad pairing");
        if (pairing == MQ_RIGHT_PAIRED) ensure_rp(m); else ensure_lp(m);
        hipcheck(hipStreamSynchronize(m->ctx->stream), "sync");
        uint64_t n = m->lay.elements();
        if (n) hipcheck(hipMemcpy(data, pairing == MQ_RIGHT_PAIRED ? m->rp : m->lp, n * sizeof(double),
                                  hipMemcpyDeviceToHost), "download");
    });
}

int mq_mps_destroy(mq_mps_t m) {
    if (!m) return MQ_OK;
    int rc = guarded(m->ctx, [&] {
        (void)hipStreamSynchronize(m->ctx->stream);
        dev_free(m->rp);
        dev_free(m->lp);
        dev_free(m->d_segs);
    });
    delete m;
    return rc;
}

// ---- growth steps ------------------------------------------------------------------------------
static double seconds_since(std::chrono::steady_clock::time_point& t) {
    auto now = std::chrono::steady_clock::now();
    double s = std::chrono::duration<double>(now - t).count();
    t = now;
    return s;
}

// t0: start of the call; the schedule tables were built since then
static void run_growth(mq_ctx_t ctx, GrowPlanHost& P, double const* bnd, double const* bra, uint64_t bra_n,
                       double const* ket, uint64_t ket_n, uint64_t bnd_n, int precision, mq_boundary_t* out,
                       std::chrono::steady_clock::time_point t0) {
    mq_growth_stats_t st{};
    auto t = t0;
    uint64_t sizes[BASE_COUNT] = {0};
    sizes[BASE_T] = P.t_elems;
    sizes[BASE_S] = P.s_elems;
    sizes[BASE_BOUNDARY] = bnd_n;
    sizes[BASE_BRA] = bra_n;
    sizes[BASE_KET] = ket_n;
    sizes[BASE_NEW] = P.out_elems;
    check_gemms(P.t_gemms, sizes, P.t_rows);
    check_gemms(P.c_gemms, sizes);
    check_lines(P.lines, P.terms, P.t_elems, P.s_elems);
    st.seconds_tables = seconds_since(t);
    SBlockTable SB = build_sblocks(P.lines, P.terms, sblock_panel(ctx));
    check_sblocks(SB, P.t_elems, P.s_elems);
    st.seconds_blocks = seconds_since(t);
    st.n_lines = P.lines.size();
    st.n_terms = P.terms.size();
    st.n_sblocks = SB.blocks.size();
    st.t_elements = P.t_elems;
    st.s_elements = P.s_elems;
    st.out_elements = P.out_elems;
    st.flops_total = P.flops.flops_total;
    st.flops_gemm = P.flops.flops_gemm;

    auto* nb = new mq_boundary_s{ctx, std::move(P.out)};
    std::unique_ptr<mq_boundary_s> guard(nb);
    nb->n = P.out_elems;
    nb->data = dev_alloc<double>(nb->n);
    nb->hold = own_device(nb->data);
    double* T = dev_alloc<double>(P.t_elems);
    double* S = dev_alloc<double>(P.s_elems);
    GemmStage tg, cg;
    SAsmStage sa;
    try {
        tg.upload(P.t_gemms, P.t_rows, ctx->stream);
        cg.upload(P.c_gemms, {}, ctx->stream);
        sa.upload(SB, ctx->stream, ctx->sblock_xcd);
        st.seconds_upload = seconds_since(t);
        Bases b{};
        b.p[BASE_T] = T;
        b.p[BASE_S] = S;
        b.p[BASE_BOUNDARY] = const_cast<double*>(bnd);
        b.p[BASE_BRA] = const_cast<double*>(bra);
        b.p[BASE_KET] = const_cast<double*>(ket);
        b.p[BASE_NEW] = nb->data;
        tg.run(b, ctx->stream, GEMM_F64);
        sa.run(T, S, ctx->stream);
        cg.run(b, ctx->stream, GEMM_F64);
        hipcheck(hipStreamSynchronize(ctx->stream), "growth step");
        st.seconds_device = seconds_since(t);
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        tg.release(); cg.release(); sa.release(); dev_free(T); dev_free(S);
        throw;
    }
    tg.release(); cg.release(); sa.release(); dev_free(T); dev_free(S);
    st.seconds_total = seconds_since(t0);
    nb->grow = st;
    *out = guard.release();
}

int mq_lstep(mq_ctx_t ctx, mq_mpo_t mpo, mq_boundary_t left, mq_mps_t bra, mq_mps_t ket, int symmetric, int precision,
             mq_boundary_t* out) {
    return guarded(ctx, [&] {
        require(mpo && left && bra && ket && out, "mq_lstep: null argument");
        auto t0 = std::chrono::steady_clock::now();
        GrowPlanHost P = build_lstep_plan(ctx->sym, mpo->mpo, left->idx, bra->lay, ket->lay, symmetric != 0);
        ensure_lp(bra);
        ensure_rp(ket);
        run_growth(ctx, P, left->data, bra->lp, bra->lay.elements(), ket->rp, ket->lay.elements(), left->n,
                   precision, out, t0);
    });
}

int mq_rstep(mq_ctx_t ctx, mq_mpo_t mpo, mq_boundary_t right, mq_mps_t bra, mq_mps_t ket, int symmetric, int precision,
             mq_boundary_t* out) {
    return guarded(ctx, [&] {
        require(mpo && right && bra && ket && out, "mq_rstep: null argument");
        auto t0 = std::chrono::steady_clock::now();
        GrowPlanHost P = build_rstep_plan(ctx->sym, mpo->mpo, right->idx, bra->lay, ket->lay, symmetric != 0);
        ensure_rp(bra);
        ensure_rp(ket);
        run_growth(ctx, P, right->data, bra->rp, bra->lay.elements(), ket->rp, ket->lay.elements(), right->n,
                   precision, out, t0);
    });
}

int mq_boundary_growth_stats(mq_boundary_t b, mq_growth_stats_t* stats) {
    if (!b || !stats) return MQ_ERR_INVALID;
    *stats = b->grow;
    return MQ_OK;
}

// ---- H*psi ------------------------------------------------------------------------------------
// L panel set-up of one site problem: copy jobs from its left boundary into the plan's panels
struct LpanelPart {
    const double* data;
    uint64_t n;
    std::vector<CopyJob> jobs;
};

// GEMM tile autotuner: each GEMM stage of the plan whose histogram key the context has not seen
// is timed under every schedule candidate (gemm_policy; one warm-up launch, then the faster of
// two timed launches, HIP events on the context stream); the fastest is recorded for the key and
// a stage with a known key takes the recorded candidate.  The stage is re-uploaded with its
// schedule unless that is the default one it already has.
static void autotune_plan(mq_ctx_t ctx, mq_hpsi_plan_s* p, std::vector<std::vector<GemmDesc>> const& tgs,
                          std::vector<std::vector<GemmRow>> const& trs, std::vector<std::vector<GemmDesc>> const& cgs) {
    hipStream_t s = ctx->stream;
    double* psi_in = dev_alloc<double>(p->psi_n);
    double* psi_out = dev_alloc<double>(p->psi_n);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto t_start = std::chrono::steady_clock::now();
    try {
        hipcheck(launch_fill(psi_in, p->psi_n, 1.0 / std::sqrt(double(std::max<uint64_t>(p->psi_n, 1))), s), "tune");
        hipcheck(hipEventCreate(&e0), "event");
        hipcheck(hipEventCreate(&e1), "event");
        Bases b{};
        b.p[BASE_PSI_IN] = psi_in;
        b.p[BASE_PSI_OUT] = psi_out;
        b.p[BASE_T] = p->T;
        b.p[BASE_S] = p->S;
        b.p[BASE_LPANEL] = p->Lpanel;
        b.p[BASE_BOUNDARY] = const_cast<double*>(p->rdata);
        auto time_stage = [&](GemmStage const& st) {
            float best = 1e30f;
            st.run(b, s, GEMM_F64);
            for (int r = 0; r < 2; ++r) {
                hipcheck(hipEventRecord(e0, s), "event");
                st.run(b, s, GEMM_F64);
                hipcheck(hipEventRecord(e1, s), "event");
                hipcheck(hipEventSynchronize(e1), "event");
                float ms = 0;
                hipcheck(hipEventElapsedTime(&ms, e0, e1), "event");
                best = std::min(best, ms);
            }
            return best;
        };
        for (size_t c = 0; c < p->chunks.size(); ++c)
            for (int kind = 0; kind < 2; ++kind) {
                std::vector<GemmDesc> const& descs = kind == 0 ? tgs[c] : cgs[c];
                static const std::vector<GemmRow> no_rows;
                std::vector<GemmRow> const& rows = kind == 0 ? trs[c] : no_rows;
                GemmStage& stage = kind == 0 ? p->chunks[c].tg : p->chunks[c].cg;
                if (descs.empty()) continue;
                const auto key = std::make_pair(gemm_stage_key(descs), kind);
                auto it = ctx->tuned.find(key);
                if (it == ctx->tuned.end()) {
                    mq_ctx_s::Tuned t{0, 1e30f, 0.0f};
                    for (int cand = 0; cand < GEMM_N_POLICIES; ++cand) {
                        GemmSchedule sch = build_gemm_schedule(descs, gemm_policy(cand));
                        GemmStage tmp;
                        tmp.upload(descs, rows, s, &sch);
                        float ms = 0;
                        try {
                            ms = time_stage(tmp);
                        } catch (...) {
                            (void)hipStreamSynchronize(s);
                            tmp.release();
                            throw;
                        }
                        tmp.release();
                        if (cand == 0) t.default_ms = ms;
                        if (ms < t.best_ms) t.best_ms = ms, t.pick = cand;
                    }
                    it = ctx->tuned.emplace(key, t).first;
                }
                const int pick = it->second.pick;
                if (c == 0) (kind == 0 ? p->policy_t : p->policy_c) = pick;
                if (pick == 0) continue;
                GemmSchedule sch = build_gemm_schedule(descs, gemm_policy(pick));
                (void)hipStreamSynchronize(s);
                stage.release();
                double* ws = kind == 1 && p->S2 ? p->gemm_ws2 : p->gemm_ws;
                const uint64_t ws_n = kind == 1 && p->S2 ? p->gemm_ws2_elems : p->gemm_ws_elems;
                stage.upload(descs, rows, s, &sch, sch.ws_elems <= ws_n ? ws : nullptr);
            }
        p->n_reduces = 0;
        for (auto const& pc : p->chunks) p->n_reduces += pc.tg.n_reduces + pc.cg.n_reduces;
    } catch (...) {
        (void)hipStreamSynchronize(s);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        dev_free(psi_in);
        dev_free(psi_out);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dev_free(psi_in);
    dev_free(psi_out);
    p->seconds_tune = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
}

// device side of an H*psi plan whose host tables are in p->host (p->psi_n, p->rdata set): checks,
// term-sum blocks, buffers, uploads, L panels; frees the host term tables
static void finish_plan(mq_ctx_t ctx, mq_hpsi_plan_s* p, std::vector<LpanelPart> const& lparts, uint64_t r_n,
                        std::vector<size_t> const& line_ranges, std::chrono::steady_clock::time_point t0) {
    auto seconds_since = [](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
    };
    HpsiPlanHost& H = p->host;
    if (H.chunks.empty()) {  // batch plans: one chunk over the merged tables
        HpsiChunk c;
        c.t_gemm1 = H.t_gemms.size();
        c.t_row1 = H.t_rows.size();
        c.line1 = H.lines.size();
        c.c_gemm1 = H.c_gemms.size();
        c.t_elems = H.t_elems;
        c.s_elems = H.s_elems;
        H.chunks.push_back(c);
        H.t_buf = H.t_elems;
        H.s_buf = H.s_elems;
    }
    uint64_t sizes[BASE_COUNT] = {0};
    sizes[BASE_PSI_IN] = p->psi_n;
    sizes[BASE_PSI_OUT] = p->psi_n;
    sizes[BASE_LPANEL] = H.lpanel_elems;
    sizes[BASE_BOUNDARY] = r_n;
    // small-tile lines, decided per site problem (line_ranges: the first line of each instance and
    // the end)
    for (size_t i = 0; i + 1 < line_ranges.size(); ++i) {
        const size_t l0 = line_ranges[i], l1 = line_ranges[i + 1];
        mark_small_lines(H.lines, l0, l1, ctx->small_tile_max == MQ_SMALL_TILE_AUTO
                                              ? auto_small_tile_max(H.lines, l0, l1) : ctx->small_tile_max);
    }
    // per chunk: its GEMM problems (T-GEMM rows re-based on the chunk's row table), checked against
    // the chunk's T and S sizes, their schedules, and the term-sum block tables of its lines
    const size_t nch = H.chunks.size();
    std::vector<std::vector<GemmDesc>> tgs(nch), cgs(nch);
    std::vector<std::vector<GemmRow>> trs(nch);
    std::vector<GemmSchedule> tsch(nch), csch(nch);
    std::vector<SBlockTable> sbs(nch);
    double t_blocks = 0;
    uint64_t ws_max = 0, ws_c_max = 0;
    for (size_t c = 0; c < nch; ++c) {
        HpsiChunk const& ch = H.chunks[c];
        sizes[BASE_T] = ch.t_elems;
        sizes[BASE_S] = ch.s_elems;
        tgs[c].assign(H.t_gemms.begin() + int64_t(ch.t_gemm0), H.t_gemms.begin() + int64_t(ch.t_gemm1));
        trs[c].assign(H.t_rows.begin() + int64_t(ch.t_row0), H.t_rows.begin() + int64_t(ch.t_row1));
        for (auto& g : tgs[c])
            if (g.flags & GEMM_ROWMAP) g.a -= int64_t(ch.t_row0);
        cgs[c].assign(H.c_gemms.begin() + int64_t(ch.c_gemm0), H.c_gemms.begin() + int64_t(ch.c_gemm1));
        check_gemms(tgs[c], sizes, trs[c]);
        check_gemms(cgs[c], sizes);
        HostTable<SLine> cl(H.lines.begin() + int64_t(ch.line0), H.lines.begin() + int64_t(ch.line1));
        check_lines(cl, H.terms, ch.t_elems, ch.s_elems);
        auto tb = std::chrono::steady_clock::now();
        sbs[c] = build_sblocks(cl, H.terms, sblock_panel(ctx));
        check_sblocks(sbs[c], ch.t_elems, ch.s_elems);
        t_blocks += seconds_since(tb);
        tsch[c] = build_gemm_schedule(tgs[c]);
        csch[c] = build_gemm_schedule(cgs[c]);
        ws_max = std::max({ws_max, tsch[c].ws_elems, csch[c].ws_elems});
        ws_c_max = std::max(ws_c_max, csch[c].ws_elems);
    }
    p->seconds_blocks = t_blocks;
    p->seconds_tables = seconds_since(t0) - t_blocks;
    for (auto const& part : lparts)
        for (auto const& j : part.jobs)
            require(j.src >= 0 && j.dst >= 0 && uint64_t(j.src) + uint64_t(j.rows) * j.cols * j.nblocks <= part.n &&
                        uint64_t(j.dst) + uint64_t(j.rows) * j.cols * j.nblocks <= H.lpanel_elems,
                    "L panel job outside its buffers");
    p->r_elems = r_n;
    p->T = dev_alloc<double>(H.t_buf);
    p->S = dev_alloc<double>(H.s_buf);
    p->Lpanel = dev_alloc<double>(H.lpanel_elems);
    p->gemm_ws = dev_alloc<double>(ws_max);
    p->gemm_ws_elems = ws_max;
    if (ctx->overlap && nch > 1) {
        // the second S buffer, workspace and stream; a device without room for them runs the
        // chunks one after the other
        try {
            p->S2 = dev_alloc<double>(H.s_buf);
            p->gemm_ws2 = dev_alloc<double>(std::max<uint64_t>(ws_c_max, 1));
            p->gemm_ws2_elems = ws_c_max;
        } catch (Error const&) {
            dev_free(p->S2);
            dev_free(p->gemm_ws2);
            p->S2 = p->gemm_ws2 = nullptr;
            p->gemm_ws2_elems = 0;
        }
        if (p->S2) {
            if (!ctx->stream2) hipcheck(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking), "stream create");
            p->ev_sa.assign(nch, nullptr);
            p->ev_cg.assign(nch, nullptr);
            hipcheck(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming), "event");
            for (size_t c = 0; c < nch; ++c) {
                hipcheck(hipEventCreateWithFlags(&p->ev_sa[c], hipEventDisableTiming), "event");
                hipcheck(hipEventCreateWithFlags(&p->ev_cg[c], hipEventDisableTiming), "event");
            }
        }
    }
    p->chunks.resize(nch);
    for (size_t c = 0; c < nch; ++c) {
        auto& pc = p->chunks[c];
        pc.tg.upload(tgs[c], trs[c], ctx->stream, &tsch[c], p->gemm_ws);
        pc.cg.upload(cgs[c], {}, ctx->stream, &csch[c], p->S2 ? p->gemm_ws2 : p->gemm_ws);
        pc.sa.upload(sbs[c], ctx->stream, ctx->sblock_xcd);
        p->n_sblocks += int64_t(sbs[c].blocks.size());
        p->union_elements += sbs[c].gathered;
        p->max_union_tiles = std::max<uint64_t>(p->max_union_tiles, sbs[c].max_union);
        p->sblock_macs += sbs[c].mac;
        p->small_lines += sbs[c].small_lines.size();
        p->small_terms += sbs[c].small_terms.size();
        p->small_gathered += sbs[c].small_gathered;
        p->n_reduces += pc.tg.n_reduces + pc.cg.n_reduces;
        sbs[c] = SBlockTable();
    }
    p->n_lines = int64_t(H.lines.size());
    if (ctx->plain_tables && p->n_instances == 1) {
        auto tp = std::chrono::steady_clock::now();
        for (HpsiChunk const& ch : H.chunks) {
            sizes[BASE_T] = ch.t_elems;
            sizes[BASE_S] = ch.s_elems;
            std::vector<GemmDesc> t(H.t_gemms_task.begin() + int64_t(ch.t_task0),
                                    H.t_gemms_task.begin() + int64_t(ch.t_task1));
            std::vector<GemmDesc> c(H.c_gemms_cohort.begin() + int64_t(ch.c_coh0),
                                    H.c_gemms_cohort.begin() + int64_t(ch.c_coh1));
            check_gemms(t, sizes);
            check_gemms(c, sizes);
            p->plain.ranges.push_back({ch.t_task0, ch.t_task1, ch.c_coh0, ch.c_coh1, ch.line0, ch.line1});
        }
        p->plain.t = H.t_gemms_task;
        p->plain.c = H.c_gemms_cohort;
        p->plain.lines = dev_upload(H.lines, ctx->stream);
        p->plain.terms = dev_upload(H.terms, ctx->stream);
        p->plain.n_lines = int64_t(H.lines.size());
        p->plain.ok = true;
        p->seconds_plain = seconds_since(tp);
    }
    if (ctx->term_sources && p->n_instances == 1) {  // batch plans keep no term sources
        require(H.term_src.size() == H.terms.size() && H.term_beta.size() == H.terms.size(),
                "MPO-term adjoint: the plan's term sources do not cover its terms");
        auto& A = p->adj;
        A.yg.resize(nch);
        for (size_t c = 0; c < nch; ++c) {
            HpsiChunk const& ch = H.chunks[c];
            sizes[BASE_T] = ch.t_elems;
            sizes[BASE_S] = ch.s_elems;
            std::vector<GemmDesc> ys = adjoint_y_gemms(cgs[c]);
            check_gemms(ys, sizes);
            A.yg[c].upload(ys, {}, ctx->stream);
            std::vector<TermGradWork> w = term_grad_work(H.lines, ch.line0, ch.line1);
            A.work.push_back(dev_upload(w, ctx->stream));
            A.n_work.push_back(int64_t(w.size()));
        }
        for (size_t k = 0; k < H.terms.size(); ++k)
            require(H.term_src[k] < p->adj.n_mpo_terms, "MPO-term adjoint: term source outside the MPO");
        A.lines = dev_upload(H.lines, ctx->stream);
        A.terms = dev_upload(H.terms, ctx->stream);
        A.src = dev_upload(H.term_src, ctx->stream);
        A.beta = dev_upload(H.term_beta, ctx->stream);
        A.grad = dev_alloc<double>(std::max<uint32_t>(A.n_mpo_terms, 1));
        A.ok = true;
        H.term_src.clear(); H.term_src.shrink_to_fit();
        H.term_beta.clear(); H.term_beta.shrink_to_fit();
    }
    std::vector<GemmDesc>().swap(H.t_gemms_task);
    std::vector<GemmDesc>().swap(H.c_gemms_cohort);
    for (auto const& part : lparts) {
        CopyJob* jobs = dev_upload(part.jobs, ctx->stream);
        hipError_t e = launch_copy_jobs(jobs, int32_t(part.jobs.size()), part.data, p->Lpanel, ctx->stream);
        hipError_t e2 = hipStreamSynchronize(ctx->stream);
        dev_free(jobs);
        hipcheck(e, "L panel copy");
        hipcheck(e2, "L panel copy");
    }
    if (ctx->autotune) autotune_plan(ctx, p, tgs, trs, cgs);
    // the host tables are on the device now; keep only what the statistics need
    for (SLine const& l : H.lines) {
        uint64_t e = uint64_t(l.rows) * uint64_t(l.cols);
        p->s_line_elements += e;
        p->gathered_elements += e * uint64_t(l.t1 - l.t0);
    }
    H.terms.clear(); H.terms.shrink_to_fit();
    H.lines.clear(); H.lines.shrink_to_fit();
    H.lpanel_jobs.clear(); H.lpanel_jobs.shrink_to_fit();
    p->build_seconds = seconds_since(t0);
}

int mq_hpsi_plan_create(mq_ctx_t ctx, mq_mpo_t mpo, mq_boundary_t left, mq_boundary_t right,
                        const mq_index_desc* left_i, const mq_index_desc* phys_i, const mq_index_desc* right_i,
                        mq_hpsi_plan_t* out) {
    return guarded(ctx, [&] {
        require(mpo && left && right && out, "mq_hpsi_plan_create: null argument");
        auto t0 = std::chrono::steady_clock::now();
        // largest T + S of one chunk (elements)
        const int64_t mb = ctx->chunk_mb == MQ_CHUNK_AUTO ? 49152 : ctx->chunk_mb;
        const uint64_t chunk_elems = uint64_t(mb) * (uint64_t(1) << 20) / sizeof(double);
        auto* p = new mq_hpsi_plan_s{ctx, build_hpsi_plan(ctx->sym, mpo->mpo, left->idx, right->idx, index_from(left_i),
                                                          index_from(phys_i), index_from(right_i), chunk_elems,
                                                          ctx->term_sources)};
        std::unique_ptr<mq_hpsi_plan_s> guard(p);
        p->adj.n_mpo_terms = uint32_t(mpo->mpo.terms.size());
        p->psi_n = p->host.mps.elements();
        p->rdata = right->data;
        p->rhold = right->hold;
        finish_plan(ctx, p, {LpanelPart{left->data, left->n, p->host.lpanel_jobs}}, right->n,
                    {0, p->host.lines.size()}, t0);
        *out = guard.release();
    });
}

// n independent site problems as one plan: the host tables of each are built as for one problem,
// then concatenated with their buffer offsets shifted (psi, T, S, L panels, right boundaries, GEMM
// row table, term list), so every stage of the batch is one grouped launch
int mq_hpsi_plan_create_batch(mq_ctx_t ctx, int n, const mq_mpo_t* mpos, const mq_boundary_t* lefts,
                              const mq_boundary_t* rights, const mq_index_desc* left_is,
                              const mq_index_desc* phys_is, const mq_index_desc* right_is, mq_hpsi_plan_t* out) {
    return guarded(ctx, [&] {
        require(n >= 1 && n <= 65535 && mpos && lefts && rights && left_is && phys_is && right_is && out,
                "mq_hpsi_plan_create_batch: null argument or more than 65535 instances");
        for (int i = 0; i < n; ++i)
            require(mpos[i] && lefts[i] && rights[i], "mq_hpsi_plan_create_batch: null instance");
        auto t0 = std::chrono::steady_clock::now();
        auto* p = new mq_hpsi_plan_s{ctx, HpsiPlanHost{}};
        std::unique_ptr<mq_hpsi_plan_s> guard(p);
        HpsiPlanHost& H = p->host;
        // every instance's tables (each build runs on all host cores), then their offsets in the
        // merged tables, then the merged tables written in parallel with the offsets applied
        std::vector<HpsiPlanHost> parts(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i)
            parts[size_t(i)] = build_hpsi_plan(ctx->sym, mpos[i]->mpo, lefts[i]->idx, rights[i]->idx,
                                               index_from(&left_is[i]), index_from(&phys_is[i]),
                                               index_from(&right_is[i]));
        H.mps = parts[0].mps;
        const size_t un = size_t(n);
        std::vector<std::array<int64_t, BASE_COUNT>> shift(un);
        std::vector<int64_t> o_tg(un + 1, 0), o_rows(un + 1, 0), o_cg(un + 1, 0), o_lines(un + 1, 0), o_terms(un + 1, 0);
        uint64_t r_n = 0;
        for (size_t i = 0; i < un; ++i) {
            HpsiPlanHost const& Q = parts[i];
            shift[i].fill(0);
            shift[i][BASE_PSI_IN] = shift[i][BASE_PSI_OUT] = int64_t(p->psi_n);
            shift[i][BASE_T] = int64_t(H.t_elems);
            shift[i][BASE_S] = int64_t(H.s_elems);
            shift[i][BASE_LPANEL] = int64_t(H.lpanel_elems);
            shift[i][BASE_BOUNDARY] = int64_t(r_n);
            o_tg[i + 1] = o_tg[i] + int64_t(Q.t_gemms.size());
            o_rows[i + 1] = o_rows[i] + int64_t(Q.t_rows.size());
            o_cg[i + 1] = o_cg[i] + int64_t(Q.c_gemms.size());
            o_lines[i + 1] = o_lines[i] + int64_t(Q.lines.size());
            o_terms[i + 1] = o_terms[i] + int64_t(Q.terms.size());
            p->inst_off.push_back(int64_t(p->psi_n));
            H.t_elems += Q.t_elems;
            H.s_elems += Q.s_elems;
            H.lpanel_elems += Q.lpanel_elems;
            H.n_cohorts += Q.n_cohorts;
            H.n_t_tasks += Q.n_t_tasks;
            H.bytes_compulsory += Q.bytes_compulsory;
            H.flops.flops_total += Q.flops.flops_total;
            H.flops.flops_gemm += Q.flops.flops_gemm;
            H.flops.flops_gemm_exec += Q.flops.flops_gemm_exec;
            H.flops.n_terms += Q.flops.n_terms;
            p->psi_n += Q.mps.elements();
            r_n += rights[i]->n;
        }
        require(o_terms[un] < int64_t(INT32_MAX), "batch: more than 2^31 terms");
        H.t_gemms.resize(size_t(o_tg[un]));
        H.t_rows.resize(size_t(o_rows[un]));
        H.c_gemms.resize(size_t(o_cg[un]));
        H.lines.resize(size_t(o_lines[un]));
        H.terms.resize(size_t(o_terms[un]));
        std::vector<LpanelPart> lparts;
        std::vector<size_t> line_ranges;
        for (size_t i = 0; i < un; ++i) {
            HpsiPlanHost& Q = parts[i];
            auto const& sh = shift[i];
            const int64_t row0 = o_rows[i];
            const int32_t term0 = int32_t(o_terms[i]);
            for (size_t k = 0; k < Q.t_gemms.size(); ++k) {
                GemmDesc g = Q.t_gemms[k];
                if (g.flags & GEMM_ROWMAP) g.a += row0;
                else g.a += sh[g.base_a];
                g.b += sh[g.base_b];
                if (!(g.flags & GEMM_ROWMAP)) g.c += sh[g.base_c];
                H.t_gemms[size_t(o_tg[i]) + k] = g;
            }
            for (size_t k = 0; k < Q.c_gemms.size(); ++k) {
                GemmDesc g = Q.c_gemms[k];
                g.a += sh[g.base_a];
                g.b += sh[g.base_b];
                g.c += sh[g.base_c];
                H.c_gemms[size_t(o_cg[i]) + k] = g;
            }
            const int64_t nr = int64_t(Q.t_rows.size()), nl = int64_t(Q.lines.size()), nt = int64_t(Q.terms.size());
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < nr; ++k) {
                GemmRow r = Q.t_rows[size_t(k)];
                r.a += sh[BASE_PSI_IN];
                r.c += sh[BASE_T];
                H.t_rows[size_t(row0 + k)] = r;
            }
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < nl; ++k) {
                SLine l = Q.lines[size_t(k)];
                l.s_off += sh[BASE_S];
                l.t0 += term0;
                l.t1 += term0;
                H.lines[size_t(o_lines[i] + k)] = l;
            }
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < nt; ++k) {
                STerm t = Q.terms[size_t(k)];
                t.t_off += sh[BASE_T];
                H.terms[size_t(o_terms[i] + k)] = t;
            }
            LpanelPart lp{lefts[i]->data, lefts[i]->n, Q.lpanel_jobs};
            for (auto& j : lp.jobs) j.dst += sh[BASE_LPANEL];
            lparts.push_back(std::move(lp));
            line_ranges.push_back(size_t(o_lines[i]));
            Q = HpsiPlanHost{};  // merged: release the instance's tables
        }
        line_ranges.push_back(size_t(o_lines[un]));
        p->inst_off.push_back(int64_t(p->psi_n));
        p->n_instances = n;
        p->d_inst_off = dev_upload(p->inst_off, ctx->stream);
        // the right boundaries of all instances in one device buffer owned by the plan
        p->rown = dev_alloc<double>(r_n);
        uint64_t roff = 0;
        for (int i = 0; i < n; ++i) {
            if (rights[i]->n)
                hipcheck(hipMemcpyAsync(p->rown + roff, rights[i]->data, rights[i]->n * sizeof(double),
                                        hipMemcpyDeviceToDevice, ctx->stream), "batch right boundaries");
            roff += rights[i]->n;
        }
        hipcheck(hipStreamSynchronize(ctx->stream), "batch right boundaries");
        p->rdata = p->rown;
        finish_plan(ctx, p, lparts, r_n, line_ranges, t0);
        *out = guard.release();
    });
}

int mq_hpsi_plan_stats(mq_hpsi_plan_t p, mq_hpsi_stats_t* s) {
    if (!p || !s) return MQ_ERR_INVALID;
    HpsiPlanHost const& H = p->host;
    s->psi_elements = p->psi_n;
    s->n_t_tasks = H.n_t_tasks;
    s->n_cohorts = H.n_cohorts;
    s->n_lines = uint64_t(p->n_lines);
    s->n_terms = H.flops.n_terms;
    s->flops_total = H.flops.flops_total;
    s->flops_gemm = H.flops.flops_gemm;
    s->flops_gemm_exec = H.flops.flops_gemm_exec;
    s->bytes_compulsory = H.bytes_compulsory;
    s->t_elements = H.t_elems;
    s->s_elements = H.s_elems;
    s->l_panel_elements = H.lpanel_elems;
    s->r_tr_elements = 0;
    s->build_seconds = p->build_seconds;
    s->s_line_elements = p->s_line_elements;
    s->gathered_elements = p->gathered_elements;
    s->n_sblocks = uint64_t(p->n_sblocks);
    s->seconds_tables = p->seconds_tables;
    s->seconds_blocks = p->seconds_blocks;
    s->union_elements = p->union_elements;
    s->seconds_plain = p->seconds_plain;
    s->small_lines = p->small_lines;
    s->small_terms = p->small_terms;
    s->small_gathered = p->small_gathered;
    s->max_union_tiles = p->max_union_tiles;
    s->sblock_macs = p->sblock_macs;
    s->split_reduces = uint64_t(p->n_reduces);
    s->n_chunks = uint64_t(p->chunks.size());
    s->ts_buffer_elements = H.t_buf + H.s_buf;
    s->gemm_policy_t = p->policy_t;
    s->gemm_policy_c = p->policy_c;
    s->seconds_tune = p->seconds_tune;
    return MQ_OK;
}

int mq_hpsi_plan_psi_elements(mq_hpsi_plan_t p, uint64_t* n) {
    if (!p || !n) return MQ_ERR_INVALID;
    *n = p->psi_n;
    return MQ_OK;
}

// FP32 copies of the plan's L panels and right boundary plus FP32 T / S / psi buffers
static void ensure_fp32(mq_hpsi_plan_t p) {
    if (p->T32) return;
    hipStream_t s = p->ctx->stream;
    HpsiPlanHost const& H = p->host;
    // FP32 T and S live in the FP64 buffers: the H*psi of both precisions run on one stream, one
    // after the other
    p->S32 = reinterpret_cast<float*>(p->S);
    p->S32b = reinterpret_cast<float*>(p->S2);
    p->Lpanel32 = dev_alloc<float>(H.lpanel_elems);
    p->R32 = dev_alloc<float>(p->r_elems);
    p->psi32 = dev_alloc<float>(p->psi_n);
    hipcheck(launch_convert(p->Lpanel, p->Lpanel32, H.lpanel_elems, s), "convert");
    hipcheck(launch_convert(p->rdata, p->R32, p->r_elems, s), "convert");
    for (auto& c : p->chunks) c.sa.ensure32(s);
    p->T32 = reinterpret_cast<float*>(p->T);
}

// H*psi on the context stream; ev (optional, 5 events) brackets the stages set-up, T-GEMMs,
// term sum and output GEMMs.  FP32: psi, L panels and R are rounded to FP32, the three stages
// run on the FP32 matrix cores with FP32 T and S, and psi' is written in FP64.
// One rocBLAS dgemm (or strided-batched dgemm) per problem on the context stream: the launch
// structure of the reference's per-task cuBLAS calls (tasks.cpp:286-344, 650-752).
// the context's library handle (made on first use) on the context stream
static rocblas_handle blas_handle(mq_ctx_t ctx) {
    if (!ctx->blas) {
        require(rocblas_create_handle(&ctx->blas) == rocblas_status_success, "rocblas_create_handle failed");
        require(rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host) == rocblas_status_success,
                "rocblas_set_pointer_mode failed");
    }
    require(rocblas_set_stream(ctx->blas, ctx->stream) == rocblas_status_success, "rocblas_set_stream failed");
    return ctx->blas;
}

static void library_gemms(mq_ctx_t ctx, const GemmDesc* first, const GemmDesc* last, Bases const& b,
                          double beta = 0.0) {
    blas_handle(ctx);
    const double one = 1.0;
    for (const GemmDesc* gp = first; gp != last; ++gp) {
        GemmDesc const& g = *gp;
        if (g.M <= 0 || g.N <= 0) continue;
        const double* A = static_cast<const double*>(b.p[g.base_a]) + g.a;
        const double* B = static_cast<const double*>(b.p[g.base_b]) + g.b;
        double* C = static_cast<double*>(b.p[g.base_c]) + g.c;
        const rocblas_operation oa = (g.flags & GEMM_TRANS_A) ? rocblas_operation_transpose : rocblas_operation_none;
        const rocblas_operation ob = (g.flags & GEMM_TRANS_B) ? rocblas_operation_transpose : rocblas_operation_none;
        rocblas_status st = g.batch == 1
            ? rocblas_dgemm(ctx->blas, oa, ob, g.M, g.N, g.K, &one, A, g.lda, B, g.ldb, &beta, C, g.ldc)
            : rocblas_dgemm_strided_batched(ctx->blas, oa, ob, g.M, g.N, g.K, &one, A, g.lda, g.sa, B, g.ldb, g.sb,
                                            &beta, C, g.ldc, g.sc, g.batch);
        require(st == rocblas_status_success, "rocblas dgemm failed");
    }
}

// The chunks of an H*psi with their output GEMMs on the second stream: chunk c's term sum writes
// the S buffer of parity c % 2 once chunk c - 2's output GEMMs (on the second stream) have read
// it; chunk c's output GEMMs start when its term sum is done.  The T buffer is written and read on
// the context stream only.  Both streams start after the work already on the context stream, and
// the context stream waits for the last output GEMMs, so callers (and graph capture) see one
// stream.
static void hpsi_overlapped(mq_hpsi_plan_t p, Bases b, bool fp32, bool bf3) {
    hipStream_t s = p->ctx->stream, s2 = p->ctx->stream2;
    const size_t n = p->chunks.size();
    hipcheck(hipEventRecord(p->ev_fork, s), "event");
    hipcheck(hipStreamWaitEvent(s2, p->ev_fork, 0), "wait");
    for (size_t c = 0; c < n; ++c) {
        auto& ch = p->chunks[c];
        const bool odd = c % 2 == 1;
        if (c >= 2) hipcheck(hipStreamWaitEvent(s, p->ev_cg[c - 2], 0), "wait");
        if (fp32) {
            b.p[BASE_S] = odd ? static_cast<void*>(p->S32b) : static_cast<void*>(p->S32);
            ch.tg.run(b, s, bf3 ? GEMM_BF3 : GEMM_F32);
            ch.sa.run32(p->T32, odd ? p->S32b : p->S32, s);
        } else {
            b.p[BASE_S] = odd ? p->S2 : p->S;
            ch.tg.run(b, s, GEMM_F64);
            ch.sa.run(p->T, odd ? p->S2 : p->S, s);
        }
        hipcheck(hipEventRecord(p->ev_sa[c], s), "event");
        hipcheck(hipStreamWaitEvent(s2, p->ev_sa[c], 0), "wait");
        ch.cg.run(b, s2, fp32 ? (bf3 ? GEMM_BF3_OUT_F64 : GEMM_F32_OUT_F64) : GEMM_F64);
        hipcheck(hipEventRecord(p->ev_cg[c], s2), "event");
    }
    hipcheck(hipStreamWaitEvent(s, p->ev_cg[n - 1], 0), "wait");
}

// ev (optional): 2 + 3 * chunks events: start, set-up done, then per chunk its T-GEMMs, term sum
// and output GEMMs done
static void hpsi_enqueue(mq_hpsi_plan_t p, const double* in, double* out, int precision, hipEvent_t* ev = nullptr) {
    hipStream_t s = p->ctx->stream;
    auto mark = [&](int i) {
        if (ev) hipcheck(hipEventRecord(ev[i], s), "event");
    };
    Bases b{};
    b.p[BASE_PSI_OUT] = out;
    if (precision == MQ_PREC_FP64) {
        b.p[BASE_PSI_IN] = const_cast<double*>(in);
        b.p[BASE_T] = p->T;
        b.p[BASE_S] = p->S;
        b.p[BASE_LPANEL] = p->Lpanel;
        b.p[BASE_BOUNDARY] = const_cast<double*>(p->rdata);
        mark(0);
        mark(1);
        if (!ev && p->S2) {
            hpsi_overlapped(p, b, false, false);
            return;
        }
        for (size_t c = 0; c < p->chunks.size(); ++c) {
            p->chunks[c].tg.run(b, s, GEMM_F64);
            mark(int(2 + 3 * c));
            p->chunks[c].sa.run(p->T, p->S, s);
            mark(int(3 + 3 * c));
            p->chunks[c].cg.run(b, s, GEMM_F64);
            mark(int(4 + 3 * c));
        }
    } else if (precision == MQ_PREC_FP32 || precision == MQ_PREC_BF16X3) {
        // BF16X3: the FP32 operands and term sum, with split-BF16 products in the GEMM stages
        const bool bf3 = precision == MQ_PREC_BF16X3;
        ensure_fp32(p);
        b.p[BASE_PSI_IN] = p->psi32;
        b.p[BASE_T] = p->T32;
        b.p[BASE_S] = p->S32;
        b.p[BASE_LPANEL] = p->Lpanel32;
        b.p[BASE_BOUNDARY] = p->R32;
        mark(0);
        hipcheck(launch_convert(in, p->psi32, p->psi_n, s), "convert");
        mark(1);
        if (!ev && p->S2) {
            hpsi_overlapped(p, b, true, bf3);
            return;
        }
        for (size_t c = 0; c < p->chunks.size(); ++c) {
            p->chunks[c].tg.run(b, s, bf3 ? GEMM_BF3 : GEMM_F32);
            mark(int(2 + 3 * c));
            p->chunks[c].sa.run32(p->T32, p->S32, s);
            mark(int(3 + 3 * c));
            p->chunks[c].cg.run(b, s, bf3 ? GEMM_BF3_OUT_F64 : GEMM_F32_OUT_F64);
            mark(int(4 + 3 * c));
        }
    } else if (precision == MQ_PREC_FP64_PLAIN) {
        require(p->plain.ok, "H*psi MQ_PREC_FP64_PLAIN: the plan was made without MQ_OPT_PLAIN_TABLES or is a batch plan");
        b.p[BASE_PSI_IN] = const_cast<double*>(in);
        b.p[BASE_T] = p->T;
        b.p[BASE_S] = p->S;
        b.p[BASE_LPANEL] = p->Lpanel;
        b.p[BASE_BOUNDARY] = const_cast<double*>(p->rdata);
        mark(0);
        // psi' = 0, then every cohort's product is added in stream order (the reference's mutex /
        // atomic accumulation into the output block)
        hipcheck(hipMemsetAsync(out, 0, p->psi_n * sizeof(double), s), "psi' clear");
        mark(1);
        for (size_t c = 0; c < p->plain.ranges.size(); ++c) {
            auto const& r = p->plain.ranges[c];
            library_gemms(p->ctx, p->plain.t.data() + r.t0, p->plain.t.data() + r.t1, b);
            mark(int(2 + 3 * c));
            hipcheck(launch_slines(p->plain.lines + r.l0, int64_t(r.l1 - r.l0), p->plain.terms, p->T, p->S, s),
                     "S lines");
            mark(int(3 + 3 * c));
            library_gemms(p->ctx, p->plain.c.data() + r.c0, p->plain.c.data() + r.c1, b, 1.0);
            mark(int(4 + 3 * c));
        }
    } else {
        throw Error(MQ_ERR_UNSUPPORTED,
                    "H*psi precision: MQ_PREC_FP64, MQ_PREC_FP32, MQ_PREC_BF16X3 and MQ_PREC_FP64_PLAIN are available");
    }
}

// H*psi through a HIP graph: the stage launches are captured once per precision on plan-owned
// input / output buffers (every allocation happens before the capture, and no memset node is
// recorded); a call copies psi in, launches the graph and copies psi' out, all on the context
// stream.
static void hpsi_graph_enqueue(mq_hpsi_plan_t p, const double* in, double* out, int precision) {
    require(precision == MQ_PREC_FP64 || precision == MQ_PREC_FP32, "H*psi graph: unsupported precision");
    hipStream_t s = p->ctx->stream;
    const uint64_t n = p->psi_n;
    auto& g = p->graphs[precision == MQ_PREC_FP64 ? 0 : 1];
    if (!g.exec) {
        if (precision == MQ_PREC_FP32) ensure_fp32(p);
        if (!g.in) g.in = dev_alloc<double>(n);
        if (!g.out) g.out = dev_alloc<double>(n);
        hipcheck(hipStreamSynchronize(s), "graph set-up");
        hipcheck(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
        hipGraph_t graph = nullptr;
        try {
            hpsi_enqueue(p, g.in, g.out, precision);
        } catch (...) {
            (void)hipStreamEndCapture(s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        hipcheck(hipStreamEndCapture(s, &graph), "end capture");
        hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        hipcheck(e, "graph instantiate");
    }
    hipcheck(hipMemcpyAsync(g.in, in, n * sizeof(double), hipMemcpyDeviceToDevice, s), "graph input");
    hipcheck(hipGraphLaunch(g.exec, s), "graph launch");
    hipcheck(hipMemcpyAsync(out, g.out, n * sizeof(double), hipMemcpyDeviceToDevice, s), "graph output");
}

static void hpsi_run(mq_hpsi_plan_t p, const double* in, double* out, int precision) {
    if (p->use_graphs && (precision == MQ_PREC_FP64 || precision == MQ_PREC_FP32)) hpsi_graph_enqueue(p, in, out, precision);
    else hpsi_enqueue(p, in, out, precision);
}

int mq_hpsi(mq_hpsi_plan_t p, const double* in, double* out, int precision) {
    if (!p) return MQ_ERR_INVALID;
    return guarded(p->ctx, [&] {
        require(in && out && in != out, "mq_hpsi: input and output must be distinct device buffers");
        hpsi_run(p, in, out, precision);
    });
}

int mq_hpsi_plan_use_graphs(mq_hpsi_plan_t p, int enable) {
    if (!p) return MQ_ERR_INVALID;
    p->use_graphs = enable != 0;
    return MQ_OK;
}

int mq_hpsi_timed(mq_hpsi_plan_t p, const double* in, double* out, int precision, float* stage_ms) {
    if (!p) return MQ_ERR_INVALID;
    return guarded(p->ctx, [&] {
        require(in && out && in != out && stage_ms, "mq_hpsi_timed: bad argument");
        hipStream_t s = p->ctx->stream;
        const size_t nch = precision == MQ_PREC_FP64_PLAIN ? p->plain.ranges.size() : p->chunks.size();
        std::vector<hipEvent_t> ev(2 + 3 * nch, nullptr);
        try {
            for (auto& e : ev) hipcheck(hipEventCreate(&e), "event create");
            hpsi_enqueue(p, in, out, precision, ev.data());
            hipcheck(hipEventSynchronize(ev.back()), "event sync");
            // stage times summed over the chunks
            auto el = [&](size_t a, size_t b) {
                float ms = 0;
                hipcheck(hipEventElapsedTime(&ms, ev[a], ev[b]), "elapsed");
                return ms;
            };
            stage_ms[0] = el(0, 1);
            stage_ms[1] = stage_ms[2] = stage_ms[3] = 0;
            for (size_t c = 0; c < nch; ++c) {
                stage_ms[1] += el(c == 0 ? 1 : 1 + 3 * c, 2 + 3 * c);
                stage_ms[2] += el(2 + 3 * c, 3 + 3 * c);
                stage_ms[3] += el(3 + 3 * c, 4 + 3 * c);
            }
        } catch (...) {
            (void)hipStreamSynchronize(s);
            for (auto e : ev)
                if (e) (void)hipEventDestroy(e);
            throw;
        }
        for (auto e : ev) (void)hipEventDestroy(e);
    });
}

// Noise term of the density-matrix split: the Gram matrices of a boundary's blocks, grouped by the
// charge of their rows (side 0: X X^T per cohort, its blocks side by side) or of their columns
// (side 1: X_b^T X_b per block), accumulated in one device buffer by rocBLAS dgemm (beta = 1,
// one stream, so the accumulations are ordered).
int mq_boundary_gram(mq_boundary_t b, int side, double alpha, int n, const int32_t* charges, const uint32_t* sizes,
                     double* out) {
    if (!b) return MQ_ERR_INVALID;
    mq_ctx_t ctx = b->ctx;
    return guarded(ctx, [&] {
        require((side == 0 || side == 1) && n >= 0 && (n == 0 || (charges && sizes && out)),
                "mq_boundary_gram: bad argument");
        BoundaryIndex const& B = b->idx;
        Index const& grouped = side == 0 ? B.bra : B.ket;
        std::vector<uint64_t> off(size_t(n) + 1, 0);
        for (int k = 0; k < n; ++k) off[size_t(k) + 1] = off[size_t(k)] + uint64_t(sizes[k]) * sizes[k];
        std::vector<int> slot(grouped.size(), -1);  // output block of every sector of the grouped index
        for (int k = 0; k < n; ++k) {
            const size_t pos = grouped.position(Charge(charges[2 * k], charges[2 * k + 1]));
            require(pos < grouped.size() && grouped.sizes[pos] == sizes[k] && slot[pos] < 0,
                    "mq_boundary_gram: charge or size not in the boundary's index");
            slot[pos] = k;
        }
        for (uint32_t ci = 0; ci < B.n_cohorts(); ++ci)
            for (uint32_t x = 0; x < B.aux_dim; ++x)
                require(B.offsets[ci][x] < 0 || (!B.tr[ci] && B.transposes[ci][x] == 0 && B.conj_scales[ci][x] == 1.0),
                        "mq_boundary_gram: a block aliases its partner's (grow with symmetric = 0 and no partners)");
        double* d = dev_alloc<double>(std::max<uint64_t>(off.back(), 1));
        try {
            hipStream_t s = ctx->stream;
            hipcheck(launch_fill(d, off.back(), 0.0, s), "clear");
            blas_handle(ctx);
            const double one = 1.0;
            for (uint32_t ci = 0; ci < B.n_cohorts(); ++ci) {
                const int k = slot[side == 0 ? B.clr[ci].first : B.clr[ci].second];
                const int ls = int(B.lsz[ci]), rs = int(B.rsz[ci]);
                if (k < 0 || B.tr[ci] || B.nblk[ci] == 0 || ls == 0 || rs == 0) continue;
                const double* X = b->data + B.data_off[ci];
                double* o = d + off[size_t(k)];
                rocblas_status st = rocblas_status_success;
                // blocks stored back to back (offsets 0, ls*rs, ...): one GEMM over the cohort
                bool packed = side == 0;
                for (uint32_t x = 0, next = 0; x < B.aux_dim && packed; ++x)
                    if (B.offsets[ci][x] >= 0) packed = B.offsets[ci][x] == int64_t(next++) * ls * rs;
                if (packed) {
                    const int64_t K = int64_t(B.nblk[ci]) * rs;
                    require(K < INT32_MAX, "mq_boundary_gram: cohort too wide");
                    st = rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_transpose, ls, ls, int(K),
                                       &alpha, X, ls, X, ls, &one, o, ls);
                } else {
                    for (uint32_t x = 0; x < B.aux_dim && st == rocblas_status_success; ++x) {
                        if (B.offsets[ci][x] < 0) continue;
                        const double* Xb = X + B.offsets[ci][x];
                        st = side == 0 ? rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_transpose,
                                                       ls, ls, rs, &alpha, Xb, ls, Xb, ls, &one, o, ls)
                                       : rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none,
                                                       rs, rs, ls, &alpha, Xb, ls, Xb, ls, &one, o, rs);
                    }
                }
                require(st == rocblas_status_success, "rocblas dgemm failed");
            }
            if (off.back())
                hipcheck(hipMemcpyAsync(out, d, off.back() * sizeof(double), hipMemcpyDeviceToHost, s), "download");
            hipcheck(hipStreamSynchronize(s), "gram");
        } catch (...) {
            (void)hipStreamSynchronize(ctx->stream);
            dev_free(d);
            throw;
        }
        dev_free(d);
    });
}

// Thin SVD of every block on the device, one rocSOLVER dgesvd per block on the context stream
// (U and V^T of the min(rows, cols) singular vectors); inputs and outputs are host arrays.
int mq_block_svd(mq_ctx_t ctx, int n, const int32_t* rows, const int32_t* cols, const double* a, double* s,
                 double* u, double* vt) {
    return guarded(ctx, [&] {
        require(n >= 0 && (n == 0 || (rows && cols && a && s && u && vt)), "mq_block_svd: null argument");
        if (n == 0) return;
        std::vector<uint64_t> oa(size_t(n) + 1, 0), os(size_t(n) + 1, 0), ou(size_t(n) + 1, 0), ov(size_t(n) + 1, 0);
        int32_t kmax = 1;
        for (int i = 0; i < n; ++i) {
            require(rows[i] > 0 && cols[i] > 0, "mq_block_svd: empty block");
            const uint64_t r = uint64_t(rows[i]), c = uint64_t(cols[i]), k = std::min(r, c);
            oa[size_t(i) + 1] = oa[size_t(i)] + r * c;
            os[size_t(i) + 1] = os[size_t(i)] + k;
            ou[size_t(i) + 1] = ou[size_t(i)] + r * k;
            ov[size_t(i) + 1] = ov[size_t(i)] + k * c;
            kmax = std::max<int32_t>(kmax, int32_t(k));
        }
        hipStream_t st = ctx->stream;
        rocblas_handle h = blas_handle(ctx);
        double *dA = nullptr, *dS = nullptr, *dU = nullptr, *dV = nullptr, *dE = nullptr;
        rocblas_int* dinfo = nullptr;
        auto release = [&] {
            (void)hipStreamSynchronize(st);
            dev_free(dA); dev_free(dS); dev_free(dU); dev_free(dV); dev_free(dE); dev_free(dinfo);
        };
        try {
            dA = dev_alloc<double>(oa[size_t(n)]);
            dS = dev_alloc<double>(os[size_t(n)]);
            dU = dev_alloc<double>(ou[size_t(n)]);
            dV = dev_alloc<double>(ov[size_t(n)]);
            dE = dev_alloc<double>(uint64_t(kmax));
            dinfo = dev_alloc<rocblas_int>(uint64_t(n));
            upload_bytes(dA, a, oa[size_t(n)] * sizeof(double), st);
            for (int i = 0; i < n; ++i) {
                const rocblas_int r = rows[i], c = cols[i], k = std::min(r, c);
                require(rocsolver_dgesvd(h, rocblas_svect_singular, rocblas_svect_singular, r, c, dA + oa[size_t(i)], r,
                                         dS + os[size_t(i)], dU + ou[size_t(i)], r, dV + ov[size_t(i)], k, dE,
                                         rocblas_outofplace, dinfo + i) == rocblas_status_success,
                        "mq_block_svd: rocsolver_dgesvd failed");
            }
            std::vector<rocblas_int> info(static_cast<size_t>(n));
            hipcheck(hipMemcpyAsync(s, dS, os[size_t(n)] * sizeof(double), hipMemcpyDeviceToHost, st), "download");
            hipcheck(hipMemcpyAsync(u, dU, ou[size_t(n)] * sizeof(double), hipMemcpyDeviceToHost, st), "download");
            hipcheck(hipMemcpyAsync(vt, dV, ov[size_t(n)] * sizeof(double), hipMemcpyDeviceToHost, st), "download");
            hipcheck(hipMemcpyAsync(info.data(), dinfo, size_t(n) * sizeof(rocblas_int), hipMemcpyDeviceToHost, st),
                     "download");
            hipcheck(hipStreamSynchronize(st), "block svd");
            for (int i = 0; i < n; ++i)
                require(info[size_t(i)] == 0, "mq_block_svd: the SVD of a block did not converge");
        } catch (...) {
            release();
            throw;
        }
        release();
    });
}

int mq_hpsi_host(mq_hpsi_plan_t p, const double* in, double* out, int precision) {
    if (!p) return MQ_ERR_INVALID;
    return guarded(p->ctx, [&] {
        require(in && out, "mq_hpsi_host: null argument");
        uint64_t n = p->psi_n;
        double* din = dev_alloc<double>(n);
        double* dout = dev_alloc<double>(n);
        try {
            hipcheck(hipMemcpyAsync(din, in, n * sizeof(double), hipMemcpyHostToDevice, p->ctx->stream), "upload");
            hpsi_run(p, din, dout, precision);
            hipcheck(hipMemcpyAsync(out, dout, n * sizeof(double), hipMemcpyDeviceToHost, p->ctx->stream), "download");
            hipcheck(hipStreamSynchronize(p->ctx->stream), "hpsi");
        } catch (...) {
            (void)hipStreamSynchronize(p->ctx->stream);
            dev_free(din);
            dev_free(dout);
            throw;
        }
        dev_free(din);
        dev_free(dout);
    });
}

int mq_hpsi_term_gradient(mq_hpsi_plan_t p, const double* psi, double* grad, int64_t n_grad) {
    if (!p) return MQ_ERR_INVALID;
    return guarded(p->ctx, [&] {
        require(psi && grad, "mq_hpsi_term_gradient: null argument");
        require(p->adj.ok, "mq_hpsi_term_gradient: the plan was made without MQ_OPT_TERM_SOURCES");
        require(n_grad == int64_t(p->adj.n_mpo_terms), "mq_hpsi_term_gradient: n_grad is not the MPO's term count");
        hipStream_t s = p->ctx->stream;
        const uint64_t n = p->psi_n;
        double* din = dev_alloc<double>(n);
        try {
            hipcheck(hipMemcpyAsync(din, psi, n * sizeof(double), hipMemcpyHostToDevice, s), "upload");
            hipcheck(launch_fill(p->adj.grad, p->adj.n_mpo_terms, 0.0, s), "clear");
            Bases b{};
            b.p[BASE_PSI_IN] = din;
            b.p[BASE_T] = p->T;
            b.p[BASE_S] = p->S;
            b.p[BASE_LPANEL] = p->Lpanel;
            b.p[BASE_BOUNDARY] = const_cast<double*>(p->rdata);
            for (size_t c = 0; c < p->chunks.size(); ++c) {
                p->chunks[c].tg.run(b, s, GEMM_F64);   // T = psi * R
                p->adj.yg[c].run(b, s, GEMM_F64);      // Y = L_panel^T psi, in the S buffer
                hipcheck(launch_term_grad(p->adj.work[c], p->adj.n_work[c], p->adj.lines, p->adj.terms, p->adj.src,
                                          p->adj.beta, p->S, p->T, p->adj.grad, s),
                         "term gradients");
            }
            hipcheck(hipMemcpyAsync(grad, p->adj.grad, p->adj.n_mpo_terms * sizeof(double), hipMemcpyDeviceToHost, s),
                     "download");
            hipcheck(hipStreamSynchronize(s), "term gradients");
        } catch (...) {
            (void)hipStreamSynchronize(s);
            dev_free(din);
            throw;
        }
        dev_free(din);
    });
}

int mq_hpsi_plan_destroy(mq_hpsi_plan_t p) {
    if (!p) return MQ_OK;
    int rc = guarded(p->ctx, [&] {
        (void)hipStreamSynchronize(p->ctx->stream);
        for (auto& y : p->adj.yg) y.release();
        for (auto* w : p->adj.work) dev_free(w);
        dev_free(p->adj.lines);
        dev_free(p->adj.terms);
        dev_free(p->adj.src);
        dev_free(p->adj.beta);
        dev_free(p->adj.grad);
        for (auto& c : p->chunks) {
            c.tg.release();
            c.cg.release();
            c.sa.release();
        }
        dev_free(p->gemm_ws);
        dev_free(p->gemm_ws2);
        dev_free(p->S2);
        if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
        for (auto e : p->ev_sa)
            if (e) (void)hipEventDestroy(e);
        for (auto e : p->ev_cg)
            if (e) (void)hipEventDestroy(e);
        dev_free(p->T);
        dev_free(p->S);
        dev_free(p->Lpanel);
        for (auto& g : p->graphs) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            dev_free(g.in);
            dev_free(g.out);
        }
        dev_free(p->Lpanel32);
        dev_free(p->R32);
        dev_free(p->psi32);
        dev_free(p->plain.lines);
        dev_free(p->plain.terms);
        dev_free(p->rown);
        dev_free(p->d_inst_off);
    });
    delete p;
    return rc;
}

}  // extern "C"

// ---- Jacobi-Davidson ----------------------------------------------------------------------------
namespace {

// relative residual at which a mixed-precision Davidson leaves FP32 H*psi for the FP64
// refinement: a decade above the FP32 contraction's relative error (about 2e-6 on the
// benchmark problems)
constexpr double MIXED_SWITCH = 1e-5;

// lowest eigenpair of a small symmetric matrix by cyclic Jacobi rotations (n <= 64)
void small_sym_lowest(std::vector<double> a, int n, double& theta, std::vector<double>& vec) {
    std::vector<double> v(size_t(n) * n, 0.0);
    for (int i = 0; i < n; ++i) v[size_t(i) * n + i] = 1.0;
    auto A = [&](int i, int j) -> double& { return a[size_t(i) * n + j]; };
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) off += A(i, j) * A(i, j);
        if (off < 1e-300) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                if (std::fabs(A(p, q)) < 1e-300) continue;
                double tau = (A(q, q) - A(p, p)) / (2.0 * A(p, q));
                double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A(k, p), akq = A(k, q);
                    A(k, p) = c * akp - s * akq;
                    A(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A(p, k), aqk = A(q, k);
                    A(p, k) = c * apk - s * aqk;
                    A(q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = v[size_t(k) * n + p], vkq = v[size_t(k) * n + q];
                    v[size_t(k) * n + p] = c * vkp - s * vkq;
                    v[size_t(k) * n + q] = s * vkp + c * vkq;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < n; ++i)
        if (A(i, i) < A(best, best)) best = i;
    theta = A(best, best);
    vec.resize(n);
    for (int k = 0; k < n; ++k) vec[k] = v[size_t(k) * n + best];
}

}  // namespace

// The Jacobi-Davidson of mq_davidson.  Its vectors come from dev_alloc, or, with an arena, are
// carved from that device array (arena_elems doubles, every vector 256-byte aligned).
static void davidson_core(mq_hpsi_plan_t p, const double* psi0, int n_ortho, const double* const* ortho, double tol,
                          int max_iter, int precision, double* eval, double* psi_out, mq_davidson_stats_t* stats,
                          double* arena, uint64_t arena_elems) {
        require(p->n_instances == 1, "mq_davidson: a batch plan holds several independent problems");
        require(psi0 && eval && psi_out && max_iter >= 1 && n_ortho >= 0, "mq_davidson: bad argument");
        require(n_ortho == 0 || ortho, "mq_davidson: null ortho vectors");
        auto t_start = std::chrono::steady_clock::now();
        hipStream_t s = p->ctx->stream;
        const uint64_t n = p->psi_n;
        require(n > 0, "mq_davidson: empty site problem");
        // the reference truncates the ortho list to at most n-1 vectors (ietl_jacobi_davidson.h:41-42)
        int no = std::min<int64_t>(n_ortho, int64_t(n) - 1);
        std::vector<double*> bufs;
        uint64_t carved = 0;
        const uint64_t stride = (n + 31) / 32 * 32;
        auto alloc = [&]() {
            if (arena) {
                require(carved + stride <= arena_elems, "mq_davidson_noalloc: search space outside the scratch array");
                double* v = arena + carved;
                carved += stride;
                return v;
            }
            bufs.push_back(dev_alloc<double>(n));
            return bufs.back();
        };
        double* sc = p->ctx->dwork + DOT_BLOCKS;  // device scalars
        double* work = p->ctx->dwork;
        double hpsi_s = 0;
        // GMRES steps of the correction equation (ietl_jcd_gmres; 0 = the plain correction)
        const int gm = p->ctx->jcd_gmres;
        require(gm == 0 || !arena, "mq_davidson_noalloc: the GMRES correction (MQ_OPT_JCD_GMRES) needs allocations");
        try {
            std::vector<double*> V(max_iter), VA(max_iter), O(no);
            for (int i = 0; i < max_iter; ++i) { V[i] = alloc(); VA[i] = alloc(); }
            double* u = alloc();
            double* r = alloc();
            // GMRES basis (gm + 1 vectors), its work vector and the projected operand
            std::vector<double*> G(gm > 0 ? gm + 1 : 0);
            for (auto& g : G) g = alloc();
            double* gw = gm > 0 ? alloc() : nullptr;
            double* gt = gm > 0 ? alloc() : nullptr;
            for (int k = 0; k < no; ++k) {
                O[k] = alloc();
                hipcheck(hipMemcpyAsync(O[k], ortho[k], n * sizeof(double), hipMemcpyHostToDevice, s), "upload");
            }
            hipcheck(hipMemcpyAsync(V[0], psi0, n * sizeof(double), hipMemcpyHostToDevice, s), "upload");
            auto dot_host = [&](const double* x, const double* y) {
                double h;
                hipcheck(launch_dot(x, y, n, work, sc, s), "dot");
                hipcheck(hipMemcpyAsync(&h, sc, sizeof(double), hipMemcpyDeviceToHost, s), "dot");
                hipcheck(hipStreamSynchronize(s), "dot");
                return h;
            };
            // out[i] = x(i) . y(i) for i < count: the products into device scalars (one batch of at
            // most DOT_BATCH at a time), then one download and one wait per batch
            constexpr int SC_BATCH = 8, DOT_BATCH = 48;
            static_assert(SC_BATCH + DOT_BATCH <= 64, "device scalars of the context");
            auto dots_host = [&](int count, auto x_of, auto y_of, double* out) {
                for (int c0 = 0; c0 < count; c0 += DOT_BATCH) {
                    const int cn = std::min(DOT_BATCH, count - c0);
                    for (int i = 0; i < cn; ++i)
                        hipcheck(launch_dot(x_of(c0 + i), y_of(c0 + i), n, work, sc + SC_BATCH + i, s), "dot");
                    hipcheck(hipMemcpyAsync(out + c0, sc + SC_BATCH, cn * sizeof(double), hipMemcpyDeviceToHost, s),
                             "dot");
                    hipcheck(hipStreamSynchronize(s), "dot");
                }
            };
            // y -= (x.y)/(x.x) x   with coefficients kept on the device
            auto subtract_along = [&](double* y, const double* x, bool normalized) {
                hipcheck(launch_dot(x, y, n, work, sc + 1, s), "dot");
                if (!normalized) hipcheck(launch_dot(x, x, n, work, sc + 2, s), "dot");
                hipcheck(launch_axpy_dev(y, x, n, sc + 1, normalized ? nullptr : sc + 2, -1.0, s), "axpy");
            };
            // Gram-Schmidt of the ortho vectors among themselves (ietl_jacobi_davidson.h:43-47)
            for (int a = 1; a < no; ++a)
                for (int b0 = 0; b0 < a; ++b0) subtract_along(O[a], O[b0], false);
            auto project = [&](double* t) {
                for (int k = 0; k < no; ++k) subtract_along(t, O[k], false);
            };
            std::vector<double> M(size_t(max_iter) * max_iter, 0.0);
            double theta = 0, rnorm = 0;
            int it = 0;
            // y = (1 - uu*)(H - theta)(1 - uu*) x  (jcd_solver_operator, jacobi.h:263-281)
            auto jcd_op = [&](const double* x, double* y, int prec, double theta_) {
                hipcheck(launch_axpby(gt, x, n, 0.0, 1.0, s), "axpby");
                hipcheck(launch_dot(u, x, n, work, sc + 3, s), "dot");
                hipcheck(launch_axpy_dev(gt, u, n, sc + 3, nullptr, -1.0, s), "axpy");
                auto h0 = std::chrono::steady_clock::now();
                hpsi_run(p, gt, y, prec);
                hipcheck(hipStreamSynchronize(s), "hpsi");
                hpsi_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - h0).count();
                hipcheck(launch_axpby(y, gt, n, 1.0, -theta_, s), "axpby");
                hipcheck(launch_dot(u, y, n, work, sc + 4, s), "dot");
                hipcheck(launch_axpy_dev(y, u, n, sc + 4, nullptr, -1.0, s), "axpy");
            };
            // x <- GMRES(op, b = -r, x0 = x) with at most gm - 1 Arnoldi steps and absolute
            // tolerance abs_tol on the rotated residual (ietl_gmres, gmres.h:89-145, restated with its
            // rule that the step which meets the tolerance is left out of the update)
            auto gmres = [&](double* x, int prec, double theta_, double abs_tol) {
                const int m = gm;
                std::vector<double> sv(size_t(m) + 1, 0.0), cs(size_t(m) + 1, 0.0), sn(size_t(m) + 1, 0.0);
                std::vector<double> H(size_t(m + 1) * size_t(m + 1), 0.0);
                auto Hat = [&](int i, int j) -> double& { return H[size_t(i) * size_t(m + 1) + size_t(j)]; };
                jcd_op(x, G[0], prec, theta_);
                hipcheck(launch_axpby(G[0], r, n, -1.0, -1.0, s), "axpby");  // -r - op(x0)
                sv[0] = std::sqrt(dot_host(G[0], G[0]));
                if (std::fabs(sv[0]) < abs_tol) return;
                hipcheck(launch_axpby(G[0], G[0], n, 0.0, 1.0 / sv[0], s), "axpby");
                int i = 0;
                auto rot = [](double& dx, double& dy, double c, double sn_) {
                    const double r0 = c * dx + sn_ * dy, r1 = -sn_ * dx + c * dy;
                    dx = r0;
                    dy = r1;
                };
                for (; i < m - 1; ++i) {
                    jcd_op(G[i], gw, prec, theta_);
                    for (int k = 0; k <= i; ++k) {
                        Hat(k, i) = dot_host(gw, G[k]);
                        hipcheck(launch_axpby(gw, G[k], n, 1.0, -Hat(k, i), s), "axpby");
                    }
                    Hat(i + 1, i) = std::sqrt(dot_host(gw, gw));
                    hipcheck(launch_axpby(G[i + 1], gw, n, 0.0, 1.0 / Hat(i + 1, i), s), "axpby");
                    for (int k = 0; k < i; ++k) rot(Hat(k, i), Hat(k + 1, i), cs[size_t(k)], sn[size_t(k)]);
                    const double dx = Hat(i, i), dy = Hat(i + 1, i);
                    if (dy == 0.0) {
                        cs[size_t(i)] = 1.0;
                        sn[size_t(i)] = 0.0;
                    } else if (std::fabs(dy) > std::fabs(dx)) {
                        const double q = dx / dy;
                        sn[size_t(i)] = 1.0 / std::sqrt(1.0 + q * q);
                        cs[size_t(i)] = q * sn[size_t(i)];
                    } else {
                        const double q = dy / dx;
                        cs[size_t(i)] = 1.0 / std::sqrt(1.0 + q * q);
                        sn[size_t(i)] = q * cs[size_t(i)];
                    }
                    rot(Hat(i, i), Hat(i + 1, i), cs[size_t(i)], sn[size_t(i)]);
                    rot(sv[size_t(i)], sv[size_t(i) + 1], cs[size_t(i)], sn[size_t(i)]);
                    if (std::fabs(sv[size_t(i) + 1]) < abs_tol) break;
                }
                // back substitution over the first i columns, then x += sum_k y_k G[k]
                std::vector<double> y(sv.begin(), sv.begin() + i);
                for (int a = i - 1; a >= 0; --a) {
                    y[size_t(a)] /= Hat(a, a);
                    for (int b = a - 1; b >= 0; --b) y[size_t(b)] -= Hat(b, a) * y[size_t(a)];
                }
                for (int k = 0; k < i; ++k) hipcheck(launch_axpby(x, G[k], n, 1.0, y[size_t(k)], s), "axpby");
            };
            // one Jacobi-Davidson run from V[0] with H*psi in precision `prec`; leaves the Ritz
            // vector in u
            auto jd = [&](int prec, double tol_run, int iters) {
            project(V[0]);
            it = 0;
            while (true) {
                double* t = V[it];
                // modified Gram-Schmidt with one refinement when the norm collapses (jacobi.h:409-415)
                double tau = std::sqrt(dot_host(t, t));
                for (int i = 0; i < it; ++i) subtract_along(t, V[i], true);
                if (std::sqrt(dot_host(t, t)) < 0.25 * tau)
                    for (int i = 0; i < it; ++i) subtract_along(t, V[i], true);
                project(t);
                hipcheck(launch_dot(t, t, n, work, sc, s), "dot");
                hipcheck(launch_scale_rsqrt(t, n, sc, s), "scale");
                auto h0 = std::chrono::steady_clock::now();
                hpsi_run(p, t, VA[it], prec);
                hipcheck(hipStreamSynchronize(s), "hpsi");
                hpsi_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - h0).count();
                {
                    std::vector<double> row(size_t(it) + 1);
                    dots_host(it + 1, [&](int i) { return V[i]; }, [&](int) { return VA[it]; }, row.data());
                    for (int i = 0; i <= it; ++i) M[size_t(i) * max_iter + it] = row[size_t(i)];
                }
                int dim = it + 1;
                std::vector<double> Ms(size_t(dim) * dim), svec;
                for (int i = 0; i < dim; ++i)
                    for (int j = i; j < dim; ++j)
                        Ms[size_t(i) * dim + j] = Ms[size_t(j) * dim + i] = M[size_t(i) * max_iter + j];
                small_sym_lowest(Ms, dim, theta, svec);
                // u = V s, uA = VA s
                hipcheck(launch_axpby(u, V[0], n, 0.0, svec[0], s), "axpby");
                hipcheck(launch_axpby(r, VA[0], n, 0.0, svec[0], s), "axpby");
                for (int j = 1; j < dim; ++j) {
                    hipcheck(launch_axpby(u, V[j], n, 1.0, svec[j], s), "axpby");
                    hipcheck(launch_axpby(r, VA[j], n, 1.0, svec[j], s), "axpby");
                }
                project(r);
                hipcheck(launch_axpby(r, u, n, 1.0, -theta, s), "axpby");  // r = uA - theta u
                ++it;
                rnorm = std::sqrt(dot_host(r, r));
                bool conv = rnorm <= tol_run * std::fabs(theta) || rnorm < tol_run;  // iteration.h:57
                if (conv || it >= iters) break;
                // correction t = -r + (r.u)/(u.u) u  (jcd_gmres_solver with no GMRES steps, jacobi.h:152-156)
                double du[2];
                dots_host(2, [&](int i) { return i == 0 ? (const double*)r : (const double*)u; },
                          [&](int) { return (const double*)u; }, du);
                const double dru = du[0], duu = du[1];
                hipcheck(launch_axpby(V[it], r, n, 0.0, -1.0, s), "axpby");
                hipcheck(launch_axpby(V[it], u, n, 1.0, dru / duu, s), "axpby");
                // with GMRES steps, that is the start of an approximate solution of the correction
                // equation to the tolerance 2^-(it+1) (jcd_gmres_solver, jacobi.h:126-164)
                if (gm > 0) gmres(V[it], prec, theta, 1.0 / std::pow(2.0, double(it + 1)));
            }
            };
            int refine = 0;
            require(precision == MQ_PREC_FP64 || precision == MQ_PREC_FP32 || precision == MQ_PREC_FP64_PLAIN,
                    "mq_davidson: precision must be MQ_PREC_FP64, MQ_PREC_FP32 or MQ_PREC_FP64_PLAIN");
            if (precision == MQ_PREC_FP64_PLAIN) {
                jd(MQ_PREC_FP64_PLAIN, tol, max_iter);
            } else if (precision == MQ_PREC_FP64 || max_iter < 2) {
                jd(MQ_PREC_FP64, tol, max_iter);
            } else {
                // mixed precision: one search space built with FP32 H*psi for up to max_iter
                // iterations, stopping once the relative residual is within MIXED_SWITCH (a decade
                // above the FP32 floor); then the Ritz vector is refined with FP64 H*psi from a
                // restart for the iterations left, at least one (the FP64 Rayleigh quotient and
                // residual of the returned vector)
                jd(MQ_PREC_FP32, std::max(tol, MIXED_SWITCH), max_iter);
                int it32 = it;
                hipcheck(hipMemcpyAsync(V[0], u, n * sizeof(double), hipMemcpyDeviceToDevice, s), "copy");
                jd(MQ_PREC_FP64, tol, std::max(1, max_iter - it32));
                refine = it;
                it += it32;
            }
            hipcheck(hipMemcpyAsync(psi_out, u, n * sizeof(double), hipMemcpyDeviceToHost, s), "download");
            hipcheck(hipStreamSynchronize(s), "davidson");
            *eval = theta;
            if (stats) {
                stats->iterations = it;
                stats->refine_iterations = refine;
                stats->residual_norm = rnorm;
                stats->hpsi_seconds = hpsi_s;
                stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
            }
        } catch (...) {
            (void)hipStreamSynchronize(s);
            for (auto b : bufs) dev_free(b);
            throw;
        }
        for (auto b : bufs) dev_free(b);
}

extern "C" int mq_davidson(mq_hpsi_plan_t p, const double* psi0, int n_ortho, const double* const* ortho, double tol,
                           int max_iter, int precision, double* eval, double* psi_out, mq_davidson_stats_t* stats) {
    if (!p) return MQ_ERR_INVALID;
    return guarded(p->ctx, [&] {
        davidson_core(p, psi0, n_ortho, ortho, tol, max_iter, precision, eval, psi_out, stats, nullptr, 0);
    });
}

extern "C" int mq_davidson_noalloc(mq_hpsi_plan_t p, const double* psi0, int n_ortho, const double* const* ortho,
                                   double tol, int max_iter, int precision, double* eval, double* psi_out,
                                   mq_davidson_stats_t* stats) {
    if (!p) return MQ_ERR_INVALID;
    return guarded(p->ctx, [&] {
        require(p->ctx->noalloc_scratch != nullptr, "mq_davidson_noalloc: no scratch array on this device");
        require(precision == MQ_PREC_FP64, "mq_davidson_noalloc: precision must be MQ_PREC_FP64");
        require(!p->use_graphs, "mq_davidson_noalloc: the plan replays HIP graphs (their buffers are allocated)");
        require(max_iter >= 1 && n_ortho >= 0, "mq_davidson_noalloc: bad argument");
        const uint64_t stride = (p->psi_n + 31) / 32 * 32;
        const uint64_t need = stride * uint64_t(2 * max_iter + 2 + std::min<int64_t>(n_ortho, int64_t(p->psi_n) - 1));
        if (need > MQ_NOALLOC_SCRATCH_ELEMS)
            throw Error(MQ_ERR_INVALID, "mq_davidson_noalloc: the search space needs " + std::to_string(need) +
                                            " doubles, the scratch array holds " +
                                            std::to_string(MQ_NOALLOC_SCRATCH_ELEMS));
        davidson_core(p, psi0, n_ortho, ortho, tol, max_iter, precision, eval, psi_out, stats, p->ctx->noalloc_scratch,
                      MQ_NOALLOC_SCRATCH_ELEMS);
    });
}

// Lock-step Jacobi-Davidson over the instances of a batch plan: the iteration of mq_davidson
// (modified Gram-Schmidt with one refinement when the norm collapses, Rayleigh matrix, lowest Ritz
// pair, residual, correction t = -r + (r.u / u.u) u) run for every instance at once, one batched
// H*psi per iteration; every vector operation is one segmented launch over all instances, and the
// per-instance scalars come back to the host together.  An instance that meets the residual rule
// stops (its segments are no longer changed); the others go on.  No ortho vectors.
extern "C" int mq_davidson_batch(mq_hpsi_plan_t p, const double* psi0, double tol, int max_iter, int precision,
                                 double* evals, double* psi_out, int32_t* iterations, mq_davidson_stats_t* stats) {
    if (!p) return MQ_ERR_INVALID;
    return guarded(p->ctx, [&] {
        require(p->n_instances > 1 || !p->inst_off.empty(), "mq_davidson_batch: not a batch plan");
        require(psi0 && evals && psi_out && iterations && max_iter >= 1, "mq_davidson_batch: bad argument");
        require(precision == MQ_PREC_FP64 || precision == MQ_PREC_FP32,
                "mq_davidson_batch: precision must be MQ_PREC_FP64 or MQ_PREC_FP32");
        auto t_start = std::chrono::steady_clock::now();
        hipStream_t s = p->ctx->stream;
        const int ns = p->n_instances;
        const uint64_t n = p->psi_n;
        const int64_t* off = p->d_inst_off;
        for (int i = 0; i < ns; ++i) require(p->inst_off[size_t(i) + 1] > p->inst_off[size_t(i)], "mq_davidson_batch: empty instance");
        std::vector<double*> bufs;
        auto alloc = [&](uint64_t k) { bufs.push_back(dev_alloc<double>(k)); return bufs.back(); };
        double hpsi_s = 0;
        try {
            std::vector<double*> V(max_iter), VA(max_iter);
            for (int i = 0; i < max_iter; ++i) { V[i] = alloc(n); VA[i] = alloc(n); }
            double* u = alloc(n);
            double* r = alloc(n);
            double* part = alloc(uint64_t(ns) * SEG_CHUNKS);
            // dot results: up to max_iter + 1 rows of ns; coefficient arrays a, b
            double* dres = alloc(uint64_t(max_iter + 1) * ns);
            double* ca = alloc(uint64_t(ns));
            double* cb = alloc(uint64_t(ns));
            std::vector<double> hres(size_t(max_iter + 1) * ns);
            hipcheck(hipMemcpyAsync(V[0], psi0, n * sizeof(double), hipMemcpyHostToDevice, s), "upload");
            hipcheck(hipStreamSynchronize(s), "upload");
            // k dots (x_j . y_j, j < k) of every instance, one download
            auto dots = [&](std::vector<std::pair<const double*, const double*>> const& xy) {
                for (size_t j = 0; j < xy.size(); ++j)
                    hipcheck(launch_seg_dot(xy[j].first, xy[j].second, off, ns, part, dres + j * ns, s), "dot");
                hipcheck(hipMemcpyAsync(hres.data(), dres, xy.size() * ns * sizeof(double), hipMemcpyDeviceToHost, s),
                         "dot");
                hipcheck(hipStreamSynchronize(s), "dot");
                std::vector<std::vector<double>> outv(xy.size());
                for (size_t j = 0; j < xy.size(); ++j) outv[j].assign(hres.begin() + j * ns, hres.begin() + (j + 1) * ns);
                return outv;
            };
            // y_i = a_i y_i + b_i x_i (the coefficient uploads complete before the host vectors go)
            auto axpby = [&](double* y, const double* x, std::vector<double> const& a, std::vector<double> const& b) {
                hipcheck(hipMemcpyAsync(ca, a.data(), ns * sizeof(double), hipMemcpyHostToDevice, s), "coef");
                hipcheck(hipMemcpyAsync(cb, b.data(), ns * sizeof(double), hipMemcpyHostToDevice, s), "coef");
                hipcheck(launch_seg_axpby(y, x, off, ns, ca, cb, s), "axpby");
                hipcheck(hipStreamSynchronize(s), "axpby");
            };
            std::vector<double> theta(ns, 0.0), rnorm(ns, 0.0);
            std::vector<int32_t> its(ns, 0);
            std::vector<char> active(ns, 1);
            std::vector<std::vector<double>> M(ns, std::vector<double>(size_t(max_iter) * max_iter, 0.0));
            const std::vector<double> one(ns, 1.0), zero(ns, 0.0);
            int it_run = 0;
            // one lock-step run from V[0] in precision prec; leaves the Ritz vectors in u
            auto jd = [&](int prec, double tol_run, int iters) {
                std::fill(active.begin(), active.end(), 1);
                std::fill(its.begin(), its.end(), 0);
                int it = 0;
                while (true) {
                    double* t = V[it];
                    std::vector<double> am(ns), bm(ns);  // masked coefficients
                    auto gs_pass = [&](std::vector<char> const& mask) {
                        for (int j = 0; j < it; ++j) {
                            auto c = dots({{V[j], t}})[0];
                            for (int i = 0; i < ns; ++i) bm[i] = mask[i] ? -c[i] : 0.0;
                            axpby(t, V[j], one, bm);
                        }
                    };
                    auto tau = dots({{t, t}})[0];
                    gs_pass(active);
                    auto nn = dots({{t, t}})[0];
                    std::vector<char> again(ns, 0);
                    bool any = false;
                    for (int i = 0; i < ns; ++i) {
                        again[i] = active[i] && std::sqrt(nn[i]) < 0.25 * std::sqrt(tau[i]);
                        any |= again[i] != 0;
                    }
                    if (any) gs_pass(again);
                    nn = dots({{t, t}})[0];
                    for (int i = 0; i < ns; ++i) {
                        am[i] = 0.0;
                        bm[i] = active[i] ? 1.0 / std::sqrt(nn[i]) : 0.0;  // inactive: t = 0, so H t = 0
                    }
                    axpby(t, t, am, bm);
                    auto h0 = std::chrono::steady_clock::now();
                    hpsi_run(p, t, VA[it], prec);
                    hipcheck(hipStreamSynchronize(s), "hpsi");
                    hpsi_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - h0).count();
                    std::vector<std::pair<const double*, const double*>> xy;
                    for (int j = 0; j <= it; ++j) xy.push_back({V[j], VA[it]});
                    auto mcol = dots(xy);
                    const int dim = it + 1;
                    std::vector<std::vector<double>> sv(ns);
                    for (int i = 0; i < ns; ++i) {
                        if (!active[i]) continue;
                        for (int j = 0; j <= it; ++j) M[i][size_t(j) * max_iter + it] = mcol[j][i];
                        std::vector<double> Ms(size_t(dim) * dim);
                        for (int a = 0; a < dim; ++a)
                            for (int b = a; b < dim; ++b)
                                Ms[size_t(a) * dim + b] = Ms[size_t(b) * dim + a] = M[i][size_t(a) * max_iter + b];
                        small_sym_lowest(Ms, dim, theta[i], sv[i]);
                    }
                    // u = V s, r = VA s (inactive instances keep their u and r)
                    for (int j = 0; j < dim; ++j) {
                        for (int i = 0; i < ns; ++i) {
                            am[i] = active[i] ? (j == 0 ? 0.0 : 1.0) : 1.0;
                            bm[i] = active[i] ? sv[i][size_t(j)] : 0.0;
                        }
                        axpby(u, V[j], am, bm);
                        axpby(r, VA[j], am, bm);
                    }
                    for (int i = 0; i < ns; ++i) bm[i] = active[i] ? -theta[i] : 0.0;
                    axpby(r, u, one, bm);  // r = uA - theta u
                    ++it;
                    auto rr = dots({{r, r}})[0];
                    bool any_active = false;
                    for (int i = 0; i < ns; ++i) {
                        if (!active[i]) continue;
                        rnorm[i] = std::sqrt(rr[i]);
                        its[i] = it;
                        if (rnorm[i] <= tol_run * std::fabs(theta[i]) || rnorm[i] < tol_run) active[i] = 0;  // iteration.h:57
                        any_active |= active[i] != 0;
                    }
                    if (!any_active || it >= iters) break;
                    // correction t = -r + (r.u)/(u.u) u of the active instances
                    auto ru = dots({{r, u}, {u, u}});
                    for (int i = 0; i < ns; ++i) {
                        am[i] = 0.0;
                        bm[i] = active[i] ? -1.0 : 0.0;
                    }
                    axpby(V[it], r, am, bm);
                    for (int i = 0; i < ns; ++i) bm[i] = active[i] ? ru[0][i] / ru[1][i] : 0.0;
                    axpby(V[it], u, one, bm);
                }
                it_run = it;
            };
            std::vector<int32_t> total(ns, 0);
            int refine = 0, lockstep = 0;
            if (precision == MQ_PREC_FP64 || max_iter < 2) {
                jd(MQ_PREC_FP64, tol, max_iter);
                total = its;
                lockstep = it_run;
            } else {
                // FP32 H*psi first (to MIXED_SWITCH), then FP64 refinement of every Ritz vector for the
                // lock-step iterations left (at least one), as mq_davidson does for one instance
                jd(MQ_PREC_FP32, std::max(tol, MIXED_SWITCH), max_iter);
                total = its;
                lockstep = it_run;
                hipcheck(hipMemcpyAsync(V[0], u, n * sizeof(double), hipMemcpyDeviceToDevice, s), "copy");
                jd(MQ_PREC_FP64, tol, std::max(1, max_iter - lockstep));
                for (int i = 0; i < ns; ++i) total[i] += its[i];
                refine = it_run;
                lockstep += it_run;
            }
            hipcheck(hipMemcpyAsync(psi_out, u, n * sizeof(double), hipMemcpyDeviceToHost, s), "download");
            hipcheck(hipStreamSynchronize(s), "davidson");
            double rmax = 0;
            for (int i = 0; i < ns; ++i) {
                evals[i] = theta[i];
                iterations[i] = total[i];
                rmax = std::max(rmax, rnorm[i]);
            }
            if (stats) {
                stats->iterations = lockstep;
                stats->refine_iterations = refine;
                stats->residual_norm = rmax;
                stats->hpsi_seconds = hpsi_s;
                stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
            }
        } catch (...) {
            (void)hipStreamSynchronize(s);
            for (auto b : bufs) dev_free(b);
            throw;
        }
        for (auto b : bufs) dev_free(b);
    });
}
