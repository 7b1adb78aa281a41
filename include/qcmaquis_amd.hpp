/*
 * qcmaquis_amd.hpp -- C++ forwarding layer over the C ABI (qcmaquis_amd.h) with the reference's
 * entry points of the contraction path, for a C++ sweep driver that calls them as a drop-in:
 *
 *   contraction::Engine<...>::overlap_mpo_left_step / overlap_mpo_right_step /
 *   contraction_schedule                       mp_tensors/contractions/engine.hpp:47-140
 *   solve<double>(out, initial, H, ortho, gmres, tol, maxiter)   solver/solver.h:64-68
 *   DavidsonVector (flat vector with per-block views)            solver/davidson_vector.h:34-82
 *
 * (paths relative to the reference tree dmrg/framework/dmrg/).  Errors of the C ABI become
 * qcmaquis_amd::Error (a std::runtime_error, what the reference throws for logic errors,
 * solver_interface.hpp:59-60).  Device objects are RAII handles; a context is one device and one
 * stream and is used by one host thread at a time.  Header-only; link libqcmaquis_amd.so.
 */
#ifndef QCMAQUIS_AMD_HPP
#define QCMAQUIS_AMD_HPP

#include <algorithm>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "qcmaquis_amd.h"

namespace qcmaquis_amd {

class Error : public std::runtime_error {
public:
    Error(int code, std::string const& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

// ---- context ---------------------------------------------------------------------------------
class Context {
public:
    Context(int device, int symmetry) {
        int rc = mq_ctx_create(device, symmetry, &h_);
        if (rc != MQ_OK) throw Error(rc, "mq_ctx_create failed");
        sym_ = symmetry;
    }
    ~Context() { if (h_) mq_ctx_destroy(h_); }
    Context(Context const&) = delete;
    Context& operator=(Context const&) = delete;
    mq_ctx_t get() const { return h_; }
    int symmetry() const { return sym_; }
    void check(int rc) const {
        if (rc != MQ_OK) throw Error(rc, mq_last_error(h_));
    }

private:
    mq_ctx_t h_ = nullptr;
    int sym_ = 0;
};

// ---- sector index: (charge, size) pairs sorted strictly descending (Index, indexing_stable.hpp) --
struct Charge {
    int32_t q0 = 0, q1 = 0;
    bool operator<(Charge const& o) const { return q0 < o.q0 || (q0 == o.q0 && q1 < o.q1); }
    bool operator==(Charge const& o) const { return q0 == o.q0 && q1 == o.q1; }
};

class Index {
public:
    Index() = default;
    Index(std::vector<std::pair<Charge, uint32_t>> sectors) : sec_(std::move(sectors)) { refresh(); }
    size_t size() const { return sec_.size(); }
    std::pair<Charge, uint32_t> const& operator[](size_t i) const { return sec_[i]; }
    uint64_t sum_of_sizes() const {
        uint64_t s = 0;
        for (auto const& p : sec_) s += p.second;
        return s;
    }
    mq_index_desc const& desc() const { return d_; }
    static Index from_desc(mq_index_desc const& d) {
        std::vector<std::pair<Charge, uint32_t>> s;
        for (int i = 0; i < d.n; ++i) s.push_back({Charge{d.charges[2 * i], d.charges[2 * i + 1]}, d.sizes[i]});
        return Index(std::move(s));
    }

private:
    void refresh() {
        for (size_t i = 1; i < sec_.size(); ++i)
            if (!(sec_[i].first < sec_[i - 1].first)) throw Error(MQ_ERR_INVALID, "Index: charges not strictly descending");
        q_.clear();
        s_.clear();
        for (auto const& p : sec_) {
            q_.push_back(p.first.q0);
            q_.push_back(p.first.q1);
            s_.push_back(p.second);
        }
        d_ = mq_index_desc{int32_t(sec_.size()), q_.data(), s_.data()};
    }
    std::vector<std::pair<Charge, uint32_t>> sec_;
    std::vector<int32_t> q_;
    std::vector<uint32_t> s_;
    mq_index_desc d_{0, nullptr, nullptr};
};

// ---- MPO tensor on the device (MPOTensor read API, mpotensor.h:45-135) --------------------------
class MPOTensor {
public:
    // desc: any mq_mpo_desc (its arrays are read during construction only)
    MPOTensor(Context const& ctx, mq_mpo_desc const& desc) : ctx_(&ctx), rows_(desc.row_dim), cols_(desc.col_dim) {
        ctx.check(mq_mpo_create(ctx.get(), &desc, &h_));
    }
    ~MPOTensor() { if (h_) mq_mpo_destroy(h_); }
    MPOTensor(MPOTensor const&) = delete;
    MPOTensor& operator=(MPOTensor const&) = delete;
    mq_mpo_t get() const { return h_; }
    uint32_t row_dim() const { return rows_; }
    uint32_t col_dim() const { return cols_; }

private:
    Context const* ctx_;
    mq_mpo_t h_ = nullptr;
    uint32_t rows_, cols_;
};

// ---- MPS site tensor (MPSTensor, mpstensor.h:45-166): indices + right-paired data --------------
class MPSTensor {
public:
    MPSTensor(Index left, Index phys, Index right, std::vector<double> right_paired)
        : left_(std::move(left)), phys_(std::move(phys)), right_(std::move(right)), data_(std::move(right_paired)) {}
    Index const& row_dim() const { return left_; }
    Index const& site_dim() const { return phys_; }
    Index const& col_dim() const { return right_; }
    std::vector<double> const& data() const { return data_; }
    std::vector<double>& data() { return data_; }
    // device copy (made on first use per context and reused)
    mq_mps_t device(Context const& ctx) const {
        if (!dev_ || ctx_ != &ctx) {
            mq_mps_t h = nullptr;
            ctx.check(mq_mps_upload(ctx.get(), &left_.desc(), &phys_.desc(), &right_.desc(), MQ_RIGHT_PAIRED,
                                    data_.data(), &h));
            dev_.reset(h, [](mq_mps_t p) { mq_mps_destroy(p); });
            ctx_ = &ctx;
        }
        return dev_.get();
    }

private:
    Index left_, phys_, right_;
    std::vector<double> data_;
    mutable std::shared_ptr<mq_mps_s> dev_;
    mutable Context const* ctx_ = nullptr;
};

// ---- boundary (Boundary, boundary.h:236-406): a device-resident handle --------------------------
class Boundary {
public:
    static Boundary initial(Context const& ctx, Index const& idx, uint32_t aux_dim) {
        mq_boundary_t h = nullptr;
        ctx.check(mq_boundary_create_initial(ctx.get(), &idx.desc(), aux_dim, &h));
        return Boundary(ctx, h);
    }
    static Boundary upload(Context const& ctx, mq_bindex_desc const& d, double const* data) {
        mq_boundary_t h = nullptr;
        ctx.check(mq_boundary_upload(ctx.get(), &d, data, &h));
        return Boundary(ctx, h);
    }
    Boundary(Context const& ctx, mq_boundary_t h) : ctx_(&ctx), h_(h, [](mq_boundary_t p) { mq_boundary_destroy(p); }) {}
    mq_boundary_t get() const { return h_.get(); }
    mq_boundary_info_t info() const {
        mq_boundary_info_t i{};
        ctx_->check(mq_boundary_info(h_.get(), &i));
        return i;
    }
    // the non-transposed cohorts' data concatenated (the reference's serialisation point,
    // boundary.h:422-429, is the only place that needs it on the host)
    std::vector<double> download() const {
        std::vector<double> d(info().total_elements);
        ctx_->check(mq_boundary_download(h_.get(), d.data()));
        return d;
    }

private:
    Context const* ctx_;
    std::shared_ptr<mq_boundary_s> h_;
};

// ---- schedule of one site problem (ScheduleNew / SuperHamil, solver.h:38-60) --------------------
class SuperHamil {
public:
    SuperHamil(Context const& ctx, mq_hpsi_plan_t h) : ctx_(&ctx), h_(h, [](mq_hpsi_plan_t p) { mq_hpsi_plan_destroy(p); }) {}
    mq_hpsi_plan_t get() const { return h_.get(); }
    Context const& context() const { return *ctx_; }
    uint64_t size() const {
        uint64_t n = 0;
        ctx_->check(mq_hpsi_plan_psi_elements(h_.get(), &n));
        return n;
    }
    mq_hpsi_stats_t stats() const {
        mq_hpsi_stats_t s{};
        ctx_->check(mq_hpsi_plan_stats(h_.get(), &s));
        return s;
    }
    // psi' = H psi on host vectors (super_hamil_mv)
    void apply(std::vector<double> const& in, std::vector<double>& out, int precision = MQ_PREC_FP64) const {
        out.resize(in.size());
        ctx_->check(mq_hpsi_host(h_.get(), in.data(), out.data(), precision));
    }

private:
    Context const* ctx_;
    std::shared_ptr<mq_hpsi_plan_s> h_;
};

// ---- the engine's static entry points (contraction::Engine, engine.hpp:47-140) ------------------
struct Engine {
    // new L' = bra^dagger (L x W x ket); symmetric: skip Hermitian partners and complete the index
    // by complement_transpose (move_boundary.hpp:147-279; the reference's default is false)
    static Boundary overlap_mpo_left_step(Context const& ctx, MPSTensor const& bra, MPSTensor const& ket,
                                          Boundary const& left, MPOTensor const& mpo, bool symmetric = false) {
        mq_boundary_t h = nullptr;
        ctx.check(mq_lstep(ctx.get(), mpo.get(), left.get(), bra.device(ctx), ket.device(ctx), symmetric,
                           MQ_PREC_FP64, &h));
        return Boundary(ctx, h);
    }
    // new R' (move_boundary.hpp:281-404; the reference's default is symmetric)
    static Boundary overlap_mpo_right_step(Context const& ctx, MPSTensor const& bra, MPSTensor const& ket,
                                           Boundary const& right, MPOTensor const& mpo, bool symmetric = true) {
        mq_boundary_t h = nullptr;
        ctx.check(mq_rstep(ctx.get(), mpo.get(), right.get(), bra.device(ctx), ket.device(ctx), symmetric,
                           MQ_PREC_FP64, &h));
        return Boundary(ctx, h);
    }
    // the H*psi schedule of the site problem (L, W, R) on mps's sector structure
    static SuperHamil contraction_schedule(Context const& ctx, MPSTensor const& mps, Boundary const& left,
                                           Boundary const& right, MPOTensor const& mpo) {
        mq_hpsi_plan_t h = nullptr;
        ctx.check(mq_hpsi_plan_create(ctx.get(), mpo.get(), left.get(), right.get(), &mps.row_dim().desc(),
                                      &mps.site_dim().desc(), &mps.col_dim().desc(), &h));
        return SuperHamil(ctx, h);
    }
};

// ---- DavidsonVector (davidson_vector.h:34-82): one flat vector with per-block views ------------
template <class T>
class DavidsonVector {
public:
    DavidsonVector() = default;
    explicit DavidsonVector(std::vector<uint64_t> const& block_sizes) {
        offsets_.push_back(0);
        for (uint64_t s : block_sizes) offsets_.push_back(offsets_.back() + s);
        data_.assign(offsets_.back(), T(0));
    }
    size_t n_blocks() const { return offsets_.empty() ? 0 : offsets_.size() - 1; }
    T* operator[](size_t b) { return data_.data() + offsets_[b]; }
    T const* operator[](size_t b) const { return data_.data() + offsets_[b]; }
    uint64_t block_size(size_t b) const { return offsets_[b + 1] - offsets_[b]; }
    std::vector<T>& data() { return data_; }
    std::vector<T> const& data() const { return data_; }

private:
    std::vector<uint64_t> offsets_;
    std::vector<T> data_;
};

// ---- solve<double> (solver.h:64-68, solver.cpp:13-34) ---------------------------------------------
// Jacobi-Davidson of the site problem H from `initial`, orthogonal to `ortho`; the eigenvector is
// written block by block into the caller's `out` pointers (one per block of `initial`, as the
// reference writes into ret.data(), solver.cpp:22-25); returns the eigenvalue.  gmres: the
// reference's GMRES steps of the correction equation (ietl_jcd_gmres, truncated to an integer as
// the reference's size_t parameter; 0 = the plain correction), set on the context for this call.
template <class T>
T solve(std::vector<T*>& out, DavidsonVector<T>& initial, SuperHamil const& H,
        std::vector<DavidsonVector<T>> const& ortho, double gmres, double tol, int maxiter,
        int precision = MQ_PREC_FP64, mq_davidson_stats_t* stats = nullptr) {
    static_assert(sizeof(T) == sizeof(double), "solve: the ABI value type is FP64");
    if (!(gmres >= 0.0 && gmres <= 64.0)) throw Error(MQ_ERR_INVALID, "solve: gmres steps must be 0..64");
    if (out.size() != initial.n_blocks()) throw Error(MQ_ERR_INVALID, "solve: one output pointer per block");
    if (initial.data().size() != H.size()) throw Error(MQ_ERR_INVALID, "solve: vector length != site problem");
    std::vector<const double*> op;
    for (auto const& v : ortho) op.push_back(v.data().data());
    std::vector<double> psi(initial.data().size());
    double e = 0;
    mq_davidson_stats_t st{};
    H.context().check(mq_ctx_set_option(H.context().get(), MQ_OPT_JCD_GMRES, int(gmres)));
    const int rc = mq_davidson(H.get(), initial.data().data(), int(op.size()), op.empty() ? nullptr : op.data(), tol,
                               maxiter, precision, &e, psi.data(), &st);
    (void)mq_ctx_set_option(H.context().get(), MQ_OPT_JCD_GMRES, 0);
    H.context().check(rc);
    for (size_t b = 0; b < out.size(); ++b)
        std::copy(psi.begin() + (initial[b] - initial[0]), psi.begin() + (initial[b] - initial[0]) + initial.block_size(b),
                  out[b]);
    if (stats) *stats = st;
    return T(e);
}

}  // namespace qcmaquis_amd

#endif
