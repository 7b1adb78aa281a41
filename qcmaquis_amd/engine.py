"""Python binding of the contraction engine (mirrors contraction::Engine / solve_site_problem).

Reference entry points (dmrg/framework/dmrg/):
  Engine::overlap_mpo_left_step / overlap_mpo_right_step   mp_tensors/contractions/engine.hpp:53-76
  Engine::contraction_schedule + super_hamil_mv             engine.hpp:117-126, solver/super_hamil_mv.hpp
  solve_site_problem (Jacobi-Davidson)                      optimize/solver_interface.hpp:30-77

Every call goes through libqcmaquis_amd.so; errors raise :class:`EngineError` with the library's
message (the reference throws std::runtime_error or exits).
"""
import ctypes as C

import numpy as np

from . import abi
from .model import BoundarySpec, Index


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """One device, one stream (mq_ctx_create)."""

    def __init__(self, device=0, symmetry=abi.MQ_SYM_TWOU1):
        self.lib = abi.lib()
        h = C.c_void_p()
        rc = self.lib.mq_ctx_create(int(device), int(symmetry), C.byref(h))
        if rc != 0:
            raise EngineError(rc, "mq_ctx_create failed (is a GPU visible?)")
        self.h = h
        self.symmetry = symmetry

    def check(self, rc):
        if rc != 0:
            raise EngineError(rc, self.lib.mq_last_error(self.h).decode())

    def sync(self):
        self.check(self.lib.mq_ctx_sync(self.h))

    def set_option(self, option, value):
        """mq_ctx_set_option (abi.MQ_OPT_PLAIN_TABLES: plans made afterwards can run MQ_PREC_FP64_PLAIN)"""
        self.check(self.lib.mq_ctx_set_option(self.h, int(option), int(value)))

    def tuned(self):
        """the GEMM autotuner's table (MQ_OPT_AUTOTUNE): [(key, kind (0 = T-GEMMs, 1 = L*S), pick,
        best ms, default ms)]"""
        n = self.lib.mq_ctx_tuned(self.h, 0, None, None, None, None, None)
        if n <= 0:
            return []
        keys = (C.c_uint64 * n)()
        kinds, picks = (C.c_int32 * n)(), (C.c_int32 * n)()
        best, dflt = (C.c_float * n)(), (C.c_float * n)()
        self.lib.mq_ctx_tuned(self.h, n, keys, kinds, picks, best, dflt)
        return [(int(keys[i]), int(kinds[i]), int(picks[i]), float(best[i]), float(dflt[i])) for i in range(n)]

    @property
    def stream(self):
        return self.lib.mq_ctx_stream(self.h)

    def close(self):
        if self.h:
            self.lib.mq_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Handle:
    _destroy = None

    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h

    def close(self):
        if self.h and self.ctx.h:
            getattr(self.ctx.lib, self._destroy)(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMPO(_Handle):
    _destroy = "mq_mpo_destroy"

    def __init__(self, ctx, spec):
        d = spec.desc()
        h = C.c_void_p()
        ctx.check(ctx.lib.mq_mpo_create(ctx.h, C.byref(d), C.byref(h)))
        super().__init__(ctx, h)
        self.spec = spec


class DeviceBoundary(_Handle):
    _destroy = "mq_boundary_destroy"

    @classmethod
    def upload(cls, ctx, spec):
        d = spec.desc()
        h = C.c_void_p()
        data = np.ascontiguousarray(spec.data, np.float64)
        ctx.check(ctx.lib.mq_boundary_upload(ctx.h, C.byref(d), _dp(data), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def initial(cls, ctx, index, aux_dim=1):
        h = C.c_void_p()
        d = index.desc()
        ctx.check(ctx.lib.mq_boundary_create_initial(ctx.h, C.byref(d), int(aux_dim), C.byref(h)))
        return cls(ctx, h)

    def info(self):
        i = abi.BoundaryInfo()
        self.ctx.check(self.ctx.lib.mq_boundary_info(self.h, C.byref(i)))
        return i

    def growth_stats(self):
        """how mq_lstep / mq_rstep built this boundary (zeros otherwise)"""
        g = abi.GrowthStats()
        self.ctx.check(self.ctx.lib.mq_boundary_growth_stats(self.h, C.byref(g)))
        return {f: getattr(g, f) for f, _ in abi.GrowthStats._fields_}

    def download(self):
        """BoundarySpec with the index and data of this boundary."""
        i = self.info()
        nb, nk, nc, aux = i.n_bra, i.n_ket, i.n_cohorts, i.aux_dim
        bch = np.zeros(2 * nb, np.int32); bsz = np.zeros(nb, np.uint32)
        kch = np.zeros(2 * nk, np.int32); ksz = np.zeros(nk, np.uint32)
        lr = np.zeros(2 * nc, np.uint32); off = np.zeros(nc * aux, np.int64)
        cs = np.zeros(nc * aux, np.float64); tp = np.zeros(nc * aux, np.uint8); tr = np.zeros(nc, np.uint8)
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        self.ctx.check(self.ctx.lib.mq_boundary_export_index(
            self.h, P(bch, C.c_int32), P(bsz, C.c_uint32), P(kch, C.c_int32), P(ksz, C.c_uint32), P(lr, C.c_uint32),
            P(off, C.c_int64), P(cs, C.c_double), P(tp, C.c_uint8), P(tr, C.c_uint8)))
        data = np.zeros(i.total_elements)
        self.ctx.check(self.ctx.lib.mq_boundary_download(self.h, _dp(data)))
        bra = Index(bch.reshape(-1, 2).tolist(), bsz.tolist())
        ket = Index(kch.reshape(-1, 2).tolist(), ksz.tolist())
        cohorts = [dict(lr=(int(lr[2 * c]), int(lr[2 * c + 1])), offsets=off[c * aux:(c + 1) * aux].tolist(),
                        conj_scales=cs[c * aux:(c + 1) * aux].tolist(), transposes=tp[c * aux:(c + 1) * aux].tolist(),
                        tr=int(tr[c])) for c in range(nc)]
        return BoundarySpec(bra, ket, aux, cohorts, data)


class DeviceMPS(_Handle):
    _destroy = "mq_mps_destroy"

    def __init__(self, ctx, layout, data, pairing=abi.MQ_RIGHT_PAIRED):
        data = np.ascontiguousarray(data, np.float64)
        if data.size != layout.n:
            raise ValueError("MPS data size does not match its layout")
        h = C.c_void_p()
        l, p, r = layout.left.desc(), layout.phys.desc(), layout.right.desc()
        ctx.check(ctx.lib.mq_mps_upload(ctx.h, C.byref(l), C.byref(p), C.byref(r), int(pairing), _dp(data), C.byref(h)))
        super().__init__(ctx, h)
        self.layout = layout

    def download(self, pairing=abi.MQ_RIGHT_PAIRED):
        out = np.zeros(self.layout.n)
        self.ctx.check(self.ctx.lib.mq_mps_download(self.h, int(pairing), _dp(out)))
        return out


def lstep(ctx, mpo, left, bra, ket, symmetric=True, precision=abi.MQ_PREC_FP64):
    """overlap_mpo_left_step (move_boundary.hpp:147-279)"""
    h = C.c_void_p()
    ctx.check(ctx.lib.mq_lstep(ctx.h, mpo.h, left.h, bra.h, ket.h, int(bool(symmetric)), int(precision), C.byref(h)))
    return DeviceBoundary(ctx, h)


def rstep(ctx, mpo, right, bra, ket, symmetric=True, precision=abi.MQ_PREC_FP64):
    """overlap_mpo_right_step (move_boundary.hpp:281-404)"""
    h = C.c_void_p()
    ctx.check(ctx.lib.mq_rstep(ctx.h, mpo.h, right.h, bra.h, ket.h, int(bool(symmetric)), int(precision), C.byref(h)))
    return DeviceBoundary(ctx, h)


def block_svd(ctx, blocks):
    """thin SVD of every block on the device (mq_block_svd, one rocSOLVER dgesvd per block; the
    decomposition of the two-site split, svd_truncate, block_matrix_algorithms.h:186-244):
    {key: X} -> {key: (U, s, Vt)} with U rows x k, s descending, Vt k x cols, k = min(rows, cols)"""
    keys = list(blocks)
    if not keys:
        return {}
    rows = np.array([blocks[k].shape[0] for k in keys], np.int32)
    cols = np.array([blocks[k].shape[1] for k in keys], np.int32)
    ks = np.minimum(rows, cols).astype(np.int64)
    a = np.ascontiguousarray(np.concatenate([np.asarray(blocks[k], np.float64).ravel(order="F") for k in keys]))
    s = np.empty(int(ks.sum()))
    u = np.empty(int((rows.astype(np.int64) * ks).sum()))
    vt = np.empty(int((ks * cols).sum()))
    ctx.check(ctx.lib.mq_block_svd(ctx.h, len(keys), rows.ctypes.data_as(C.POINTER(C.c_int32)),
                                   cols.ctypes.data_as(C.POINTER(C.c_int32)), _dp(a), _dp(s), _dp(u), _dp(vt)))
    out, o_s, o_u, o_v = {}, 0, 0, 0
    for key, r, c, k in zip(keys, rows.tolist(), cols.tolist(), ks.tolist()):
        out[key] = (u[o_u:o_u + r * k].reshape(k, r).T, s[o_s:o_s + k], vt[o_v:o_v + k * c].reshape(c, k).T)
        o_s, o_u, o_v = o_s + k, o_u + r * k, o_v + k * c
    return out


def boundary_gram(boundary, side, alpha, blocks):
    """noise term of the density-matrix split (mq_boundary_gram; alpha_dm_direct and
    alpha_dm_direct_right, move_boundary.hpp:38-145): for a boundary grown with an identity bra and
    symmetric = False, {charge: alpha * sum X X^T} over the cohorts of each row charge (side 0) or
    {charge: alpha * sum X_b^T X_b} over the blocks of each column charge (side 1); blocks: list of
    (charge, size) of the grouped index"""
    if not blocks:
        return {}
    charges = np.ascontiguousarray(np.array([c for c, _ in blocks], np.int32).reshape(-1, 2))
    sizes = np.array([n for _, n in blocks], np.uint32)
    out = np.empty(int((sizes.astype(np.int64) ** 2).sum()))
    ctx = boundary.ctx
    ctx.check(ctx.lib.mq_boundary_gram(boundary.h, int(side), float(alpha), len(blocks),
                                       charges.ctypes.data_as(C.POINTER(C.c_int32)),
                                       sizes.ctypes.data_as(C.POINTER(C.c_uint32)), _dp(out)))
    res, o = {}, 0
    for (c, n) in blocks:
        res[c] = out[o:o + n * n].reshape(n, n).T
        o += n * n
    return res


class HpsiPlan(_Handle):
    """create_contraction_schedule + the device work lists of one site (mq_hpsi_plan_create)."""

    _destroy = "mq_hpsi_plan_destroy"

    def __init__(self, ctx, mpo, left, right, layout):
        h = C.c_void_p()
        l, p, r = layout.left.desc(), layout.phys.desc(), layout.right.desc()
        ctx.check(ctx.lib.mq_hpsi_plan_create(ctx.h, mpo.h, left.h, right.h, C.byref(l), C.byref(p), C.byref(r),
                                              C.byref(h)))
        super().__init__(ctx, h)
        self.layout = layout
        self._refs = (mpo, left, right)

    def stats(self):
        s = abi.HpsiStats()
        self.ctx.check(self.ctx.lib.mq_hpsi_plan_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in abi.HpsiStats._fields_}

    def apply(self, psi, precision=abi.MQ_PREC_FP64):
        """H*psi for a host vector (right-paired blocks concatenated)."""
        psi = np.ascontiguousarray(psi, np.float64)
        if psi.size != self.layout.n:
            raise ValueError("psi size does not match the plan")
        out = np.zeros_like(psi)
        self.ctx.check(self.ctx.lib.mq_hpsi_host(self.h, _dp(psi), _dp(out), int(precision)))
        return out

    def use_graphs(self, enable=True):
        """run H*psi (apply, apply_device, davidson) as a captured HIP graph"""
        self.ctx.check(self.ctx.lib.mq_hpsi_plan_use_graphs(self.h, int(bool(enable))))

    def apply_device(self, in_ptr, out_ptr, precision=abi.MQ_PREC_FP64):
        """Stream-ordered H*psi on device pointers (context stream)."""
        self.ctx.check(self.ctx.lib.mq_hpsi(self.h, C.c_void_p(in_ptr), C.c_void_p(out_ptr), int(precision)))

    def davidson(self, psi0, ortho=(), tol=1e-8, max_iter=8, precision=abi.MQ_PREC_FP64, gmres=None):
        """Jacobi-Davidson lowest eigenpair (solve<double>, solver/solver.cpp:13-28); gmres: GMRES
        steps of the correction equation for this call (ietl_jcd_gmres; None: the context's
        MQ_OPT_JCD_GMRES)"""
        if gmres is not None:
            self.ctx.set_option(abi.MQ_OPT_JCD_GMRES, int(gmres))
            try:
                return self.davidson(psi0, ortho, tol, max_iter, precision)
            finally:
                self.ctx.set_option(abi.MQ_OPT_JCD_GMRES, 0)
        psi0 = np.ascontiguousarray(psi0, np.float64)
        out = np.zeros_like(psi0)
        ev = C.c_double()
        st = abi.DavidsonStats()
        orth = [np.ascontiguousarray(o, np.float64) for o in ortho]
        arr = (C.POINTER(C.c_double) * max(1, len(orth)))(*[_dp(o) for o in orth])
        self.ctx.check(self.ctx.lib.mq_davidson(self.h, _dp(psi0), len(orth), arr, float(tol), int(max_iter),
                                                int(precision), C.byref(ev), _dp(out), C.byref(st)))
        return ev.value, out, {f: getattr(st, f) for f, _ in abi.DavidsonStats._fields_}


class HpsiPlanBatch(_Handle):
    """Independent site problems in one plan (mq_hpsi_plan_create_batch): psi is the concatenation
    of the instances' psi and one H*psi over all of them is one grouped launch per stage.
    items: (mpo, left, right, layout) per instance."""

    _destroy = "mq_hpsi_plan_destroy"

    def __init__(self, ctx, items):
        n = len(items)
        h = C.c_void_p()
        vp = C.c_void_p * n
        descs = [[it[3].left.desc() for it in items], [it[3].phys.desc() for it in items],
                 [it[3].right.desc() for it in items]]
        arr = [(abi.IndexDesc * n)(*d) for d in descs]
        ctx.check(ctx.lib.mq_hpsi_plan_create_batch(ctx.h, n, vp(*[it[0].h for it in items]),
                                                    vp(*[it[1].h for it in items]), vp(*[it[2].h for it in items]),
                                                    arr[0], arr[1], arr[2], C.byref(h)))
        super().__init__(ctx, h)
        self.layouts = [it[3] for it in items]
        self.offsets = np.concatenate([[0], np.cumsum([lay.n for lay in self.layouts])]).astype(np.int64)
        self._refs = items

    @property
    def n(self):
        return int(self.offsets[-1])

    def split(self, vec):
        """the instances' segments of a concatenated vector"""
        return [vec[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.layouts))]

    def stats(self):
        s = abi.HpsiStats()
        self.ctx.check(self.ctx.lib.mq_hpsi_plan_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in abi.HpsiStats._fields_}

    def apply(self, psi, precision=abi.MQ_PREC_FP64):
        """H*psi of every instance for a host vector (instances concatenated)"""
        psi = np.ascontiguousarray(psi, np.float64)
        if psi.size != self.n:
            raise ValueError("psi size does not match the batch")
        out = np.zeros_like(psi)
        self.ctx.check(self.ctx.lib.mq_hpsi_host(self.h, _dp(psi), _dp(out), int(precision)))
        return out

    def apply_device(self, in_ptr, out_ptr, precision=abi.MQ_PREC_FP64):
        self.ctx.check(self.ctx.lib.mq_hpsi(self.h, C.c_void_p(in_ptr), C.c_void_p(out_ptr), int(precision)))

    def davidson(self, psi0s, tol=1e-8, max_iter=8, precision=abi.MQ_PREC_FP64):
        """lowest eigenpair of every instance, lock-step on the device (mq_davidson_batch):
        (energies, vectors, iterations per instance, stats)"""
        psi0 = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64) for x in psi0s]))
        if psi0.size != self.n:
            raise ValueError("start vectors do not match the batch")
        k = len(self.layouts)
        out = np.zeros_like(psi0)
        ev = np.zeros(k)
        its = np.zeros(k, np.int32)
        st = abi.DavidsonStats()
        self.ctx.check(self.ctx.lib.mq_davidson_batch(self.h, _dp(psi0), float(tol), int(max_iter), int(precision),
                                                      _dp(ev), _dp(out), its.ctypes.data_as(C.POINTER(C.c_int32)),
                                                      C.byref(st)))
        return ev, self.split(out), its, {f: getattr(st, f) for f, _ in abi.DavidsonStats._fields_}

