"""CPU oracle of the contraction path -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package, and
only as the checker / CPU baseline.  The product path never routes through it.

``liboracle.so`` (oracle.cpp) restates the reference's CPU schedule builders and kernels
(solver/tasks.cpp, contractions/create_schedule/*.hpp, boundary_ops/move_boundary.hpp);
``dense`` is an independent dense-tensor restatement for tiny abelian problems.
"""
import ctypes as C
import glob
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-B", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        from qcmaquis_amd import abi
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        dp = P(C.c_double)
        idx = P(abi.IndexDesc)
        L.oracle_set_blas.argtypes = [C.c_char_p, C.c_int]
        L.oracle_blas_threads.argtypes = [C.c_int]
        L.oracle_ninej.restype = C.c_double
        L.oracle_ninej.argtypes = [C.c_int] * 9
        L.oracle_sixj.restype = C.c_double
        L.oracle_sixj.argtypes = [C.c_int] * 6
        L.oracle_hpsi.argtypes = [C.c_int, P(abi.MPODesc), P(abi.BIndexDesc), dp, P(abi.BIndexDesc), dp, idx, idx, idx,
                                  dp, dp, C.c_int, dp]
        L.oracle_hpsi_sample.argtypes = [C.c_int, P(abi.MPODesc), P(abi.BIndexDesc), dp, P(abi.BIndexDesc), dp, idx,
                                         idx, idx, dp, dp, C.c_int, dp, P(C.c_uint8), dp]
        for f in (L.oracle_lstep, L.oracle_rstep):
            f.restype = C.c_void_p
            f.argtypes = [C.c_int, P(abi.MPODesc), P(abi.BIndexDesc), dp, idx, idx, idx, dp, idx, idx, idx, dp, C.c_int,
                          C.c_int]
        L.oracle_boundary_info.argtypes = [C.c_void_p, P(C.c_int), P(C.c_uint), P(C.c_int), P(C.c_int), P(C.c_uint64)]
        L.oracle_boundary_export.argtypes = [C.c_void_p, P(C.c_uint32), P(C.c_int64), dp, P(C.c_uint8), P(C.c_uint8),
                                             P(C.c_uint32), dp]
        L.oracle_boundary_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def openblas_path():
    """The OpenBLAS shipped with scipy (exports scipy_dgemm_), or None."""
    try:
        import scipy
        d = os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs")
        c = sorted(glob.glob(os.path.join(d, "libscipy_openblas*.so")))
        return os.path.abspath(c[0]) if c else None
    except Exception:
        return None


def use_blas(nthreads=1):
    """dlopen OpenBLAS for the oracle's dgemm; returns the library path or None (triple loop)."""
    p = openblas_path()
    if p is None:
        lib().oracle_set_blas(None, 0)
        return None
    rc = lib().oracle_set_blas(p.encode(), int(nthreads))
    return p if rc == 0 else None


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def hpsi(sym, mpo, L, R, layout, psi, nthreads=0, return_flops=False):
    psi = np.ascontiguousarray(psi, np.float64)
    out = np.zeros_like(psi)
    md, Ld, Rd = mpo.desc(), L.desc(), R.desc()
    li, pi, ri = layout.left.desc(), layout.phys.desc(), layout.right.desc()
    fl = C.c_double()
    Ldat = np.ascontiguousarray(L.data)
    Rdat = np.ascontiguousarray(R.data)
    rc = lib().oracle_hpsi(sym, C.byref(md), C.byref(Ld), _dp(Ldat), C.byref(Rd), _dp(Rdat), C.byref(li), C.byref(pi),
                           C.byref(ri), _dp(psi), _dp(out), int(nthreads), C.byref(fl))
    if rc != 0:
        raise RuntimeError("oracle_hpsi failed")
    return (out, fl.value) if return_flops else out


def hpsi_sample(sym, mpo, L, R, layout, psi, mask, nthreads=0):
    """the H*psi of the input blocks with mask[lb] set only: returns (partial output, their flops in
    the reference's model, seconds of the contraction loop, schedule build excluded)"""
    psi = np.ascontiguousarray(psi, np.float64)
    out = np.zeros_like(psi)
    md, Ld, Rd = mpo.desc(), L.desc(), R.desc()
    li, pi, ri = layout.left.desc(), layout.phys.desc(), layout.right.desc()
    fl, sec = C.c_double(), C.c_double()
    m = np.ascontiguousarray(np.asarray(mask, np.uint8))
    Ldat = np.ascontiguousarray(L.data)
    Rdat = np.ascontiguousarray(R.data)
    rc = lib().oracle_hpsi_sample(sym, C.byref(md), C.byref(Ld), _dp(Ldat), C.byref(Rd), _dp(Rdat), C.byref(li),
                                  C.byref(pi), C.byref(ri), _dp(psi), _dp(out), int(nthreads), C.byref(fl),
                                  m.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(sec))
    if rc != 0:
        raise RuntimeError("oracle_hpsi_sample failed")
    return out, fl.value, sec.value


def _export(h, bra, ket):
    from qcmaquis_amd.model import BoundarySpec
    L = lib()
    nc, aux, nb, nk, ne = C.c_int(), C.c_uint(), C.c_int(), C.c_int(), C.c_uint64()
    L.oracle_boundary_info(h, C.byref(nc), C.byref(aux), C.byref(nb), C.byref(nk), C.byref(ne))
    n, a = nc.value, aux.value
    lr = np.zeros(2 * n, np.uint32); off = np.zeros(n * a, np.int64); cs = np.zeros(n * a)
    tp = np.zeros(n * a, np.uint8); tr = np.zeros(n, np.uint8); nbl = np.zeros(n, np.uint32)
    data = np.zeros(ne.value)
    P = lambda x, t: x.ctypes.data_as(C.POINTER(t))
    L.oracle_boundary_export(h, P(lr, C.c_uint32), P(off, C.c_int64), _dp(cs), P(tp, C.c_uint8), P(tr, C.c_uint8),
                             P(nbl, C.c_uint32), _dp(data))
    L.oracle_boundary_free(h)
    cohorts = [dict(lr=(int(lr[2 * c]), int(lr[2 * c + 1])), offsets=off[c * a:(c + 1) * a].tolist(),
                    conj_scales=cs[c * a:(c + 1) * a].tolist(), transposes=tp[c * a:(c + 1) * a].tolist(),
                    tr=int(tr[c])) for c in range(n)]
    return BoundarySpec(bra, ket, a, cohorts, data)


def lstep(sym, mpo, L, bra_layout, bra_rp, ket_layout, ket_rp, symmetric=True, nthreads=0):
    """overlap_mpo_left_step; MPS data given right-paired (the bra is left-paired here)."""
    bra_lp = np.ascontiguousarray(bra_layout.to_left(np.asarray(bra_rp)))
    ket_rp = np.ascontiguousarray(ket_rp, np.float64)
    md, Ld = mpo.desc(), L.desc()
    d = [x.desc() for x in (bra_layout.left, bra_layout.phys, bra_layout.right, ket_layout.left, ket_layout.phys,
                            ket_layout.right)]
    Ldat = np.ascontiguousarray(L.data)
    h = lib().oracle_lstep(sym, C.byref(md), C.byref(Ld), _dp(Ldat), C.byref(d[0]), C.byref(d[1]), C.byref(d[2]),
                           _dp(bra_lp), C.byref(d[3]), C.byref(d[4]), C.byref(d[5]), _dp(ket_rp), int(bool(symmetric)),
                           int(nthreads))
    if not h:
        raise RuntimeError("oracle_lstep failed")
    return _export(h, bra_layout.right, ket_layout.right)


def rstep(sym, mpo, R, bra_layout, bra_rp, ket_layout, ket_rp, symmetric=True, nthreads=0):
    """overlap_mpo_right_step; MPS data right-paired."""
    bra_rp = np.ascontiguousarray(bra_rp, np.float64)
    ket_rp = np.ascontiguousarray(ket_rp, np.float64)
    md, Rd = mpo.desc(), R.desc()
    d = [x.desc() for x in (bra_layout.left, bra_layout.phys, bra_layout.right, ket_layout.left, ket_layout.phys,
                            ket_layout.right)]
    Rdat = np.ascontiguousarray(R.data)
    h = lib().oracle_rstep(sym, C.byref(md), C.byref(Rd), _dp(Rdat), C.byref(d[0]), C.byref(d[1]), C.byref(d[2]),
                           _dp(bra_rp), C.byref(d[3]), C.byref(d[4]), C.byref(d[5]), _dp(ket_rp), int(bool(symmetric)),
                           int(nthreads))
    if not h:
        raise RuntimeError("oracle_rstep failed")
    return _export(h, ket_layout.left, bra_layout.left)


def ninej(*a):
    return lib().oracle_ninej(*[int(x) for x in a])


def sixj(*a):
    return lib().oracle_sixj(*[int(x) for x in a])
