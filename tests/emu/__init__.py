"""ctypes binding of the host-side table emulator (plan_emulator.cpp), built on first use with g++
from the product's host sources."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "qcmaquis_amd", "csrc")
LIB = os.path.join(HERE, "libplanemu.so")
_lib = None


def build():
    srcs = [os.path.join(HERE, "plan_emulator.cpp")] + [os.path.join(CSRC, f) for f in
                                                         ("plans.cpp", "coupling.cpp", "descriptors.cpp")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
                    "-o", LIB + ".tmp"] + srcs, check=True)
    os.replace(LIB + ".tmp", LIB)


def lib():
    global _lib
    if _lib is None:
        build()
        from qcmaquis_amd import abi
        L = C.CDLL(LIB)
        P = C.POINTER
        dp = P(C.c_double)
        idx = P(abi.IndexDesc)
        L.emu_hpsi.argtypes = [C.c_int, P(abi.MPODesc), P(abi.BIndexDesc), dp, P(abi.BIndexDesc), dp, idx, idx, idx, dp, dp,
                               C.c_int, P(C.c_uint64), C.c_uint64, P(C.c_uint64), P(C.c_uint64), C.c_int]
        L.emu_set_sblock_panel.argtypes = [C.c_int]
        L.emu_set_gemm_policy.argtypes = [C.c_int]
        L.emu_term_grad.argtypes = [C.c_int, P(abi.MPODesc), P(abi.BIndexDesc), dp, P(abi.BIndexDesc), dp, idx, idx,
                                    idx, dp, dp, C.c_uint64]
        for f in (L.emu_lstep, L.emu_rstep):
            f.restype = C.c_void_p
            f.argtypes = [C.c_int, P(abi.MPODesc), P(abi.BIndexDesc), dp, idx, idx, idx, dp, idx, idx, idx, dp, C.c_int]
        L.emu_boundary_info.argtypes = [C.c_void_p, P(C.c_int), P(C.c_uint), P(C.c_uint64)]
        L.emu_boundary_export.argtypes = [C.c_void_p, P(C.c_uint32), P(C.c_int64), dp, P(C.c_uint8), P(C.c_uint8), dp]
        L.emu_boundary_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


last_small_lines = 0
last_chunks = 0
last_ts = 0


def hpsi(sym, mpo, L, R, layout, psi, small_max=-1, chunk_elems=0, plain=False):
    """H*psi through the product's tables; small_max: tiles of at most this many elements take the
    small-tile term sum (MQ_OPT_SMALL_TILE_MAX; -1 = the automatic rule), their count is left in
    last_small_lines; chunk_elems: largest T + S of one chunk (0 = one chunk), the number of chunks
    is left in last_chunks and the plan's T + S elements in last_ts; plain: run the reference-style
    tables (MQ_PREC_FP64_PLAIN) instead of the grouped ones"""
    global last_small_lines, last_chunks, last_ts
    psi = np.ascontiguousarray(psi, np.float64)
    out = np.zeros_like(psi)
    md, Ld, Rd = mpo.desc(), L.desc(), R.desc()
    li, pi, ri = layout.left.desc(), layout.phys.desc(), layout.right.desc()
    Ldat, Rdat = np.ascontiguousarray(L.data), np.ascontiguousarray(R.data)
    ns, nc, nts = C.c_uint64(), C.c_uint64(), C.c_uint64()
    rc = lib().emu_hpsi(sym, C.byref(md), C.byref(Ld), _dp(Ldat), C.byref(Rd), _dp(Rdat), C.byref(li), C.byref(pi),
                        C.byref(ri), _dp(psi), _dp(out), int(small_max), C.byref(ns), int(chunk_elems), C.byref(nc),
                        C.byref(nts), int(bool(plain)))
    if rc != 0:
        raise RuntimeError("emu_hpsi failed")
    last_small_lines = ns.value
    last_chunks = nc.value
    last_ts = nts.value
    return out


def term_grad(sym, mpo, L, R, layout, psi, chunk_elems=0):
    """dE/d(scale of every MPO term) of E = <psi|H_eff|psi> through the product's adjoint tables"""
    psi = np.ascontiguousarray(psi, np.float64)
    md, Ld, Rd = mpo.desc(), L.desc(), R.desc()
    grad = np.zeros(md.n_terms)
    li, pi, ri = layout.left.desc(), layout.phys.desc(), layout.right.desc()
    Ldat, Rdat = np.ascontiguousarray(L.data), np.ascontiguousarray(R.data)
    rc = lib().emu_term_grad(sym, C.byref(md), C.byref(Ld), _dp(Ldat), C.byref(Rd), _dp(Rdat), C.byref(li),
                             C.byref(pi), C.byref(ri), _dp(psi), _dp(grad), int(chunk_elems))
    if rc != 0:
        raise RuntimeError("emu_term_grad failed")
    return grad


def _grow(fn, sym, mpo, B, bra_layout, bra_rp, ket_layout, ket_rp, symmetric, bra_idx, ket_idx):
    from qcmaquis_amd.model import BoundarySpec
    md, Bd = mpo.desc(), B.desc()
    d = [x.desc() for x in (bra_layout.left, bra_layout.phys, bra_layout.right, ket_layout.left, ket_layout.phys,
                            ket_layout.right)]
    Bdat = np.ascontiguousarray(B.data)
    bra_rp = np.ascontiguousarray(bra_rp, np.float64)
    ket_rp = np.ascontiguousarray(ket_rp, np.float64)
    h = fn(sym, C.byref(md), C.byref(Bd), _dp(Bdat), C.byref(d[0]), C.byref(d[1]), C.byref(d[2]), _dp(bra_rp),
           C.byref(d[3]), C.byref(d[4]), C.byref(d[5]), _dp(ket_rp), int(bool(symmetric)))
    if not h:
        raise RuntimeError("emulated growth step failed")
    L = lib()
    nc, aux, ne = C.c_int(), C.c_uint(), C.c_uint64()
    L.emu_boundary_info(h, C.byref(nc), C.byref(aux), C.byref(ne))
    n, a = nc.value, aux.value
    lr = np.zeros(2 * n, np.uint32); off = np.zeros(n * a, np.int64); cs = np.zeros(n * a)
    tp = np.zeros(n * a, np.uint8); tr = np.zeros(n, np.uint8); data = np.zeros(ne.value)
    P = lambda x, t: x.ctypes.data_as(C.POINTER(t))
    L.emu_boundary_export(h, P(lr, C.c_uint32), P(off, C.c_int64), _dp(cs), P(tp, C.c_uint8), P(tr, C.c_uint8), _dp(data))
    L.emu_boundary_free(h)
    cohorts = [dict(lr=(int(lr[2 * c]), int(lr[2 * c + 1])), offsets=off[c * a:(c + 1) * a].tolist(),
                    conj_scales=cs[c * a:(c + 1) * a].tolist(), transposes=tp[c * a:(c + 1) * a].tolist(),
                    tr=int(tr[c])) for c in range(n)]
    return BoundarySpec(bra_idx, ket_idx, a, cohorts, data)


def lstep(sym, mpo, L, bra_layout, bra_rp, ket_layout, ket_rp, symmetric=True):
    return _grow(lib().emu_lstep, sym, mpo, L, bra_layout, bra_rp, ket_layout, ket_rp, symmetric, bra_layout.right,
                 ket_layout.right)


def rstep(sym, mpo, R, bra_layout, bra_rp, ket_layout, ket_rp, symmetric=True):
    return _grow(lib().emu_rstep, sym, mpo, R, bra_layout, bra_rp, ket_layout, ket_rp, symmetric, ket_layout.left,
                 bra_layout.left)
