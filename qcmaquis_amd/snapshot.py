"""Site-problem snapshots: the inputs of one two-site Davidson solve (left and right boundaries,
the two-site MPO, the start vector and its layout) in one HDF5 file, and their replay.

Reference: the `snapshot` parameter of the two-site sweep (ts_optimize.hpp:185-213) saves, at the
listed (2 sweep + direction, site) pairs, the left and right boundaries (`save_boundary`,
boundary.h:422-429), the initial two-site tensor and the two-site MPO; the `shtm` tool
(applications/tools/shtm.cpp) loads them and runs solve_ietl_jcd with no GMRES steps, tol 1e-6
and 9 iterations.  The file layout here is this package's own: HDF5 arrays of the C-ABI
descriptors (the reference writes Boost binary archives, which are tied to its class layouts).

    python -m qcmaquis_amd.snapshot FILE [--tol 1e-6] [--max-iter 9] [--precision fp64|fp32]

replays a snapshot on GPU 0 and prints the plan's statistics and the energy.
"""
import ctypes as C

import numpy as np

from . import abi, results
from .model import BoundarySpec, Index, MPSLayout

_U32_NONE = 0xFFFFFFFF


def _arr(p, n, dtype):
    """copy of n elements behind a ctypes pointer (None for a null pointer)"""
    if not p:
        return None
    return np.ctypeslib.as_array(p, shape=(int(n),)).astype(dtype, copy=True) if n else np.zeros(0, dtype)


class ArrayMPO:
    """An MPO tensor held as the arrays of its C descriptor (mq_mpo_desc)."""

    FIELDS = ("op_spin", "op_block_ptr", "block_charges", "block_entry_ptr", "entry_rc", "entry_spins",
              "entry_coeff", "term_b", "term_op", "term_scale", "left_spin", "right_spin", "left_herm",
              "left_phase", "right_herm", "right_phase")
    TYPES = {"op_spin": np.int32, "op_block_ptr": np.int64, "block_charges": np.int32,
             "block_entry_ptr": np.int64, "entry_rc": np.uint32, "entry_spins": np.int32,
             "entry_coeff": np.float64, "term_b": np.uint32, "term_op": np.uint32, "term_scale": np.float64,
             "left_spin": np.int32, "right_spin": np.int32, "left_herm": np.uint32, "left_phase": np.int32,
             "right_herm": np.uint32, "right_phase": np.int32}

    def __init__(self, row_dim, col_dim, arrays):
        self.row_dim, self.col_dim = int(row_dim), int(col_dim)
        self.a = {k: (None if arrays.get(k) is None else np.ascontiguousarray(arrays[k], self.TYPES[k]))
                  for k in self.FIELDS}

    @classmethod
    def from_desc(cls, d):
        n_ops, nb = d.n_ops, d.op_block_ptr[d.n_ops] if d.n_ops else 0
        ne = d.block_entry_ptr[nb] if nb else 0
        nt, rd, cd = d.n_terms, d.row_dim, d.col_dim
        sizes = {"op_spin": n_ops, "op_block_ptr": n_ops + 1, "block_charges": 4 * nb, "block_entry_ptr": nb + 1,
                 "entry_rc": 2 * ne, "entry_spins": 2 * ne, "entry_coeff": ne, "term_b": 2 * nt, "term_op": nt,
                 "term_scale": nt, "left_spin": rd, "right_spin": cd, "left_herm": rd, "left_phase": rd,
                 "right_herm": cd, "right_phase": cd}
        arrays = {k: _arr(getattr(d, k), sizes[k], cls.TYPES[k]) for k in cls.FIELDS}
        return cls(rd, cd, arrays)

    def desc(self):
        d = abi.MPODesc()
        keep = []

        def ptr(name, ctype):
            a = self.a[name]
            if a is None:
                return C.POINTER(ctype)()
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))
        d.n_ops = int(self.a["op_spin"].size)
        d.op_spin = ptr("op_spin", C.c_int32)
        d.op_block_ptr = ptr("op_block_ptr", C.c_int64)
        d.block_charges = ptr("block_charges", C.c_int32)
        d.block_entry_ptr = ptr("block_entry_ptr", C.c_int64)
        d.entry_rc = ptr("entry_rc", C.c_uint32)
        d.entry_spins = ptr("entry_spins", C.c_int32)
        d.entry_coeff = ptr("entry_coeff", C.c_double)
        d.row_dim, d.col_dim = self.row_dim, self.col_dim
        d.n_terms = int(self.a["term_op"].size)
        d.term_b = ptr("term_b", C.c_uint32)
        d.term_op = ptr("term_op", C.c_uint32)
        d.term_scale = ptr("term_scale", C.c_double)
        d.left_spin = ptr("left_spin", C.c_int32)
        d.right_spin = ptr("right_spin", C.c_int32)
        d.left_herm = ptr("left_herm", C.c_uint32)
        d.left_phase = ptr("left_phase", C.c_int32)
        d.right_herm = ptr("right_herm", C.c_uint32)
        d.right_phase = ptr("right_phase", C.c_int32)
        d._keep = keep
        return d


def _write_boundary(f, path, b):
    """a BoundarySpec (index, cohort tables, data); offsets are stored + 1 (-1 = absent block)"""
    results._write_index(f, path + "/bra", b.bra)
    results._write_index(f, path + "/ket", b.ket)
    n, aux = len(b.cohorts), b.aux_dim
    results.write(f, path + "/aux_dim", np.uint64(aux), "u64")
    results.write(f, path + "/n_cohorts", np.uint64(n), "u64")
    results.write(f, path + "/cohort_lr", np.array([c["lr"] for c in b.cohorts], np.uint64).reshape(-1), "u64")
    results.write(f, path + "/offsets_plus_one",
                  (np.array([c["offsets"] for c in b.cohorts], np.int64).reshape(-1) + 1).astype(np.uint64), "u64")
    results.write(f, path + "/conj_scales", np.array([c.get("conj_scales", [1.0] * aux) for c in b.cohorts],
                                                     np.float64).reshape(-1), "f64")
    results.write(f, path + "/transposes", np.array([c.get("transposes", [0] * aux) for c in b.cohorts],
                                                    np.uint64).reshape(-1), "u64")
    results.write(f, path + "/tr", np.array([c.get("tr", 0) for c in b.cohorts], np.uint64), "u64")
    results.write(f, path + "/data", b.data, "f64")


def _read_boundary(f, path):
    bra, ket = results._read_index(f, path + "/bra"), results._read_index(f, path + "/ket")
    aux, n = int(results.read(f, path + "/aux_dim")), int(results.read(f, path + "/n_cohorts"))
    lr = np.asarray(results.read(f, path + "/cohort_lr"), np.int64).reshape(n, 2) if n else np.zeros((0, 2), int)
    offs = np.asarray(results.read(f, path + "/offsets_plus_one"), np.int64).reshape(n, aux) - 1
    cs = np.asarray(results.read(f, path + "/conj_scales"), np.float64).reshape(n, aux)
    tp = np.asarray(results.read(f, path + "/transposes"), np.int64).reshape(n, aux)
    tr = np.asarray(results.read(f, path + "/tr"), np.int64).reshape(n)
    cohorts = [{"lr": (int(lr[i, 0]), int(lr[i, 1])), "offsets": offs[i].tolist(), "conj_scales": cs[i].tolist(),
                "transposes": tp[i].tolist(), "tr": int(tr[i])} for i in range(n)]
    return BoundarySpec(bra, ket, aux, cohorts, np.asarray(results.read(f, path + "/data"), np.float64).reshape(-1))


def save(path, sym, mpo, left, right, layout, psi, site=-1, twosweep=-1, ecore=0.0):
    """write a site problem: mpo (anything with desc(), e.g. the two-site MPO of a QCHamiltonian),
    left / right (BoundarySpec, or DeviceBoundary to download), layout (MPSLayout of psi), psi
    (host vector in the MPO's basis), the site, 2 sweep + direction (0 left to right, 1 back) and
    the constant energy shift"""
    import os
    if os.path.exists(path):
        os.remove(path)
    left = left if isinstance(left, BoundarySpec) else left.download()
    right = right if isinstance(right, BoundarySpec) else right.download()
    m = mpo if isinstance(mpo, ArrayMPO) else ArrayMPO.from_desc(mpo.desc())
    results.write(path, "/meta/symmetry", np.int32(sym), "i32")
    results.write(path, "/meta/site", np.int32(site), "i32")
    results.write(path, "/meta/twosweep", np.int32(twosweep), "i32")
    results.write(path, "/meta/ecore", np.float64(ecore), "f64")
    results.write(path, "/mpo/row_dim", np.uint64(m.row_dim), "u64")
    results.write(path, "/mpo/col_dim", np.uint64(m.col_dim), "u64")
    for k in ArrayMPO.FIELDS:
        a = m.a[k]
        if a is None:
            continue
        if a.dtype == np.float64:
            results.write(path, "/mpo/" + k, a, "f64")
        elif a.dtype == np.uint32 or a.dtype == np.int64:
            results.write(path, "/mpo/" + k, a.astype(np.uint64), "u64")
        else:
            results.write(path, "/mpo/" + k, a, "i32")
    _write_boundary(path, "/left", left)
    _write_boundary(path, "/right", right)
    for name, idx in (("left_i", layout.left), ("phys_i", layout.phys), ("right_i", layout.right)):
        results._write_index(path, "/initial/" + name, idx)
    results.write(path, "/initial/data", np.asarray(psi, np.float64).reshape(-1), "f64")


def load(path):
    """dict(sym, site, twosweep, ecore, mpo (ArrayMPO), left, right (BoundarySpec), layout, psi)"""
    arrays = {}
    for k in ArrayMPO.FIELDS:
        if results.exists(path, "/mpo/" + k):
            arrays[k] = np.asarray(results.read(path, "/mpo/" + k))
    mpo = ArrayMPO(int(results.read(path, "/mpo/row_dim")), int(results.read(path, "/mpo/col_dim")), arrays)
    layout = MPSLayout(*(results._read_index(path, "/initial/" + n) for n in ("left_i", "phys_i", "right_i")))
    psi = np.asarray(results.read(path, "/initial/data"), np.float64).reshape(-1)
    if psi.size != layout.n:
        raise results.H5Error("snapshot %s: %d start-vector elements for a layout of %d" % (path, psi.size, layout.n))
    return {"sym": int(results.read(path, "/meta/symmetry")), "site": int(results.read(path, "/meta/site")),
            "twosweep": int(results.read(path, "/meta/twosweep")), "ecore": float(results.read(path, "/meta/ecore")),
            "mpo": mpo, "left": _read_boundary(path, "/left"), "right": _read_boundary(path, "/right"),
            "layout": layout, "psi": psi}


def replay(ctx, snap, tol=1e-6, max_iter=9, precision=abi.MQ_PREC_FP64):
    """the snapshot's Davidson solve on the device (shtm: no GMRES steps, tol 1e-6, 9 iterations):
    (energy with the shift, eigenvector, Davidson statistics, plan statistics)"""
    from .engine import DeviceBoundary, DeviceMPO, HpsiPlan
    mpo = DeviceMPO(ctx, snap["mpo"])
    left = DeviceBoundary.upload(ctx, snap["left"])
    right = DeviceBoundary.upload(ctx, snap["right"])
    plan = HpsiPlan(ctx, mpo, left, right, snap["layout"])
    try:
        e, vec, st = plan.davidson(snap["psi"], tol=tol, max_iter=max_iter, precision=precision)
        pst = plan.stats()
    finally:
        plan.close()
    return e + snap["ecore"], vec, st, pst


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="replay a site-problem snapshot on GPU 0")
    ap.add_argument("file")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iter", type=int, default=9)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    args = ap.parse_args(argv)
    from .engine import Context
    snap = load(args.file)
    ctx = Context(0, snap["sym"])
    prec = abi.MQ_PREC_FP64 if args.precision == "fp64" else abi.MQ_PREC_FP32
    e, _, st, pst = replay(ctx, snap, args.tol, args.max_iter, prec)
    print("site %d, two-sweep %d: %d psi elements, %d T tasks, %d cohorts, %d lines, %d terms, %.3e flops per H*psi"
          % (snap["site"], snap["twosweep"], pst["psi_elements"], pst["n_t_tasks"], pst["n_cohorts"],
             pst["n_lines"], pst["n_terms"], pst["flops_total"]))
    print("plan %.3f s; davidson %d iterations (%d FP64 refinement), %.3f s (H*psi %.3f s), residual %.3e"
          % (pst["build_seconds"], st["iterations"], st["refine_iterations"], st["seconds"], st["hpsi_seconds"],
             st["residual_norm"]))
    print("Energy %.12f" % e)
    ctx.close()


if __name__ == "__main__":
    main()
