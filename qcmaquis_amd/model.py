"""Host-side description of the contraction inputs: sector indices, MPO tensors, boundaries and MPS
site tensors, as numpy arrays that pack into the C-ABI descriptors.

Conventions follow the reference (paths relative to dmrg/framework/dmrg/):
  * Index: (charge, size) pairs sorted strictly descending (block_matrix/indexing_stable.hpp:97-305)
  * right-paired MPS block of left sector lc: ls x sum_sigma d_sigma*rs(lc+sigma), columns grouped
    by (sigma, rc) in ProductBasis(phys, right, -sigma+rc) order (indexing_stable.hpp:331-345)
  * left-paired MPS block of right sector rc: sum_sigma d_sigma*ls(rc-sigma) x rs
  * element map between the two: right(ll, roff + ss*rs + rr) = left(loff + ss*ls + ll, rr)
    (block_matrix/detail/alps_detail.hpp:159-168)
"""
import ctypes as C

import numpy as np

from . import abi


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None and a.size else C.POINTER(ctype)()


class Index:
    """Sector index: charges (n x 2 int) and sizes, sorted strictly descending."""

    def __init__(self, charges, sizes):
        ch = [tuple(int(x) for x in c) for c in charges]
        sz = [int(s) for s in sizes]
        order = sorted(range(len(ch)), key=lambda i: ch[i], reverse=True)
        self.charges = [ch[i] for i in order]
        self.sizes = [sz[i] for i in order]
        for a, b in zip(self.charges, self.charges[1:]):
            if not a > b:
                raise ValueError("duplicate charge in index")
        self._pos = {c: i for i, c in enumerate(self.charges)}
        self._ch = np.ascontiguousarray(np.array(self.charges, dtype=np.int32).reshape(-1, 2))
        self._sz = np.ascontiguousarray(np.array(self.sizes, dtype=np.uint32))

    @classmethod
    def from_dict(cls, d):
        return cls(list(d.keys()), list(d.values()))

    def __len__(self):
        return len(self.charges)

    def __eq__(self, o):
        return self.charges == o.charges and self.sizes == o.sizes

    def position(self, c):
        return self._pos.get(tuple(c), len(self.charges))

    def size_of(self, c):
        p = self.position(c)
        return self.sizes[p] if p < len(self.sizes) else 0

    def total(self):
        return int(sum(self.sizes))

    def desc(self):
        return abi.IndexDesc(len(self.charges), _ptr(self._ch, C.c_int32), _ptr(self._sz, C.c_uint32))


def qadd(a, b):
    return (a[0] + b[0], a[1] + b[1])


def qsub(a, b):
    return (a[0] - b[0], a[1] - b[1])


class MPSLayout:
    """Block offsets of one MPS site tensor in both pairings (pure numpy restatement)."""

    def __init__(self, left, phys, right):
        self.left, self.phys, self.right = left, phys, right
        # right-paired column offsets: ProductBasis(phys, right, -sigma + rc)
        self.rpb, rsize = {}, {}
        for s, ds in zip(phys.charges, phys.sizes):
            for rc, rs in zip(right.charges, right.sizes):
                pc = qsub(rc, s)
                self.rpb[(s, rc)] = rsize.get(pc, 0)
                rsize[pc] = rsize.get(pc, 0) + ds * rs
        self.lpb, lsize = {}, {}
        for s, ds in zip(phys.charges, phys.sizes):
            for lc, ls in zip(left.charges, left.sizes):
                pc = qadd(s, lc)
                self.lpb[(s, lc)] = lsize.get(pc, 0)
                lsize[pc] = lsize.get(pc, 0) + ds * ls
        self.rp_cols = [rsize.get(lc, 0) for lc in left.charges]
        self.rp_off = np.concatenate([[0], np.cumsum([ls * c for ls, c in zip(left.sizes, self.rp_cols)])]).astype(np.int64)
        self.lp_rows = [lsize.get(rc, 0) for rc in right.charges]
        self.lp_off = np.concatenate([[0], np.cumsum([r * rs for r, rs in zip(self.lp_rows, right.sizes)])]).astype(np.int64)
        self.n = int(self.rp_off[-1])
        assert self.n == int(self.lp_off[-1])

    def segments(self):
        """(lb, s, rb) for every left sector / sigma with lc + sigma in the right index."""
        for lb, lc in enumerate(self.left.charges):
            for s, sc in enumerate(self.phys.charges):
                rb = self.right.position(qadd(lc, sc))
                if rb < len(self.right):
                    yield lb, s, rb

    def to_left(self, rp):
        lp = np.zeros(self.n)
        for lb, s, rb in self.segments():
            ls, rs, d = self.left.sizes[lb], self.right.sizes[rb], self.phys.sizes[s]
            sc, lc, rc = self.phys.charges[s], self.left.charges[lb], self.right.charges[rb]
            blk_r = rp[self.rp_off[lb]:self.rp_off[lb + 1]].reshape(self.rp_cols[lb], ls).T
            blk_l = lp[self.lp_off[rb]:self.lp_off[rb + 1]].reshape(rs, self.lp_rows[rb]).T
            ro, lo = self.rpb[(sc, rc)], self.lpb[(sc, lc)]
            for ss in range(d):
                blk_l[lo + ss * ls: lo + (ss + 1) * ls, :] = blk_r[:, ro + ss * rs: ro + (ss + 1) * rs]
        return lp

    def to_right(self, lp):
        rp = np.zeros(self.n)
        for lb, s, rb in self.segments():
            ls, rs, d = self.left.sizes[lb], self.right.sizes[rb], self.phys.sizes[s]
            sc, lc, rc = self.phys.charges[s], self.left.charges[lb], self.right.charges[rb]
            blk_r = rp[self.rp_off[lb]:self.rp_off[lb + 1]].reshape(self.rp_cols[lb], ls).T
            blk_l = lp[self.lp_off[rb]:self.lp_off[rb + 1]].reshape(rs, self.lp_rows[rb]).T
            ro, lo = self.rpb[(sc, rc)], self.lpb[(sc, lc)]
            for ss in range(d):
                blk_r[:, ro + ss * rs: ro + (ss + 1) * rs] = blk_l[lo + ss * ls: lo + (ss + 1) * ls, :]
        return rp

    def rp_block(self, data, lb):
        """Right-paired block lb as an (ls x stripe) view (column-major storage)."""
        return data[self.rp_off[lb]:self.rp_off[lb + 1]].reshape(self.rp_cols[lb], self.left.sizes[lb]).T


class SiteOp:
    """Block-sparse site operator.  blocks: list of (left charge, right charge, entries) with
    entries (row, col, coeff[, row_spin, col_spin]); spin = 2K (reference SiteOperator::spin)."""

    def __init__(self, blocks, spin=0):
        self.blocks = [(tuple(a), tuple(b), [tuple(e) for e in ents]) for a, b, ents in blocks]
        self.spin = int(spin)


class MPOSpec:
    """One MPO tensor: terms[(b1, b2)] = [(op index, scale), ...] plus bond spins and Hermitian
    partners (reference MPOTensor + BondProperty, mpotensor.h:45-135, mpotensor_detail.h:139-239)."""

    def __init__(self, row_dim, col_dim, ops, terms, left_spin=None, right_spin=None, left_herm=None,
                 left_phase=None, right_herm=None, right_phase=None):
        self.row_dim, self.col_dim = int(row_dim), int(col_dim)
        self.ops = list(ops)
        self.terms = {tuple(k): list(v) for k, v in terms.items()}
        self.left_spin = None if left_spin is None else np.asarray(left_spin, np.int32)
        self.right_spin = None if right_spin is None else np.asarray(right_spin, np.int32)
        self.left_herm = None if left_herm is None else np.asarray(left_herm, np.uint32)
        self.left_phase = None if left_phase is None else np.asarray(left_phase, np.int32)
        self.right_herm = None if right_herm is None else np.asarray(right_herm, np.uint32)
        self.right_phase = None if right_phase is None else np.asarray(right_phase, np.int32)

    def desc(self):
        keep = []
        op_spin = np.array([o.spin for o in self.ops], np.int32)
        bptr, bch, eptr, erc, esp, eco = [0], [], [0], [], [], []
        has_spins = False
        for o in self.ops:
            for lc, rc, ents in o.blocks:
                bch += [lc[0], lc[1], rc[0], rc[1]]
                for e in ents:
                    erc += [e[0], e[1]]
                    eco.append(e[2])
                    if len(e) > 3:
                        has_spins = True
                        esp += [e[3], e[4]]
                    else:
                        esp += [0, 0]
                eptr.append(len(eco))
            bptr.append(len(bch) // 4)
        keys = sorted(self.terms)
        tb, top, tsc = [], [], []
        for k in keys:
            for op, sc in self.terms[k]:
                tb += [k[0], k[1]]
                top.append(op)
                tsc.append(sc)
        arrs = dict(
            op_spin=op_spin, bptr=np.array(bptr, np.int64), bch=np.array(bch, np.int32), eptr=np.array(eptr, np.int64),
            erc=np.array(erc, np.uint32), esp=np.array(esp, np.int32), eco=np.array(eco, np.float64),
            tb=np.array(tb, np.uint32), top=np.array(top, np.uint32), tsc=np.array(tsc, np.float64))
        keep.append(arrs)
        d = abi.MPODesc()
        d.n_ops = len(self.ops)
        d.op_spin = _ptr(arrs["op_spin"], C.c_int32)
        d.op_block_ptr = _ptr(arrs["bptr"], C.c_int64)
        d.block_charges = _ptr(arrs["bch"], C.c_int32)
        d.block_entry_ptr = _ptr(arrs["eptr"], C.c_int64)
        d.entry_rc = _ptr(arrs["erc"], C.c_uint32)
        d.entry_spins = _ptr(arrs["esp"], C.c_int32) if has_spins else C.POINTER(C.c_int32)()
        d.entry_coeff = _ptr(arrs["eco"], C.c_double)
        d.row_dim, d.col_dim = self.row_dim, self.col_dim
        d.n_terms = len(top)
        d.term_b = _ptr(arrs["tb"], C.c_uint32)
        d.term_op = _ptr(arrs["top"], C.c_uint32)
        d.term_scale = _ptr(arrs["tsc"], C.c_double)
        d.left_spin = _ptr(self.left_spin, C.c_int32)
        d.right_spin = _ptr(self.right_spin, C.c_int32)
        d.left_herm = _ptr(self.left_herm, C.c_uint32)
        d.left_phase = _ptr(self.left_phase, C.c_int32)
        d.right_herm = _ptr(self.right_herm, C.c_uint32)
        d.right_phase = _ptr(self.right_phase, C.c_int32)
        d._keep = keep
        return d


class BoundarySpec:
    """Boundary index plus data on the host.  cohorts: list of dicts with keys lr (bra pos, ket
    pos), offsets (aux_dim int, -1 absent), conj_scales, transposes, tr.  data: non-transposed
    cohorts concatenated in cohort order (each n_blocks x ls x rs, column-major blocks)."""

    def __init__(self, bra, ket, aux_dim, cohorts, data):
        self.bra, self.ket, self.aux_dim = bra, ket, int(aux_dim)
        self.cohorts = cohorts
        self.data = np.ascontiguousarray(np.asarray(data, np.float64))

    def cohort_index(self, lc, rc):
        lb, rb = self.bra.position(lc), self.ket.position(rc)
        for i, c in enumerate(self.cohorts):
            if tuple(c["lr"]) == (lb, rb):
                return i
        return len(self.cohorts)

    def cohort_offsets(self):
        """element offset of each cohort's storage (transposed cohorts alias their partner)"""
        off, pos = [], 0
        sizes = []
        for c in self.cohorts:
            ls, rs = self.bra.sizes[c["lr"][0]], self.ket.sizes[c["lr"][1]]
            nb = int(np.sum(np.asarray(c["offsets"]) > -1))
            sizes.append(0 if c.get("tr", 0) else nb * ls * rs)
            off.append(pos)
            pos += sizes[-1]
        for i, c in enumerate(self.cohorts):
            if c.get("tr", 0):
                l, r = c["lr"]
                j = self.cohort_index(self.ket.charges[r], self.bra.charges[l])
                off[i] = off[j]
        return off

    def desc(self):
        n = len(self.cohorts)
        lr = np.array([c["lr"] for c in self.cohorts], np.uint32).reshape(-1)
        offs = np.array([c["offsets"] for c in self.cohorts], np.int64).reshape(-1)
        cs = np.array([c.get("conj_scales", [1.0] * self.aux_dim) for c in self.cohorts], np.float64).reshape(-1)
        tp = np.array([c.get("transposes", [0] * self.aux_dim) for c in self.cohorts], np.uint8).reshape(-1)
        tr = np.array([c.get("tr", 0) for c in self.cohorts], np.uint8)
        d = abi.BIndexDesc()
        d.bra = self.bra.desc()
        d.ket = self.ket.desc()
        d.aux_dim = self.aux_dim
        d.n_cohorts = n
        d.cohort_lr = _ptr(lr, C.c_uint32)
        d.offsets = _ptr(offs, C.c_int64)
        d.conj_scales = _ptr(cs, C.c_double)
        d.transposes = _ptr(tp, C.c_uint8)
        d.tr = _ptr(tr, C.c_uint8)
        d._keep = [lr, offs, cs, tp, tr, self.bra, self.ket]
        return d


def initial_boundary(index, aux_dim=1):
    """All-ones boundary with one diagonal 1x1... cohort per sector (reference Boundary ctor,
    mp_tensors/boundary.h:257-283)."""
    cohorts, data = [], []
    for i, s in enumerate(index.sizes):
        bs = s * s
        cohorts.append(dict(lr=(i, i), offsets=[b * bs for b in range(aux_dim)]))
        data.append(np.ones(bs * aux_dim))
    return BoundarySpec(index, index, aux_dim, cohorts, np.concatenate(data) if data else np.zeros(0))
