"""Random but charge-consistent DMRG contraction problems.

A "chain" is a sequence of MPO site tensors and MPS site indices with consistent bond charges:
  * MPO bond index b carries an operator charge q_b; the operator of cell (b1, b2) maps physical
    sector s to s + q_b2 - q_b1, so the left boundary of bond b has cohorts (a + q_b, a) and the
    right boundary cohorts (r, r + q_b) -- the selection rules of the reference's schedule
    builders (contractions/create_schedule/*.hpp).
  * Non-zero-charge bond indices come in Hermitian pairs (q, -q); zero-charge indices are
    self-adjoint, so every index is registered and symmetric growth steps give "b-index complete"
    cohorts (mpotensor_detail.h:147-157).
  * SU2U1 chains carry bond spins A and operator spins K with triangle(A, K, A') and entry spins
    equal to the spins of their sectors.
Boundaries are then grown with the growth steps from the all-ones edge boundaries, exactly as a
sweep initialises them (optimize/optimize.h:158-195).
"""
import numpy as np

from . import abi
from .model import Index, MPOSpec, MPSLayout, SiteOp, qadd, qsub

TWOU1_SITE = Index([(1, 1), (1, 0), (0, 1), (0, 0)], [1, 1, 1, 1])
# (1,-1) is the singly occupied state coupled downwards: S(rc) = S(lc) - 1/2 under plain charge
# addition (reference chem/su2u1/model.hpp:47-66)
SU2U1_SITE = Index([(2, 0), (1, 1), (1, -1), (0, 0)], [1, 1, 1, 1])


def fuse_index(a, b):
    """product index a x b (reference operator*, indexing_stable.hpp:425-448)"""
    d = {}
    for ca, sa in zip(a.charges, a.sizes):
        for cb, sb in zip(b.charges, b.sizes):
            c = qadd(ca, cb)
            d[c] = d.get(c, 0) + sa * sb
    return Index.from_dict(d)


def _spin_ok(sym, a, k, c):
    if sym != abi.MQ_SYM_SU2U1:
        return True
    sa, sc = abs(a[1]), abs(c[1])
    return abs(sa - k) <= sc <= sa + k and (sa + k + sc) % 2 == 0


def random_mps_indices(rng, sym, phys_list, total, max_m, min_m=1):
    """Bond indices of an MPS with the given site physical indices and total charge.  Sector sizes
    are random in [min_m, max_m] for charges reachable from both ends."""
    n = len(phys_list)
    fwd = [{(0, 0)}]
    for p in phys_list:
        nxt = set()
        for c in fwd[-1]:
            for s in p.charges:
                q = qadd(c, s)
                if sym == abi.MQ_SYM_SU2U1 and q[1] < 0:
                    continue
                nxt.add(q)
        fwd.append(nxt)
    bwd = [None] * (n + 1)
    bwd[n] = {tuple(total)}
    for k in range(n - 1, -1, -1):
        prev = set()
        for c in bwd[k + 1]:
            for s in phys_list[k].charges:
                prev.add(qsub(c, s))
        bwd[k] = prev
    bonds = []
    for k in range(n + 1):
        cs = sorted(fwd[k] & bwd[k])
        if sym == abi.MQ_SYM_SU2U1:
            cs = [c for c in cs if c[1] >= 0 and (c[0] + c[1]) % 2 == 0]
        if k in (0, n):
            sizes = [1] * len(cs)
        else:
            sizes = [int(rng.integers(min_m, max_m + 1)) for _ in cs]
        bonds.append(Index(cs, sizes))
    # drop sectors that cannot connect to a neighbour bond (keeps every block non-empty)
    for _ in range(2):
        for k in range(1, n):
            keep = []
            for c, s in zip(bonds[k].charges, bonds[k].sizes):
                left_ok = any(bonds[k - 1].position(qsub(c, p)) < len(bonds[k - 1]) for p in phys_list[k - 1].charges)
                right_ok = any(bonds[k + 1].position(qadd(c, p)) < len(bonds[k + 1]) for p in phys_list[k].charges)
                if left_ok and right_ok:
                    keep.append((c, s))
            bonds[k] = Index([c for c, _ in keep], [s for _, s in keep])
    return bonds


def _bond(rng, sym, D, edge):
    """charges, spins, herm, phase of one MPO bond"""
    if edge:
        return [(0, 0)], [0], [0], [1]
    if sym == abi.MQ_SYM_SU2U1:
        pool = [((1, 0), 1), ((2, 0), 0), ((2, 0), 2), ((0, 0), 2), ((1, 0), 1)]
    else:
        pool = [(1, 0), (0, 1), (1, -1), (2, 0), (1, 1), (0, 2)]
    ch, sp, herm, ph = [(0, 0)], [0], [0], [1]
    while len(ch) < D:
        if rng.random() < 0.25 or len(ch) + 2 > D:
            ch.append((0, 0)); sp.append(2 if (sym == abi.MQ_SYM_SU2U1 and rng.random() < 0.5) else 0)
            herm.append(len(ch) - 1); ph.append(1)
            continue
        if sym == abi.MQ_SYM_SU2U1:
            q, A = pool[int(rng.integers(len(pool)))]
        else:
            q, A = pool[int(rng.integers(len(pool)))], 0
        i = len(ch)
        ch += [q, (-q[0], -q[1])]
        sp += [A, A]
        herm += [i + 1, i]
        p = int(rng.choice([1, -1]))
        ph += [p, p]
    return ch, sp, herm, ph


def random_mpo_chain(rng, sym, phys_list, D, density=0.35, row_cells=None):
    """list of MPOSpec, one per site, with bond dimension D in the bulk; row_cells caps the number
    of occupied cells per MPO row (quantum-chemistry MPOs have O(1-10) per row)"""
    n = len(phys_list)
    bonds = [_bond(rng, sym, D, edge=(k == 0 or k == n)) for k in range(n + 1)]
    mpos = []
    for k in range(n):
        phys = phys_list[k]
        lch, lsp, lh, lph = bonds[k]
        rch, rsp, rh, rph = bonds[k + 1]
        ops, terms = [], {}
        cells = []
        for b1 in range(len(lch)):
            for b2 in range(len(rch)):
                dq = qsub(rch[b2], lch[b1])
                for K in ([0, 1, 2] if sym == abi.MQ_SYM_SU2U1 else [0]):
                    if sym == abi.MQ_SYM_SU2U1:
                        if not _spin_ok(sym, (0, lsp[b1]), K, (0, rsp[b2])) or (dq[0] % 2) != (K % 2):
                            continue
                    blocks = []
                    for s, ds in zip(phys.charges, phys.sizes):
                        for t, dt in zip(phys.charges, phys.sizes):
                            if t[0] - s[0] != dq[0]:
                                continue
                            if sym != abi.MQ_SYM_SU2U1 and t != qadd(s, dq):
                                continue
                            if not _spin_ok(sym, s, K, t):
                                continue
                            ents = []
                            for i in range(ds):
                                for j in range(dt):
                                    if rng.random() < 0.8:
                                        e = (i, j, float(rng.normal()))
                                        if sym == abi.MQ_SYM_SU2U1:
                                            e = e + (abs(s[1]), abs(t[1]))
                                        ents.append(e)
                            if ents:
                                blocks.append((s, t, ents))
                    if blocks:
                        cells.append((b1, b2, SiteOp(blocks, spin=K)))
                    break
        rng.shuffle(cells)
        used_l, used_r = set(), set()
        per_row = {}
        chosen = []
        for b1, b2, op in cells:
            full = row_cells is not None and per_row.get(b1, 0) >= row_cells
            if (rng.random() < density and not full) or b1 not in used_l or b2 not in used_r:
                chosen.append((b1, b2, op))
                used_l.add(b1)
                used_r.add(b2)
                per_row[b1] = per_row.get(b1, 0) + 1
        for b1, b2, op in chosen:
            ops.append(op)
            terms.setdefault((b1, b2), []).append((len(ops) - 1, float(rng.uniform(0.5, 1.5))))
        mpos.append(MPOSpec(len(lch), len(rch), ops, terms, left_spin=lsp, right_spin=rsp, left_herm=lh,
                            left_phase=lph, right_herm=rh, right_phase=rph))
    return mpos


def random_mps_data(rng, layout):
    return rng.normal(size=layout.n)


class Chain:
    """MPO chain + MPS site layouts/data with consistent indices."""

    def __init__(self, rng, sym, n_sites, total, max_m, D, phys=None, density=0.35, row_cells=None, min_m=1):
        self.sym = sym
        phys = phys or (SU2U1_SITE if sym == abi.MQ_SYM_SU2U1 else TWOU1_SITE)
        self.phys = [phys] * n_sites
        self.bonds = random_mps_indices(rng, sym, self.phys, total, max_m, min_m)
        self.mpos = random_mpo_chain(rng, sym, self.phys, D, density, row_cells)
        self.layouts = [MPSLayout(self.bonds[k], self.phys[k], self.bonds[k + 1]) for k in range(n_sites)]
        self.data = [random_mps_data(rng, l) for l in self.layouts]
