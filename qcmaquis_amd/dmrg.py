"""Two-site DMRG sweeps over the HIP contraction engine (restating optimize/ts_optimize.hpp).

Per two-site step (reference ts_optimize::sweep, ts_optimize.hpp:114-362):
  1. psi2 = TwoSiteTensor(mps[i], mps[i+1])                        (twositetensor.hpp:28-120)
  2. Jacobi-Davidson on the device with the H*psi plan of (L_i, W_i W_i+1, R_i+2)
                                                                   (solver_interface.hpp:30-77)
  3. SVD split with the reference truncation rule (estimate_truncation / svd_truncate,
     block_matrix_algorithms.h:135-244; twositetensor.hpp:153-200)
  4. boundary growth on the device (overlap_mpo_left/right_step)
The SVD split runs on the host (numpy/LAPACK) in this version; boundaries and the contraction live
in HBM.  MPS site tensors are kept as right-paired vectors with their MPSLayout.
"""
import ctypes as C
import os
import time

import numpy as np

from . import abi
from .engine import Context, DeviceBoundary, DeviceMPO, DeviceMPS, HpsiPlan, block_svd, boundary_gram, lstep, rstep
from .model import Index, MPSLayout, qadd, qsub


# ---- block helpers ---------------------------------------------------------------------------
def rp_block(lay, data, lb):
    ls = lay.left.sizes[lb]
    return data[lay.rp_off[lb]:lay.rp_off[lb + 1]].reshape(lay.rp_cols[lb], ls).T


def twosite_offsets(phys):
    """(s1, s2) -> (fused charge, offset inside the fused sector); ProductBasis(phys, phys) order"""
    off, acc = {}, {}
    for s1, d1 in zip(phys.charges, phys.sizes):
        for s2, d2 in zip(phys.charges, phys.sizes):
            c = qadd(s1, s2)
            off[(s1, s2)] = (c, acc.get(c, 0))
            acc[c] = acc.get(c, 0) + d1 * d2
    return off


def merge_two(la, a, lb, b, phys2, off12):
    """psi2(l, s1 s2, r) = sum_m A(l, s1, m) B(m, s2, r) as a right-paired vector of
    MPSLayout(la.left, phys2, lb.right)"""
    lay2 = MPSLayout(la.left, phys2, lb.right)
    out = np.zeros(lay2.n)
    for l_i, lc in enumerate(la.left.charges):
        ls = la.left.sizes[l_i]
        blk2 = out[lay2.rp_off[l_i]:lay2.rp_off[l_i + 1]].reshape(lay2.rp_cols[l_i], ls).T
        A = rp_block(la, a, l_i)
        for s1 in la.phys.charges:
            m = qadd(lc, s1)
            mb = la.right.position(m)
            if mb >= len(la.right):
                continue
            ms = la.right.sizes[mb]
            Aslice = A[:, la.rpb[(s1, m)]: la.rpb[(s1, m)] + ms]
            B = rp_block(lb, b, mb)
            for s2 in lb.phys.charges:
                r = qadd(m, s2)
                rb = lb.right.position(r)
                if rb >= len(lb.right):
                    continue
                rs = lb.right.sizes[rb]
                c, o = off12[(s1, s2)]
                col = lay2.rpb[(c, r)] + o * rs
                blk2[:, col:col + rs] += Aslice @ B[:, lb.rpb[(s2, r)]: lb.rpb[(s2, r)] + rs]
    return lay2, out


def split_matrices(lay2, psi2, phys, off12, sym=abi.MQ_SYM_TWOU1):
    """X_m[(s1, l), (s2, r)] for every middle charge m, rows in ProductBasis(phys, left) order and
    columns in ProductBasis(phys, right, -s2 + r) order; SU2U1 middle charges of negative spin
    are not states and are left out"""
    left, right = lay2.left, lay2.right
    mids = {}
    for lc in left.charges:
        for s1 in phys.charges:
            m = qadd(lc, s1)
            if sym == abi.MQ_SYM_SU2U1 and m[1] < 0:
                continue
            mids.setdefault(m, None)
    mats = {}
    for m in mids:
        rows, rows_n = [], 0
        for s1 in phys.charges:
            lb = left.position(qsub(m, s1))
            if lb < len(left):
                rows.append((s1, lb, rows_n))
                rows_n += left.sizes[lb]
        cols, cols_n = [], 0
        for s2 in phys.charges:
            rb = right.position(qadd(m, s2))
            if rb < len(right):
                cols.append((s2, rb, cols_n))
                cols_n += right.sizes[rb]
        if not rows or not cols:
            continue
        X = np.zeros((rows_n, cols_n))
        for s1, lb, r0 in rows:
            ls = left.sizes[lb]
            blk = rp_block(lay2, psi2, lb)
            for s2, rb, c0 in cols:
                rs = right.sizes[rb]
                c, o = off12[(s1, s2)]
                col = lay2.rpb[(c, right.charges[rb])] + o * rs
                X[r0:r0 + ls, c0:c0 + rs] = blk[:, col:col + rs]
        mats[m] = (X, rows, cols)
    return mats


def estimate_truncation(svals, Mmax, cutoff):
    """keeps per block, truncated weight (estimate_truncation, block_matrix_algorithms.h:135-181):
    all values >= max(cutoff * s_max, s_[Mmax]) are kept"""
    allv = np.sort(np.concatenate([s for s in svals.values()]))[::-1]
    cut = cutoff * allv[0]
    if allv.size > Mmax:
        cut = max(cut, allv[Mmax])
    keeps = {m: int(np.sum(s >= cut)) for m, s in svals.items()}
    tw = float(np.sum(allv[allv < cut] ** 2) / np.sum(allv ** 2))
    return keeps, tw


def truncation_details(svals, Mmax, cutoff):
    """the reference's truncation record (estimate_truncation, block_matrix_algorithms.h:135-181;
    svd_truncate): kept states, truncated weight (discarded s^2 / all s^2), truncated fraction
    (discarded s / all s) and smallest kept ratio (cut / s_max)"""
    keeps, tw = estimate_truncation(svals, Mmax, cutoff)
    allv = np.sort(np.concatenate([s for s in svals.values()]))[::-1]
    cut = cutoff * allv[0]
    if allv.size > Mmax:
        cut = max(cut, allv[Mmax])
    return {"BondDimension": int(sum(keeps.values())), "TruncatedWeight": tw,
            "TruncatedFraction": float(np.sum(allv[allv < cut]) / np.sum(allv)), "SmallestEV": float(cut / allv[0])}


_svd_pool = None


def block_svds(blocks):
    """thin SVD of every block: the blocks in parallel on the host cores, each with one BLAS
    thread (the blocks are at most a few hundred rows, too small for a threaded LAPACK call)"""
    global _svd_pool
    if len(blocks) < 2:
        return {m: np.linalg.svd(X, full_matrices=False) for m, X in blocks.items()}
    from concurrent.futures import ThreadPoolExecutor
    from threadpoolctl import threadpool_limits
    if _svd_pool is None:
        _svd_pool = ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1))
    order = sorted(blocks, key=lambda m: -blocks[m].size)  # largest first
    with threadpool_limits(1):
        res = list(_svd_pool.map(lambda m: np.linalg.svd(blocks[m], full_matrices=False), order))
    out = dict(zip(order, res))
    return {m: out[m] for m in blocks}


def split_two(lay2, psi2, phys, off12, Mmax, cutoff, to_right, details=None, sym=abi.MQ_SYM_TWOU1, svd_fn=None):
    """SVD split of a two-site tensor (sequentially coupled basis for SU2U1); to_right: A = U
    left-normalised, B = S V^T (split_mps_l2r), else A = U S, B = V^T right-normalised
    (split_mps_r2l); details (a dict) receives the truncation record; svd_fn: the block SVD
    ({m: X} -> {m: (U, s, Vt)}; default: LAPACK on the host cores, engine.block_svd on the device)"""
    mats = split_matrices(lay2, psi2, phys, off12, sym)
    svd = (svd_fn or block_svds)({m: X for m, (X, _, _) in mats.items()})
    keeps, tw = estimate_truncation({m: s[1] for m, s in svd.items()}, Mmax, cutoff)
    if details is not None:
        details.update(truncation_details({m: s[1] for m, s in svd.items()}, Mmax, cutoff))
    mid = Index([m for m in keeps if keeps[m] > 0], [keeps[m] for m in keeps if keeps[m] > 0])
    la = MPSLayout(lay2.left, phys, mid)
    lb = MPSLayout(mid, phys, lay2.right)
    a_lp = np.zeros(la.n)
    b_rp = np.zeros(lb.n)
    for m, (U, S, Vt) in svd.items():
        k = keeps[m]
        if k == 0:
            continue
        Uk, Sk, Vk = U[:, :k], S[:k], Vt[:k, :]
        if to_right:
            Am, Bm = Uk, Sk[:, None] * Vk
        else:
            Am, Bm = Uk * Sk[None, :], Vk
        _, rows, cols = mats[m]
        mb = mid.position(m)
        # A left-paired block of right sector m: rows (s1, l) in the same order as X
        a_blk = a_lp[la.lp_off[mb]:la.lp_off[mb + 1]].reshape(k, la.lp_rows[mb]).T
        for s1, lbi, r0 in rows:
            ls = lay2.left.sizes[lbi]
            o = la.lpb[(s1, lay2.left.charges[lbi])]
            a_blk[o:o + ls, :] = Am[r0:r0 + ls, :]
        b_blk = b_rp[lb.rp_off[mb]:lb.rp_off[mb + 1]].reshape(lb.rp_cols[mb], k).T
        for s2, rbi, c0 in cols:
            rs = lay2.right.sizes[rbi]
            o = lb.rpb[(s2, lay2.right.charges[rbi])]
            b_blk[:, o:o + rs] = Bm[:, c0:c0 + rs]
    return la, la.to_right(a_lp), lb, b_rp, tw


def heev_truncation(evals, Mmax, cutoff):
    """keeps per block and the truncation record of heev_truncate (block_matrix_algorithms.h:
    298-354): the positive eigenvalues of the density matrix are replaced by their square roots
    (the others are kept as they are), estimate_truncation runs on those, and the record's
    truncated weight is the truncated fraction (sum of the discarded roots / sum of all)"""
    roots = {m: np.where(w > 0.0, np.sqrt(np.maximum(w, 0.0)), w) for m, w in evals.items()}
    keeps, _ = estimate_truncation(roots, Mmax, cutoff)
    rec = truncation_details(roots, Mmax, cutoff)
    rec["TruncatedWeight"] = rec["TruncatedFraction"]
    return keeps, rec


class _Unpaired:
    """an MPO tensor with its Hermitian-partner tables removed: growth steps then compute every
    MPO index and alias no block to its partner's transpose (the noise term needs all S tiles of
    a bra that differs from the ket; alpha_dm_direct_right schedules with skip = false)"""

    def __init__(self, spec):
        self.spec = spec
        self.row_dim, self.col_dim = spec.row_dim, spec.col_dim

    def desc(self):
        d = self.spec.desc()
        out = type(d).from_buffer_copy(d)
        out.left_herm = out.right_herm = C.POINTER(C.c_uint32)()
        out.left_phase = out.right_phase = C.POINTER(C.c_int32)()
        out._keep = d
        return out


def predict_split(ctx, lay2, psi2, phys, off12, Mmax, cutoff, alpha, to_right, boundary, mpo, details=None,
                  sym=abi.MQ_SYM_TWOU1):
    """density-matrix split of a two-site tensor with the noise term (Engine::predict_split_l2r /
    predict_split_r2l, boundary_ops/prediction.hpp:164-235; twosite_truncation = heev).

    l2r (boundary = left boundary of site i, mpo = the one-site MPO of site i): X_m = psi2 with rows
    (s1, l) and columns (s2, r) per middle charge m; the density matrix is X_m X_m^T + alpha
    sum_b S_b S_b^T, S_b the S tiles of the left growth of the tensor with left index l, physical
    index s1 and the fused right index (s2, r) (alpha_dm, prediction.hpp:40-53 with
    alpha_dm_direct).  Growing that tensor with an identity bra gives S_b as the new boundary's
    blocks; mq_boundary_gram sums their Gram matrices on the device.  heev_truncate keeps its
    leading eigenvectors U: site i = U (left-normalised), site i+1 = U^T X.
    r2l (boundary = right boundary of site i+2, mpo = the one-site MPO of site i+1): the density
    matrix is X_m^T X_m + alpha sum_b S_b^T S_b from the right growth of the tensor with the fused
    left index (s1, l); site i+1 = U^T (right-normalised), site i = X U.
    alpha = 0 is the SVD split (the reference returns split_mps_l2r / split_mps_r2l).
    Returns (layout i, right-paired data i, layout i+1, data i+1, truncated weight)."""
    if alpha == 0.0:
        return split_two(lay2, psi2, phys, off12, Mmax, cutoff, to_right, details, sym)
    mats = split_matrices(lay2, psi2, phys, off12, sym)
    keys = list(mats)
    rows_n = {m: mats[m][0].shape[0] for m in keys}
    cols_n = {m: mats[m][0].shape[1] for m in keys}
    left, right = lay2.left, lay2.right
    if to_right:
        fused = Index(keys, [cols_n[m] for m in keys])  # (s2, r) fused
        ident = Index(keys, [rows_n[m] for m in keys])
        lk, lb_ = MPSLayout(left, phys, fused), MPSLayout(left, phys, ident)
        ket_lp, bra_lp = np.zeros(lk.n), np.zeros(lb_.n)
        for m in keys:
            X, rows, _ = mats[m]
            kb, bb = lk.right.position(m), lb_.right.position(m)
            kblk = ket_lp[lk.lp_off[kb]:lk.lp_off[kb + 1]].reshape(cols_n[m], lk.lp_rows[kb]).T
            bblk = bra_lp[lb_.lp_off[bb]:lb_.lp_off[bb + 1]].reshape(rows_n[m], lb_.lp_rows[bb]).T
            for s1, lbi, r0 in rows:
                ls = left.sizes[lbi]
                o = lk.lpb[(s1, left.charges[lbi])]
                kblk[o:o + ls, :] = X[r0:r0 + ls, :]
                bblk[o:o + ls, r0:r0 + ls] = np.eye(ls)
        bra, ket = DeviceMPS(ctx, lb_, lb_.to_right(bra_lp)), DeviceMPS(ctx, lk, lk.to_right(ket_lp))
        grown = lstep(ctx, mpo, boundary, bra, ket, symmetric=False)
        noise = boundary_gram(grown, 0, alpha, [(m, rows_n[m]) for m in keys])
        dms = {m: mats[m][0] @ mats[m][0].T + noise[m] for m in keys}
    else:
        fused = Index(keys, [rows_n[m] for m in keys])  # (s1, l) fused
        ident = Index(keys, [cols_n[m] for m in keys])
        lk, lb_ = MPSLayout(fused, phys, right), MPSLayout(ident, phys, right)
        ket_rp, bra_rp = np.zeros(lk.n), np.zeros(lb_.n)
        for m in keys:
            X, _, cols = mats[m]
            kb, bb = lk.left.position(m), lb_.left.position(m)
            kblk = lk.rp_block(ket_rp, kb)
            bblk = lb_.rp_block(bra_rp, bb)
            for s2, rbi, c0 in cols:
                rs = right.sizes[rbi]
                o = lk.rpb[(s2, right.charges[rbi])]
                kblk[:, o:o + rs] = X[:, c0:c0 + rs]
                bblk[c0:c0 + rs, o:o + rs] = np.eye(rs)
        bra, ket = DeviceMPS(ctx, lb_, bra_rp), DeviceMPS(ctx, lk, ket_rp)
        # the right step completes partner cohorts in any case: grow with the unpaired MPO
        unpaired = DeviceMPO(ctx, _Unpaired(mpo.spec))
        grown = rstep(ctx, unpaired, boundary, bra, ket, symmetric=False)
        unpaired.close()
        noise = boundary_gram(grown, 1, alpha, [(m, cols_n[m]) for m in keys])
        dms = {m: mats[m][0].T @ mats[m][0] + noise[m] for m in keys}
    grown.close()
    bra.close()
    ket.close()
    eig = {}
    for m, dm in dms.items():
        w, V = np.linalg.eigh(0.5 * (dm + dm.T))
        eig[m] = (w[::-1], V[:, ::-1])
    keeps, rec = heev_truncation({m: w for m, (w, _) in eig.items()}, Mmax, cutoff)
    if details is not None:
        details.update(rec)
    mid = Index([m for m in keys if keeps[m] > 0], [keeps[m] for m in keys if keeps[m] > 0])
    la = MPSLayout(left, phys, mid)
    lb = MPSLayout(mid, phys, right)
    a_lp = np.zeros(la.n)
    b_rp = np.zeros(lb.n)
    for m in keys:
        k = keeps[m]
        if k == 0:
            continue
        X, rows, cols = mats[m]
        U = eig[m][1][:, :k]
        Am, Bm = (U, U.T @ X) if to_right else (X @ U, U.T)
        mb = mid.position(m)
        a_blk = a_lp[la.lp_off[mb]:la.lp_off[mb + 1]].reshape(k, la.lp_rows[mb]).T
        for s1, lbi, r0 in rows:
            ls = left.sizes[lbi]
            o = la.lpb[(s1, left.charges[lbi])]
            a_blk[o:o + ls, :] = Am[r0:r0 + ls, :]
        b_blk = b_rp[lb.rp_off[mb]:lb.rp_off[mb + 1]].reshape(lb.rp_cols[mb], k).T
        for s2, rbi, c0 in cols:
            rs = right.sizes[rbi]
            o = lb.rpb[(s2, right.charges[rbi])]
            b_blk[:, o:o + rs] = Bm[:, c0:c0 + rs]
    return la, la.to_right(a_lp), lb, b_rp, rec["TruncatedWeight"]


# ---- MPS initialisation ------------------------------------------------------------------------
def allowed_bonds(phys_list, total, m_cap, full_right=False):
    """bond indices with per-sector sizes min(#states from the left, #states to the right, m_cap)
    (reference allowed_sectors, mp_tensors/mps_sectors.h); full_right: sizes = #states to the
    right, so right-normalised tensors span the whole right Hilbert space"""
    from math import comb
    n = len(phys_list)
    bonds = []
    for k in range(n + 1):
        d = {}
        for nu in range(0, min(k, total[0]) + 1):
            for nd in range(0, min(k, total[1]) + 1):
                left = comb(k, nu) * comb(k, nd)
                ru, rd = total[0] - nu, total[1] - nd
                if ru > n - k or rd > n - k:
                    continue
                right = comb(n - k, ru) * comb(n - k, rd)
                s = (right if k > 0 else 1) if full_right else min(left, right, m_cap)
                if s > 0:
                    d[(nu, nd)] = s
        bonds.append(Index.from_dict(d))
    return bonds


def allowed_bonds_su2(phys_list, total, m_cap, full_right=False):
    """SU2U1 bond indices: sectors (N, 2S) with 2S >= 0 reachable from both ends under plain charge
    addition, sized by the number of spin-coupling paths (configuration state functions) from the
    left end, the right end or m_cap, as allowed_bonds"""
    n = len(phys_list)
    fwd = [{(0, 0): 1}]
    for p in phys_list:
        nxt = {}
        for c, m in fwd[-1].items():
            for s in p.charges:
                q = qadd(c, s)
                if q[1] >= 0:
                    nxt[q] = nxt.get(q, 0) + m
        fwd.append(nxt)
    bwd = [None] * (n + 1)
    bwd[n] = {tuple(total): 1}
    for k in range(n - 1, -1, -1):
        prev = {}
        for c, m in bwd[k + 1].items():
            for s in phys_list[k].charges:
                q = qsub(c, s)
                if q[1] >= 0:
                    prev[q] = prev.get(q, 0) + m
        bwd[k] = prev
    bonds = []
    for k in range(n + 1):
        d = {}
        for c, left in fwd[k].items():
            right = bwd[k].get(c, 0)
            if right == 0:
                continue
            d[c] = (right if k > 0 else 1) if full_right else min(left, right, m_cap)
        bonds.append(Index.from_dict(d))
    return bonds


def right_normalize(layouts, data, stop=1):
    """right-canonical form of sites stop .. n-1 by QR sweeps from the right (MPS::canonize); the
    remainder of each QR is multiplied into the left neighbour, and with stop == 1 site 0 is
    normalised"""
    n = len(layouts)
    for k in range(n - 1, stop - 1, -1):
        lay, x = layouts[k], data[k]
        new_sizes, Rts = [], []
        blocks = []
        for lb in range(len(lay.left)):
            X = rp_block(lay, x, lb)
            Q, R = np.linalg.qr(X.T)
            blocks.append(Q.T)
            Rts.append(R.T)
            new_sizes.append(Q.shape[1])
        mid = Index(lay.left.charges, new_sizes)
        nl = MPSLayout(mid, lay.phys, lay.right)
        nx = np.zeros(nl.n)
        for lb in range(len(mid)):
            nx[nl.rp_off[lb]:nl.rp_off[lb + 1]] = blocks[lb].T.reshape(-1)
        layouts[k], data[k] = nl, nx
        # absorb R^T into the left neighbour's right bond
        pl, px = layouts[k - 1], data[k - 1]
        lp = pl.to_left(px)
        npl = MPSLayout(pl.left, pl.phys, mid)
        nlp = np.zeros(npl.n)
        for rb in range(len(pl.right)):
            rows = pl.lp_rows[rb]
            blk = lp[pl.lp_off[rb]:pl.lp_off[rb + 1]].reshape(pl.right.sizes[rb], rows).T
            mb = mid.position(pl.right.charges[rb])
            out = blk @ Rts[mb]
            nlp[npl.lp_off[mb]:npl.lp_off[mb + 1]] = out.T.reshape(-1)
        layouts[k - 1], data[k - 1] = npl, npl.to_right(nlp)
    if stop == 1:
        data[0] = data[0] / np.linalg.norm(data[0])
    return layouts, data


def initial_bonds(sym, phys, L, n_up, n_dn, m_cap, full_right=False):
    """bond indices of a state with n_up + n_dn electrons: TwoU1 total charge (n_up, n_dn), SU2U1
    total charge (N, 2S) = (n_up + n_dn, n_up - n_dn)"""
    if sym == abi.MQ_SYM_SU2U1:
        return allowed_bonds_su2([phys] * L, (n_up + n_dn, n_up - n_dn), m_cap, full_right)
    return allowed_bonds([phys] * L, (n_up, n_dn), m_cap, full_right)


def left_normalize(layouts, data, upto):
    """left-canonical form of sites 0 .. upto-1 by QR sweeps from the left (MPS::canonize); the
    remainder of each QR is multiplied into the right neighbour"""
    for k in range(upto):
        lay, x = layouts[k], data[k]
        lp = lay.to_left(x)
        new_sizes, Rs, blocks = [], [], []
        for rb in range(len(lay.right)):
            rows = lay.lp_rows[rb]
            X = lp[lay.lp_off[rb]:lay.lp_off[rb + 1]].reshape(lay.right.sizes[rb], rows).T
            Q, R = np.linalg.qr(X)
            blocks.append(Q)
            Rs.append(R)
            new_sizes.append(Q.shape[1])
        mid = Index(lay.right.charges, new_sizes)
        nl = MPSLayout(lay.left, lay.phys, mid)
        nlp = np.zeros(nl.n)
        for mb in range(len(mid)):
            nlp[nl.lp_off[mb]:nl.lp_off[mb + 1]] = blocks[mb].T.reshape(-1)
        layouts[k], data[k] = nl, nl.to_right(nlp)
        layouts[k + 1], data[k + 1] = _absorb_into_right_neighbour(layouts[k + 1], data[k + 1], mid, Rs)
    return layouts, data


class TwoSiteDMRG:
    """Two-site DMRG for a quantum-chemistry Hamiltonian (QCHamiltonian, TwoU1 or SU2U1).  With
    SU2U1 the two-site vector is recoupled into the pair-coupled basis of the two-site MPO before
    the Davidson solve and back before the split (su2.TwoSiteReduction)."""

    def __init__(self, ham, n_up, n_dn, M, ctx=None, init_m=8, seed=42, tol=1e-8, max_iter=8, cutoff=1e-16,
                 verbose=False, precision=abi.MQ_PREC_FP64, mps=None, resultfile=None, first_sweep=0,
                 device_split=False, snapshots=(), snapshot_dir=".", gmres=0, twosite_truncation="svd",
                 alpha_initial=1e-2, alpha_main=1e-4, alpha_final=1e-8, ngrowsweeps=0, nmainsweeps=0):
        """precision: arithmetic of the Davidson H*psi (MQ_PREC_FP32 = FP32 contraction with an FP64
        refinement of each eigenvector, see mq_davidson); mps: (layouts, data) to start from (e.g.
        results.load_checkpoint) instead of a random state; resultfile: HDF5 result file that every
        sweep appends to (results.ResultArchive); first_sweep: number of the first sweep;
        device_split: the block SVDs of the two-site split on the device (mq_block_svd); snapshots:
        (2 sweep + direction, site) pairs whose site problem is saved before its solve, as
        snapshot_dir/snapshot_<twosweep>_<site>.h5 (snapshot.save; ts_optimize.hpp:185-213); gmres:
        GMRES steps of the Davidson correction equation (ietl_jcd_gmres, default 0);
        twosite_truncation: "svd" (split_two) or "heev" (predict_split: the density matrix with the
        noise term alpha, alpha_initial in the first ngrowsweeps sweeps, alpha_main in the next
        nmainsweeps, alpha_final after; ts_optimize.hpp:275-296, defaults of DmrgParameters.h)"""
        if twosite_truncation not in ("svd", "heev"):
            raise ValueError("twosite_truncation must be 'svd' or 'heev'")
        self.twosite_truncation = twosite_truncation
        self.alphas = (float(alpha_initial), float(alpha_main), float(alpha_final), int(ngrowsweeps), int(nmainsweeps))
        self.gmres = int(gmres)
        self.device_split = device_split
        self.snapshots = {(int(a), int(b)) for a, b in snapshots}
        self.snapshot_dir = snapshot_dir
        self.ham, self.M, self.tol, self.max_iter, self.cutoff = ham, M, tol, max_iter, cutoff
        self.precision = precision
        self.verbose = verbose
        self.sym = getattr(ham, "sym", abi.MQ_SYM_TWOU1)
        self.ctx = ctx or Context(0, self.sym)
        self.L = ham.L
        self.phys = ham.phys()
        self.phys2 = ham.phys(True)
        self.off12 = twosite_offsets(self.phys)
        if mps is None:
            rng = np.random.default_rng(seed)
            bonds = initial_bonds(self.sym, self.phys, self.L, n_up, n_dn, init_m)
            self.lay = [MPSLayout(bonds[k], self.phys, bonds[k + 1]) for k in range(self.L)]
            self.data = [rng.normal(size=l.n) for l in self.lay]
        else:
            self.lay = list(mps[0])
            self.data = [np.array(d, np.float64) for d in mps[1]]
            if len(self.lay) != self.L:
                raise ValueError("the MPS has %d sites, the Hamiltonian %d" % (len(self.lay), self.L))
        right_normalize(self.lay, self.data)
        self.resultfile = resultfile
        self.sweep_number = first_sweep
        self.iteration_results = {}
        self.mpo1 = [DeviceMPO(self.ctx, ham.mpo(k)) for k in range(self.L)]
        self.mpo2 = {}
        # boundaries: left[k] of bond k, right[k] of bond k
        self.left = {0: DeviceBoundary.initial(self.ctx, self.lay[0].left, self.mpo1[0].spec.row_dim)}
        self.right = {self.L: DeviceBoundary.initial(self.ctx, self.lay[-1].right, self.mpo1[-1].spec.col_dim)}
        for k in range(self.L - 1, 1, -1):
            m = DeviceMPS(self.ctx, self.lay[k], self.data[k])
            self.right[k] = rstep(self.ctx, self.mpo1[k], self.right[k + 1], m, m)
        self.energies, self.trunc = [], []
        self.last_residual = None  # Davidson residual norm of the last step
        self.timing = {"davidson": 0.0, "hpsi": 0.0, "plan": 0.0, "svd": 0.0, "grow": 0.0, "merge": 0.0,
                       "mpo": 0.0, "plan_tables": 0.0, "plan_blocks": 0.0, "plan_total": 0.0}
        self.hpsi_flops = 0

    def ts_mpo(self, i):
        if i not in self.mpo2:
            self.mpo2[i] = DeviceMPO(self.ctx, self.ham.mpo(i, True))
        return self.mpo2[i]

    def _merge(self, i):
        """the two-site vector of sites (i, i+1) in the basis the two-site MPO acts on, and the
        SU2U1 recoupling (None for TwoU1)"""
        lay2, psi2 = merge_two(self.lay[i], self.data[i], self.lay[i + 1], self.data[i + 1], self.phys2, self.off12)
        red = None
        if self.sym == abi.MQ_SYM_SU2U1:
            from .su2 import TwoSiteReduction
            red = TwoSiteReduction(lay2, self.phys, self.off12)
            psi2 = red.reduce(psi2)
        return lay2, psi2, red

    def step(self, i, to_right):
        t0 = time.perf_counter()
        lay2, psi2, red = self._merge(i)
        tm = time.perf_counter()
        mpo = self.ts_mpo(i)
        t1 = time.perf_counter()
        self.timing["mpo"] += t1 - tm
        twosweep = 2 * self.sweep_number + (0 if to_right else 1)
        if (twosweep, i) in self.snapshots:
            from . import snapshot
            snapshot.save(os.path.join(self.snapshot_dir, "snapshot_%d_%d.h5" % (twosweep, i)), self.sym, mpo.spec,
                          self.left[i], self.right[i + 2], lay2, psi2, i, twosweep, self.ham.ecore)
        plan = HpsiPlan(self.ctx, mpo, self.left[i], self.right[i + 2], lay2)
        t2 = time.perf_counter()
        e, psi, st = plan.davidson(psi2, tol=self.tol, max_iter=self.max_iter, precision=self.precision,
                                   gmres=self.gmres or None)
        t3 = time.perf_counter()
        pst = plan.stats()
        self.last_residual = st["residual_norm"]
        self.hpsi_flops += pst["flops_total"] * st["iterations"]
        self.timing["plan_tables"] += pst["seconds_tables"]
        self.timing["plan_blocks"] += pst["seconds_blocks"]
        self.timing["plan_total"] += pst["build_seconds"]
        self.timing["plan_plain"] = self.timing.get("plan_plain", 0.0) + pst["seconds_plain"]
        self.timing["hpsi"] += st["hpsi_seconds"]
        plan.close()
        tw, lb, t4, t5 = self.accept(i, to_right, e, psi, lay2, red)
        # boundaries this sweep direction has passed are grown again on the way back: free them
        # (the edge boundaries stay), so the device holds about one boundary per bond
        if to_right and i + 2 < self.L:
            self.right.pop(i + 2, None)
        elif not to_right and i > 0:
            self.left.pop(i, None)
        for k, v in (("merge", tm - t0), ("plan", t2 - t1), ("davidson", t3 - t2), ("svd", t4 - t3), ("grow", t5 - t4)):
            self.timing[k] += v
        if self.verbose:
            print(f"  site {i} {'->' if to_right else '<-'} E={e + self.ham.ecore:.12f} M={sum(lb.left.sizes)} "
                  f"tw={tw:.2e} it={st['iterations']} plan {t2 - t1:.3f} s (tables {pst['seconds_tables']:.3f}, "
                  f"blocks {pst['seconds_blocks']:.3f}, terms {pst['n_terms']}) davidson {t3 - t2:.3f} s "
                  f"svd {t4 - t3:.3f} s grow {t5 - t4:.3f} s", flush=True)
        return e + self.ham.ecore

    def site_problem(self, i):
        """the two-site problem at sites (i, i+1): (two-site MPO, left boundary, right boundary,
        two-site layout, start vector; SU2U1: in the pair-coupled basis)"""
        lay2, psi2, _ = self._merge(i)
        return self.ts_mpo(i), self.left[i], self.right[i + 2], lay2, psi2

    def accept(self, i, to_right, e, psi, lay2, red=None):
        """the second half of a step: split the solved two-site vector psi (energy e, without the
        core energy; in the basis of the two-site MPO, recoupled back for SU2U1) into sites i, i+1,
        grow the boundary in the sweep direction and record the step's results; returns (truncated
        weight, layout of site i+1, split end, growth end)"""
        if self.sym == abi.MQ_SYM_SU2U1:
            if red is None:
                from .su2 import TwoSiteReduction
                red = TwoSiteReduction(lay2, self.phys, self.off12)
            psi = red.unreduce(psi)
        trunc = {}
        if self.twosite_truncation == "heev":
            a0, a1, a2, ngs, nms = self.alphas
            sw = self.sweep_number
            alpha = a0 if sw < ngs else (a1 if sw < ngs + nms else a2)
            boundary, mpo = (self.left[i], self.mpo1[i]) if to_right else (self.right[i + 2], self.mpo1[i + 1])
            la, a, lb, b, tw = predict_split(self.ctx, lay2, psi, self.phys, self.off12, self.M, self.cutoff, alpha,
                                             to_right, boundary, mpo, trunc, self.sym)
        else:
            svd_fn = (lambda blocks: block_svd(self.ctx, blocks)) if self.device_split else None
            la, a, lb, b, tw = split_two(lay2, psi, self.phys, self.off12, self.M, self.cutoff, to_right, trunc,
                                         self.sym, svd_fn)
        t4 = time.perf_counter()
        self.lay[i], self.data[i], self.lay[i + 1], self.data[i + 1] = la, a, lb, b
        if to_right:
            m = DeviceMPS(self.ctx, la, a)
            self.left[i + 1] = lstep(self.ctx, self.mpo1[i], self.left[i], m, m)
        else:
            m = DeviceMPS(self.ctx, lb, b)
            self.right[i + 1] = rstep(self.ctx, self.mpo1[i + 1], self.right[i + 2], m, m)
        t5 = time.perf_counter()
        g = (self.left[i + 1] if to_right else self.right[i + 1]).growth_stats()
        for k in ("tables", "blocks", "upload", "device"):
            self.timing["grow_" + k] = self.timing.get("grow_" + k, 0.0) + g["seconds_" + k]
        self.energies.append(e + self.ham.ecore)
        self.trunc.append(tw)
        # per-step results of the sweep (ts_optimize.hpp:272, 348-351)
        self.iteration_results.setdefault("Energy", []).append(e + self.ham.ecore)
        for k, v in trunc.items():
            self.iteration_results.setdefault(k, []).append(v)
        return tw, lb, t4, t5

    def sweep(self):
        """one full sweep (2L-2 two-site steps, ts_optimize.hpp:128-131); with a result file, the
        sweep's per-step results and the lowest energy are written (dmrg_sim.hpp:88-100)"""
        t0 = time.perf_counter()
        self.iteration_results = {}
        for i in range(0, self.L - 1):
            self.step(i, True)
        for i in range(self.L - 2, -1, -1):
            self.step(i, False)
        dt = time.perf_counter() - t0
        if self.resultfile:
            from .results import ResultArchive
            ar = ResultArchive(self.resultfile)
            ar.append_iteration(self.sweep_number, self.iteration_results)
            emin = min(self.iteration_results["Energy"])
            prev = getattr(self, "_emin_prev", None)
            # the first sweep's change is measured from the largest double, as in the reference
            ar.write_sweep_energy(emin, abs(emin - (prev if prev is not None else np.finfo(np.float64).max)))
            self._emin_prev = emin
        self.sweep_number += 1
        return self.energies[-1], dt

    def checkpoint(self, directory, site=-1):
        """the MPS and the status (last finished sweep) in the reference's checkpoint layout
        (sim.hpp:223-235; results.save_checkpoint)"""
        from .results import save_checkpoint
        save_checkpoint(directory, self.lay, self.data, self.sweep_number - 1, site)


def _absorb_into_right_neighbour(lay, data, mid, Rs):
    """site right of a moved centre: its right-paired block of left sector m becomes
    R_m (k x rs) @ block (rs x cols) (MPSTensor::multiply_from_left, mpstensor.hpp:345)"""
    nl = MPSLayout(mid, lay.phys, lay.right)
    out = np.zeros(nl.n)
    for mb in range(len(mid)):
        lb = lay.left.position(mid.charges[mb])
        blk = rp_block(lay, data, lb)
        out[nl.rp_off[mb]:nl.rp_off[mb + 1]] = (Rs[mb] @ blk).T.reshape(-1)
    return nl, out


def _absorb_into_left_neighbour(lay, data, mid, Ls):
    """site left of a moved centre: its left-paired block of right sector m becomes
    block (rows x ls) @ L_m (ls x k) (MPSTensor::multiply_from_right)"""
    lp = lay.to_left(data)
    nl = MPSLayout(lay.left, lay.phys, mid)
    nlp = np.zeros(nl.n)
    for mb in range(len(mid)):
        rb = lay.right.position(mid.charges[mb])
        rows = lay.lp_rows[rb]
        blk = lp[lay.lp_off[rb]:lay.lp_off[rb + 1]].reshape(lay.right.sizes[rb], rows).T
        nlp[nl.lp_off[mb]:nl.lp_off[mb + 1]] = (blk @ Ls[mb]).T.reshape(-1)
    return nl, nl.to_right(nlp)


def grow_sweep_step(ctx, lay, psi, Mmax, cutoff, alpha, to_right, boundary, mpo, details=None):
    """single-site subspace expansion (Engine::grow_l2r_sweep / grow_r2l_sweep, engine.hpp:82-118):
    l2r: the density matrix of the solved site tensor psi, left-paired (rows (s, l), columns the
    right bond), X_m X_m^T + alpha sum_b S_b S_b^T with S_b the S tiles of the left growth of psi
    (predict_new_state_l2r_sweep, prediction.hpp:70-92, alpha_dm); heev_truncate keeps its leading
    eigenvectors U, which span up to all (s, l) rows of a sector, so the bond can grow; the site
    becomes U and the right neighbour takes U^T psi (predict_lanczos_l2r_sweep,
    prediction.hpp:94-108).  r2l: X_m^T X_m + alpha sum_b S_b^T S_b per left sector of the
    right-paired psi from the right growth, the site becomes U^T and the left neighbour takes
    psi U (prediction.hpp:110-148).  The S tiles come from a growth step with an identity bra and
    mq_boundary_gram, as in predict_split.  Returns (new layout, right-paired data, middle index,
    carry per middle sector, truncated weight)."""
    if to_right:
        lp = lay.to_left(psi)
        keys = [(rb, c) for rb, c in enumerate(lay.right.charges) if lay.lp_rows[rb] and lay.right.sizes[rb]]
        Xs = {c: lp[lay.lp_off[rb]:lay.lp_off[rb + 1]].reshape(lay.right.sizes[rb], lay.lp_rows[rb]).T
              for rb, c in keys}
        dims = {c: lay.lp_rows[rb] for rb, c in keys}
    else:
        keys = [(lb, c) for lb, c in enumerate(lay.left.charges) if lay.rp_cols[lb] and lay.left.sizes[lb]]
        Xs = {c: rp_block(lay, psi, lb) for lb, c in keys}
        dims = {c: lay.rp_cols[lb] for lb, c in keys}
    charges = [c for _, c in keys]
    if alpha != 0.0 and charges:
        ident = Index(charges, [dims[c] for c in charges])
        if to_right:
            bl = MPSLayout(lay.left, lay.phys, ident)
            b_lp = np.zeros(bl.n)
            for mb, c in enumerate(bl.right.charges):
                b_lp[bl.lp_off[mb]:bl.lp_off[mb + 1]] = np.eye(dims[c]).reshape(-1)
            bra, ket = DeviceMPS(ctx, bl, bl.to_right(b_lp)), DeviceMPS(ctx, lay, psi)
            grown = lstep(ctx, mpo, boundary, bra, ket, symmetric=False)
        else:
            bl = MPSLayout(ident, lay.phys, lay.right)
            b_rp = np.zeros(bl.n)
            for mb, c in enumerate(bl.left.charges):
                b_rp[bl.rp_off[mb]:bl.rp_off[mb + 1]] = np.eye(dims[c]).reshape(-1)
            bra, ket = DeviceMPS(ctx, bl, b_rp), DeviceMPS(ctx, lay, psi)
            unpaired = DeviceMPO(ctx, _Unpaired(mpo.spec))
            grown = rstep(ctx, unpaired, boundary, bra, ket, symmetric=False)
            unpaired.close()
        noise = boundary_gram(grown, 0 if to_right else 1, alpha, [(c, dims[c]) for c in charges])
        grown.close()
        bra.close()
        ket.close()
    else:
        noise = {c: 0.0 for c in charges}
    eig = {}
    for c in charges:
        X = Xs[c]
        dm = (X @ X.T if to_right else X.T @ X) + noise[c]
        w, V = np.linalg.eigh(0.5 * (dm + dm.T))
        eig[c] = (w[::-1], V[:, ::-1])
    keeps, rec = heev_truncation({c: w for c, (w, _) in eig.items()}, Mmax, cutoff)
    if details is not None:
        details.update(rec)
    mid = Index([c for c in charges if keeps[c] > 0], [keeps[c] for c in charges if keeps[c] > 0])
    carry = []
    if to_right:
        na = MPSLayout(lay.left, lay.phys, mid)
        a_lp = np.zeros(na.n)
        for mb, c in enumerate(mid.charges):
            U = eig[c][1][:, :keeps[c]]
            a_lp[na.lp_off[mb]:na.lp_off[mb + 1]] = U.T.reshape(-1)
            carry.append(U.T @ Xs[c])
        return na, na.to_right(a_lp), mid, carry, rec["TruncatedWeight"]
    nb = MPSLayout(mid, lay.phys, lay.right)
    b_rp = np.zeros(nb.n)
    for mb, c in enumerate(mid.charges):
        U = eig[c][1][:, :keeps[c]]
        b_rp[nb.rp_off[mb]:nb.rp_off[mb + 1]] = U.reshape(-1)  # block U^T (k x cols), column-major
        carry.append(Xs[c] @ U)
    return nb, b_rp, mid, carry, rec["TruncatedWeight"]


class SingleSiteDMRG:
    """Single-site DMRG (ss_optimize.hpp): each step solves the one-site problem
    H_eff(L_i, W_i, R_{i+1}) with the device Jacobi-Davidson, then moves the orthogonality centre
    and grows the left or right boundary of the moved bond on the device.  The move is the
    reference's subspace expansion (grow_sweep_step: the density matrix with the noise term alpha,
    truncated to M states, so bonds grow from init_m up to M; alpha_initial in the first ngrowsweeps
    sweeps, alpha_main in the next nmainsweeps, alpha_final after, ss_optimize.hpp:147-190) or, with
    expansion=False, an SVD of the site tensor without truncation (normalize_left /
    normalize_right, mpstensor.hpp:242), for which the bonds the state starts from are the
    variational space."""

    def __init__(self, ham, n_up, n_dn, M, ctx=None, seed=42, tol=1e-8, max_iter=8, precision=abi.MQ_PREC_FP64,
                 expansion=False, init_m=None, cutoff=1e-16, alpha_initial=1e-2, alpha_main=1e-4, alpha_final=1e-8,
                 ngrowsweeps=0, nmainsweeps=0):
        self.M, self.cutoff, self.expansion = M, cutoff, bool(expansion)
        self.alphas = (float(alpha_initial), float(alpha_main), float(alpha_final), int(ngrowsweeps), int(nmainsweeps))
        self.sweep_number = 0
        self.trunc = []
        self.ham, self.tol, self.max_iter, self.precision = ham, tol, max_iter, precision
        self.sym = getattr(ham, "sym", abi.MQ_SYM_TWOU1)
        self.ctx = ctx or Context(0, self.sym)
        self.L = ham.L
        self.phys = ham.phys()
        rng = np.random.default_rng(seed)
        bonds = initial_bonds(self.sym, self.phys, self.L, n_up, n_dn, M if init_m is None else init_m)
        self.lay = [MPSLayout(bonds[k], self.phys, bonds[k + 1]) for k in range(self.L)]
        self.data = [rng.normal(size=l.n) for l in self.lay]
        right_normalize(self.lay, self.data)
        self.mpo1 = [DeviceMPO(self.ctx, ham.mpo(k)) for k in range(self.L)]
        self.left = {0: DeviceBoundary.initial(self.ctx, self.lay[0].left, self.mpo1[0].spec.row_dim)}
        self.right = {self.L: DeviceBoundary.initial(self.ctx, self.lay[-1].right, self.mpo1[-1].spec.col_dim)}
        for k in range(self.L - 1, 0, -1):
            m = DeviceMPS(self.ctx, self.lay[k], self.data[k])
            self.right[k] = rstep(self.ctx, self.mpo1[k], self.right[k + 1], m, m)
        self.energies = []

    def step(self, i, to_right):
        plan = HpsiPlan(self.ctx, self.mpo1[i], self.left[i], self.right[i + 1], self.lay[i])
        e, psi, _ = plan.davidson(self.data[i], tol=self.tol, max_iter=self.max_iter, precision=self.precision)
        plan.close()
        lay = self.lay[i]
        if self.expansion:
            a0, a1, a2, ngs, nms = self.alphas
            sw = self.sweep_number
            alpha = a0 if sw < ngs else (a1 if sw < ngs + nms else a2)
            bnd = self.left[i] if to_right else self.right[i + 1]
            nl, nd, mid, carry, tw = grow_sweep_step(self.ctx, lay, psi, self.M, self.cutoff, alpha, to_right, bnd,
                                                     self.mpo1[i])
            self.trunc.append(tw)
            self.lay[i], self.data[i] = nl, nd
            m = DeviceMPS(self.ctx, nl, nd)
            if to_right:
                self.lay[i + 1], self.data[i + 1] = _absorb_into_right_neighbour(self.lay[i + 1], self.data[i + 1], mid,
                                                                                 carry)
                self.left[i + 1] = lstep(self.ctx, self.mpo1[i], self.left[i], m, m)
            else:
                self.lay[i - 1], self.data[i - 1] = _absorb_into_left_neighbour(self.lay[i - 1], self.data[i - 1], mid,
                                                                                carry)
                self.right[i] = rstep(self.ctx, self.mpo1[i], self.right[i + 1], m, m)
        elif to_right:
            lp = lay.to_left(psi)
            us, carry, sizes = {}, [], []
            for rb in range(len(lay.right)):
                rows, rs = lay.lp_rows[rb], lay.right.sizes[rb]
                X = lp[lay.lp_off[rb]:lay.lp_off[rb + 1]].reshape(rs, rows).T
                U, S, Vt = np.linalg.svd(X, full_matrices=False)
                us[rb] = U
                carry.append(S[:, None] * Vt)
                sizes.append(U.shape[1])
            mid = Index(lay.right.charges, sizes)
            na = MPSLayout(lay.left, lay.phys, mid)
            a_lp = np.zeros(na.n)
            for mb in range(len(mid)):
                a_lp[na.lp_off[mb]:na.lp_off[mb + 1]] = us[mb].T.reshape(-1)
            self.lay[i], self.data[i] = na, na.to_right(a_lp)
            self.lay[i + 1], self.data[i + 1] = _absorb_into_right_neighbour(self.lay[i + 1], self.data[i + 1], mid, carry)
            m = DeviceMPS(self.ctx, self.lay[i], self.data[i])
            self.left[i + 1] = lstep(self.ctx, self.mpo1[i], self.left[i], m, m)
        else:
            vts, carry, sizes = {}, [], []
            for lb in range(len(lay.left)):
                X = rp_block(lay, psi, lb)
                U, S, Vt = np.linalg.svd(X, full_matrices=False)
                vts[lb] = Vt
                carry.append(U * S[None, :])
                sizes.append(Vt.shape[0])
            mid = Index(lay.left.charges, sizes)
            nb = MPSLayout(mid, lay.phys, lay.right)
            b_rp = np.zeros(nb.n)
            for mb in range(len(mid)):
                b_rp[nb.rp_off[mb]:nb.rp_off[mb + 1]] = vts[mb].T.reshape(-1)
            self.lay[i], self.data[i] = nb, b_rp
            self.lay[i - 1], self.data[i - 1] = _absorb_into_left_neighbour(self.lay[i - 1], self.data[i - 1], mid, carry)
            m = DeviceMPS(self.ctx, self.lay[i], self.data[i])
            self.right[i] = rstep(self.ctx, self.mpo1[i], self.right[i + 1], m, m)
        self.energies.append(e + self.ham.ecore)
        return e + self.ham.ecore

    def sweep(self):
        """left to right over sites 0..L-2, then right to left over L-1..1 (ss_optimize.hpp:100-190)"""
        t0 = time.perf_counter()
        for i in range(0, self.L - 1):
            self.step(i, True)
        for i in range(self.L - 1, 0, -1):
            self.step(i, False)
        self.sweep_number += 1
        return self.energies[-1], time.perf_counter() - t0
