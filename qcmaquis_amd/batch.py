"""Many independent site problems at once: lock-step Jacobi-Davidson over a batch plan, and
independent two-site DMRG runs swept concurrently (BatchTwoSiteDMRG).

Every iteration applies H to the new search vector of every unconverged instance with one H*psi
of the batch plan (one grouped launch per stage for all instances, engine.HpsiPlanBatch); the
small subspace algebra of each instance runs on the host.  The iteration restates the device
solver (mq_davidson; reference jacobi_davidson::calculate_eigenvalue, ietl/jacobi.h:386-474, with
the gmres = 0 correction t = -r + (r.u / u.u) u, ietl/iteration.h:41-77 stop rule): modified
Gram-Schmidt against the subspace with one refinement when the norm drops below 0.25 of its
start, the Rayleigh matrix, its lowest eigenpair, the residual r = Hu - theta u; an instance stops
when |r| <= tol |theta| or |r| < tol, or after max_iter iterations.
"""
import time

import numpy as np

from . import abi
from .engine import HpsiPlanBatch


def lockstep_davidson(batch, psi0s, tol=1e-8, max_iter=8, precision=abi.MQ_PREC_FP64):
    """lowest eigenpair of every instance of batch (engine.HpsiPlanBatch) from the start vectors
    psi0s; returns (energies, vectors, iterations per instance, number of batched H*psi)"""
    n = len(psi0s)
    V = [[] for _ in range(n)]
    AV = [[] for _ in range(n)]
    t = [np.array(p, np.float64) for p in psi0s]
    theta = np.zeros(n)
    u = [None] * n
    its = np.zeros(n, np.int64)
    active = np.ones(n, bool)
    n_hpsi = 0
    for _ in range(max_iter):
        vec = np.zeros(batch.n)
        seg = batch.split(vec)
        for i in np.flatnonzero(active):
            x = t[i].copy()
            tau = np.linalg.norm(x)
            for v in V[i]:
                x -= (v @ x) * v
            if np.linalg.norm(x) < 0.25 * tau:
                for v in V[i]:
                    x -= (v @ x) * v
            x /= np.linalg.norm(x)
            V[i].append(x)
            seg[i][:] = x
        out = batch.split(batch.apply(vec, precision))
        n_hpsi += 1
        for i in np.flatnonzero(active):
            AV[i].append(out[i].copy())
            Vm, AVm = np.array(V[i]), np.array(AV[i])
            Mr = Vm @ AVm.T
            w, s = np.linalg.eigh(0.5 * (Mr + Mr.T))
            theta[i], sv = w[0], s[:, 0]
            u[i] = sv @ Vm
            ua = sv @ AVm
            r = ua - theta[i] * u[i]
            its[i] += 1
            rn = np.linalg.norm(r)
            if rn <= tol * abs(theta[i]) or rn < tol:
                active[i] = False
                continue
            t[i] = -r + (r @ u[i]) / (u[i] @ u[i]) * u[i]
        if not active.any():
            break
    return theta.copy(), u, its, n_hpsi


class BatchTwoSiteDMRG:
    """Independent two-site DMRG runs (dmrg.TwoSiteDMRG, all on one context) swept concurrently.
    At every step the current two-site problems of all runs that still have one form one batch
    plan; the device lock-step Davidson (mq_davidson_batch: one batched H*psi per iteration) solves
    them together, then each run splits its vector and grows its boundary as TwoSiteDMRG.step does.
    A sweep of a run is its 2L-2 steps in the reference's order (ts_optimize.hpp:128-131); runs of
    different lengths finish their sweeps at different steps."""

    def __init__(self, runs, precision=abi.MQ_PREC_FP64):
        if not runs:
            raise ValueError("no runs")
        ctx = runs[0].ctx
        if any(r.ctx is not ctx for r in runs):
            raise ValueError("the runs of a batch share one context")
        self.runs, self.ctx, self.precision = list(runs), ctx, precision
        self.timing = {"plan": 0.0, "davidson": 0.0, "accept": 0.0}

    def sweep(self):
        """one sweep of every run; returns the runs' last energies"""
        seqs = [[(i, True) for i in range(r.L - 1)] + [(i, False) for i in range(r.L - 2, -1, -1)] for r in self.runs]
        for r in self.runs:
            r.iteration_results = {}
        tol, max_iter = self.runs[0].tol, self.runs[0].max_iter
        for t in range(max(len(q) for q in seqs)):
            act = [k for k in range(len(self.runs)) if t < len(seqs[k])]
            probs = [self.runs[k].site_problem(seqs[k][t][0]) for k in act]
            t0 = time.perf_counter()
            batch = HpsiPlanBatch(self.ctx, [p[:4] for p in probs])
            t1 = time.perf_counter()
            e, vecs, _, _ = batch.davidson([p[4] for p in probs], tol=tol, max_iter=max_iter, precision=self.precision)
            batch.close()
            t2 = time.perf_counter()
            for k, ek, v, p in zip(act, e, vecs, probs):
                i, to_right = seqs[k][t]
                self.runs[k].accept(i, to_right, ek, v, p[3])
            self.timing["plan"] += t1 - t0
            self.timing["davidson"] += t2 - t1
            self.timing["accept"] += time.perf_counter() - t2
        for r in self.runs:
            r.sweep_number += 1
        return [r.energies[-1] for r in self.runs]


def scaled_bonds(phys_list, total, M):
    """bond indices of a random MPS of bond dimension about M: every sector allowed by the total
    quantum numbers, sizes proportional to the number of states the sector can hold (min of the
    left and right counts), scaled down to about M states per bond"""
    from .dmrg import allowed_bonds
    from .model import Index
    out = []
    for b in allowed_bonds(phys_list, total, 1 << 40):
        sizes = np.array(b.sizes, np.float64)
        if sizes.sum() > M:
            sizes = np.maximum(1.0, np.floor(sizes * M / sizes.sum()))
        out.append(Index.from_dict({tuple(int(x) for x in q): int(m) for q, m in zip(b.charges, sizes)}))
    return out


def config4_instance(ctx, L, M, seed, with_host=False):
    """one instance of configuration 4 (many independent site problems): seeded integrals
    (PCG64(seed), L-rank 16), half filling, TwoU1, a right-normalised random MPS of bond dimension
    about M, boundaries grown on the device to the middle two-site problem (sites L/2-1, L/2);
    returns its (mpo, left, right, layout) and a normalised start vector, and with with_host also
    (two-site MPO spec, left boundary, right boundary) on the host"""
    from .dmrg import merge_two, right_normalize, twosite_offsets
    from .engine import DeviceBoundary, DeviceMPO, DeviceMPS, lstep, rstep
    from .model import MPSLayout
    from .qc import QCHamiltonian, make_integrals
    h, eri, ec = make_integrals(L, seed, 16)
    ham = QCHamiltonian(h, eri, ec)
    phys = ham.phys()
    bonds = scaled_bonds([phys] * L, (L // 2, L // 2), M)
    lay = [MPSLayout(bonds[k], phys, bonds[k + 1]) for k in range(L)]
    rng = np.random.default_rng(seed)
    data = [rng.normal(size=x.n) for x in lay]
    right_normalize(lay, data)
    mid = L // 2 - 1
    R = DeviceBoundary.initial(ctx, lay[-1].right, ham.mpo(L - 1).col_dim)
    for k in range(L - 1, mid + 1, -1):
        m = DeviceMPS(ctx, lay[k], data[k])
        R = rstep(ctx, DeviceMPO(ctx, ham.mpo(k)), R, m, m)
    Lb = DeviceBoundary.initial(ctx, lay[0].left, ham.mpo(0).row_dim)
    for k in range(mid):
        m = DeviceMPS(ctx, lay[k], data[k])
        Lb = lstep(ctx, DeviceMPO(ctx, ham.mpo(k)), Lb, m, m)
    lay2, psi2 = merge_two(lay[mid], data[mid], lay[mid + 1], data[mid + 1], ham.phys(True), twosite_offsets(phys))
    item = (DeviceMPO(ctx, ham.mpo(mid, True)), Lb, R, lay2)
    psi0 = psi2 / np.linalg.norm(psi2)
    if with_host:
        return item, psi0, (ham.mpo(mid, True).to_spec(), Lb.download(), R.download())
    return item, psi0

