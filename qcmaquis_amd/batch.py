"""Many independent site problems at once: lock-step Jacobi-Davidson over a batch plan.

Every iteration applies H to the new search vector of every unconverged instance with one H*psi
of the batch plan (one grouped launch per stage for all instances, engine.HpsiPlanBatch); the
small subspace algebra of each instance runs on the host.  The iteration restates the device
solver (mq_davidson; reference jacobi_davidson::calculate_eigenvalue, ietl/jacobi.h:386-474, with
the gmres = 0 correction t = -r + (r.u / u.u) u, ietl/iteration.h:41-77 stop rule): modified
Gram-Schmidt against the subspace with one refinement when the norm drops below 0.25 of its
start, the Rayleigh matrix, its lowest eigenpair, the residual r = Hu - theta u; an instance stops
when |r| <= tol |theta| or |r| < tol, or after max_iter iterations.
"""
import numpy as np

from . import abi


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
