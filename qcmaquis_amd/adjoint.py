"""Integral gradients of the energy of an MPS: dE/dh_pq and dE/d(pq|rs), i.e. the 1- and 2-RDMs
that orbital optimisation consumes, from stored boundaries in one pass over the chain.

E = <psi|H|psi> is linear in every scale of every site MPO.  With the left boundaries grown from
the same site tensors as the right ones, the site problem (L_p, W_p, R_p+1, M_p) of every site p
gives the energy, and the derivative of E with respect to the scale of each term of W_p is that
term's share of <M_p| H_eff |M_p> (mq_hpsi_term_gradient: the H*psi plan's T-GEMMs, the product
Y = L_panel^T M_p, and one dot product of Y with the T tile of every S term, on the device).
The QC builder's gradient layout (QCHamiltonian(..., gradient_layout=True)) then maps these
derivatives back to its integrals (mq_qc_integral_gradient).  At an eigenstate these are the
Hellmann-Feynman forces, dE/dh_pq = gamma_pq + gamma_qp and dE/d(pq|rs) = 1/2 sum over the images
of (pq|rs) of Gamma; the reference measures the same RDMs one expectation value per element
(measurements/tagged_nrankrdm.h:413-661, expval of mps_mpo_ops.h:53-70).

Checkpointed recompute: the left-to-right pass needs the right boundaries R_L .. R_1 in the
reverse of the order they are grown.  Kept right boundaries may hold at most a given fraction of
the bytes of all of them (mem_fraction); the others are grown again from the nearest kept one to
their right when the pass reaches them, and kept on the way when they fit.
"""
import time

import numpy as np

from . import abi
from .dmrg import right_normalize
from .engine import DeviceBoundary, DeviceMPO, DeviceMPS, HpsiPlan, _dp, lstep, rstep
from .rdm import eri_images


def term_gradient(plan, psi):
    """dE/d(scale of every term of the plan's MPO) for E = <psi|H_eff|psi> (mq_hpsi_term_gradient;
    the plan must have been made with the context option MQ_OPT_TERM_SOURCES)"""
    psi = np.ascontiguousarray(psi, np.float64)
    if psi.size != plan.layout.n:
        raise ValueError("psi size does not match the plan")
    n = plan._refs[0].spec.desc().n_terms
    out = np.zeros(n)
    plan.ctx.check(plan.ctx.lib.mq_hpsi_term_gradient(plan.h, _dp(psi), _dp(out), int(n)))
    return out


def _bytes(b):
    return int(b.info().total_elements) * 8


class _RightStore:
    """right boundaries R_k of the chain, kept under a byte budget and grown again on demand;
    peak: the most bytes of boundaries alive at once (kept ones, those in flight and `extra`)"""

    def __init__(self, ctx, mpos, lay, mps, budget, sizes):
        self.ctx, self.mpos, self.lay, self.mps = ctx, mpos, lay, mps
        self.L = len(lay)
        self.budget = budget
        self.sizes = sizes
        self.kept = {self.L: DeviceBoundary.initial(ctx, lay[-1].right, mpos[-1].spec.col_dim)}
        self.sizes[self.L] = _bytes(self.kept[self.L])
        self.kept_bytes = self.sizes[self.L]
        self.peak_kept = self.kept_bytes
        self.peak = self.kept_bytes
        self.grown = 0

    def _held(self, extra, *flight):
        self.peak = max(self.peak, self.kept_bytes + extra + sum(self.sizes[k] for k, b in flight
                                                                   if self.kept.get(k) is not b))

    def _keep(self, k, b):
        if self.kept_bytes + self.sizes[k] > self.budget:
            return False
        self.kept[k] = b
        self.kept_bytes += self.sizes[k]
        self.peak_kept = max(self.peak_kept, self.kept_bytes)
        return True

    def _grow(self, k, right):
        self.grown += 1
        m = self.mps(k)
        return rstep(self.ctx, self.mpos[k], right, m, m)

    def walk(self, q, k, extra, keep):
        """R_k grown from the kept R_q (q > k); keep(j) says whether R_j (k < j < q) may be kept"""
        cur = self.kept[q]
        for j in range(q - 1, k - 1, -1):
            nxt = self._grow(j, cur)
            self.sizes.setdefault(j, _bytes(nxt))
            self._held(extra, (j + 1, cur), (j, nxt))
            if cur is not self.kept.get(j + 1):
                cur.close()
            if j > k and keep(j):
                self._keep(j, nxt)
            cur = nxt
        return cur

    def close(self):
        for b in self.kept.values():
            b.close()
        self.kept.clear()
        self.kept_bytes = 0

    def get(self, k, extra):
        """R_k (and whether the caller owns it); kept boundaries left of k are released (the pass
        moves right)"""
        for j in [j for j in self.kept if j < k]:
            self.kept_bytes -= self.sizes[j]
            self.kept.pop(j).close()
        if k in self.kept:
            return self.kept[k], False
        q = min(j for j in self.kept if j > k)
        return self.walk(q, k, extra, lambda j: True), True


def energy_gradient(ctx, ham, layouts, data, mem_fraction=None, stride=None):
    """(E, dh, deri, stats) of the MPS (layouts, data) for the chain ham (a QCHamiltonian built with
    gradient_layout=True): E with the core energy; dh (L x L) and deri (L^4) hold dE/d(integral) at
    the canonical positions the builder reads (h[i, j], j <= i; eri[i, j, k, l], i >= j, k >= l,
    (i, j) >= (k, l)), zero elsewhere.  The state is brought to right-canonical form and normalised
    first (a copy; the energy is the Rayleigh quotient).  mem_fraction: the kept right boundaries
    hold at most this fraction of the bytes of all right boundaries R_1 .. R_L (None: all are
    kept); their sizes come from a first growth pass that keeps none.  stride: the first keeping
    pass keeps every stride-th boundary that fits (default about 1 / mem_fraction)."""
    if not getattr(ham, "gradient_layout", False):
        raise ValueError("energy_gradient needs QCHamiltonian(..., gradient_layout=True)")
    t0 = time.perf_counter()
    L = ham.L
    lay = list(layouts)
    dat = [np.array(d, np.float64) for d in data]
    right_normalize(lay, dat)
    mpos = [DeviceMPO(ctx, ham.mpo(k)) for k in range(L)]
    cache = {}

    def mps(k):
        if k not in cache:
            cache[k] = DeviceMPS(ctx, lay[k], dat[k])
        return cache[k]

    sizes = {}
    if mem_fraction is None:
        store = _RightStore(ctx, mpos, lay, mps, float("inf"), sizes)
        keep = lambda j: True
    else:
        # sizes first (nothing kept), then the keeping pass under the budget
        probe = _RightStore(ctx, mpos, lay, mps, 0.0, sizes)
        probe.walk(L, 1, 0, lambda j: False).close()
        probe.close()
        stride = stride or max(1, int(round(1.0 / mem_fraction)))
        store = _RightStore(ctx, mpos, lay, mps, mem_fraction * sum(sizes.values()), sizes)
        store.grown = probe.grown
        store.peak = probe.peak
        keep = lambda j: (L - j) % stride == 0
    r1 = store.walk(L, 1, 0, keep)
    if not (keep(1) and store._keep(1, r1)):
        r1.close()
    all_right = sum(sizes.values())
    first_pass_grown = store.grown
    ctx.set_option(abi.MQ_OPT_TERM_SOURCES, 1)
    grads, site_e = [], []
    left = DeviceBoundary.initial(ctx, lay[0].left, mpos[0].spec.row_dim)
    try:
        for p in range(L):
            right, owned = store.get(p + 1, _bytes(left))
            store._held(_bytes(left), (p + 1, right))
            plan = HpsiPlan(ctx, mpos[p], left, right, lay[p])
            g = term_gradient(plan, dat[p])
            plan.close()
            if owned:
                right.close()
            grads.append(g)
            site_e.append(float(g @ ham.term_scales(p)))
            if p + 1 < L:
                m = mps(p)
                nxt = lstep(ctx, mpos[p], left, m, m)
                store._held(_bytes(left) + _bytes(nxt))
                left.close()
                left = nxt
                cache.pop(p).close()
    finally:
        ctx.set_option(abi.MQ_OPT_TERM_SOURCES, 0)
        left.close()
        store.close()
    dh, deri = ham.integral_gradient(grads)
    stats = {"seconds": time.perf_counter() - t0, "all_right_bytes": all_right, "peak_kept_bytes": store.peak_kept,
             "peak_boundary_bytes": store.peak, "right_steps": store.grown, "regrown": store.grown - first_pass_grown,
             "site_energies": np.array(site_e) + ham.ecore, "mem_fraction": mem_fraction}
    return float(np.mean(site_e)) + ham.ecore, dh, deri, stats


def rdms(dh, deri):
    """(gamma, Gamma_sym) from the integral gradients: the spin-summed 1-RDM and the 2-RDM averaged
    over the eight permutations of (pq|rs) (the part the energy of real integrals sees; the
    convention of rdm.two_rdm_sym)"""
    L = dh.shape[0]
    gamma = np.zeros((L, L))
    for i in range(L):
        for j in range(i + 1):
            gamma[i, j] = gamma[j, i] = dh[i, j] if i == j else 0.5 * dh[i, j]
    gam = np.zeros((L,) * 4)
    for i in range(L):
        for j in range(i + 1):
            for k in range(i + 1):
                for l in range(k + 1):
                    if k == i and l > j:
                        continue
                    imgs = eri_images(i, j, k, l)
                    v = 2.0 * deri[i, j, k, l] / len(imgs)
                    for x in imgs:
                        gam[x] = v
    return gamma, gam
