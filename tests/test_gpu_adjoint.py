"""The MPO-term adjoint and the integral gradients (1- and 2-RDMs) on the device.

* mq_hpsi_term_gradient at every site of an 8-orbital SU2U1 chain against the table emulator (all
  terms, 1e-11 relative) and the oracle's H*psi of single-term MPOs (sampled terms); the terms'
  gradients recombine to <psi|H_eff|psi>;
* a converged 6-orbital SU2U1 state (untruncated, = FCI): the integral gradients equal the FCI
  RDMs' (dE/dh_pq = gamma_pq + gamma_qp, dE/d(pq|rs) = 1/2 sum over images of Gamma) and the
  energy is linear-homogeneous in the integrals (sum theta dE/dtheta = E - E_core);
* 12 orbitals, M = 128: through a test-only torch.autograd.Function, the gradients match central
  finite differences of E(h, eri) = <psi|H(h, eri)|psi> (32 random integral entries, step 1e-5,
  1e-6 absolute), the derivative-MPO expectation values of qcmaquis_amd.rdm, and the gradients
  computed with the kept right boundaries capped at 25 % of their total bytes (the rest regrown).
"""
import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.adjoint import energy_gradient, rdms, term_gradient
from qcmaquis_amd.dmrg import TwoSiteDMRG, initial_bonds, right_normalize
from qcmaquis_amd.engine import DeviceBoundary, DeviceMPO, DeviceMPS, HpsiPlan, lstep, rstep
from qcmaquis_amd.model import MPOSpec, MPSLayout
from qcmaquis_amd.qc import QCHamiltonian, make_integrals
from qcmaquis_amd.rdm import eri_gradient, expval, h_gradient
from tests import emu
from tests.fci_rdm import fci_ground

pytestmark = pytest.mark.gpu
SU2 = abi.MQ_SYM_SU2U1


def single_term(spec, cell, k):
    op, _ = spec.terms[cell][k]
    return MPOSpec(spec.row_dim, spec.col_dim, spec.ops, {cell: [(op, 1.0)]}, spec.left_spin, spec.right_spin,
                   spec.left_herm, spec.left_phase, spec.right_herm, spec.right_phase)


def canonical(L):
    hs = [(i, j) for i in range(L) for j in range(i + 1)]
    es = [(i, j, k, l) for i in range(L) for j in range(i + 1) for k in range(i + 1) for l in range(k + 1)
          if not (k == i and l > j)]
    return hs, es


def test_term_gradient_vs_emulator_and_oracle(ctxsu2, oracle_lib):
    L = 8
    print("term gradients, 8 orbitals SU2U1, M=24", flush=True)
    h, eri, _ = make_integrals(L, 1001, 16)
    ham = QCHamiltonian(h, eri, symmetry=SU2, gradient_layout=True)
    phys = ham.phys()
    bonds = initial_bonds(SU2, phys, L, 4, 4, 24)
    lay = [MPSLayout(bonds[k], phys, bonds[k + 1]) for k in range(L)]
    rng = np.random.default_rng(9)
    data = [rng.normal(size=l.n) for l in lay]
    right_normalize(lay, data)
    mpos = [DeviceMPO(ctxsu2, ham.mpo(k)) for k in range(L)]
    rights = {L: DeviceBoundary.initial(ctxsu2, lay[-1].right, 1)}
    for k in range(L - 1, 0, -1):
        m = DeviceMPS(ctxsu2, lay[k], data[k])
        rights[k] = rstep(ctxsu2, mpos[k], rights[k + 1], m, m)
    left = DeviceBoundary.initial(ctxsu2, lay[0].left, 1)
    ctxsu2.set_option(abi.MQ_OPT_TERM_SOURCES, 1)
    try:
        for p in range(L):
            plan = HpsiPlan(ctxsu2, mpos[p], left, rights[p + 1], lay[p])
            g = term_gradient(plan, data[p])
            e = float(data[p] @ plan.apply(data[p]))
            plan.close()
            Lo, Ro = left.download(), rights[p + 1].download()
            want = emu.term_grad(SU2, ham.mpo(p), Lo, Ro, lay[p], data[p])
            err = np.max(np.abs(g - want)) / max(np.max(np.abs(want)), 1e-300)
            spec = ham.mpo(p).to_spec()
            cells = [(c, k) for c, lst in spec.terms.items() for k in range(len(lst))]
            pick = rng.choice(len(cells), size=min(8, len(cells)), replace=False)
            for t in pick:
                hv = oracle_lib.hpsi(SU2, single_term(spec, *cells[t]), Lo, Ro, lay[p], data[p], nthreads=1)
                assert abs(g[t] - float(data[p] @ hv)) <= 1e-11 * max(1.0, np.max(np.abs(want)))
            de = abs(float(g @ ham.term_scales(p)) - e)
            print(f"  site {p}: {g.size} terms, rel err vs emulator {err:.2e}, |sum scale*grad - <psi|H psi>| {de:.2e}",
                  flush=True)
            assert err < 1e-11 and de < 1e-10 * max(1.0, abs(e))
            if p + 1 < L:
                m = DeviceMPS(ctxsu2, lay[p], data[p])
                left = lstep(ctxsu2, mpos[p], left, m, m)
    finally:
        ctxsu2.set_option(abi.MQ_OPT_TERM_SOURCES, 0)


def rdm_gradients(gamma, Gamma, L):
    from qcmaquis_amd.rdm import eri_images
    hs, es = canonical(L)
    dh = np.zeros((L, L))
    deri = np.zeros((L,) * 4)
    for i, j in hs:
        dh[i, j] = gamma[i, j] + gamma[j, i] if i != j else gamma[i, i]
    for e in es:
        deri[e] = 0.5 * sum(Gamma[x] for x in eri_images(*e))
    return dh, deri


def test_adjoint_rdms_equal_fci(ctxsu2):
    L, N = 6, 6
    print("adjoint RDMs, 6 orbitals SU2U1, untruncated", flush=True)
    h, eri, _ = make_integrals(L, 17, 6)
    ecore = 0.5
    ham = QCHamiltonian(h, eri, ecore, symmetry=SU2, gradient_layout=True)
    d = TwoSiteDMRG(ham, N // 2, N // 2, M=400, ctx=ctxsu2, tol=1e-12, max_iter=20)
    for s in range(5):
        e, _ = d.sweep()
    E, dh, deri, st = energy_gradient(ctxsu2, ham, d.lay, d.data)
    e_fci, gamma, Gamma = fci_ground(h, eri, N // 2, N // 2, ecore)
    rh, re = rdm_gradients(gamma, Gamma, L)
    euler = float(np.sum(dh * h) + np.sum(deri * eri))
    print(f"  E {E:.12f} (sweep {e:.12f}, FCI {e_fci:.12f}); max |dE/dh - FCI| {np.max(np.abs(dh - rh)):.2e}, "
          f"max |dE/deri - FCI| {np.max(np.abs(deri - re)):.2e}; Euler residual {abs(euler - (E - ecore)):.2e}; "
          f"{st['seconds']:.2f} s", flush=True)
    assert abs(E - e_fci) < 1e-9
    assert np.max(np.abs(st["site_energies"] - E)) < 1e-9
    assert abs(euler - (E - ecore)) < 1e-9
    assert np.max(np.abs(dh - rh)) < 1e-6 and np.max(np.abs(deri - re)) < 1e-6
    gam, Gam = rdms(dh, deri)
    assert abs(np.trace(gam) - N) < 1e-8


class _State:
    def __init__(self, ctx, lay, data):
        self.ctx, self.lay = ctx, list(lay)
        self.data = [np.array(x) for x in data]
        right_normalize(self.lay, self.data)


def _energy_function():
    import torch

    class QCEnergy(torch.autograd.Function):
        """test-only binding: E(h, eri) = <psi|H(h, eri)|psi> of a fixed normalised SU2U1 MPS
        (forward: one left-boundary growth over the chain; backward: the adjoint)"""

        @staticmethod
        def forward(fctx, h, eri, state, mem_fraction):
            ham = QCHamiltonian(h.detach().numpy(), eri.detach().numpy(), symmetry=SU2)
            fctx.state, fctx.mem_fraction = state, mem_fraction
            fctx.save_for_backward(h, eri)
            return torch.tensor(expval(state.ctx, state.lay, state.data, ham), dtype=torch.float64)

        @staticmethod
        def backward(fctx, grad_e):
            h, eri = fctx.saved_tensors
            st = fctx.state
            ham = QCHamiltonian(h.detach().numpy(), eri.detach().numpy(), symmetry=SU2, gradient_layout=True)
            _, dh, deri, info = energy_gradient(st.ctx, ham, st.lay, st.data, mem_fraction=fctx.mem_fraction)
            st.last_info = info
            return grad_e * torch.from_numpy(dh), grad_e * torch.from_numpy(deri), None, None

    return QCEnergy


def test_adjoint_fd_cap_and_rdm_paths(ctxsu2):
    import torch
    L, N, M = 12, 12, 128
    print(f"adjoint, {L} orbitals SU2U1, M={M}: finite differences, 25% boundary cap", flush=True)
    h, eri, _ = make_integrals(L, 2024, 16)
    ham = QCHamiltonian(h, eri, symmetry=SU2)
    d = TwoSiteDMRG(ham, N // 2, N // 2, M=M, ctx=ctxsu2, tol=1e-8, max_iter=8)
    for s in range(2):
        e, dt = d.sweep()
        print(f"  sweep {s}: E={e:.10f} ({dt:.2f} s)", flush=True)
    state = _State(ctxsu2, d.lay, d.data)
    QCEnergy = _energy_function()
    ht = torch.tensor(h, requires_grad=True)
    et = torch.tensor(eri, requires_grad=True)
    E = QCEnergy.apply(ht, et, state, 0.25)
    gh, ge = torch.autograd.grad(E, (ht, et))
    gh, ge = gh.numpy(), ge.numpy()
    capped = state.last_info
    print(f"  E {E.item():.10f}; kept right boundaries at most {capped['peak_kept_bytes'] / 2**20:.1f} MiB of "
          f"{capped['all_right_bytes'] / 2**20:.1f} MiB, {capped['regrown']} regrown; {capped['seconds']:.2f} s",
          flush=True)
    assert capped["peak_kept_bytes"] <= 0.25 * capped["all_right_bytes"] and capped["regrown"] > 0
    # the same gradients with every right boundary kept
    _, dh, deri, full = energy_gradient(ctxsu2, QCHamiltonian(h, eri, symmetry=SU2, gradient_layout=True), state.lay,
                                        state.data)
    scale = max(np.max(np.abs(dh)), np.max(np.abs(deri)))
    assert np.max(np.abs(dh - gh)) <= 1e-11 * scale and np.max(np.abs(deri - ge)) <= 1e-11 * scale
    assert full["regrown"] == 0
    # central finite differences on 32 random canonical entries
    hs, es = canonical(L)
    rng = np.random.default_rng(5)
    picks = [("h", hs[i]) for i in rng.choice(len(hs), 12, replace=False)] + \
            [("eri", es[i]) for i in rng.choice(len(es), 20, replace=False)]
    step = 1e-5
    worst = 0.0
    for kind, idx in picks:
        vals = []
        for sgn in (1.0, -1.0):
            hp, ep = h.copy(), eri.copy()
            (hp if kind == "h" else ep)[idx] += sgn * step
            vals.append(float(QCEnergy.apply(torch.tensor(hp), torch.tensor(ep), state, None)))
        fd = (vals[0] - vals[1]) / (2 * step)
        an = (gh if kind == "h" else ge)[idx]
        worst = max(worst, abs(fd - an))
    print(f"  32 entries: max |FD - adjoint| {worst:.2e}", flush=True)
    assert worst < 1e-6
    # the derivative-MPO expectation values (one chain growth per entry)
    hp = [(i, j) for i, j in hs[:3]]
    ep = es[len(es) // 2:len(es) // 2 + 3]
    ref_h = h_gradient(ctxsu2, state.lay, state.data, hp, SU2)
    ref_e = eri_gradient(ctxsu2, state.lay, state.data, ep, SU2)
    assert np.allclose(ref_h, [gh[x] for x in hp], rtol=0, atol=1e-9)
    assert np.allclose(ref_e, [ge[x] for x in ep], rtol=0, atol=1e-9)
