"""Integral gradients of the SU2U1 chain on the CPU: the gradient layout of the QC builder
(mq_qc_build_ex(MQ_QC_GRADIENT)) and its reverse map mq_qc_integral_gradient, pinned by
mathematics and by determinant FCI:

* the layout's structure (bond dimensions, term cells and operators) does not depend on the
  integral values, and its energy is the one of the default chain (one-site effective Hamiltonian
  of an untruncated state = FCI spin spectrum);
* the reverse map is the transpose of the builder's linear map from integrals to term scales
  (dot-product test against the scales of a chain built from other integrals);
* with dE/d(term scale) taken from the oracle's H*psi of single-term MPOs at every site of the FCI
  ground state, the reverse map gives the FCI 1- and 2-RDMs (Hellmann-Feynman):
  dE/dh_pq = gamma_pq + gamma_qp (p != q), dE/d(pq|rs) = 1/2 sum over the distinct images of
  (pq|rs) of Gamma."""
import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.dmrg import allowed_bonds_su2, right_normalize
from qcmaquis_amd.model import MPOSpec, MPSLayout, initial_boundary
from qcmaquis_amd.qc import QCHamiltonian, fci_spin_spectrum, make_integrals
from qcmaquis_amd.rdm import eri_images
from tests.fci_rdm import fci_ground

SU2 = abi.MQ_SYM_SU2U1


def structure(ham):
    out = [list(ham.bond_dims)]
    for p in range(ham.L):
        d = ham.mpo(p).desc()
        out.append((d.row_dim, d.col_dim, [d.term_b[i] for i in range(2 * d.n_terms)],
                    [d.term_op[i] for i in range(d.n_terms)]))
    return out


def canonical(L):
    """the canonical positions the builder reads: h[i, j] (j <= i), eri[i, j, k, l]"""
    hs = [(i, j) for i in range(L) for j in range(i + 1)]
    es = [(i, j, k, l) for i in range(L) for j in range(i + 1) for k in range(i + 1) for l in range(k + 1)
          if not (k == i and l > j)]
    return hs, es


def test_gradient_layout_structure_is_value_free():
    L = 5
    h, eri, _ = make_integrals(L, 3, rank=6)
    h2, eri2, _ = make_integrals(L, 4, rank=6)
    a = QCHamiltonian(h, eri, symmetry=SU2, gradient_layout=True)
    b = QCHamiltonian(h2, eri2, symmetry=SU2, gradient_layout=True)
    z = QCHamiltonian(np.zeros_like(h), np.zeros_like(eri), symmetry=SU2, gradient_layout=True)
    assert structure(a) == structure(b) == structure(z)
    # the default chain drops nothing here either (all integrals nonzero), so its bond dimensions agree
    assert QCHamiltonian(h, eri, symmetry=SU2).bond_dims == a.bond_dims


@pytest.mark.parametrize("L,seed", [(5, 11), (6, 12)])
def test_reverse_map_is_the_transpose(L, seed):
    h, eri, _ = make_integrals(L, seed, rank=6)
    dh_in, deri_in, _ = make_integrals(L, seed + 100, rank=6)
    base = QCHamiltonian(h, eri, symmetry=SU2, gradient_layout=True)
    pert = QCHamiltonian(dh_in, deri_in, symmetry=SU2, gradient_layout=True)
    zero = QCHamiltonian(np.zeros_like(h), np.zeros_like(eri), symmetry=SU2, gradient_layout=True)
    rng = np.random.default_rng(seed)
    e = [rng.normal(size=base.term_scales(p).size) for p in range(L)]
    dh, deri = base.integral_gradient(e)
    # the scales' linear part (attach and fill terms are constant ones)
    lhs = sum(float(e[p] @ (pert.term_scales(p) - zero.term_scales(p))) for p in range(L))
    rhs = float(np.sum(dh * dh_in) + np.sum(deri * deri_in))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))
    # nothing outside the canonical positions
    hs, es = canonical(L)
    mh = np.zeros((L, L), bool)
    mh[tuple(np.array(hs).T)] = True
    me = np.zeros((L,) * 4, bool)
    me[tuple(np.array(es).T)] = True
    assert not np.any(dh[~mh]) and not np.any(deri[~me])
    # the reverse map is linear in the term gradients
    dh2, deri2 = base.integral_gradient([2.0 * x for x in e])
    assert np.allclose(dh2, 2 * dh, rtol=0, atol=1e-13) and np.allclose(deri2, 2 * deri, rtol=0, atol=1e-13)


def untruncated_ground(ham, oracle_lib, L, N, two_s):
    """the exact lowest (N, S) state as an untruncated MPS: right-normalised sites 1..L-1, site 0 the
    lowest eigenvector of its one-site effective Hamiltonian (= the FCI multiplet)"""
    phys = ham.phys()
    bonds = allowed_bonds_su2([phys] * L, (N, two_s), 10 ** 6, full_right=True)
    lay = [MPSLayout(bonds[k], phys, bonds[k + 1]) for k in range(L)]
    rng = np.random.default_rng(7)
    data = [rng.normal(size=l.n) for l in lay]
    right_normalize(lay, data)
    R = initial_boundary(lay[-1].right, 1)
    for k in range(L - 1, 0, -1):
        R = oracle_lib.rstep(SU2, ham.mpo(k).to_spec(), R, lay[k], data[k], lay[k], data[k])
    spec0 = ham.mpo(0).to_spec()
    L0 = initial_boundary(lay[0].left, 1)
    n = lay[0].n
    H = np.zeros((n, n))
    for c in range(n):
        v = np.zeros(n)
        v[c] = 1.0
        H[:, c] = oracle_lib.hpsi(SU2, spec0, L0, R, lay[0], v, nthreads=1)
    w, V = np.linalg.eigh(0.5 * (H + H.T))
    data[0] = V[:, 0].copy()
    return lay, data, w[0]


def single_term(spec, cell, k):
    """the MPO tensor with term k of cell (b1, b2) alone, at scale one"""
    op, _ = spec.terms[cell][k]
    return MPOSpec(spec.row_dim, spec.col_dim, spec.ops, {cell: [(op, 1.0)]}, spec.left_spin, spec.right_spin,
                   spec.left_herm, spec.left_phase, spec.right_herm, spec.right_phase)


def oracle_term_gradients(ham, oracle_lib, lay, data):
    """dE/d(term scale) for every term of every site MPO: <psi| H_eff(term alone) |psi> at each
    site, with the left boundaries grown from the same tensors as the right ones"""
    L = ham.L
    Rs = {L: initial_boundary(lay[-1].right, 1)}
    for k in range(L - 1, 0, -1):
        Rs[k] = oracle_lib.rstep(SU2, ham.mpo(k).to_spec(), Rs[k + 1], lay[k], data[k], lay[k], data[k])
    Lb = initial_boundary(lay[0].left, 1)
    grads, energies = [], []
    for p in range(L):
        spec = ham.mpo(p).to_spec()
        g = []
        for cell, lst in spec.terms.items():
            for k in range(len(lst)):
                hv = oracle_lib.hpsi(SU2, single_term(spec, cell, k), Lb, Rs[p + 1], lay[p], data[p], nthreads=1)
                g.append(float(data[p] @ hv))
        g = np.array(g)
        grads.append(g)
        energies.append(float(g @ ham.term_scales(p)))
        if p + 1 < L:
            Lb = oracle_lib.lstep(SU2, spec, Lb, lay[p], data[p], lay[p], data[p])
    return grads, energies


def rdm_gradients(gamma, Gamma, L):
    """the derivatives the reverse map returns, from FCI RDMs"""
    hs, es = canonical(L)
    dh = np.zeros((L, L))
    deri = np.zeros((L,) * 4)
    for i, j in hs:
        dh[i, j] = gamma[i, j] + gamma[j, i] if i != j else gamma[i, i]
    for e in es:
        deri[e] = 0.5 * sum(Gamma[x] for x in eri_images(*e))
    return dh, deri


def test_gradient_layout_energy_is_fci(oracle_lib):
    L, N = 4, 4
    h, eri, _ = make_integrals(L, 21, rank=6)
    ham = QCHamiltonian(h, eri, 0.25, symmetry=SU2, gradient_layout=True)
    lay, data, e0 = untruncated_ground(ham, oracle_lib, L, N, 0)
    ref = fci_spin_spectrum(h, eri, N, 0, 0.25)
    assert abs(e0 + ham.ecore - ref[0]) < 1e-10


@pytest.mark.parametrize("L,N", [(4, 4), (5, 4)])
def test_oracle_term_gradients_give_fci_rdms(oracle_lib, L, N):
    h, eri, _ = make_integrals(L, 31 + L, rank=6)
    ham = QCHamiltonian(h, eri, symmetry=SU2, gradient_layout=True)
    lay, data, e0 = untruncated_ground(ham, oracle_lib, L, N, 0)
    grads, energies = oracle_term_gradients(ham, oracle_lib, lay, data)
    # every site's term gradients recombine to the energy
    assert np.allclose(energies, e0, rtol=0, atol=1e-10)
    dh, deri = ham.integral_gradient(grads)
    # Euler: the energy is linear and homogeneous in the integrals
    assert abs(np.sum(dh * h) + np.sum(deri * eri) - e0) < 1e-10
    e_fci, gamma, Gamma = fci_ground(h, eri, N // 2, N // 2)
    assert abs(e_fci - e0) < 1e-10
    rh, re = rdm_gradients(gamma, Gamma, L)
    assert np.max(np.abs(dh - rh)) < 1e-9
    assert np.max(np.abs(deri - re)) < 1e-9


@pytest.mark.parametrize("sym,chunk", [(SU2, 0), (SU2, 40), (abi.MQ_SYM_TWOU1, 0), (abi.MQ_SYM_TWOU1, 40)])
def test_emulated_adjoint_tables_match_single_term_oracle(oracle_lib, sym, chunk):
    """the plan's term sources, Y products and term work items (run by the table emulator) give
    every MPO term's dE/d(scale) of the oracle's single-term H*psi, in one chunk and in several"""
    from tests import emu
    L = 5
    h, eri, _ = make_integrals(L, 41, rank=6)
    if sym == SU2:
        ham = QCHamiltonian(h, eri, symmetry=SU2, gradient_layout=True)
        lay, data, _ = untruncated_ground(ham, oracle_lib, L, 4, 0)
    else:
        from qcmaquis_amd.dmrg import allowed_bonds
        ham = QCHamiltonian(h, eri, symmetry=sym)
        phys = ham.phys()
        bonds = allowed_bonds([phys] * L, (2, 2), 10 ** 6, full_right=True)
        lay = [MPSLayout(bonds[k], phys, bonds[k + 1]) for k in range(L)]
        rng = np.random.default_rng(2)
        data = [rng.normal(size=l.n) for l in lay]
        right_normalize(lay, data)
    Rs = {L: initial_boundary(lay[-1].right, 1)}
    for k in range(L - 1, 0, -1):
        Rs[k] = oracle_lib.rstep(sym, ham.mpo(k).to_spec(), Rs[k + 1], lay[k], data[k], lay[k], data[k])
    Lb = initial_boundary(lay[0].left, 1)
    emu.last_chunks = 0
    for p in range(L):
        spec = ham.mpo(p).to_spec()
        got = emu.term_grad(sym, ham.mpo(p), Lb, Rs[p + 1], lay[p], data[p], chunk_elems=chunk)
        want = []
        for cell, lst in spec.terms.items():
            for k in range(len(lst)):
                hv = oracle_lib.hpsi(sym, single_term(spec, cell, k), Lb, Rs[p + 1], lay[p], data[p], nthreads=1)
                want.append(float(data[p] @ hv))
        assert np.allclose(got, want, rtol=0, atol=1e-11 * max(1.0, np.max(np.abs(want))))
        full = oracle_lib.hpsi(sym, spec, Lb, Rs[p + 1], lay[p], data[p], nthreads=1)
        assert abs(float(got @ ham.term_scales(p)) - float(data[p] @ full)) < 1e-10
        if p + 1 < L:
            Lb = oracle_lib.lstep(sym, spec, Lb, lay[p], data[p], lay[p], data[p])
