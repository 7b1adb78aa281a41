"""SU2U1 quantum-chemistry MPO and two-site recoupling on the CPU, pinned by mathematics:

* for an untruncated right-canonical SU2U1 MPS the one-site effective Hamiltonian of site 0 spans
  every spin-coupled state of (N, S), so its spectrum is the FCI spectrum of the spin-S
  multiplets (determinant FCI at M_s = S minus the one at M_s = S + 1);
* the two-site effective Hamiltonian of sites (0, 1), restricted to the image of the recoupling of
  the sequentially coupled product basis, has the same spectrum, and annihilates the rest;
* growth with Hermitian-compressed boundaries (complement_transpose) gives every boundary block
  equal to the one computed explicitly.
Integrals come from the benchmark generator with all index patterns of (ij|kl) present.  H*psi
columns come from the oracle and, for the one-site case, from the product's host-built plan
tables run by the table emulator."""
import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.dmrg import allowed_bonds_su2, merge_two, right_normalize, split_two, twosite_offsets
from qcmaquis_amd.model import MPSLayout, initial_boundary
from qcmaquis_amd.qc import QCHamiltonian, fci_spin_spectrum, make_integrals
from qcmaquis_amd.su2 import TwoSiteReduction
from tests import emu

SU2 = abi.MQ_SYM_SU2U1


def untruncated(ham, L, N, two_s, seed):
    phys = ham.phys()
    bonds = allowed_bonds_su2([phys] * L, (N, two_s), 10 ** 6, full_right=True)
    lay = [MPSLayout(bonds[k], phys, bonds[k + 1]) for k in range(L)]
    rng = np.random.default_rng(seed)
    data = [rng.normal(size=l.n) for l in lay]
    right_normalize(lay, data)
    return lay, data


def dense_heff(hpsi, spec, Lb, R, lay):
    n = lay.n
    H = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        H[:, c] = hpsi(SU2, spec, Lb, R, lay, e)
    return H


@pytest.mark.parametrize("L,N,two_s,backend", [(4, 4, 0, "oracle"), (4, 3, 1, "oracle"), (5, 4, 0, "oracle"),
                                               (5, 6, 2, "oracle"), (5, 5, 3, "oracle"), (5, 4, 0, "emulator"),
                                               (5, 5, 1, "emulator")])
def test_one_site_heff_is_fci_spin_spectrum(oracle_lib, L, N, two_s, backend):
    h, eri, _ = make_integrals(L, 21, rank=6)
    ham = QCHamiltonian(h, eri, 0.25, symmetry=SU2)
    lay, data = untruncated(ham, L, N, two_s, 0)
    rstep = oracle_lib.rstep if backend == "oracle" else emu.rstep
    hpsi = oracle_lib.hpsi if backend == "oracle" else emu.hpsi
    R = initial_boundary(lay[-1].right, 1)
    for k in range(L - 1, 0, -1):
        R = rstep(SU2, ham.mpo(k).to_spec(), R, lay[k], data[k], lay[k], data[k])
    H = dense_heff(hpsi, ham.mpo(0).to_spec(), initial_boundary(lay[0].left, 1), R, lay[0])
    assert np.allclose(H, H.T, atol=1e-10)
    w = np.linalg.eigvalsh(H) + ham.ecore
    ref = fci_spin_spectrum(h, eri, N, two_s, ecore=0.25)
    assert len(w) == len(ref)
    assert np.max(np.abs(w - ref)) < 1e-10


@pytest.mark.parametrize("L,N,two_s", [(4, 4, 0), (4, 3, 1), (5, 6, 2)])
def test_two_site_heff_on_recoupled_states_is_fci(oracle_lib, L, N, two_s):
    h, eri, _ = make_integrals(L, 22, rank=6)
    ham = QCHamiltonian(h, eri, 0.0, symmetry=SU2)
    phys, phys2 = ham.phys(), ham.phys(True)
    off12 = twosite_offsets(phys)
    lay, data = untruncated(ham, L, N, two_s, 1)
    R = initial_boundary(lay[-1].right, 1)
    for k in range(L - 1, 1, -1):
        R = oracle_lib.rstep(SU2, ham.mpo(k).to_spec(), R, lay[k], data[k], lay[k], data[k])
    Lb = initial_boundary(lay[0].left, 1)
    lay2, psi2 = merge_two(lay[0], data[0], lay[1], data[1], phys2, off12)
    red = TwoSiteReduction(lay2, phys, off12)
    H = dense_heff(oracle_lib.hpsi, ham.mpo(0, True).to_spec(), Lb, R, lay2)
    assert np.allclose(H, H.T, atol=1e-10)
    n = lay2.n
    P = np.array([red.reduce(np.eye(n)[c]) for c in range(n)]).T
    U, sv, _ = np.linalg.svd(P)
    r = int(np.sum(sv > 1e-10))
    ref = fci_spin_spectrum(h, eri, N, two_s)
    assert r == len(ref)
    w = np.linalg.eigvalsh(U[:, :r].T @ H @ U[:, :r])
    assert np.max(np.abs(w - ref)) < 1e-10
    assert np.linalg.norm(H @ U[:, r:]) < 1e-10
    # the state itself: the recoupled two-site energy equals the one-site energy of site 0
    R1 = oracle_lib.rstep(SU2, ham.mpo(1).to_spec(), R, lay[1], data[1], lay[1], data[1])
    e1 = data[0] @ oracle_lib.hpsi(SU2, ham.mpo(0).to_spec(), Lb, R1, lay[0], data[0])
    x = red.reduce(psi2)
    e2 = x @ oracle_lib.hpsi(SU2, ham.mpo(0, True).to_spec(), Lb, R, lay2, x)
    assert abs(e1 - e2) < 1e-11 * max(1.0, abs(e1))


def materialize(B):
    """every block of a boundary as a dense (ls x rs) array keyed (b, bra charge, ket charge),
    aliased (transposed) blocks resolved"""
    offs = B.cohort_offsets()
    out = {}
    for ci, c in enumerate(B.cohorts):
        l, r = c["lr"]
        ls, rs = B.bra.sizes[l], B.ket.sizes[r]
        tp = c.get("transposes", [0] * B.aux_dim)
        cs = c.get("conj_scales", [1.0] * B.aux_dim)
        for b, o in enumerate(c["offsets"]):
            if o < 0:
                continue
            if c.get("tr", 0) or tp[b]:
                a = B.cohort_index(B.ket.charges[r], B.bra.charges[l])
                blk = cs[b] * B.data[offs[a] + o: offs[a] + o + ls * rs].reshape(ls, rs)
            else:
                blk = cs[b] * B.data[offs[ci] + o: offs[ci] + o + ls * rs].reshape(rs, ls).T
            out[(b, B.bra.charges[l], B.ket.charges[r])] = blk
    return out


@pytest.mark.parametrize("L,N,two_s,seed", [(5, 4, 0, 1), (6, 6, 0, 3), (6, 5, 1, 4), (6, 8, 2, 5)])
def test_hermitian_compressed_boundaries(oracle_lib, L, N, two_s, seed):
    h, eri, _ = make_integrals(L, seed, rank=6)
    ham = QCHamiltonian(h, eri, 0.0, symmetry=SU2)
    phys = ham.phys()
    bonds = allowed_bonds_su2([phys] * L, (N, two_s), 10 ** 6, full_right=True)
    lay = [MPSLayout(bonds[k], phys, bonds[k + 1]) for k in range(L)]
    rng = np.random.default_rng(seed)
    data = [rng.normal(size=l.n) for l in lay]

    def spec(k, plain):
        s = ham.mpo(k).to_spec()
        if plain:  # no Hermitian partners: every block computed
            s.left_herm = s.right_herm = None
        return s

    for left in (True, False):
        blocks = {}
        for plain in (True, False):
            out = []
            if left:
                B = initial_boundary(lay[0].left, 1)
                for k in range(L - 1):
                    B = oracle_lib.lstep(SU2, spec(k, plain), B, lay[k], data[k], lay[k], data[k])
                    out.append(materialize(B))
            else:
                B = initial_boundary(lay[-1].right, 1)
                for k in range(L - 1, 0, -1):
                    B = oracle_lib.rstep(SU2, spec(k, plain), B, lay[k], data[k], lay[k], data[k])
                    out.append(materialize(B))
            blocks[plain] = out
        n_compressed = 0
        for ref, got in zip(blocks[True], blocks[False]):
            assert set(ref) == set(got)
            for key, v in ref.items():
                assert np.max(np.abs(got[key] - v)) <= 1e-11 * max(1.0, np.max(np.abs(v)))
        for k in range(1, L - 1):
            n_compressed += int(np.sum(np.asarray(ham.mpo(k).to_spec().left_herm) != abi.MQ_NO_PARTNER))
        assert n_compressed > 0


def test_recoupling_round_trip_and_split(oracle_lib):
    L = 6
    ham = QCHamiltonian(*make_integrals(L, 7, rank=4)[:2], symmetry=SU2)
    phys, phys2 = ham.phys(), ham.phys(True)
    off12 = twosite_offsets(phys)
    bonds = allowed_bonds_su2([phys] * L, (6, 0), 10 ** 6)
    lay = [MPSLayout(bonds[k], phys, bonds[k + 1]) for k in range(L)]
    rng = np.random.default_rng(3)
    data = [rng.normal(size=l.n) for l in lay]
    right_normalize(lay, data)
    lay2, psi2 = merge_two(lay[2], data[2], lay[3], data[3], phys2, off12)
    red = TwoSiteReduction(lay2, phys, off12)
    x = red.reduce(psi2)
    assert abs(np.linalg.norm(x) - np.linalg.norm(psi2)) < 1e-12 * np.linalg.norm(psi2)
    assert np.max(np.abs(red.unreduce(x) - psi2)) < 1e-13
    la, a, lb, b, tw = split_two(lay2, psi2, phys, off12, 10 ** 6, 0.0, True, sym=SU2)
    assert all(c[1] >= 0 for c in lb.left.charges) and tw < 1e-20
    _, back = merge_two(la, a, lb, b, phys2, off12)
    assert np.max(np.abs(back - psi2)) < 1e-12


def test_su2_mpo_structure():
    L = 8
    h, eri, _ = make_integrals(L, 1008, rank=16)
    su2 = QCHamiltonian(h, eri, symmetry=SU2)
    u1 = QCHamiltonian(h, eri)
    assert su2.bond_dims[0] == su2.bond_dims[-1] == 1
    # spin adaptation merges the spin components of every operator string
    assert max(su2.bond_dims) < max(u1.bond_dims)
    for k in range(L):
        s = su2.mpo(k).to_spec()
        # bond spins are 0, 1/2 or 1 and every Hermitian partner has the same spin
        for sp, hm in ((s.left_spin, s.left_herm), (s.right_spin, s.right_herm)):
            assert set(np.unique(sp)) <= {0, 1, 2}
            for b, p in enumerate(hm):
                if p != abi.MQ_NO_PARTNER:
                    assert sp[p] == sp[b] and hm[p] == b
