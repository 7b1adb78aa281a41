"""The TwoU1 quantum-chemistry MPO (mq_qc_build) against an independent determinant FCI:
the MPO chain contracted to a dense operator must have the FCI spectrum in every charge sector.
Also: Hermitian partners of every bulk bond index, two-site MPO = product of site MPOs, FCIDUMP
round trip, and the integral generator's symmetries."""
import os

import numpy as np
import pytest

from oracle import dense
from qcmaquis_amd.qc import QCHamiltonian, fci_energy, make_integrals, read_fcidump, write_fcidump

SITE_Q = [(1, 1), (1, 0), (0, 1), (0, 0)]


def dense_chain(ham, L):
    phys = ham.phys()
    op = None
    for k in range(L):
        W = dense.dense_mpo(ham.mpo(k).to_spec(), phys)  # (b1, b2, in, out)
        if op is None:
            op = W[0]  # (b2, in, out)
        else:
            op = np.einsum("aij,abkl->bikjl", op, W).reshape(W.shape[1], op.shape[1] * 4, op.shape[2] * 4)
    return op[0].T  # H[out, in]


def sector_mask(L, nu, nd):
    idx = np.arange(4 ** L)
    cu = np.zeros(4 ** L, int)
    cd = np.zeros(4 ** L, int)
    for k in range(L):
        s = (idx // 4 ** (L - 1 - k)) % 4
        cu += np.array([q[0] for q in SITE_Q])[s]
        cd += np.array([q[1] for q in SITE_Q])[s]
    return (cu == nu) & (cd == nd)


@pytest.mark.parametrize("L,seed", [(3, 11), (4, 12), (5, 13)])
def test_spectrum_matches_fci(L, seed):
    h, eri, _ = make_integrals(L, seed, rank=4)
    ham = QCHamiltonian(h, eri, ecore=0.7)
    H = dense_chain(ham, L)
    assert np.allclose(H, H.T, atol=1e-12)
    for nu, nd in [(1, 1), (2, 1), (L // 2, L // 2), (L - 1, 1)]:
        m = sector_mask(L, nu, nd)
        w = np.linalg.eigvalsh(H[np.ix_(m, m)])[:3] + ham.ecore
        ref = fci_energy(h, eri, nu, nd, ecore=0.7, k=min(3, m.sum()))
        assert np.allclose(w[:len(ref)], ref, atol=1e-10), (nu, nd, w, ref)


def test_hermitian_partners_and_bond_dims():
    L = 8
    h, eri, _ = make_integrals(L, 1001, rank=16)
    ham = QCHamiltonian(h, eri)
    dims = ham.bond_dims
    assert dims[0] == 1 and dims[-1] == 1
    assert max(dims) <= 16 * L * L  # O(L^2)
    for k in range(L - 1):
        spec = ham.mpo(k).to_spec()
        herm = spec.right_herm
        for b in range(2, spec.col_dim):
            assert herm[b] != 0xFFFFFFFF, (k, b)
            assert herm[herm[b]] == b


def test_twosite_is_product():
    L = 4
    h, eri, _ = make_integrals(L, 5, rank=4)
    ham = QCHamiltonian(h, eri)
    p1, p2 = ham.phys(), ham.phys(True)
    for k in range(L - 1):
        A = dense.dense_mpo(ham.mpo(k).to_spec(), p1)
        B = dense.dense_mpo(ham.mpo(k + 1).to_spec(), p1)
        T = dense.dense_mpo(ham.mpo(k, True).to_spec(), p2)
        prod = np.einsum("abij,bckl->acikjl", A, B).reshape(A.shape[0], B.shape[1], 16, 16)
        # fused two-site states: sectors of (s1, s2) in ProductBasis order
        order = []
        offs = {}
        for s1 in range(4):
            for s2 in range(4):
                c = (SITE_Q[s1][0] + SITE_Q[s2][0], SITE_Q[s1][1] + SITE_Q[s2][1])
                order.append((c, offs.get(c, 0), s1 * 4 + s2))
                offs[c] = offs.get(c, 0) + 1
        starts, acc = {}, 0
        for c, sz in zip(p2.charges, p2.sizes):
            starts[c] = acc
            acc += sz
        perm = np.zeros(16, int)
        for c, o, st in order:
            perm[starts[c] + o] = st
        assert np.allclose(T, prod[:, :, perm][:, :, :, perm], atol=1e-13)


def test_fcidump_roundtrip(tmp_path):
    h, eri, _ = make_integrals(5, 1000, rank=16)
    assert np.allclose(eri, eri.transpose(1, 0, 2, 3)) and np.allclose(eri, eri.transpose(2, 3, 0, 1))
    assert np.linalg.eigvalsh(eri.reshape(25, 25)).min() > -1e-12
    assert np.allclose(np.diag(h), np.linspace(-10, 2, 5))
    p = os.path.join(tmp_path, "FCIDUMP")
    write_fcidump(p, h, eri, 1.25, nelec=5)
    h2, eri2, e2 = read_fcidump(p)
    assert np.allclose(h, h2, atol=1e-15) and np.allclose(eri, eri2, atol=1e-15) and e2 == 1.25
