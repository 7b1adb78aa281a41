"""Determinant FCI ground state with its spin-summed 1- and 2-RDMs (test reference for small L).

Built from the same spin-summed generators as qcmaquis_amd.qc.fci_energy: E_pq on the alpha and
beta strings with fermionic signs; gamma_pq = <E_pq>, Gamma_pqrs = <E_pq E_rs> - delta_qr <E_ps>,
so that E = sum h_pq gamma_pq + 1/2 sum (pq|rs) Gamma_pqrs + E_core.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from qcmaquis_amd.qc import _excitation_tables, _strings


def fci_ground(h, eri, n_up, n_dn, ecore=0.0):
    """(E0, gamma, Gamma) of the lowest state in the (n_up, n_dn) sector"""
    L = h.shape[0]
    sa, sb = _strings(L, n_up), _strings(L, n_dn)
    Ea, Eb = _excitation_tables(L, sa), _excitation_tables(L, sb)
    Ia, Ib = sp.identity(len(sa), format="csr"), sp.identity(len(sb), format="csr")
    E = [[(sp.kron(sp.csr_matrix(Ea[p, q]), Ib) + sp.kron(Ia, sp.csr_matrix(Eb[p, q]))).tocsr() for q in range(L)]
         for p in range(L)]
    dim = len(sa) * len(sb)
    H = sp.csr_matrix((dim, dim))
    hp = h - 0.5 * np.einsum("pqqs->ps", eri)
    for p in range(L):
        for q in range(L):
            H = H + hp[p, q] * E[p][q]
            G = sp.csr_matrix((dim, dim))
            for r in range(L):
                for s in range(L):
                    if eri[p, q, r, s] != 0.0:
                        G = G + eri[p, q, r, s] * E[r][s]
            H = H + 0.5 * (E[p][q] @ G)
    if dim <= 400:
        w, V = np.linalg.eigh(H.toarray())
    else:
        w, V = sla.eigsh(H, k=1, which="SA", tol=1e-13)
    v = V[:, np.argmin(w)]
    Ev = [[E[p][q] @ v for q in range(L)] for p in range(L)]
    gamma = np.array([[v @ Ev[p][q] for q in range(L)] for p in range(L)])
    Gamma = np.zeros((L, L, L, L))
    for p in range(L):
        for q in range(L):
            for r in range(L):
                for s in range(L):
                    Gamma[p, q, r, s] = Ev[q][p] @ Ev[r][s] - (gamma[p, s] if q == r else 0.0)
    return float(np.min(w)) + ecore, gamma, Gamma
