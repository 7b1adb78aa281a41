"""Quantum-chemistry inputs: seeded integrals, FCIDUMP files and the Hamiltonian MPO chain.

Integral generator (the benchmark configurations):
  h_pp = linear ramp -10 -> +2 Eh; h_pq (p != q) = N(0, 0.1) exp(-|p-q|/4), symmetrised;
  (pq|rs) = sum_k L^k_pq L^k_rs with L^k_pq = N(0, 0.3) exp(-|p-q|/4) symmetrised in pq, so all
  8 permutational symmetries hold and the ERI matrix is positive semidefinite.
FCIDUMP: the reference's parser skips a 4-line header and reads "value i j k l" lines, 1-based,
with the core energy on "0 0 0 0" (models/chem/parse_integrals.h:36-156).
"""
import ctypes as C

import numpy as np

from . import abi
from .model import Index, MPOSpec, SiteOp


def make_integrals(L, seed, rank=16):
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.arange(L)
    decay = np.exp(-np.abs(p[:, None] - p[None, :]) / 4.0)
    off = rng.normal(0.0, 0.1, (L, L)) * decay
    h = 0.5 * (off + off.T)
    np.fill_diagonal(h, np.linspace(-10.0, 2.0, L))
    Lk = rng.normal(0.0, 0.3, (rank, L, L)) * decay[None]
    Lk = 0.5 * (Lk + Lk.transpose(0, 2, 1))
    eri = np.einsum("kpq,krs->pqrs", Lk, Lk)
    return h, eri, 0.0


def write_fcidump(path, h, eri, ecore=0.0, nelec=None, ms2=0, tol=1e-14):
    L = h.shape[0]
    nelec = L if nelec is None else nelec
    with open(path, "w") as f:
        f.write(f" &FCI NORB={L},NELEC={nelec},MS2={ms2},\n")
        f.write("  ORBSYM=" + ",".join(["1"] * L) + ",\n")
        f.write("  ISYM=1,\n")
        f.write(" &END\n")
        for i in range(L):
            for j in range(i + 1):
                for k in range(L):
                    for l in range(k + 1):
                        if (i * (i + 1) // 2 + j) < (k * (k + 1) // 2 + l):
                            continue
                        v = eri[i, j, k, l]
                        if abs(v) > tol:
                            f.write(f"{v: .16e} {i + 1} {j + 1} {k + 1} {l + 1}\n")
        for i in range(L):
            for j in range(i + 1):
                if abs(h[i, j]) > tol:
                    f.write(f"{h[i, j]: .16e} {i + 1} {j + 1} 0 0\n")
        f.write(f"{ecore: .16e} 0 0 0 0\n")


def read_fcidump(path):
    """(h, eri, ecore); the first 4 lines are the header the reference parser skips"""
    lines = open(path).read().splitlines()
    head = " ".join(lines[:4])
    L = int(head.split("NORB=")[1].split(",")[0])
    h = np.zeros((L, L))
    eri = np.zeros((L, L, L, L))
    ecore = 0.0
    for ln in lines[4:]:
        parts = ln.split()
        if len(parts) != 5:
            continue
        v = float(parts[0])
        i, j, k, l = (int(x) for x in parts[1:])
        if i == j == k == l == 0:
            ecore = v
        elif k == 0 and l == 0:
            h[i - 1, j - 1] = h[j - 1, i - 1] = v
        else:
            i, j, k, l = i - 1, j - 1, k - 1, l - 1
            for a, b, c, d in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                               (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
                eri[a, b, c, d] = v
    return h, eri, ecore


class RawMPO:
    """An MPO tensor whose descriptor arrays live in the QC builder (mq_qc_site_mpo)."""

    def __init__(self, owner, desc):
        self._owner = owner
        self._desc = desc
        self.row_dim, self.col_dim = desc.row_dim, desc.col_dim

    def desc(self):
        return self._desc

    def to_spec(self):
        """MPOSpec copy (for the dense test checker)"""
        d = self._desc
        ops = []
        for o in range(d.n_ops):
            blocks = []
            for bl in range(d.op_block_ptr[o], d.op_block_ptr[o + 1]):
                lc = (d.block_charges[4 * bl], d.block_charges[4 * bl + 1])
                rc = (d.block_charges[4 * bl + 2], d.block_charges[4 * bl + 3])
                if d.entry_spins:
                    ents = [(d.entry_rc[2 * e], d.entry_rc[2 * e + 1], d.entry_coeff[e], d.entry_spins[2 * e],
                             d.entry_spins[2 * e + 1]) for e in range(d.block_entry_ptr[bl], d.block_entry_ptr[bl + 1])]
                else:
                    ents = [(d.entry_rc[2 * e], d.entry_rc[2 * e + 1], d.entry_coeff[e])
                            for e in range(d.block_entry_ptr[bl], d.block_entry_ptr[bl + 1])]
                blocks.append((lc, rc, ents))
            ops.append(SiteOp(blocks, d.op_spin[o] if d.op_spin else 0))
        terms = {}
        for t in range(d.n_terms):
            terms.setdefault((d.term_b[2 * t], d.term_b[2 * t + 1]), []).append((d.term_op[t], d.term_scale[t]))
        arr = lambda p, n: None if not p else [p[i] for i in range(n)]
        return MPOSpec(d.row_dim, d.col_dim, ops, terms, arr(d.left_spin, d.row_dim), arr(d.right_spin, d.col_dim),
                       arr(d.left_herm, d.row_dim), arr(d.left_phase, d.row_dim), arr(d.right_herm, d.col_dim),
                       arr(d.right_phase, d.col_dim))


class QCHamiltonian:
    """Quantum-chemistry MPO chain (mq_qc_build): TwoU1 (charges (N_up, N_down)) or SU2U1
    (charges (N, 2S), spin-adapted operators)."""

    def __init__(self, h, eri, ecore=0.0, cutoff=1e-20, symmetry=abi.MQ_SYM_TWOU1, gradient_layout=False):
        """gradient_layout (SU2U1): the chain of mq_qc_build_ex(MQ_QC_GRADIENT), whose structure
        does not depend on the integral values and which maps MPO-term gradients back to the
        integrals (integral_gradient)"""
        self.lib = abi.lib()
        self.sym = int(symmetry)
        self.L = h.shape[0]
        self.h = np.ascontiguousarray(h, np.float64)
        self.eri = np.ascontiguousarray(eri, np.float64)
        self.gradient_layout = bool(gradient_layout)
        hd = C.c_void_p()
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        rc = self.lib.mq_qc_build_ex(self.sym, self.L, dp(self.h), dp(self.eri), float(ecore), float(cutoff),
                                     abi.MQ_QC_GRADIENT if gradient_layout else 0, C.byref(hd))
        if rc != 0:
            raise RuntimeError(f"mq_qc_build failed ({rc})")
        self.h_qc = hd
        self.ecore = float(ecore)
        dims = (C.c_uint32 * (self.L + 1))()
        self.lib.mq_qc_info(hd, None, dims, None)
        self.bond_dims = list(dims)
        self._cache = {}

    def phys(self, twosite=False):
        d = abi.IndexDesc()
        self.lib.mq_qc_phys(self.h_qc, int(twosite), C.byref(d))
        return Index([(d.charges[2 * i], d.charges[2 * i + 1]) for i in range(d.n)], [d.sizes[i] for i in range(d.n)])

    def mpo(self, site, twosite=False):
        key = (site, twosite)
        if key not in self._cache:
            d = abi.MPODesc()
            rc = self.lib.mq_qc_site_mpo(self.h_qc, int(site), int(twosite), C.byref(d))
            if rc != 0:
                raise RuntimeError("mq_qc_site_mpo failed")
            self._cache[key] = RawMPO(self, d)
        return self._cache[key]

    def term_scales(self, site):
        """the term scales of the single-site MPO of one site (a copy)"""
        d = self.mpo(site).desc()
        return np.ctypeslib.as_array(d.term_scale, shape=(d.n_terms,)).copy() if d.n_terms else np.zeros(0)

    def integral_gradient(self, term_grads):
        """mq_qc_integral_gradient: (dE/dh, dE/deri) at the canonical integral positions (h[i, j],
        j <= i; eri[i, j, k, l], i >= j, k >= l, (i, j) >= (k, l)), zero elsewhere, from
        term_grads[p] = dE/d(term scales of site p's single-site MPO)"""
        if not self.gradient_layout:
            raise ValueError("integral gradients need a chain built with gradient_layout=True")
        L = self.L
        arrs = [np.ascontiguousarray(g, np.float64) for g in term_grads]
        for p, a in enumerate(arrs):
            if a.size != self.mpo(p).desc().n_terms:
                raise ValueError("site %d: %d term gradients for %d terms" % (p, a.size, self.mpo(p).desc().n_terms))
        ptrs = (C.POINTER(C.c_double) * L)(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
        dh = np.zeros((L, L))
        deri = np.zeros((L, L, L, L))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        rc = self.lib.mq_qc_integral_gradient(self.h_qc, ptrs, dp(dh), dp(deri))
        if rc != 0:
            raise RuntimeError(f"mq_qc_integral_gradient failed ({rc})")
        return dh, deri

    def __del__(self):
        try:
            self._cache.clear()
            self.lib.mq_qc_destroy(self.h_qc)
        except Exception:
            pass


# ---- independent determinant FCI (spin-summed E_pq form) for small L -------------------------
def _strings(L, n):
    from itertools import combinations
    return [sum(1 << i for i in c) for c in combinations(range(L), n)]


def _excitation_tables(L, strings):
    """E_pq on one spin's strings: dense (ns x ns) per (p, q) with fermionic signs"""
    pos = {s: i for i, s in enumerate(strings)}
    ns = len(strings)
    E = np.zeros((L, L, ns, ns))
    for i, s in enumerate(strings):
        for q in range(L):
            if not (s >> q) & 1:
                continue
            t = s & ~(1 << q)
            sign_q = (-1) ** bin(s & ((1 << q) - 1)).count("1")
            for p in range(L):
                if (t >> p) & 1:
                    continue
                u = t | (1 << p)
                sign_p = (-1) ** bin(t & ((1 << p) - 1)).count("1")
                E[p, q, pos[u], i] += sign_p * sign_q
    return E


def fci_energy(h, eri, n_up, n_dn, ecore=0.0, k=1):
    """Lowest eigenvalue(s) of H in the (n_up, n_dn) sector by a determinant CI built from the
    spin-summed generators: H = sum h_pq E_pq + 1/2 sum (pq|rs) (E_pq E_rs - delta_qr E_ps)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    L = h.shape[0]
    sa, sb = _strings(L, n_up), _strings(L, n_dn)
    Ea, Eb = _excitation_tables(L, sa), _excitation_tables(L, sb)
    Ia, Ib = sp.identity(len(sa), format="csr"), sp.identity(len(sb), format="csr")
    E = [[sp.kron(sp.csr_matrix(Ea[p, q]), Ib) + sp.kron(Ia, sp.csr_matrix(Eb[p, q])) for q in range(L)] for p in range(L)]
    dim = len(sa) * len(sb)
    H = sp.csr_matrix((dim, dim))
    hp = h - 0.5 * np.einsum("pqqs->ps", eri)
    for p in range(L):
        for q in range(L):
            H = H + hp[p, q] * E[p][q]
    for p in range(L):
        for q in range(L):
            G = sp.csr_matrix((dim, dim))
            for r in range(L):
                for s in range(L):
                    if eri[p, q, r, s] != 0.0:
                        G = G + eri[p, q, r, s] * E[r][s]
            H = H + 0.5 * (E[p][q] @ G)
    if dim <= 400:
        w = np.linalg.eigvalsh(H.toarray())[:k]
    else:
        w = np.sort(sla.eigsh(H, k=k, which="SA", tol=1e-12)[0])
    return w + ecore


def fci_spectrum(h, eri, n_up, n_dn, ecore=0.0):
    """all eigenvalues of H in the (n_up, n_dn) determinant sector (dense, small L only)"""
    dim = len(_strings(h.shape[0], n_up)) * len(_strings(h.shape[0], n_dn))
    return fci_energy(h, eri, n_up, n_dn, ecore=ecore, k=dim) if dim <= 400 else None


def fci_spin_spectrum(h, eri, n, two_s, ecore=0.0, tol=1e-8):
    """eigenvalues of the spin-S multiplets of n electrons (one per multiplet): the M_s = S
    spectrum with the M_s = S + 1 spectrum taken out (every multiplet with S' > S has both)"""
    lo = np.sort(fci_spectrum(h, eri, (n + two_s) // 2, (n - two_s) // 2, ecore))
    if (n + two_s) // 2 + 1 > h.shape[0] or (n - two_s) // 2 - 1 < 0:
        return lo
    hi = list(np.sort(fci_spectrum(h, eri, (n + two_s) // 2 + 1, (n - two_s) // 2 - 1, ecore)))
    out = []
    for e in lo:
        j = next((x for x, v in enumerate(hi) if abs(v - e) < tol), None)
        if j is None:
            out.append(e)
        else:
            hi.pop(j)
    assert not hi, "unmatched higher-spin eigenvalues"
    return np.array(out)


def ladder_hubbard(legs_len, t=1.0, t_rung=1.0, U=8.0):
    """integrals of the fermion Hubbard model on an open two-leg ladder of legs_len rungs, in the
    quantum-chemistry form: h_ij = -t on leg bonds, -t_rung on rung bonds, (ii|ii) = U; site of
    (x, y) = 2 x + y (rung partners adjacent)"""
    n = 2 * legs_len
    h = np.zeros((n, n))
    for x in range(legs_len):
        a, b = 2 * x, 2 * x + 1
        h[a, b] = h[b, a] = -t_rung
        if x + 1 < legs_len:
            for y in (0, 1):
                i, j = 2 * x + y, 2 * (x + 1) + y
                h[i, j] = h[j, i] = -t
    eri = np.zeros((n, n, n, n))
    for i in range(n):
        eri[i, i, i, i] = U
    return h, eri
