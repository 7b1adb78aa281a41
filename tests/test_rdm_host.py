"""Host checks of the RDM / integral-derivative definitions used by qcmaquis_amd.rdm.

* the FCI reference's RDMs reproduce its energy (E = sum h gamma + 1/2 sum eri Gamma + E_core),
  trace gamma = N, and Gamma's pair symmetry Gamma_pqrs = Gamma_rspq;
* Hellmann-Feynman on FCI itself: the unit-integral derivatives that rdm.py evaluates as
  expectation values (sum over the images of an ERI entry, both triangles of h) equal central
  finite differences of the FCI energy, so the linear combinations rdm.py reports are the
  gradients a caller differentiates;
* the permutational images of an ERI index tuple.
"""
import numpy as np

from qcmaquis_amd.qc import fci_energy, make_integrals
from qcmaquis_amd.rdm import eri_images, unit_eri, unit_h
from tests.fci_rdm import fci_ground


def test_eri_images():
    assert len(eri_images(0, 1, 2, 3)) == 8
    assert len(eri_images(0, 0, 1, 1)) == 2
    assert len(eri_images(1, 1, 1, 1)) == 1
    assert len(eri_images(0, 1, 0, 1)) == 4
    e = unit_eri(4, 0, 1, 2, 3)
    assert e.sum() == 8 and np.array_equal(e, e.transpose(1, 0, 2, 3)) and np.array_equal(e, e.transpose(2, 3, 0, 1))


def test_fci_rdms_reproduce_energy_and_gradients():
    L, nu, nd = 5, 2, 3
    h, eri, _ = make_integrals(L, 11, rank=6)
    e0, g, G = fci_ground(h, eri, nu, nd)
    assert abs(e0 - fci_energy(h, eri, nu, nd)[0]) < 1e-10
    assert abs(np.trace(g) - (nu + nd)) < 1e-10
    assert abs(np.einsum("pq,pq", h, g) + 0.5 * np.einsum("pqrs,pqrs", eri, G) - e0) < 1e-9
    assert np.allclose(G, G.transpose(2, 3, 0, 1), atol=1e-10)
    eps = 1e-5
    rng = np.random.default_rng(3)
    for _ in range(4):
        p, q = rng.integers(0, L, 2)
        dh = unit_h(L, p, q)
        fd = (fci_energy(h + eps * dh, eri, nu, nd)[0] - fci_energy(h - eps * dh, eri, nu, nd)[0]) / (2 * eps)
        assert abs(fd - np.einsum("pq,pq", dh, g)) < 1e-6
        p, q, r, s = rng.integers(0, L, 4)
        de = unit_eri(L, p, q, r, s)
        fd = (fci_energy(h, eri + eps * de, nu, nd)[0] - fci_energy(h, eri - eps * de, nu, nd)[0]) / (2 * eps)
        assert abs(fd - 0.5 * np.einsum("pqrs,pqrs", de, G)) < 1e-6


def test_eri_orbits_cover_every_index_once():
    """the classes of (pq|rs) under the eight permutations: m (m + 1) / 2 of them with
    m = L (L + 1) / 2, and their images partition all L^4 index tuples"""
    import itertools

    from qcmaquis_amd.rdm import eri_images, eri_orbits
    for L in (1, 2, 3, 5):
        orbits = eri_orbits(L)
        m = L * (L + 1) // 2
        assert len(orbits) == m * (m + 1) // 2
        seen = [img for e in orbits for img in eri_images(*e)]
        assert sorted(seen) == sorted(itertools.product(range(L), repeat=4))
