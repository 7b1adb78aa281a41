"""Reduced density matrices and integral derivatives of an MPS through derivative MPOs.

The Hamiltonian is linear in its integrals,

    H(h, eri) = sum_pq h_pq E_pq + 1/2 sum_pqrs (pq|rs) (E_pq E_rs - delta_qr E_ps),

so for a normalised state the derivative of E = <psi|H|psi> with respect to one independent
integral is the expectation value of the MPO that the QC builder makes from the unit integral set
of that entry (the entry and its permutational images set to one, everything else zero):

    dE/dh_pq       = gamma_pq + gamma_qp            (p != q; gamma_pp for p == q)
    dE/d(pq|rs)    = 1/2 sum over the distinct images (ab|cd) of (pq|rs) of Gamma_abcd

(two_rdm_sym collects the second for every class of (pq|rs): Gamma averaged over the images.)

with gamma_pq = <E_pq> the spin-summed 1-RDM and Gamma_pqrs = <E_pq E_rs> - delta_qr <E_ps>.
At an eigenstate these are the Hellmann-Feynman forces on the integrals (the gradients orbital
optimisation consumes).  Each expectation value is one left-to-right boundary growth over all
sites on the device (reference: expval, mp_tensors/mps_mpo_ops.h:53-70, which grows the left
boundary with contraction::Engine::overlap_mpo_left_step and takes its trace), so the path is the
product's own growth kernels; the 1-RDM measurement it replaces is
measurements/tagged_nrankrdm.h:413-661.
"""
import itertools

import numpy as np

from . import abi
from .engine import DeviceBoundary, DeviceMPO, DeviceMPS, lstep
from .qc import QCHamiltonian


def expval(ctx, layouts, data, ham):
    """<psi|H|psi> (without the core energy) of the MPS (layouts, data) for a QCHamiltonian chain:
    the left boundary grown over every site holds the single number."""
    L = len(layouts)
    b = DeviceBoundary.initial(ctx, layouts[0].left, ham.mpo(0).row_dim)
    for k in range(L):
        m = DeviceMPS(ctx, layouts[k], data[k])
        b = lstep(ctx, DeviceMPO(ctx, ham.mpo(k)), b, m, m)
    out = b.download()
    if out.data.size != 1:
        raise RuntimeError("expval: the last boundary holds %d numbers, expected 1" % out.data.size)
    return float(out.data[0])


def unit_h(L, p, q):
    h = np.zeros((L, L))
    h[p, q] = h[q, p] = 1.0
    return h


def unit_eri(L, p, q, r, s):
    eri = np.zeros((L, L, L, L))
    for a, b, c, d in eri_images(p, q, r, s):
        eri[a, b, c, d] = 1.0
    return eri


def eri_images(p, q, r, s):
    """the distinct index tuples (pq|rs) = (qp|rs) = (pq|sr) = (rs|pq) = ... maps to"""
    out = set()
    for a, b in ((p, q), (q, p)):
        for c, d in ((r, s), (s, r)):
            out.add((a, b, c, d))
            out.add((c, d, a, b))
    return sorted(out)


def h_gradient(ctx, layouts, data, entries, symmetry=abi.MQ_SYM_TWOU1):
    """dE/dh_pq for the (p, q) entries (h symmetric: h_pq and h_qp are one parameter)"""
    L = len(layouts)
    zero = np.zeros((L, L, L, L))
    return np.array([expval(ctx, layouts, data, QCHamiltonian(unit_h(L, p, q), zero, symmetry=symmetry))
                     for p, q in entries])


def eri_gradient(ctx, layouts, data, entries, symmetry=abi.MQ_SYM_TWOU1):
    """dE/d(pq|rs) for the (p, q, r, s) entries (all eight permutational images one parameter)"""
    L = len(layouts)
    zero = np.zeros((L, L))
    return np.array([expval(ctx, layouts, data, QCHamiltonian(zero, unit_eri(L, *e), symmetry=symmetry))
                     for e in entries])


def eri_orbits(L):
    """one representative (the smallest index tuple) of every class of (pq|rs) under the eight
    permutational symmetries"""
    out = set()
    for e in itertools.product(range(L), repeat=4):
        out.add(eri_images(*e)[0])
    return sorted(out)


def two_rdm_sym(ctx, layouts, data):
    """the spin-summed 2-RDM averaged over the eight permutations of (pq|rs):
    Gamma^sym_pqrs = 2 dE/d(pq|rs) / (number of distinct images), so that the two-body energy is
    1/2 sum (pq|rs) Gamma^sym_pqrs for any integrals with the eight-fold symmetry (the part of
    Gamma that real integrals see); one expectation value per class"""
    L = len(layouts)
    orbits = eri_orbits(L)
    g = eri_gradient(ctx, layouts, data, orbits)
    gam = np.zeros((L, L, L, L))
    for e, v in zip(orbits, g):
        imgs = eri_images(*e)
        for a, b, c, d in imgs:
            gam[a, b, c, d] = 2.0 * v / len(imgs)
    return gam


def one_rdm(ctx, layouts, data):
    """spin-summed 1-RDM gamma_pq = <E_pq> (real symmetric for a real state)"""
    L = len(layouts)
    pairs = [(p, q) for p, q in itertools.combinations_with_replacement(range(L), 2)]
    g = h_gradient(ctx, layouts, data, pairs)
    gamma = np.zeros((L, L))
    for (p, q), v in zip(pairs, g):
        gamma[p, q] = gamma[q, p] = v if p == q else 0.5 * v
    return gamma
