"""Helpers shared by the parity tests: boundaries of a synthetic chain grown with the oracle or
with the HIP engine, exactly as a sweep initialises them (optimize/optimize.h:158-195)."""
import numpy as np

from qcmaquis_amd import abi
from qcmaquis_amd.model import initial_boundary
from qcmaquis_amd.synthetic import Chain


def make_chain(seed, sym=abi.MQ_SYM_TWOU1, n_sites=4, total=(2, 2), max_m=4, D=6, density=0.4, phys=None):
    return Chain(np.random.default_rng(seed), sym, n_sites, total, max_m, D, phys=phys, density=density)


def oracle_left(oracle, ch, upto, symmetric=True):
    """[L_0, ..., L_upto] with L_k the boundary of bond k"""
    L = [initial_boundary(ch.bonds[0], ch.mpos[0].row_dim)]
    for k in range(upto):
        L.append(oracle.lstep(ch.sym, ch.mpos[k], L[-1], ch.layouts[k], ch.data[k], ch.layouts[k], ch.data[k],
                              symmetric=symmetric))
    return L


def oracle_right(oracle, ch, downto, symmetric=True):
    """{k: R_k} for k = n ... downto, R_k the boundary of bond k"""
    n = len(ch.mpos)
    R = {n: initial_boundary(ch.bonds[n], ch.mpos[n - 1].col_dim)}
    for k in range(n - 1, downto - 1, -1):
        R[k] = oracle.rstep(ch.sym, ch.mpos[k], R[k + 1], ch.layouts[k], ch.data[k], ch.layouts[k], ch.data[k],
                            symmetric=symmetric)
    return R


def same_index(a, b):
    """bit-exact comparison of two BoundarySpec indices (integers, scales)"""
    assert a.bra == b.bra and a.ket == b.ket and a.aux_dim == b.aux_dim
    assert len(a.cohorts) == len(b.cohorts)
    for x, y in zip(a.cohorts, b.cohorts):
        assert tuple(x["lr"]) == tuple(y["lr"])
        assert list(x["offsets"]) == list(y["offsets"])
        assert int(x.get("tr", 0)) == int(y.get("tr", 0))
        assert list(x.get("transposes", [0] * a.aux_dim)) == list(y.get("transposes", [0] * a.aux_dim))
        assert np.array_equal(np.asarray(x.get("conj_scales", [1.0] * a.aux_dim)),
                              np.asarray(y.get("conj_scales", [1.0] * a.aux_dim)))


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = max(np.max(np.abs(b)), 1e-300) if b.size else 1.0
    return float(np.max(np.abs(a - b)) / den) if a.size else 0.0
