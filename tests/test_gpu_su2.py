"""SU2U1 quantum chemistry on the device.

* one two-site H*psi (pair-coupled basis) and left / right growth steps of the SU2U1 QC MPO
  through the C ABI vs the CPU oracle: FP64 to 1e-12, FP32 to 1e-5 (relative);
* SU2U1 two-site DMRG sweeps with untruncated bond dimension reproduce the determinant FCI
  energies of tests/golden/fci_golden.json: the S = 0 ground states of the 6- and 8-orbital
  problems (configuration 1 integrals) and, with 2S = 2, the lowest triplet, to 1e-9 Eh;
* SU2U1 single-site DMRG on the same problem converges to the singlet FCI energy.
"""
import json
import os

import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.dmrg import SingleSiteDMRG, TwoSiteDMRG, initial_bonds, right_normalize
from qcmaquis_amd.engine import DeviceBoundary, DeviceMPO, DeviceMPS, HpsiPlan, lstep, rstep
from qcmaquis_amd.model import MPSLayout, initial_boundary
from qcmaquis_amd.qc import QCHamiltonian, make_integrals

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fci_golden.json")))
SU2 = abi.MQ_SYM_SU2U1


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_su2_qc_hpsi_and_growth_vs_oracle(ctxsu2, oracle_lib):
    g = GOLD["config1"]
    L = g["L"]
    print("su2 qc parity, 8 orbitals, M=24", flush=True)
    h, eri, _ = make_integrals(L, g["seed"], g["rank"])
    ham = QCHamiltonian(h, eri, symmetry=SU2)
    phys = ham.phys()
    bonds = initial_bonds(SU2, phys, L, 4, 4, 24)
    lay = [MPSLayout(bonds[k], phys, bonds[k + 1]) for k in range(L)]
    rng = np.random.default_rng(5)
    data = [rng.normal(size=l.n) for l in lay]
    right_normalize(lay, data)
    mpos = [DeviceMPO(ctxsu2, ham.mpo(k)) for k in range(L)]
    specs = [ham.mpo(k).to_spec() for k in range(L)]
    mid = L // 2 - 1
    Ld = DeviceBoundary.initial(ctxsu2, lay[0].left, 1)
    Lo = initial_boundary(lay[0].left, 1)
    for k in range(mid):
        m = DeviceMPS(ctxsu2, lay[k], data[k])
        Ld = lstep(ctxsu2, mpos[k], Ld, m, m)
        Lo = oracle_lib.lstep(SU2, specs[k], Lo, lay[k], data[k], lay[k], data[k])
        err = rel(Ld.download().data, Lo.data)
        print(f"  lstep {k}: rel err {err:.2e}", flush=True)
        assert err < 1e-12
    Rd = DeviceBoundary.initial(ctxsu2, lay[-1].right, 1)
    Ro = initial_boundary(lay[-1].right, 1)
    for k in range(L - 1, mid + 1, -1):
        m = DeviceMPS(ctxsu2, lay[k], data[k])
        Rd = rstep(ctxsu2, mpos[k], Rd, m, m)
        Ro = oracle_lib.rstep(SU2, specs[k], Ro, lay[k], data[k], lay[k], data[k])
        err = rel(Rd.download().data, Ro.data)
        print(f"  rstep {k}: rel err {err:.2e}", flush=True)
        assert err < 1e-12
    from qcmaquis_amd.dmrg import merge_two, twosite_offsets
    from qcmaquis_amd.su2 import TwoSiteReduction
    off12 = twosite_offsets(phys)
    lay2, psi2 = merge_two(lay[mid], data[mid], lay[mid + 1], data[mid + 1], ham.phys(True), off12)
    x = TwoSiteReduction(lay2, phys, off12).reduce(psi2)
    plan = HpsiPlan(ctxsu2, DeviceMPO(ctxsu2, ham.mpo(mid, True)), Ld, Rd, lay2)
    ref = oracle_lib.hpsi(SU2, ham.mpo(mid, True).to_spec(), Lo, Ro, lay2, x)
    out = plan.apply(x)
    err = rel(out, ref)
    out32 = plan.apply(x, precision=abi.MQ_PREC_FP32)
    err32 = rel(out32, ref)
    print(f"  two-site H*psi n={lay2.n}: FP64 rel err {err:.2e}, FP32 {err32:.2e}", flush=True)
    assert err < 1e-12 and err32 < 1e-5
    plan.close()


@pytest.mark.parametrize("name,n_up,n_dn,key,M", [("small6", 3, 3, "e0", 400), ("small6", 4, 2, "e1", 400),
                                                  ("config1", 4, 4, "e0", 300)])
def test_su2_sweeps_equal_fci(ctxsu2, name, n_up, n_dn, key, M):
    g = GOLD[name]
    print(f"su2 sweeps {name} 2S={n_up - n_dn} M={M}", flush=True)
    h, eri, ec = make_integrals(g["L"], g["seed"], g["rank"])
    ham = QCHamiltonian(h, eri, ec, symmetry=SU2)
    d = TwoSiteDMRG(ham, n_up, n_dn, M=M, ctx=ctxsu2, tol=1e-10, max_iter=12)
    for s in range(4):
        e, dt = d.sweep()
        print(f"  sweep {s}: E={e:.12f} ({dt:.2f} s) max M {max(sum(l.right.sizes) for l in d.lay)}", flush=True)
    assert abs(e - g[key]) < 1e-9, (e, g[key])


def test_su2_single_site_sweeps(ctxsu2):
    g = GOLD["small6"]
    print("su2 single-site sweeps small6", flush=True)
    h, eri, ec = make_integrals(g["L"], g["seed"], g["rank"])
    ham = QCHamiltonian(h, eri, ec, symmetry=SU2)
    d = SingleSiteDMRG(ham, 3, 3, M=400, ctx=ctxsu2, tol=1e-10, max_iter=20)
    for s in range(4):
        e, _ = d.sweep()
        print(f"  sweep {s}: E={e:.12f}", flush=True)
    assert abs(e - g["e0"]) < 1e-8
