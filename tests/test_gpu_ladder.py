"""The reference's Hubbard-ladder regression problem on the device (open ladder of 6 rungs,
t = t' = 1, U = 8, N_up = N_down = 10; integrals qc.ladder_hubbard through the QC MPO builder),
against the reference's own exact-diagonalisation results (tests/golden/hubbard_ladder.json):
two-site DMRG at M = 400 in TwoU1 and in SU2U1 (the ground state is a singlet) reaches the ED
energy to 1e-6, and the spin-summed one-body density matrix of the TwoU1 state (derivative-MPO
expectation values) equals the reference's up + down matrices to 1e-5."""
import json
import os

import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.dmrg import TwoSiteDMRG
from qcmaquis_amd.qc import QCHamiltonian, ladder_hubbard
from qcmaquis_amd.rdm import one_rdm

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "hubbard_ladder.json")))


def reference_onebody():
    n = GOLD["model"]["sites"]
    g = np.zeros((n, n))
    for spin in ("up", "down"):
        for i, j, v in GOLD["ed_onebody_" + spin]:
            g[i, j] += v
    return g


@pytest.mark.parametrize("sym", [abi.MQ_SYM_TWOU1, abi.MQ_SYM_SU2U1])
def test_ladder_dmrg_matches_reference_ed(ctx2u1, ctxsu2, sym):
    m = GOLD["model"]
    print(f"hubbard ladder, symmetry {sym}", flush=True)
    h, eri = ladder_hubbard(m["L"], m["t"], m["t_rung"], m["U"])
    ham = QCHamiltonian(h, eri, symmetry=sym)
    ctx = ctx2u1 if sym == abi.MQ_SYM_TWOU1 else ctxsu2
    d = TwoSiteDMRG(ham, m["n_up"], m["n_dn"], M=400, ctx=ctx, tol=1e-10, max_iter=12)
    for s in range(4):
        e, dt = d.sweep()
        print(f"  sweep {s}: E={e:.12f} ({dt:.2f} s)", flush=True)
    assert abs(e - GOLD["ed_energy"]) < 1e-6, (e, GOLD["ed_energy"])
    if sym == abi.MQ_SYM_TWOU1:
        gamma = one_rdm(ctx, d.lay, d.data)
        err = np.max(np.abs(gamma - reference_onebody()))
        print(f"  one-body density matrix max |diff| {err:.2e}", flush=True)
        assert err < 1e-5
