"""Density-matrix split with the noise term on the device (dmrg.predict_split: the identity-bra
growth step and mq_boundary_gram; the reference's predict_split_l2r / predict_split_r2l):

* TwoU1 chain problem, both directions: the split equals the one of the dense density matrix
  X X^T + alpha sum_b S_b S_b^T (X^T X + ... right to left), S_b from dense L (R), W and psi;
* SU2U1 quantum chemistry (8 orbitals) after a sweep with the noise split: at the middle bond the
  device split equals the split with the oracle's growth steps and a host Gram sum, both
  directions (same kept states, record to 1e-7, A B to 1e-10 relative);
* SU2U1 sweeps with twosite_truncation = heev (a growth sweep at alpha 1e-2, then alpha 1e-8) and
  untruncated bond dimension reach the determinant FCI energy, as the SVD sweeps do;
* the single-site subspace expansion (dmrg.grow_sweep_step) on the device against the dense
  density matrices (TwoU1, both directions, truncating and growing), and single-site sweeps with
  the expansion from 4 states per sector reach the determinant FCI energy (SU2U1).
"""
import json
import os

import numpy as np
import pytest

from qcmaquis_amd import abi, dmrg
from qcmaquis_amd.dmrg import TwoSiteDMRG, heev_truncation, merge_two, predict_split, split_matrices
from qcmaquis_amd.engine import DeviceBoundary, DeviceMPO
from qcmaquis_amd.qc import QCHamiltonian, make_integrals
from tests.chains import oracle_left, oracle_right
from tests.test_predict_split import (FakeMPS, Wrap, _expansion_expected, _expected_dense, _one_site, _problem,
                                     check_expansion, spec_gram)

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fci_golden.json")))
SU2 = abi.MQ_SYM_SU2U1


@pytest.mark.parametrize("to_right", [True, False])
def test_device_noise_split_matches_dense_definition(ctx2u1, oracle_lib, to_right):
    print(f"noise split on the device, TwoU1 chain, to_right={to_right}", flush=True)
    ch, phys, off12, lay2, psi2 = _problem(abi.MQ_SYM_TWOU1, 31, 4)
    alpha, Mmax = 0.3, 5
    mats, dms = _expected_dense(ch, oracle_lib, phys, off12, lay2, psi2, alpha, to_right)
    if to_right:
        b, mpo = DeviceBoundary.upload(ctx2u1, oracle_left(oracle_lib, ch, 1)[1]), DeviceMPO(ctx2u1, ch.mpos[1])
    else:
        b, mpo = DeviceBoundary.upload(ctx2u1, oracle_right(oracle_lib, ch, 3)[3]), DeviceMPO(ctx2u1, ch.mpos[2])
    rec = {}
    la, a, lb, bb, _ = predict_split(ctx2u1, lay2, psi2, phys, off12, Mmax, 0.0, alpha, to_right, b, mpo, rec,
                                     abi.MQ_SYM_TWOU1)
    eig = {m: np.linalg.eigh(dm) for m, dm in dms.items()}
    keeps, want = heev_truncation({m: w[::-1] for m, (w, _) in eig.items()}, Mmax, 0.0)
    assert rec["BondDimension"] == want["BondDimension"] and abs(rec["TruncatedFraction"] - want["TruncatedFraction"]) < 1e-7
    _, got = merge_two(la, a, lb, bb, lay2.phys, off12)
    got_m = split_matrices(lay2, got, phys, off12)
    for m, (X, _, _) in mats.items():
        U = eig[m][1][:, ::-1][:, :keeps[m]]
        want_m = U @ (U.T @ X) if to_right else (X @ U) @ U.T
        assert np.max(np.abs(got_m[m][0] - want_m)) < 1e-12
    b.close()
    mpo.close()


def test_su2_noise_split_device_vs_oracle_growth(ctxsu2, oracle_lib, monkeypatch):
    L, M = 8, 24
    h, eri, ec = make_integrals(L, 21, rank=8)
    ham = QCHamiltonian(h, eri, ec, symmetry=SU2)
    print("su2 noise split, 8 orbitals, M=24: one sweep with the noise split", flush=True)
    d = TwoSiteDMRG(ham, 4, 4, M=M, ctx=ctxsu2, tol=1e-8, max_iter=8, twosite_truncation="heev",
                    alpha_initial=1e-2, ngrowsweeps=1)
    d.sweep()
    i = L // 2 - 1
    for k in range(i):
        d.step(k, True)
    lay2, psi2, red = d._merge(i)
    psi = red.unreduce(psi2)
    psi /= np.linalg.norm(psi)
    for to_right in (True, False):
        bnd, k = (d.left[i], i) if to_right else (d.right[i + 2], i + 1)
        rec_d, rec_o = {}, {}
        dev = predict_split(ctxsu2, lay2, psi, d.phys, d.off12, 12, 1e-16, 0.05, to_right, bnd, d.mpo1[k], rec_d, SU2)
        with monkeypatch.context() as mp:
            def grow(fn):
                return lambda ctx, mpo, b, bra, ket, symmetric=True: Wrap(
                    fn(SU2, mpo.spec, b.spec, bra.layout, bra.data, ket.layout, ket.data, symmetric=False, nthreads=1))
            mp.setattr(dmrg, "DeviceMPS", FakeMPS)
            mp.setattr(dmrg, "DeviceMPO", lambda ctx, spec: Wrap(spec))
            mp.setattr(dmrg, "lstep", grow(oracle_lib.lstep))
            mp.setattr(dmrg, "rstep", grow(oracle_lib.rstep))
            mp.setattr(dmrg, "boundary_gram", spec_gram)
            ref = predict_split(None, lay2, psi, d.phys, d.off12, 12, 1e-16, 0.05, to_right, Wrap(bnd.download()),
                                Wrap(ham.mpo(k).to_spec()), rec_o, SU2)
        _, got = merge_two(dev[0], dev[1], dev[2], dev[3], lay2.phys, d.off12)
        _, want = merge_two(ref[0], ref[1], ref[2], ref[3], lay2.phys, d.off12)
        err = np.max(np.abs(got - want)) / np.max(np.abs(want))
        print(f"  to_right={to_right}: kept {rec_d['BondDimension']} (oracle {rec_o['BondDimension']}), "
              f"fraction {rec_d['TruncatedFraction']:.6e} ({rec_o['TruncatedFraction']:.6e}), A B rel err {err:.2e}",
              flush=True)
        assert dev[0].right == ref[0].right and rec_d["BondDimension"] == rec_o["BondDimension"]
        assert abs(rec_d["TruncatedFraction"] - rec_o["TruncatedFraction"]) < 1e-7
        assert err < 1e-10


def test_su2_noise_sweeps_equal_fci(ctxsu2):
    g = GOLD["small6"]
    print("su2 sweeps small6 with the noise split (heev)", flush=True)
    h, eri, ec = make_integrals(g["L"], g["seed"], g["rank"])
    ham = QCHamiltonian(h, eri, ec, symmetry=SU2)
    d = TwoSiteDMRG(ham, 3, 3, M=400, ctx=ctxsu2, tol=1e-10, max_iter=12, twosite_truncation="heev",
                    alpha_initial=1e-2, ngrowsweeps=1)
    for s in range(4):
        e, dt = d.sweep()
        tw = d.iteration_results["TruncatedWeight"]
        print(f"  sweep {s}: E={e:.12f} ({dt:.2f} s) max truncated fraction {max(tw):.2e}", flush=True)
    assert abs(e - g["e0"]) < 1e-9, (e, g["e0"])


@pytest.mark.parametrize("Mmax", [4, 40])
@pytest.mark.parametrize("to_right", [True, False])
def test_device_expansion_matches_dense_definition(ctx2u1, oracle_lib, to_right, Mmax):
    print(f"single-site expansion on the device, TwoU1, to_right={to_right}, Mmax={Mmax}", flush=True)
    ch, lay, psi = _one_site(abi.MQ_SYM_TWOU1, 12, 4)
    Xs, dms, spec = _expansion_expected(ch, oracle_lib, lay, psi, 0.2, to_right)
    b, mpo = DeviceBoundary.upload(ctx2u1, spec), DeviceMPO(ctx2u1, ch.mpos[1])
    rec = {}
    out = dmrg.grow_sweep_step(ctx2u1, lay, psi, Mmax, 1e-6, 0.2, to_right, b, mpo, rec)
    check_expansion(Xs, dms, Mmax, 1e-6, to_right, out, rec)
    b.close()
    mpo.close()


def test_su2_single_site_expansion_sweeps_equal_fci(ctxsu2):
    from qcmaquis_amd.dmrg import SingleSiteDMRG
    g = GOLD["small6"]
    print("su2 single-site sweeps with the subspace expansion, small6, from 4 states per sector", flush=True)
    h, eri, ec = make_integrals(g["L"], g["seed"], g["rank"])
    ham = QCHamiltonian(h, eri, ec, symmetry=SU2)
    d = SingleSiteDMRG(ham, 3, 3, M=400, ctx=ctxsu2, tol=1e-10, max_iter=20, expansion=True, init_m=4,
                       alpha_initial=1e-2, ngrowsweeps=2)
    for s in range(8):
        e, dt = d.sweep()
        print(f"  sweep {s}: E={e:.12f} ({dt:.2f} s) bonds {[sum(l.right.sizes) for l in d.lay]}", flush=True)
    assert abs(e - g["e0"]) < 1e-8, (e, g["e0"])
