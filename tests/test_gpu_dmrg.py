"""Device Jacobi-Davidson and full two-site sweeps on the quantum-chemistry Hamiltonian.

* Davidson's lowest eigenvalue vs a dense diagonalisation of H_eff assembled column by column
  with the CPU oracle (H_eff symmetric to 1e-12: a Hermitian MPO with compressed boundaries);
* two-site DMRG on the seeded 6- and 8-orbital problems vs the determinant FCI energies committed
  in tests/golden/fci_golden.json: untruncated bond dimension reproduces FCI to 1e-9 Eh, M = 64
  is variational (above FCI) and within 1e-3 Eh (configuration 1: 8 orbitals, 8 electrons).
"""
import json
import os

import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.dmrg import TwoSiteDMRG, merge_two
from qcmaquis_amd.engine import HpsiPlan
from qcmaquis_amd.qc import QCHamiltonian, make_integrals

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fci_golden.json")))


def test_davidson_vs_dense(ctx2u1, oracle_lib):
    print("davidson vs dense H_eff", flush=True)
    h, eri, _ = make_integrals(5, 3, rank=6)
    ham = QCHamiltonian(h, eri)
    d = TwoSiteDMRG(ham, 2, 3, M=6, ctx=ctx2u1, init_m=3, seed=1)
    i = 0
    lay2, psi2 = merge_two(d.lay[i], d.data[i], d.lay[i + 1], d.data[i + 1], d.phys2, d.off12)
    plan = HpsiPlan(ctx2u1, d.ts_mpo(i), d.left[i], d.right[i + 2], lay2)
    n = lay2.n
    Ls, Rs = d.left[i].download(), d.right[i + 2].download()
    spec = ham.mpo(i, True).to_spec()
    H = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        H[:, c] = oracle_lib.hpsi(abi.MQ_SYM_TWOU1, spec, Ls, Rs, lay2, e)
    assert np.allclose(H, H.T, atol=1e-10 * np.abs(H).max())
    w = np.linalg.eigvalsh(0.5 * (H + H.T))
    ev, vec, st = plan.davidson(psi2, tol=1e-10, max_iter=min(n, 60))
    print(f"n={n} davidson {ev:.12f} dense {w[0]:.12f} iters {st['iterations']}", flush=True)
    assert abs(ev - w[0]) < 1e-8 * max(1.0, abs(w[0]))
    r = plan.apply(vec) - ev * vec
    assert np.linalg.norm(r) < 1e-6
    # ortho vectors: the next state is found when the ground state is projected out
    ev1, vec1, _ = plan.davidson(psi2, ortho=[vec], tol=1e-10, max_iter=min(n, 60))
    assert abs(ev1 - w[1]) < 1e-7 * max(1.0, abs(w[1]))
    assert abs(vec1 @ vec) < 1e-8
    # mixed precision: FP32 H*psi to the FP32 floor, then FP64 refinement to the same tolerance
    ev2, vec2, st2 = plan.davidson(psi2, tol=1e-10, max_iter=min(n, 60), precision=abi.MQ_PREC_FP32)
    print(f"mixed davidson {ev2:.12f} iters {st2['iterations']} refine {st2['refine_iterations']}", flush=True)
    assert abs(ev2 - w[0]) < 1e-8 * max(1.0, abs(w[0]))
    assert st2["refine_iterations"] >= 1


def test_davidson_gmres_matches_restatement(ctx2u1, oracle_lib):
    """mq_davidson with and without the GMRES correction (MQ_OPT_JCD_GMRES) against the oracle's
    restatement of the reference's Jacobi-Davidson on the dense H_eff: the same Ritz value after
    every budget, the same iteration counts, the lowest eigenvalue at convergence"""
    from oracle.jcd import jcd
    print("davidson with GMRES correction vs the restatement", flush=True)
    h, eri, _ = make_integrals(6, 4, rank=6)
    ham = QCHamiltonian(h, eri)
    d = TwoSiteDMRG(ham, 3, 3, M=12, ctx=ctx2u1, init_m=6, seed=2)
    i = 2
    for k in range(i):  # the left boundaries up to the site problem
        d.step(k, True)
    lay2, psi2 = merge_two(d.lay[i], d.data[i], d.lay[i + 1], d.data[i + 1], d.phys2, d.off12)
    plan = HpsiPlan(ctx2u1, d.ts_mpo(i), d.left[i], d.right[i + 2], lay2)
    n = lay2.n
    H = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        H[:, c] = plan.apply(e)
    H = 0.5 * (H + H.T)
    w = np.linalg.eigvalsh(H)
    try:
        for steps in (0, 3, 6):
            ctx2u1.set_option(abi.MQ_OPT_JCD_GMRES, steps)
            for budget, tol in ((4, 1e-12), (40, 1e-10)):
                ev, vec, st = plan.davidson(psi2, tol=tol, max_iter=min(n, budget))
                th, u, it, _ = jcd(H, psi2, tol=tol, max_iter=min(n, budget), gmres_steps=steps)
                print(f"  gmres {steps} budget {budget}: device {ev:.12f} ({st['iterations']} it), restated "
                      f"{th:.12f} ({it} it), lowest {w[0]:.12f}", flush=True)
                assert st["iterations"] == it
                assert abs(ev - th) < 1e-9 * max(1.0, abs(th))
                if budget == 40:
                    assert abs(ev - w[0]) < 1e-8 * max(1.0, abs(w[0]))
    finally:
        ctx2u1.set_option(abi.MQ_OPT_JCD_GMRES, 0)
    plan.close()


@pytest.mark.parametrize("M,tol", [(400, 1e-9), (16, None)])
def test_sweeps_small6(ctx2u1, M, tol):
    g = GOLD["small6"]
    print(f"small6 sweeps M={M}", flush=True)
    h, eri, ec = make_integrals(g["L"], g["seed"], g["rank"])
    ham = QCHamiltonian(h, eri, ec)
    d = TwoSiteDMRG(ham, g["n_up"], g["n_dn"], M=M, ctx=ctx2u1, tol=1e-10, max_iter=12)
    for s in range(4):
        e, dt = d.sweep()
        print(f"  sweep {s}: E={e:.12f} ({dt:.2f} s)", flush=True)
    if tol is not None:
        assert abs(e - g["e0"]) < tol
    else:
        assert e > g["e0"] - 1e-10 and e - g["e0"] < 5e-2


def test_sweeps_config1(ctx2u1):
    """configuration 1: 8 orbitals, 8 electrons, S_z = 0, M = 64, 4 sweeps, tol 1e-8"""
    g = GOLD["config1"]
    print("config1 sweeps M=64", flush=True)
    h, eri, ec = make_integrals(g["L"], g["seed"], g["rank"])
    ham = QCHamiltonian(h, eri, ec)
    d = TwoSiteDMRG(ham, g["n_up"], g["n_dn"], M=64, ctx=ctx2u1, tol=1e-8, max_iter=8)
    for s in range(4):
        e, dt = d.sweep()
        print(f"  sweep {s}: E={e:.12f} dE_fci={e - g['e0']:.3e} ({dt:.2f} s)", flush=True)
    assert e > g["e0"] - 1e-10
    assert e - g["e0"] < 1e-3


def test_single_site_sweeps_small6(ctx2u1):
    """single-site DMRG (ss_optimize without noise) on bonds that span the whole space reaches the
    FCI energy; the energy never rises from one site problem to the next (variational)"""
    from qcmaquis_amd.dmrg import SingleSiteDMRG
    g = GOLD["small6"]
    print("small6 single-site sweeps", flush=True)
    h, eri, ec = make_integrals(g["L"], g["seed"], g["rank"])
    d = SingleSiteDMRG(QCHamiltonian(h, eri, ec), g["n_up"], g["n_dn"], M=400, ctx=ctx2u1, tol=1e-11, max_iter=20)
    for s in range(6):
        e, dt = d.sweep()
        print(f"  sweep {s}: E={e:.12f} ({dt:.2f} s)", flush=True)
    assert abs(e - g["e0"]) < 1e-8
    assert np.all(np.diff(d.energies) < 1e-9)
    # truncated bonds (at most 4 states per sector): variational, monotone, above FCI
    d = SingleSiteDMRG(QCHamiltonian(h, eri, ec), g["n_up"], g["n_dn"], M=4, ctx=ctx2u1, tol=1e-10, max_iter=20)
    for s in range(4):
        e, _ = d.sweep()
    print(f"  M=4 per sector: E={e:.12f} (FCI {g['e0']:.12f})", flush=True)
    assert e > g["e0"] - 1e-10 and e - g["e0"] < 1.0
    assert np.all(np.diff(d.energies) < 1e-9)


def test_result_file_and_restart_small6(ctx2u1, tmp_path):
    """sweeps write the reference's result layout (per-step Energy, BondDimension and truncation
    records per sweep, the lowest energy); a checkpoint after three sweeps restarts the run, whose
    next sweep continues the numbering and stays at the FCI energy"""
    from qcmaquis_amd import results as R
    g = GOLD["small6"]
    h, eri, ec = make_integrals(g["L"], g["seed"], g["rank"])
    ham = QCHamiltonian(h, eri, ec)
    rf, chk = str(tmp_path / "res.h5"), str(tmp_path / "chkp")
    print("small6 result file and restart", flush=True)
    d = TwoSiteDMRG(ham, g["n_up"], g["n_dn"], M=400, ctx=ctx2u1, tol=1e-10, max_iter=12, resultfile=rf)
    for _ in range(3):
        d.sweep()
    d.checkpoint(chk)
    ar = R.ResultArchive(rf)
    assert ar.sweeps() == [0, 1, 2]
    for k in range(3):
        e = ar.iteration(k, "Energy")
        assert e.size == 2 * (g["L"] - 1)
        assert R.info(rf, "/spectrum/iteration/%d/results/BondDimension/mean/value" % k)[0] == R.INTEGER
        assert ar.iteration(k, "TruncatedWeight").size == e.size
    assert ar.energy() == ar.iteration(2, "Energy").min()
    lay, data, status = R.load_checkpoint(chk)
    assert status == {"sweep": 2, "site": -1}
    d2 = TwoSiteDMRG(ham, g["n_up"], g["n_dn"], M=400, ctx=ctx2u1, tol=1e-10, max_iter=12, mps=(lay, data),
                     resultfile=rf, first_sweep=status["sweep"] + 1)
    e, _ = d2.sweep()
    print(f"  restarted sweep: E={e:.12f} (FCI {g['e0']:.12f})", flush=True)
    assert abs(e - g["e0"]) < 1e-9
    assert R.ResultArchive(rf).sweeps() == [0, 1, 2, 3]


def test_batch_sweeps_match_individual_runs(ctx2u1):
    """three runs of different lengths swept concurrently (BatchTwoSiteDMRG: one batch plan and one
    device lock-step Davidson per step) give, step by step, the energies of the same runs swept one
    by one (same start states, same iteration); the 6-orbital run reaches its FCI energy"""
    from qcmaquis_amd.batch import BatchTwoSiteDMRG
    g = GOLD["small6"]
    specs = [(g["L"], g["n_up"], g["n_dn"], g["seed"], g["rank"]), (5, 3, 2, 11, 6), (4, 2, 2, 12, 5)]

    def runs():
        out = []
        for L, nu, nd, seed, rank in specs:
            h, eri, ec = make_integrals(L, seed, rank)
            out.append(TwoSiteDMRG(QCHamiltonian(h, eri, ec), nu, nd, M=400, ctx=ctx2u1, tol=1e-10, max_iter=12,
                                   seed=seed))
        return out

    print("batch sweeps vs individual runs", flush=True)
    single = runs()
    for r in single:
        for _ in range(3):
            r.sweep()
    batch = BatchTwoSiteDMRG(runs())
    for _ in range(3):
        batch.sweep()
    for a, b in zip(single, batch.runs):
        ea, eb = np.array(a.iteration_results["Energy"]), np.array(b.iteration_results["Energy"])
        print(f"  L={a.L}: last sweep max |dE| {np.abs(ea - eb).max():.2e}, E={eb[-1]:.12f}", flush=True)
        assert ea.shape == eb.shape and np.abs(ea - eb).max() < 1e-9
        assert b.sweep_number == 3
    assert abs(batch.runs[0].energies[-1] - g["e0"]) < 1e-9
