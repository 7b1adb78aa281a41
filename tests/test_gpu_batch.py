"""Batch plans (mq_hpsi_plan_create_batch): independent site problems of different sizes in one
plan.

* every instance's segment of the batched H*psi equals that instance's own plan's H*psi (FP64,
  1e-12 relative; FP32 to 1e-5 of the FP64 result), and the batch's tables and flop counts are the
  sums of the instances';
* mq_davidson refuses a batch plan (the solver works on one instance);
* SU2U1: a batch of every site of two random chains, with small-tile lines, equals the CPU
  oracle's H*psi of each instance to 1e-12;
* the lock-step Davidson over the batch, on the host (batch.py) and on the device
  (mq_davidson_batch), follows every instance's own device Davidson (same iteration): the same
  eigenvalue and vector, the same iteration counts, and the residual rule holds where it stopped an
  instance.
"""
import ctypes as C

import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.batch import lockstep_davidson
from qcmaquis_amd.dmrg import TwoSiteDMRG, merge_two
from qcmaquis_amd.engine import DeviceMPS, EngineError, HpsiPlan, HpsiPlanBatch, _dp, lstep
from qcmaquis_amd.qc import QCHamiltonian, make_integrals

pytestmark = pytest.mark.gpu

# (orbitals, n_up, n_dn, M, seed, site): uneven sizes, including a tiny one
CASES = [(6, 3, 3, 16, 1, 1), (5, 3, 2, 8, 2, 0), (4, 1, 1, 4, 3, 2), (7, 4, 3, 24, 4, 2)]


def _instances(ctx):
    out = []
    for L, nu, nd, M, seed, i in CASES:
        h, eri, _ = make_integrals(L, seed, rank=8)
        d = TwoSiteDMRG(QCHamiltonian(h, eri), nu, nd, M=M, ctx=ctx, init_m=M, seed=seed)
        for k in range(i):          # grow the left boundaries up to site i
            m = DeviceMPS(ctx, d.lay[k], d.data[k])
            d.left[k + 1] = lstep(ctx, d.mpo1[k], d.left[k], m, m)
        lay2, psi2 = merge_two(d.lay[i], d.data[i], d.lay[i + 1], d.data[i + 1], d.phys2, d.off12)
        out.append(dict(d=d, item=(d.ts_mpo(i), d.left[i], d.right[i + 2], lay2), psi=psi2))
    return out


def test_batch_hpsi_matches_instances(ctx2u1):
    inst = _instances(ctx2u1)
    plans = [HpsiPlan(ctx2u1, *x["item"]) for x in inst]
    batch = HpsiPlanBatch(ctx2u1, [x["item"] for x in inst])
    sizes = [p.layout.n for p in plans]
    print(f"batch of {len(plans)}: psi sizes {sizes}", flush=True)
    assert batch.n == sum(sizes)
    rng = np.random.default_rng(5)
    psi = rng.normal(size=batch.n)
    got = batch.split(batch.apply(psi))
    got32 = batch.split(batch.apply(psi, abi.MQ_PREC_FP32))
    got3 = batch.split(batch.apply(psi, abi.MQ_PREC_BF16X3))  # split-BF16 GEMM products
    for p, x, g, g32, g3 in zip(plans, batch.split(psi), got, got32, got3):
        ref = p.apply(x)
        err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
        err32 = np.linalg.norm(g32 - ref) / np.linalg.norm(ref)
        err3 = np.linalg.norm(g3 - ref) / np.linalg.norm(ref)
        print(f"  n={x.size}: fp64 {err:.2e} fp32 {err32:.2e} bf16x3 {err3:.2e}", flush=True)
        assert err < 1e-12 and err32 < 1e-5 and err3 < 5e-4
    bs = batch.stats()
    for f in ("psi_elements", "n_t_tasks", "n_cohorts", "n_lines", "n_terms", "flops_total", "flops_gemm",
              "bytes_compulsory", "t_elements", "s_elements", "l_panel_elements"):
        assert bs[f] == sum(p.stats()[f] for p in plans), f
    # term-sum blocks are packed over the concatenated line list, so a batch needs at most as many
    assert bs["n_sblocks"] <= sum(p.stats()["n_sblocks"] for p in plans)
    # the device solver takes one instance only
    ev, st = C.c_double(), abi.DavidsonStats()
    z = np.zeros(batch.n)
    out = np.zeros(batch.n)
    arr = (C.POINTER(C.c_double) * 1)()
    with pytest.raises(EngineError):
        ctx2u1.check(ctx2u1.lib.mq_davidson(batch.h, _dp(z), 0, arr, 1e-8, 4, abi.MQ_PREC_FP64, C.byref(ev),
                                            _dp(out), C.byref(st)))


def test_lockstep_davidson(ctx2u1):
    inst = _instances(ctx2u1)
    batch = HpsiPlanBatch(ctx2u1, [x["item"] for x in inst])
    psi0 = [x["psi"] for x in inst]
    e, vecs, its, n_hpsi = lockstep_davidson(batch, psi0, tol=1e-10, max_iter=40)
    print(f"lock-step: energies {e}, iterations {its}, batched H*psi {n_hpsi}", flush=True)
    assert n_hpsi == its.max()
    for x, ei, v, it in zip(inst, e, vecs, its):
        plan = HpsiPlan(ctx2u1, *x["item"])
        ref, vref, st = plan.davidson(x["psi"], tol=1e-10, max_iter=40)
        print(f"  n={v.size}: {ei:.12f} device {ref:.12f}, iterations {it} / {st['iterations']}", flush=True)
        assert abs(ei - ref) < 1e-8 * max(1.0, abs(ref))
        assert abs(v @ vref) > 1 - 1e-6
        if it < 40:         # stopped on the residual rule
            r = plan.apply(v) - ei * v
            assert np.linalg.norm(r) <= 1e-10 * max(1.0, abs(ei)) + 1e-11


@pytest.mark.parametrize("precision", [abi.MQ_PREC_FP64, abi.MQ_PREC_FP32])
def test_device_lockstep_davidson(ctx2u1, precision):
    """mq_davidson_batch: every instance's eigenvalue, vector and iteration count equal those of
    its own plan's mq_davidson (same iteration; FP64, and FP32 contraction with FP64 refinement)"""
    inst = _instances(ctx2u1)
    batch = HpsiPlanBatch(ctx2u1, [x["item"] for x in inst])
    e, vecs, its, st = batch.davidson([x["psi"] for x in inst], tol=1e-10, max_iter=40, precision=precision)
    print(f"device lock-step (precision {precision}): {e}, iterations {its}, lock-step {st['iterations']}", flush=True)
    assert st["iterations"] == (its.max() if precision == abi.MQ_PREC_FP64 else st["iterations"])
    for x, ei, v, it in zip(inst, e, vecs, its):
        plan = HpsiPlan(ctx2u1, *x["item"])
        ref, vref, rst = plan.davidson(x["psi"], tol=1e-10, max_iter=40, precision=precision)
        print(f"  n={v.size}: {ei:.12f} device {ref:.12f}, iterations {it} / {rst['iterations']}", flush=True)
        assert abs(ei - ref) < 1e-9 * max(1.0, abs(ref))
        assert abs(abs(v @ vref) - 1) < 1e-6
        if precision == abi.MQ_PREC_FP64:
            assert it == rst["iterations"]


def test_batch_su2u1_chains_vs_oracle(ctxsu2, oracle_lib):
    """a batch of SU2U1 site problems from two random chains (every site of each): each segment
    of the batched H*psi equals the oracle's H*psi of its instance to 1e-12 (FP64), with the
    small-tile term sum forced on for tiles of up to 8 elements"""
    from tests.chains import make_chain, oracle_left, oracle_right, rel_err
    from qcmaquis_amd.engine import DeviceBoundary, DeviceMPO
    items, refs = [], []
    for seed, max_m, D in ((1, 3, 6), (4, 8, 9)):
        ch = make_chain(seed, sym=abi.MQ_SYM_SU2U1, n_sites=5, total=(4, 0), max_m=max_m, D=D)
        n = len(ch.mpos)
        L_or, R_or = oracle_left(oracle_lib, ch, n - 1), oracle_right(oracle_lib, ch, 1)
        for k in range(n):
            items.append((DeviceMPO(ctxsu2, ch.mpos[k]), DeviceBoundary.upload(ctxsu2, L_or[k]),
                          DeviceBoundary.upload(ctxsu2, R_or[k + 1]), ch.layouts[k]))
            refs.append((ch, k, L_or[k], R_or[k + 1]))
    ctxsu2.set_option(abi.MQ_OPT_SMALL_TILE_MAX, 8)
    try:
        batch = HpsiPlanBatch(ctxsu2, items)
    finally:
        ctxsu2.set_option(abi.MQ_OPT_SMALL_TILE_MAX, abi.MQ_SMALL_TILE_AUTO)
    print(f"su2u1 batch of {len(items)}: psi {batch.n}, small-tile lines {batch.stats()['small_lines']}", flush=True)
    assert batch.stats()["small_lines"] > 0
    psi = np.random.default_rng(3).normal(size=batch.n)
    for x, g, (ch, k, L, R) in zip(batch.split(psi), batch.split(batch.apply(psi)), refs):
        ref = oracle_lib.hpsi(ch.sym, ch.mpos[k], L, R, ch.layouts[k], x)
        assert rel_err(g, ref) < 1e-12, k
