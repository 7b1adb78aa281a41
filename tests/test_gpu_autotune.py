"""GEMM tile autotuner (MQ_OPT_AUTOTUNE) on the device: plans whose GEMM stages are scheduled by
the timed pick of the candidate tile rules give the oracle's H*psi (1e-12 relative) at every site of
a chain with sectors above 64 states; every stage key is recorded once with a pick among the
candidates, and a second plan of the same site problem takes the recorded picks without timing."""
import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.engine import Context, HpsiPlan
from qcmaquis_amd.synthetic import TWOU1_SITE, fuse_index
from tests.chains import make_chain, rel_err
from tests.test_gpu_parity import _grow_both

pytestmark = pytest.mark.gpu


def test_autotuned_plans_match_oracle(oracle_lib):
    print("autotuned plans, 4-site two-site chain, sectors up to 90", flush=True)
    ctx = Context(0, abi.MQ_SYM_TWOU1)
    ctx.set_option(abi.MQ_OPT_AUTOTUNE, 1)
    ch = make_chain(5, n_sites=4, total=(4, 4), max_m=90, D=6, phys=fuse_index(TWOU1_SITE, TWOU1_SITE), density=0.3)
    mpos, L_dev, L_or, R_dev, R_or = _grow_both(ctx, oracle_lib, ch)
    rng = np.random.default_rng(3)
    for k in range(len(ch.mpos)):
        lay = ch.layouts[k]
        plan = HpsiPlan(ctx, mpos[k], L_dev[k], R_dev[k + 1], lay)
        st = plan.stats()
        psi = rng.normal(size=lay.n)
        got = plan.apply(psi)
        ref = oracle_lib.hpsi(ch.sym, ch.mpos[k], L_or[k], R_or[k + 1], lay, psi)
        n_keys = len(ctx.tuned())
        again = HpsiPlan(ctx, mpos[k], L_dev[k], R_dev[k + 1], lay)
        st2 = again.stats()
        got2 = again.apply(psi)
        print(f"  site {k}: picks T {st['gemm_policy_t']} L*S {st['gemm_policy_c']}, tuning {st['seconds_tune']:.3f} s "
              f"(again {st2['seconds_tune']:.3f} s), rel err {rel_err(got, ref):.2e}", flush=True)
        assert rel_err(got, ref) < 1e-12 and np.array_equal(got, got2)
        assert len(ctx.tuned()) == n_keys
        assert (st2["gemm_policy_t"], st2["gemm_policy_c"]) == (st["gemm_policy_t"], st["gemm_policy_c"])
        plan.close()
        again.close()
    table = ctx.tuned()
    assert table and all(0 <= pick < 7 and best <= dflt * 1.0001 for _, kind, pick, best, dflt in table)
    for key, kind, pick, best, dflt in table:
        print(f"  key {key:#x} kind {kind}: pick {pick}, {best:.4f} ms (default {dflt:.4f} ms)", flush=True)
