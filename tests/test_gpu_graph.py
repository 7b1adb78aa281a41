"""H*psi as a captured HIP graph (mq_hpsi_plan_use_graphs) against the eager launches.

Every replay is checked against the eager result of the same psi, with eager H*psi work and host
synchronisations between the replays (graph nodes recorded once, replayed with other work
interleaved); FP64 replays are bitwise equal to the eager path (same kernels, same order), FP32
replays equal the eager FP32 result.  The Jacobi-Davidson eigenpair is unchanged with graphs.
"""
import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.engine import DeviceBoundary, DeviceMPO, DeviceMPS, HpsiPlan, lstep, rstep
from qcmaquis_amd.synthetic import TWOU1_SITE, fuse_index
from tests.chains import make_chain

pytestmark = pytest.mark.gpu


def _site_plans(ctx):
    ch = make_chain(5, n_sites=4, total=(4, 4), max_m=90, D=6, phys=fuse_index(TWOU1_SITE, TWOU1_SITE), density=0.3)
    mpos = [DeviceMPO(ctx, m) for m in ch.mpos]
    mps = [DeviceMPS(ctx, lay, d) for lay, d in zip(ch.layouts, ch.data)]
    n = len(ch.mpos)
    L = [DeviceBoundary.initial(ctx, ch.bonds[0], ch.mpos[0].row_dim)]
    for k in range(n - 1):
        L.append(lstep(ctx, mpos[k], L[-1], mps[k], mps[k]))
    R = {n: DeviceBoundary.initial(ctx, ch.bonds[n], ch.mpos[n - 1].col_dim)}
    for k in range(n - 1, 0, -1):
        R[k] = rstep(ctx, mpos[k], R[k + 1], mps[k], mps[k])
    k = 1
    return ch, (HpsiPlan(ctx, mpos[k], L[k], R[k + 1], ch.layouts[k]),
                HpsiPlan(ctx, mpos[k], L[k], R[k + 1], ch.layouts[k])), ch.layouts[k]


@pytest.mark.parametrize("precision", [abi.MQ_PREC_FP64, abi.MQ_PREC_FP32])
def test_graph_replays_match_eager(ctx2u1, precision):
    print(f"graph replays precision={precision}", flush=True)
    ch, (gplan, eplan), lay = _site_plans(ctx2u1)
    rng = np.random.default_rng(3)
    psis = [rng.normal(size=lay.n) for _ in range(4)]
    eager = [eplan.apply(p, precision) for p in psis]
    gplan.use_graphs(True)
    for r in range(3):
        for i, p in enumerate(psis):
            got = gplan.apply(p, precision)
            if precision == abi.MQ_PREC_FP64:
                assert np.array_equal(got, eager[i]), (r, i)
            else:
                assert np.allclose(got, eager[i], rtol=0, atol=1e-6 * np.abs(eager[i]).max()), (r, i)
            # eager work between the replays
            other = eplan.apply(psis[(i + 1) % len(psis)], precision)
            assert np.allclose(other, eager[(i + 1) % len(psis)], rtol=0, atol=1e-6 * np.abs(other).max())
        print(f"replay round {r} ok", flush=True)


def test_graph_davidson(ctx2u1):
    print("graph davidson", flush=True)
    ch, (gplan, eplan), lay = _site_plans(ctx2u1)
    psi0 = np.random.default_rng(4).normal(size=lay.n)
    e0, v0, s0 = eplan.davidson(psi0, tol=1e-10, max_iter=20)
    gplan.use_graphs(True)
    e1, v1, s1 = gplan.davidson(psi0, tol=1e-10, max_iter=20)
    assert s0["iterations"] == s1["iterations"]
    assert e0 == e1
    assert np.array_equal(v0, v1)
