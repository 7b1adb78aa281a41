"""mq_davidson_noalloc (the Jacobi-Davidson that allocates no device memory) against mq_davidson
through the C ABI: two calls in a row, the second on a smaller site problem, each equal to the
allocating solver's eigenvalue and eigenvector (the same arithmetic in the same order: 1e-13), and
a search space beyond the scratch array refused with MQ_ERR_INVALID before any device work."""
import ctypes as C

import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.engine import DeviceBoundary, DeviceMPO, DeviceMPS, HpsiPlan, lstep, rstep
from qcmaquis_amd.synthetic import TWOU1_SITE, Chain, fuse_index

pytestmark = pytest.mark.gpu


def _problem(ctx, seed, max_m):
    rng = np.random.default_rng(seed)
    ch = Chain(rng, abi.MQ_SYM_TWOU1, 4, (4, 4), max_m, 8, phys=fuse_index(TWOU1_SITE, TWOU1_SITE))
    mpos = [DeviceMPO(ctx, m) for m in ch.mpos]
    mps = [DeviceMPS(ctx, l, d) for l, d in zip(ch.layouts, ch.data)]
    Lb = DeviceBoundary.initial(ctx, ch.bonds[0], 1)
    for k in range(2):
        Lb = lstep(ctx, mpos[k], Lb, mps[k], mps[k])
    Rb = DeviceBoundary.initial(ctx, ch.bonds[4], 1)
    Rb = rstep(ctx, mpos[3], Rb, mps[3], mps[3])
    lay = ch.layouts[2]
    plan = HpsiPlan(ctx, mpos[2], Lb, Rb, lay)
    psi0 = rng.normal(size=lay.n)
    return plan, psi0, (mpos, mps, Lb, Rb)


def _solve(ctx, plan, psi0, entry, max_iter=8, tol=1e-10):
    psi0 = np.ascontiguousarray(psi0, np.float64)
    out = np.zeros_like(psi0)
    ev = C.c_double()
    st = abi.DavidsonStats()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rc = getattr(ctx.lib, entry)(plan.h, dp(psi0), 0, None, float(tol), int(max_iter), abi.MQ_PREC_FP64,
                                 C.byref(ev), dp(out), C.byref(st))
    return rc, ev.value, out, st


def test_davidson_noalloc_equals_davidson(ctx2u1):
    keep = []
    for seed, max_m in ((5, 24), (6, 10)):  # the second problem is the smaller one
        plan, psi0, objs = _problem(ctx2u1, seed, max_m)
        keep.append(objs)
        print(f"noalloc davidson: seed {seed}, n={psi0.size}", flush=True)
        rc0, e0, v0, s0 = _solve(ctx2u1, plan, psi0, "mq_davidson")
        rc1, e1, v1, s1 = _solve(ctx2u1, plan, psi0, "mq_davidson_noalloc")
        assert rc0 == 0 and rc1 == 0, ctx2u1.lib.mq_last_error(ctx2u1.h)
        print(f"  E {e0:.14f} / {e1:.14f}, iterations {s0.iterations} / {s1.iterations}", flush=True)
        assert s0.iterations == s1.iterations
        assert abs(e0 - e1) <= 1e-13 * max(1.0, abs(e0))
        assert np.max(np.abs(v0 - v1)) <= 1e-13 * np.max(np.abs(v0))
        plan.close()


def test_davidson_noalloc_refuses_oversized_search_space(ctx2u1):
    plan, psi0, objs = _problem(ctx2u1, 7, 24)
    n = psi0.size
    max_iter = abi.MQ_NOALLOC_SCRATCH_ELEMS // (((n + 31) // 32) * 32) // 2 + 1  # 2*max_iter+2 vectors > array
    rc, _, _, _ = _solve(ctx2u1, plan, psi0, "mq_davidson_noalloc", max_iter=max_iter)
    assert rc == abi.MQ_ERR_INVALID
    assert b"scratch array" in ctx2u1.lib.mq_last_error(ctx2u1.h)
    rc, _, _, _ = _solve(ctx2u1, plan, psi0, "mq_davidson_noalloc", max_iter=8)
    assert rc == 0
    plan.close()
