"""Site-problem snapshots (qcmaquis_amd.snapshot; the reference's `snapshot` parameter of the
two-site sweep and its shtm replay tool): a snapshot written and read back holds the same MPO
descriptor arrays, boundary indices and data, layout and start vector (CPU, bit-exact; the
oracle's H*psi of the read-back problem equals the original's), and a snapshot taken during a
GPU sweep replays to the live site problem's H*psi and Davidson result."""
import os

import numpy as np
import pytest

from qcmaquis_amd import abi, snapshot
from tests.chains import make_chain, oracle_left, oracle_right, same_index


@pytest.mark.parametrize("sym", [abi.MQ_SYM_TWOU1, abi.MQ_SYM_SU2U1])
def test_snapshot_roundtrip(tmp_path, oracle_lib, sym):
    ch = make_chain(3, sym=sym, n_sites=5, total=(4, 2) if sym == abi.MQ_SYM_TWOU1 else (4, 0), max_m=6, D=8)
    k = 2
    L = oracle_left(oracle_lib, ch, k)[k]
    R = oracle_right(oracle_lib, ch, k + 1)[k + 1]
    f = str(tmp_path / "snap.h5")
    snapshot.save(f, sym, ch.mpos[k], L, R, ch.layouts[k], ch.data[k], site=k, twosweep=3, ecore=1.25)
    s = snapshot.load(f)
    assert (s["sym"], s["site"], s["twosweep"], s["ecore"]) == (sym, k, 3, 1.25)
    a = snapshot.ArrayMPO.from_desc(ch.mpos[k].desc())
    for name in snapshot.ArrayMPO.FIELDS:
        x, y = a.a[name], s["mpo"].a[name]
        assert (x is None) == (y is None), name
        if x is not None:
            assert x.dtype == y.dtype and np.array_equal(x, y), name
    for orig, back in ((L, s["left"]), (R, s["right"])):
        same_index(orig, back)
        for c0, c1 in zip(orig.cohorts, back.cohorts):
            assert list(c0.get("conj_scales", [1.0] * orig.aux_dim)) == list(c1["conj_scales"])
            assert list(c0.get("transposes", [0] * orig.aux_dim)) == list(c1["transposes"])
        assert np.array_equal(orig.data, back.data)
    lay = s["layout"]
    assert (lay.left, lay.phys, lay.right) == (ch.layouts[k].left, ch.layouts[k].phys, ch.layouts[k].right)
    assert np.array_equal(s["psi"], ch.data[k])
    # one oracle thread: its multi-threaded accumulation order is not fixed
    x = np.random.default_rng(1).normal(size=lay.n)
    want = oracle_lib.hpsi(sym, ch.mpos[k], L, R, ch.layouts[k], x, nthreads=1)
    got = oracle_lib.hpsi(sym, s["mpo"], s["left"], s["right"], lay, x, nthreads=1)
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_snapshot_replays_live_problem(tmp_path, ctx2u1):
    from qcmaquis_amd.dmrg import TwoSiteDMRG
    from qcmaquis_amd.engine import HpsiPlan
    from qcmaquis_amd.qc import QCHamiltonian, make_integrals
    L = 8
    h, eri, ec = make_integrals(L, 7, rank=6)
    print("snapshot: sweep with snapshots at (0, 3) and (1, 2)", flush=True)
    d = TwoSiteDMRG(QCHamiltonian(h, eri, ec), 4, 4, M=32, ctx=ctx2u1, tol=1e-8, max_iter=8,
                    snapshots=[(0, 3), (1, 2)], snapshot_dir=str(tmp_path))
    d.sweep()
    files = sorted(os.listdir(tmp_path))
    assert files == ["snapshot_0_3.h5", "snapshot_1_2.h5"], files
    # a live site problem against its snapshot: H*psi and Davidson bit for bit
    for i in range(3):
        d.step(i, True)
    mpo, Ld, Rd, lay2, psi2 = d.site_problem(3)
    f = str(tmp_path / "live.h5")
    snapshot.save(f, abi.MQ_SYM_TWOU1, mpo.spec, Ld, Rd, lay2, psi2, site=3, ecore=d.ham.ecore)
    s = snapshot.load(f)
    plan = HpsiPlan(ctx2u1, mpo, Ld, Rd, lay2)
    x = np.random.default_rng(2).normal(size=lay2.n)
    want = plan.apply(x)
    e_live, v_live, _ = plan.davidson(psi2, tol=1e-6, max_iter=9)
    plan.close()
    from qcmaquis_amd.engine import DeviceBoundary, DeviceMPO
    p2 = HpsiPlan(ctx2u1, DeviceMPO(ctx2u1, s["mpo"]), DeviceBoundary.upload(ctx2u1, s["left"]),
                  DeviceBoundary.upload(ctx2u1, s["right"]), s["layout"])
    got = p2.apply(x)
    p2.close()
    assert np.array_equal(got, want)
    e, v, st, pst = snapshot.replay(ctx2u1, s)
    print(f"  live E {e_live + d.ham.ecore:.12f}, replay E {e:.12f}, {st['iterations']} iterations", flush=True)
    assert e == e_live + d.ham.ecore and np.array_equal(v, v_live)
