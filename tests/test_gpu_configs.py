"""Parity with the CPU oracle on the benchmark configurations' own site problems.

* configuration 2 (24 orbitals, 24 electrons, TwoU1, M = 512, integrals PCG64(1002) rank 16):
  after one warm sweep, the middle two-site H*psi (FP64 to 1e-12, FP32 to 1e-5, BF16x3 split-BF16
  GEMM products to 5e-4 relative), the same with the plan's output columns in chunks
  (MQ_OPT_CHUNK_MB 256: more than four chunks run on buffers of less than half of T + S, also as a captured graph replayed several times and with the
  reference-style tables; the chunks overlapped on two streams equal the chunks in turn bit for
  bit) and the left and right growth
  steps of the accepted tensors (1e-12).  The plan statistics show that the
  paths this problem exercises ran: term-sum blocks with more than 64 union tiles (several union
  chunks per block) and split-K GEMM reductions (K up to ~9e4);
* configuration 4 (independent instances of 10-20 orbitals, M from {128, 256, 512}, the generator
  qcmaquis_amd.batch.config4_instance): one batch plan over three instances that span those
  ranges, its H*psi split per instance against the oracle's H*psi of each instance (1e-12).
"""
import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.batch import config4_instance
from qcmaquis_amd.dmrg import TwoSiteDMRG
from qcmaquis_amd.engine import DeviceMPS, HpsiPlan, HpsiPlanBatch, lstep, rstep
from qcmaquis_amd.qc import QCHamiltonian, make_integrals

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_config2_site_problem_vs_oracle(ctx2u1, oracle_lib):
    L, M = 24, 512
    print("configuration 2: warm sweep", flush=True)
    h, eri, ec = make_integrals(L, 1002, 16)
    ham = QCHamiltonian(h, eri, ec)
    d = TwoSiteDMRG(ham, L // 2, L // 2, M=M, ctx=ctx2u1, tol=1e-8, max_iter=8)
    d.sweep()
    mid = L // 2 - 1
    for i in range(mid):
        d.step(i, True)
    mpo, Ld, Rd, lay2, psi2 = d.site_problem(mid)
    plan = HpsiPlan(ctx2u1, mpo, Ld, Rd, lay2)
    st = plan.stats()
    print(f"  plan: n={lay2.n} lines={st['n_lines']} terms={st['n_terms']} max union tiles "
          f"{st['max_union_tiles']} split-K reduces {st['split_reduces']}", flush=True)
    assert st["max_union_tiles"] > 64 and st["split_reduces"] > 0
    Ls, Rs = Ld.download(), Rd.download()
    spec = ham.mpo(mid, True).to_spec()
    x = np.random.default_rng(2).normal(size=lay2.n)
    ref = oracle_lib.hpsi(abi.MQ_SYM_TWOU1, spec, Ls, Rs, lay2, x)
    err = rel(plan.apply(x), ref)
    err32 = rel(plan.apply(x, precision=abi.MQ_PREC_FP32), ref)
    # split-BF16 GEMM products keep 16 bits of each operand (|x - hi - lo| <= 2^-16 |x|, lo lo left
    # out): about 1e-4 of the largest element
    err3 = rel(plan.apply(x, precision=abi.MQ_PREC_BF16X3), ref)
    print(f"  H*psi FP64 rel err {err:.2e}, FP32 {err32:.2e}, BF16x3 {err3:.2e}", flush=True)
    assert err < 1e-12 and err32 < 1e-5 and err3 < 5e-4
    plan.close()
    # the same site problem with its output columns in chunks of at most 256 MiB of T + S
    ctx2u1.set_option(abi.MQ_OPT_CHUNK_MB, 256)
    ctx2u1.set_option(abi.MQ_OPT_PLAIN_TABLES, 1)
    try:
        cplan = HpsiPlan(ctx2u1, mpo, Ld, Rd, lay2)
    finally:
        ctx2u1.set_option(abi.MQ_OPT_CHUNK_MB, abi.MQ_CHUNK_AUTO)
        ctx2u1.set_option(abi.MQ_OPT_PLAIN_TABLES, 0)
    cst = cplan.stats()
    yc = cplan.apply(x)
    yc32 = cplan.apply(x, precision=abi.MQ_PREC_FP32)
    errc, errc32 = rel(yc, ref), rel(yc32, ref)
    errc3 = rel(cplan.apply(x, precision=abi.MQ_PREC_BF16X3), ref)  # overlapped chunks, split-BF16 GEMMs
    assert errc3 < 5e-4
    errp = rel(cplan.apply(x, precision=abi.MQ_PREC_FP64_PLAIN), ref)
    # the chunks overlap on two streams (MQ_OPT_OVERLAP, default): the same result as one chunk after
    # the other, bit for bit
    ctx2u1.set_option(abi.MQ_OPT_CHUNK_MB, 256)
    ctx2u1.set_option(abi.MQ_OPT_OVERLAP, 0)
    try:
        splan = HpsiPlan(ctx2u1, mpo, Ld, Rd, lay2)
    finally:
        ctx2u1.set_option(abi.MQ_OPT_CHUNK_MB, abi.MQ_CHUNK_AUTO)
        ctx2u1.set_option(abi.MQ_OPT_OVERLAP, 1)
    assert np.array_equal(splan.apply(x), yc) and np.array_equal(splan.apply(x, precision=abi.MQ_PREC_FP32), yc32)
    splan.close()
    cplan.use_graphs(True)
    errg = rel(cplan.apply(x), ref)
    # graph replays, with eager work of the other precision between them, equal the eager result
    for _ in range(3):
        assert np.array_equal(cplan.apply(x), yc)
        assert np.array_equal(cplan.apply(x, precision=abi.MQ_PREC_FP32), yc32)
    print(f"  chunked plan: {cst['n_chunks']} chunks, T+S buffers {cst['ts_buffer_elements']} of "
          f"{cst['t_elements'] + cst['s_elements']} elements; H*psi FP64 rel err {errc:.2e}, FP32 {errc32:.2e}, "
          f"graph {errg:.2e}, reference-style {errp:.2e}", flush=True)
    assert cst["n_chunks"] > 4 and cst["ts_buffer_elements"] < (cst["t_elements"] + cst["s_elements"]) // 2
    assert errc < 1e-12 and errc32 < 1e-5 and errg < 1e-12 and errp < 1e-12
    cplan.close()
    # growth of the middle tensors, left and right
    lay, data = d.lay[mid], d.data[mid]
    m = DeviceMPS(ctx2u1, lay, data)
    got = lstep(ctx2u1, d.mpo1[mid], Ld, m, m).download()
    want = oracle_lib.lstep(abi.MQ_SYM_TWOU1, ham.mpo(mid).to_spec(), Ls, lay, data, lay, data)
    err_l = rel(got.data, want.data)
    lay1, data1 = d.lay[mid + 1], d.data[mid + 1]
    m1 = DeviceMPS(ctx2u1, lay1, data1)
    got_r = rstep(ctx2u1, d.mpo1[mid + 1], Rd, m1, m1).download()
    want_r = oracle_lib.rstep(abi.MQ_SYM_TWOU1, ham.mpo(mid + 1).to_spec(), Rs, lay1, data1, lay1, data1)
    err_r = rel(got_r.data, want_r.data)
    print(f"  lstep rel err {err_l:.2e} ({got.data.size} el), rstep {err_r:.2e} ({got_r.data.size} el)", flush=True)
    assert err_l < 1e-12 and err_r < 1e-12


def test_config4_batch_vs_oracle(ctx2u1, oracle_lib):
    items, psis, host = [], [], []
    for i, (L, M) in enumerate([(10, 128), (15, 256), (20, 512)]):
        print(f"configuration 4 instance L={L} M={M}", flush=True)
        item, psi0, h = config4_instance(ctx2u1, L, M, 1004 + 1 + i, with_host=True)
        items.append(item)
        psis.append(psi0)
        host.append(h)
    batch = HpsiPlanBatch(ctx2u1, items)
    outs = batch.split(batch.apply(np.concatenate(psis)))
    for (mpo, Ld, Rd, lay2), psi, (spec, Ls, Rs), out in zip(items, psis, host, outs):
        ref = oracle_lib.hpsi(abi.MQ_SYM_TWOU1, spec, Ls, Rs, lay2, psi)
        err = rel(out, ref)
        print(f"  n={lay2.n}: batch H*psi rel err {err:.2e}", flush=True)
        assert err < 1e-12
    batch.close()
