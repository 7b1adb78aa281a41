"""HIP path vs the CPU oracle, through the C-ABI.

* boundary growth: new boundary index bit-exact (cohorts, offsets, scales, transposes), data to
  1e-12 relative (FP64, different summation order in the GEMMs only);
* H*psi at every site of random TwoU1 and SU2U1 chains, 1e-12 relative;
* multi-tile GEMMs (sector dims > 64), transposed cohorts, empty / size-1 sectors.
"""
import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.engine import DeviceBoundary, DeviceMPO, DeviceMPS, HpsiPlan, lstep, rstep
from qcmaquis_amd.model import initial_boundary
from tests.chains import make_chain, oracle_left, oracle_right, rel_err, same_index

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _grow_both(ctx, oracle, ch, symmetric=True):
    """grow all left and right boundaries with the engine and the oracle, comparing each step"""
    n = len(ch.mpos)
    mpos = [DeviceMPO(ctx, m) for m in ch.mpos]
    mps = [DeviceMPS(ctx, lay, d) for lay, d in zip(ch.layouts, ch.data)]
    L_dev = [DeviceBoundary.initial(ctx, ch.bonds[0], ch.mpos[0].row_dim)]
    L_or = [initial_boundary(ch.bonds[0], ch.mpos[0].row_dim)]
    for k in range(n - 1):
        L_dev.append(lstep(ctx, mpos[k], L_dev[-1], mps[k], mps[k], symmetric))
        L_or.append(oracle.lstep(ch.sym, ch.mpos[k], L_or[-1], ch.layouts[k], ch.data[k], ch.layouts[k], ch.data[k],
                                 symmetric=symmetric))
        got = L_dev[-1].download()
        print(f"lstep site {k}: cohorts={len(got.cohorts)} elems={got.data.size}", flush=True)
        same_index(got, L_or[-1])
        assert rel_err(got.data, L_or[-1].data) < TOL
        g = L_dev[-1].growth_stats()
        assert g["out_elements"] == got.data.size and g["seconds_total"] > 0
        assert g["seconds_total"] >= g["seconds_tables"] + g["seconds_blocks"] + g["seconds_device"]
    R_dev = {n: DeviceBoundary.initial(ctx, ch.bonds[n], ch.mpos[n - 1].col_dim)}
    R_or = {n: initial_boundary(ch.bonds[n], ch.mpos[n - 1].col_dim)}
    for k in range(n - 1, 0, -1):
        R_dev[k] = rstep(ctx, mpos[k], R_dev[k + 1], mps[k], mps[k], symmetric)
        R_or[k] = oracle.rstep(ch.sym, ch.mpos[k], R_or[k + 1], ch.layouts[k], ch.data[k], ch.layouts[k], ch.data[k],
                               symmetric=symmetric)
        got = R_dev[k].download()
        print(f"rstep site {k}: cohorts={len(got.cohorts)} elems={got.data.size}", flush=True)
        same_index(got, R_or[k])
        assert rel_err(got.data, R_or[k].data) < TOL
    return mpos, L_dev, L_or, R_dev, R_or


def _hpsi_all_sites(ctx, oracle, ch, mpos, L_dev, L_or, R_dev, R_or):
    n = len(ch.mpos)
    rng = np.random.default_rng(11)
    for k in range(n):
        lay = ch.layouts[k]
        plan = HpsiPlan(ctx, mpos[k], L_dev[k], R_dev[k + 1], lay)
        psi = rng.normal(size=lay.n)
        got = plan.apply(psi)
        ref = oracle.hpsi(ch.sym, ch.mpos[k], L_or[k], R_or[k + 1], lay, psi)
        st = plan.stats()
        print(f"hpsi site {k}: n={lay.n} cohorts={st['n_cohorts']} lines={st['n_lines']} "
              f"err={rel_err(got, ref):.2e}", flush=True)
        assert rel_err(got, ref) < TOL
        # applying twice reuses the plan's buffers
        assert np.array_equal(plan.apply(psi), got)
        plan.close()


@pytest.mark.parametrize("seed,max_m,D", [(1, 3, 6), (2, 5, 8), (3, 12, 10)])
def test_twou1_chain(ctx2u1, oracle_lib, seed, max_m, D):
    print(f"twou1 chain seed={seed} max_m={max_m} D={D}", flush=True)
    ch = make_chain(seed, n_sites=5, total=(3, 2), max_m=max_m, D=D)
    mpos, L_dev, L_or, R_dev, R_or = _grow_both(ctx2u1, oracle_lib, ch)
    _hpsi_all_sites(ctx2u1, oracle_lib, ch, mpos, L_dev, L_or, R_dev, R_or)


def test_twou1_large_sectors(ctx2u1, oracle_lib):
    """sector dims above the 64x64 GEMM tile and K > 16 slices; two-site physical index"""
    from qcmaquis_amd.synthetic import TWOU1_SITE, fuse_index
    print("twou1 large sectors", flush=True)
    ch = make_chain(5, n_sites=4, total=(4, 4), max_m=90, D=6, phys=fuse_index(TWOU1_SITE, TWOU1_SITE), density=0.3)
    mpos, L_dev, L_or, R_dev, R_or = _grow_both(ctx2u1, oracle_lib, ch)
    _hpsi_all_sites(ctx2u1, oracle_lib, ch, mpos, L_dev, L_or, R_dev, R_or)


@pytest.mark.parametrize("sym", [abi.MQ_SYM_TWOU1, abi.MQ_SYM_SU2U1])
def test_fp32_hpsi(ctx2u1, ctxsu2, oracle_lib, sym):
    """mixed precision: FP32 operands and intermediates on the FP32 matrix cores, psi' in FP64;
    max-norm relative error against the FP64 oracle below 1e-5 (FP32 rounding of psi, L, R, T
    and S with K-long FP32 sums); the split-BF16 products (MQ_PREC_BF16X3: each FP32 operand as
    hi + lo BF16 parts, |x - hi - lo| <= 2^-16 |x|, lo lo left out, so about 3 * 2^-16 per
    product) below 5e-4, over sector dims 1..90 (ragged K, every tile shape)"""
    from qcmaquis_amd.synthetic import TWOU1_SITE, fuse_index
    ctx = ctx2u1 if sym == abi.MQ_SYM_TWOU1 else ctxsu2
    print(f"fp32 hpsi sym={sym}", flush=True)
    if sym == abi.MQ_SYM_TWOU1:
        ch = make_chain(5, n_sites=4, total=(4, 4), max_m=90, D=6, phys=fuse_index(TWOU1_SITE, TWOU1_SITE),
                        density=0.3)
    else:
        ch = make_chain(4, sym=sym, n_sites=5, total=(4, 0), max_m=8, D=9)
    mpos, L_dev, L_or, R_dev, R_or = _grow_both(ctx, oracle_lib, ch)
    rng = np.random.default_rng(5)
    for k in range(len(ch.mpos)):
        lay = ch.layouts[k]
        plan = HpsiPlan(ctx, mpos[k], L_dev[k], R_dev[k + 1], lay)
        psi = rng.normal(size=lay.n)
        ref = oracle_lib.hpsi(ch.sym, ch.mpos[k], L_or[k], R_or[k + 1], lay, psi)
        got = plan.apply(psi, precision=abi.MQ_PREC_FP32)
        err = rel_err(got, ref)
        err3 = rel_err(plan.apply(psi, precision=abi.MQ_PREC_BF16X3), ref)
        print(f"fp32 hpsi site {k}: n={lay.n} err={err:.2e} bf16x3 {err3:.2e}", flush=True)
        assert err < 1e-5 and err3 < 5e-4
        # FP64 after FP32 on the same plan is unchanged
        assert rel_err(plan.apply(psi), ref) < TOL
        plan.close()


@pytest.mark.parametrize("small_max", [0, 8, 256])
def test_small_tile_limits(ctx2u1, oracle_lib, small_max):
    """MPO-term sum with every S tile in 16-line blocks (0), split between blocks and the small-tile
    term sum (8), or every tile of at most 256 elements term by term (256): FP64 to 1e-12 and FP32
    to 1e-5 of the oracle at every site of a chain with sector dims 1..90"""
    from qcmaquis_amd.synthetic import TWOU1_SITE, fuse_index
    print(f"small-tile limit {small_max}", flush=True)
    ch = make_chain(5, n_sites=4, total=(4, 4), max_m=90, D=6, phys=fuse_index(TWOU1_SITE, TWOU1_SITE), density=0.3)
    mpos, L_dev, L_or, R_dev, R_or = _grow_both(ctx2u1, oracle_lib, ch)
    rng = np.random.default_rng(6)
    ctx2u1.set_option(abi.MQ_OPT_SMALL_TILE_MAX, small_max)
    try:
        n_small = 0
        for k in range(len(ch.mpos)):
            lay = ch.layouts[k]
            plan = HpsiPlan(ctx2u1, mpos[k], L_dev[k], R_dev[k + 1], lay)
            st = plan.stats()
            n_small += st["small_lines"]
            assert st["small_lines"] <= st["n_lines"]
            psi = rng.normal(size=lay.n)
            ref = oracle_lib.hpsi(ch.sym, ch.mpos[k], L_or[k], R_or[k + 1], lay, psi)
            err, err32 = rel_err(plan.apply(psi), ref), rel_err(plan.apply(psi, abi.MQ_PREC_FP32), ref)
            print(f"  site {k}: n={lay.n} lines {st['n_lines']} small {st['small_lines']} fp64 {err:.2e} "
                  f"fp32 {err32:.2e}", flush=True)
            assert err < TOL and err32 < 1e-5
            plan.close()
        assert (n_small == 0) == (small_max == 0)
    finally:
        ctx2u1.set_option(abi.MQ_OPT_SMALL_TILE_MAX, abi.MQ_SMALL_TILE_AUTO)


@pytest.mark.parametrize("sym", [abi.MQ_SYM_TWOU1, abi.MQ_SYM_SU2U1])
def test_plain_hpsi(ctx2u1, ctxsu2, oracle_lib, sym):
    """the reference-style launch structure (one rocBLAS dgemm per T task and per output block,
    one gather per term) gives the oracle's H*psi to the FP64 tolerance; a plan made without the
    plain tables refuses MQ_PREC_FP64_PLAIN with an error code"""
    from qcmaquis_amd.engine import EngineError
    from qcmaquis_amd.synthetic import TWOU1_SITE, fuse_index
    ctx = ctx2u1 if sym == abi.MQ_SYM_TWOU1 else ctxsu2
    print(f"plain hpsi sym={sym}", flush=True)
    if sym == abi.MQ_SYM_TWOU1:
        ch = make_chain(5, n_sites=4, total=(4, 4), max_m=90, D=6, phys=fuse_index(TWOU1_SITE, TWOU1_SITE),
                        density=0.3)
    else:
        ch = make_chain(4, sym=sym, n_sites=5, total=(4, 0), max_m=8, D=9)
    mpos, L_dev, L_or, R_dev, R_or = _grow_both(ctx, oracle_lib, ch)
    rng = np.random.default_rng(6)
    k = len(ch.mpos) // 2
    lay = ch.layouts[k]
    psi = rng.normal(size=lay.n)
    ref = oracle_lib.hpsi(ch.sym, ch.mpos[k], L_or[k], R_or[k + 1], lay, psi)
    plan = HpsiPlan(ctx, mpos[k], L_dev[k], R_dev[k + 1], lay)
    with pytest.raises(EngineError):
        plan.apply(psi, precision=abi.MQ_PREC_FP64_PLAIN)
    plan.close()
    ctx.set_option(abi.MQ_OPT_PLAIN_TABLES, 1)
    try:
        plan = HpsiPlan(ctx, mpos[k], L_dev[k], R_dev[k + 1], lay)
        got = plan.apply(psi, precision=abi.MQ_PREC_FP64_PLAIN)
        print(f"plain hpsi site {k}: n={lay.n} err={rel_err(got, ref):.2e}", flush=True)
        assert rel_err(got, ref) < TOL
        assert rel_err(plan.apply(psi), ref) < TOL
        e_p, _, _ = plan.davidson(psi, tol=1e-10, max_iter=30, precision=abi.MQ_PREC_FP64_PLAIN)
        e_g, _, _ = plan.davidson(psi, tol=1e-10, max_iter=30)
        assert abs(e_p - e_g) < 1e-9 * max(1.0, abs(e_g))
        plan.close()
    finally:
        ctx.set_option(abi.MQ_OPT_PLAIN_TABLES, 0)


@pytest.mark.parametrize("seed,max_m,D", [(1, 3, 6), (4, 8, 9)])
def test_su2u1_chain(ctxsu2, oracle_lib, seed, max_m, D):
    print(f"su2u1 chain seed={seed} max_m={max_m} D={D}", flush=True)
    ch = make_chain(seed, sym=abi.MQ_SYM_SU2U1, n_sites=5, total=(4, 0), max_m=max_m, D=D)
    mpos, L_dev, L_or, R_dev, R_or = _grow_both(ctxsu2, oracle_lib, ch)
    _hpsi_all_sites(ctxsu2, oracle_lib, ch, mpos, L_dev, L_or, R_dev, R_or)


def test_nonsymmetric_left_growth(ctx2u1, oracle_lib):
    print("twou1 non-symmetric left growth", flush=True)
    ch = make_chain(6, n_sites=4, max_m=4, D=6)
    mpos = [DeviceMPO(ctx2u1, m) for m in ch.mpos]
    mps = [DeviceMPS(ctx2u1, lay, d) for lay, d in zip(ch.layouts, ch.data)]
    Ld = DeviceBoundary.initial(ctx2u1, ch.bonds[0], 1)
    Lo = initial_boundary(ch.bonds[0], 1)
    for k in range(3):
        Ld = lstep(ctx2u1, mpos[k], Ld, mps[k], mps[k], symmetric=False)
        Lo = oracle_lib.lstep(ch.sym, ch.mpos[k], Lo, ch.layouts[k], ch.data[k], ch.layouts[k], ch.data[k],
                              symmetric=False)
        got = Ld.download()
        same_index(got, Lo)
        assert rel_err(got.data, Lo.data) < TOL


def test_mps_pairing_roundtrip(ctx2u1):
    ch = make_chain(8, n_sites=3, max_m=6, D=4)
    lay = ch.layouts[1]
    m = DeviceMPS(ctx2u1, lay, ch.data[1])
    lp = m.download(abi.MQ_LEFT_PAIRED)
    assert np.array_equal(lp, lay.to_left(ch.data[1]))
    m2 = DeviceMPS(ctx2u1, lay, lp, abi.MQ_LEFT_PAIRED)
    assert np.array_equal(m2.download(abi.MQ_RIGHT_PAIRED), ch.data[1])


def test_errors_are_codes(ctx2u1):
    """mismatched indices are rejected on the host before any launch"""
    ch = make_chain(9, n_sites=4, max_m=4, D=6)
    mpo = DeviceMPO(ctx2u1, ch.mpos[1])
    L = DeviceBoundary.initial(ctx2u1, ch.bonds[0], 1)  # wrong aux dim / index for site 1
    R = DeviceBoundary.initial(ctx2u1, ch.bonds[4], 1)
    from qcmaquis_amd.engine import EngineError
    with pytest.raises(EngineError):
        HpsiPlan(ctx2u1, mpo, L, R, ch.layouts[1])


def test_plan_outlives_its_right_boundary(ctx2u1, oracle_lib):
    """the plan reads the right boundary in place and shares its storage: destroying the boundary
    handle and allocating new boundaries of the same size does not change the plan's H*psi"""
    ch = make_chain(21, n_sites=4, max_m=6, D=8)
    mpos, L_dev, L_or, R_dev, R_or = _grow_both(ctx2u1, oracle_lib, ch)
    k = 1
    lay = ch.layouts[k]
    R = rstep(ctx2u1, mpos[k + 1], R_dev[k + 2], DeviceMPS(ctx2u1, ch.layouts[k + 1], ch.data[k + 1]),
              DeviceMPS(ctx2u1, ch.layouts[k + 1], ch.data[k + 1]))
    plan = HpsiPlan(ctx2u1, mpos[k], L_dev[k], R, lay)
    psi = np.random.default_rng(3).normal(size=lay.n)
    before = plan.apply(psi)
    R.close()
    R_dev[k + 1].close()
    others = [rstep(ctx2u1, mpos[k + 1], R_dev[k + 2], DeviceMPS(ctx2u1, ch.layouts[k + 1], 2.0 * ch.data[k + 1]),
                    DeviceMPS(ctx2u1, ch.layouts[k + 1], ch.data[k + 1])) for _ in range(3)]
    after = plan.apply(psi)
    assert np.array_equal(before, after)
    ref = oracle_lib.hpsi(ch.sym, ch.mpos[k], L_or[k], R_or[k + 1], lay, psi)
    assert rel_err(after, ref) < TOL
    plan.close()
    for o in others:
        o.close()
