#!/usr/bin/env python3
"""Benchmark of the Davidson H*psi step (Boundary x MPOTensor x MPS) on one MI355X.

One "step" = one H*psi of the site problem with psi, psi', boundaries and work lists resident in
HBM (one Davidson iteration's operator application, reference super_hamil_mv).  The value is the
reference's own flop model (ScheduleNew::total_flops, solver/tasks.cpp:389-398, 778-791) per
second.  Also reported: the roofline of the dominant kernel (the block-form MPO-term sum,
sblock_kernel<double>) and of the grouped FP64 GEMMs, from HIP events recorded on the launch stream,
and the CPU oracle (restated reference CPU path, OpenMP + OpenBLAS) timed on the host cores on the
same site problem.

  python bench.py [--steps K] [--warmup W] [--workload synthetic|batch|batch-sweep] [--no-cpu]

--workload batch measures configuration 4 (many independent site problems in one batch plan) and
prints its own line; --workload batch-sweep sweeps such runs concurrently against one by one.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sweep time & H\u00b7\u03c8 TFLOP/s at M=2048: mixed+batched vs FP64 plain, |\u0394E|\u22641e-6 Eh"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
GEMM_FAMILY = "gemm_kernel<*, double, double> (T-GEMM and L*S GEMM launches, split-K reduce)"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def synthetic_workload(ctx, seed=7):
    """Two-site TwoU1 site problem in the middle of a synthetic chain: MPS sector structure of an
    8-site two-site chain (16 states / 9 sectors per site), sparse MPO rows (<= 6 cells per row,
    bond dimension 256), boundaries grown on the device with the engine's growth steps."""
    from qcmaquis_amd import abi
    from qcmaquis_amd.engine import DeviceBoundary, DeviceMPO, DeviceMPS, lstep, rstep
    from qcmaquis_amd.synthetic import TWOU1_SITE, Chain, fuse_index
    rng = np.random.default_rng(seed)
    two = fuse_index(TWOU1_SITE, TWOU1_SITE)
    n, mid = 8, 4
    ch = Chain(rng, abi.MQ_SYM_TWOU1, n, (8, 8), 40, 256, phys=two, density=0.02, row_cells=6, min_m=8)
    mpos = [DeviceMPO(ctx, m) for m in ch.mpos]
    mps = [DeviceMPS(ctx, l, d) for l, d in zip(ch.layouts, ch.data)]
    L = DeviceBoundary.initial(ctx, ch.bonds[0], ch.mpos[0].row_dim)
    for k in range(mid):
        L = lstep(ctx, mpos[k], L, mps[k], mps[k])
    R = DeviceBoundary.initial(ctx, ch.bonds[n], ch.mpos[-1].col_dim)
    for k in range(n - 1, mid, -1):
        R = rstep(ctx, mpos[k], R, mps[k], mps[k])
    lay = ch.layouts[mid]
    desc = {"workload": "synthetic two-site TwoU1 site problem (random sparse MPO, D=256, 8-site chain, "
                        "sector dims 8-40)",
            "symmetry": "TwoU1", "M_left": lay.left.total(), "M_right": lay.right.total(),
            "sectors_left": len(lay.left), "sectors_right": len(lay.right),
            "D_left": ch.mpos[mid].row_dim, "D_right": ch.mpos[mid].col_dim, "psi_elements": lay.n}
    return dict(mpo=mpos[mid], mpo_spec=ch.mpos[mid], L=L, R=R, layout=lay, sym=abi.MQ_SYM_TWOU1, desc=desc)


def qc_workload(ctx, L=24, M=512, seed=1002, warm_sweeps=1, dump_bonds=None):
    """Configuration 2: 24 orbitals, 24 electrons (S_z = 0), TwoU1, M = 512, seeded integrals
    (PCG64(1002), L-rank 16).  Two-site DMRG from a random MPS: `warm_sweeps` full sweeps at
    M = 512, then the left-to-right half sweep up to the middle two-site problem, whose H*psi is
    the benchmark step."""
    from qcmaquis_amd import abi
    from qcmaquis_amd.dmrg import TwoSiteDMRG, merge_two
    from qcmaquis_amd.qc import QCHamiltonian, make_integrals
    h, eri, ec = make_integrals(L, seed, 16)
    t0 = time.perf_counter()
    ham = QCHamiltonian(h, eri, ec)
    log("MPO built in %.1f s, bond dims max %d" % (time.perf_counter() - t0, max(ham.bond_dims)))
    d = TwoSiteDMRG(ham, L // 2, L // 2, M=M, ctx=ctx, tol=1e-8, max_iter=8)
    sweeps = []
    for s in range(warm_sweeps):
        e, dt = d.sweep()
        sweeps.append(dt)
        log("sweep %d: E = %.10f, %.1f s, timing %s" % (s, e, dt, {k: round(v, 2) for k, v in d.timing.items()}))
    mid = L // 2 - 1
    for i in range(0, mid):
        d.step(i, True)
    if dump_bonds:
        bonds = [d.lay[k].left for k in range(L)] + [d.lay[-1].right]
        json.dump([{"charges": [list(map(int, q)) for q in b.charges], "sizes": [int(x) for x in b.sizes]}
                   for b in bonds], open(dump_bonds, "w"))
    lay2, psi2 = merge_two(d.lay[mid], d.data[mid], d.lay[mid + 1], d.data[mid + 1], d.phys2, d.off12)
    mpo = d.ts_mpo(mid)
    desc = {"workload": f"configuration 2: {L} orbitals, {L} electrons, TwoU1, M={M}, seeded FCIDUMP "
                        f"(PCG64({seed}), rank 16); two-site H*psi at sites ({mid},{mid + 1}) after {warm_sweeps} sweep(s)",
            "symmetry": "TwoU1", "L": L, "M": M, "M_left": lay2.left.total(), "M_right": lay2.right.total(),
            "sectors_left": len(lay2.left), "sectors_right": len(lay2.right), "D_left": mpo.spec.row_dim,
            "D_right": mpo.spec.col_dim, "psi_elements": lay2.n, "energy": d.energies[-1],
            "sweep_seconds": sweeps, "sweep_timing_s": {k: round(v, 3) for k, v in d.timing.items()}}
    return dict(mpo=mpo, mpo_spec=ham.mpo(mid, True), L=d.left[mid], R=d.right[mid + 2], layout=lay2,
                sym=abi.MQ_SYM_TWOU1, desc=desc, keep=(d, ham))


def scale_to(bonds, caps, M):
    """sector sizes of `bonds` scaled so that every bond holds about M states, each sector at most
    its number of states (caps: the untruncated bond indices)"""
    from qcmaquis_amd.model import Index
    out = []
    for b, f in zip(bonds, caps):
        cap = {c: s for c, s in zip(f.charges, f.sizes)}
        sizes = {c: s for c, s in zip(b.charges, b.sizes) if c in cap}
        for _ in range(8):
            tot = sum(sizes.values())
            free = [c for c in sizes if sizes[c] < cap[c]]
            if tot >= M or not free:
                break
            g = 1.0 + (M - tot) / sum(sizes[c] for c in free)
            for c in free:
                sizes[c] = min(cap[c], max(1, int(round(sizes[c] * g))))
        out.append(Index.from_dict(sizes))
    return out


BONDS3 = os.path.join(ROOT, "qcmaquis_amd", "data", "config3_bonds.json")


def su2_workload(ctx, L=40, M=2048, seed=1003, rank=80, warm_M=128, warm_iter=2, regen=False):
    """Configuration 3: 40 orbitals, 40 electrons, SU2U1, S = 0, seeded integrals (PCG64(1003),
    L-rank 80), bond dimension M = 2048 at the middle two-site problem (sites L/2-1, L/2).
    Sector structure: a left-to-right half sweep at M = warm_M (FP32 Davidson, warm_iter iterations
    per site) from a random MPS gives the sectors a DMRG state of this Hamiltonian occupies at
    every bond; their sizes are scaled to about M states per bond (each at most its number of
    spin-coupled states), filled with random numbers, left-normalised up to the middle problem and
    right-normalised after it; its boundaries are grown on the device.  The half sweep's bond
    indices are kept in qcmaquis_amd/data/config3_bonds.json (written by the first run, or with
    regen) so that later runs start from them."""
    from qcmaquis_amd import abi
    from qcmaquis_amd.dmrg import TwoSiteDMRG, allowed_bonds_su2, left_normalize, right_normalize
    from qcmaquis_amd.engine import DeviceBoundary, DeviceMPO, DeviceMPS, lstep, rstep
    from qcmaquis_amd.model import MPSLayout
    from qcmaquis_amd.qc import QCHamiltonian, make_integrals
    SU2 = abi.MQ_SYM_SU2U1
    h, eri, ec = make_integrals(L, seed, rank)
    t0 = time.perf_counter()
    ham = QCHamiltonian(h, eri, ec, symmetry=SU2)
    t_mpo = time.perf_counter() - t0
    log("SU2U1 MPO built in %.1f s, bond dims max %d" % (t_mpo, max(ham.bond_dims)))
    from qcmaquis_amd.model import Index
    tag = {"L": L, "seed": seed, "rank": rank, "warm_M": warm_M, "warm_iter": warm_iter}
    warm = None
    t_warm = 0.0
    if not regen and os.path.exists(BONDS3):
        rec = json.load(open(BONDS3))
        if rec.get("tag") == tag:
            warm = [Index([tuple(c) for c in b["charges"]], b["sizes"]) for b in rec["bonds"]]
            log("warm-up bond indices from %s" % BONDS3)
    if warm is None:
        t0 = time.perf_counter()
        d = TwoSiteDMRG(ham, L // 2, L // 2, M=warm_M, ctx=ctx, init_m=8, tol=1e-6, max_iter=warm_iter,
                        precision=abi.MQ_PREC_FP32)
        for i in range(L - 1):
            d.step(i, True)
        t_warm = time.perf_counter() - t0
        log("warm-up half sweep at M=%d: %.1f s, E = %.8f" % (warm_M, t_warm, d.energies[-1]))
        warm = [d.lay[k].left for k in range(L)] + [d.lay[-1].right]
        del d
        os.makedirs(os.path.dirname(BONDS3), exist_ok=True)
        json.dump({"tag": tag, "bonds": [{"charges": [list(map(int, q)) for q in b.charges],
                                          "sizes": [int(x) for x in b.sizes]} for b in warm]}, open(BONDS3, "w"))
    phys = ham.phys()
    bonds = scale_to(warm, allowed_bonds_su2([phys] * L, (L, 0), 1 << 30), M)
    mid = L // 2 - 1
    lay = [MPSLayout(bonds[k], phys, bonds[k + 1]) for k in range(L)]
    rng = np.random.default_rng(seed)
    data = [rng.normal(size=x.n) for x in lay]
    left_normalize(lay, data, mid)
    right_normalize(lay, data, stop=mid + 2)
    t0 = time.perf_counter()
    Lb = DeviceBoundary.initial(ctx, lay[0].left, 1)
    for k in range(mid):
        m = DeviceMPS(ctx, lay[k], data[k])
        Lb = lstep(ctx, DeviceMPO(ctx, ham.mpo(k)), Lb, m, m)
    R = DeviceBoundary.initial(ctx, lay[-1].right, 1)
    for k in range(L - 1, mid + 1, -1):
        m = DeviceMPS(ctx, lay[k], data[k])
        R = rstep(ctx, DeviceMPO(ctx, ham.mpo(k)), R, m, m)
    ctx.sync()
    t_grow = time.perf_counter() - t0
    from qcmaquis_amd.model import Index  # noqa: F401
    lay2 = MPSLayout(lay[mid].left, ham.phys(True), lay[mid + 1].right)
    mpo = DeviceMPO(ctx, ham.mpo(mid, True))
    li, ri = Lb.info(), R.info()
    log("boundaries grown in %.1f s: left %d elements, right %d elements" % (
        t_grow, li.total_elements, ri.total_elements))
    desc = {"workload": f"configuration 3: {L} orbitals, {L} electrons, SU2U1, S=0, M={M}, seeded FCIDUMP "
                        f"(PCG64({seed}), rank {rank}); two-site H*psi at sites ({mid},{mid + 1}) of an MPS with the "
                        f"sector structure of a DMRG state (M={warm_M} half sweep) scaled to M={M}",
            "symmetry": "SU2U1", "L": L, "M": M, "M_left": lay2.left.total(), "M_right": lay2.right.total(),
            "bond_dims": [b.total() for b in bonds], "sectors_left": len(lay2.left),
            "sectors_right": len(lay2.right), "D_left": mpo.spec.row_dim, "D_right": mpo.spec.col_dim,
            "mpo_bond_dims_max": int(max(ham.bond_dims)), "psi_elements": lay2.n,
            "boundary_elements": [int(li.total_elements), int(ri.total_elements)],
            "setup_seconds": {"mpo": round(t_mpo, 2), "warm_half_sweep": round(t_warm, 2),
                              "boundaries": round(t_grow, 2)}}
    return dict(mpo=mpo, mpo_spec=ham.mpo(mid, True), L=Lb, R=R, layout=lay2, sym=SU2, desc=desc, keep=(ham,))


def qc3_sweep_bench(args, ctx):
    """Real two-site sweeps at configuration 3 (40 orbitals, SU2U1, S=0, M=2048): the MPS of the
    H*psi workload (sectors of the warm-up half sweep scaled to M=2048, seeded random elements) is
    the common start of one full sweep per precision (FP64 grouped, mixed; with --plain-sweep also
    FP64 with the reference's launch structure), Davidson 8 iterations / tol 1e-8 per step.  Every
    run plans, solves, splits (truncating to M) and grows its own boundaries; prints one JSON line
    with the sweep times, their breakdown and the final energies."""
    from qcmaquis_amd import abi
    from qcmaquis_amd.dmrg import TwoSiteDMRG, allowed_bonds_su2
    from qcmaquis_amd.model import Index, MPSLayout
    from qcmaquis_amd.qc import QCHamiltonian, make_integrals
    import gc
    L, M, seed, rank = 40, args.M or 2048, 1003, 80
    h, eri, ec = make_integrals(L, seed, rank)
    ham = QCHamiltonian(h, eri, ec, symmetry=abi.MQ_SYM_SU2U1)
    rec = json.load(open(BONDS3))
    warm = [Index([tuple(c) for c in b["charges"]], b["sizes"]) for b in rec["bonds"]]
    phys = ham.phys()
    bonds = scale_to(warm, allowed_bonds_su2([phys] * L, (L, 0), 1 << 30), M)
    lay = [MPSLayout(bonds[k], phys, bonds[k + 1]) for k in range(L)]
    rng = np.random.default_rng(seed)
    data = [rng.normal(size=x.n) for x in lay]
    t0 = time.perf_counter()
    for k in range(L - 1):  # every two-site MPO once (cached by the Hamiltonian for all runs)
        ham.mpo(k, True)
    t_mpo = time.perf_counter() - t0
    log("qc3 sweeps: %d two-site MPOs built in %.1f s" % (L - 1, t_mpo))
    runs = {}
    allp = {"fp64": abi.MQ_PREC_FP64, "mixed": abi.MQ_PREC_FP32, "fp64_plain": abi.MQ_PREC_FP64_PLAIN}
    names = args.sweep_runs.split(",") + (["fp64_plain"] if args.plain_sweep else [])
    precs = [(n, allp[n]) for n in names]
    for name, prec in precs:
        ctx.set_option(abi.MQ_OPT_PLAIN_TABLES, int(prec == abi.MQ_PREC_FP64_PLAIN))
        t0 = time.perf_counter()
        d = TwoSiteDMRG(ham, L // 2, L // 2, M=M, ctx=ctx, tol=1e-8, max_iter=8, precision=prec,
                        mps=(lay, data), verbose=True)
        t_init = time.perf_counter() - t0
        log("qc3 sweeps %s: start state normalised and right boundaries grown in %.1f s" % (name, t_init))
        e, dt = d.sweep()
        r = {"sweep_seconds": round(dt, 3), "init_seconds": round(t_init, 3), "final_energy": e,
             "lowest_energy": min(d.energies), "max_truncated_weight": float(max(d.trunc)),
             "final_residual_norm": d.last_residual, "hpsi_flops": int(d.hpsi_flops),
             "timing_s": {k: round(v, 3) for k, v in d.timing.items()}}
        r["hpsi_tflops_in_davidson"] = round(d.hpsi_flops / max(d.timing["hpsi"], 1e-9) / 1e12, 3)
        runs[name] = r
        log("qc3 sweeps %s: %.1f s, E %.10f, timing %s" % (name, dt, e, r["timing_s"]))
        del d
        gc.collect()
    ctx.set_option(abi.MQ_OPT_PLAIN_TABLES, 0)
    ref = runs.get("fp64")
    main_run = "mixed" if "mixed" in runs else names[0]
    line = {"metric": METRIC, "value": runs[main_run]["sweep_seconds"], "unit": f"s (one full sweep, {main_run})",
            "n_gpus": 1, "higher_is_better": False, "dtype": "f64 (FP32 contraction in the mixed run)",
            "data": "synthetic (seeded rank-80 integrals; MPS sectors from a warm-up half sweep at M=128, scaled "
                    "to M=2048, seeded random elements)",
            "config": {"workload": f"configuration 3 sweeps: {L} orbitals, {L} electrons, SU2U1, S=0, M={M}, "
                                   "one full two-site sweep (78 steps) per run from one common start, Davidson 8 "
                                   "iterations / tol 1e-8",
                       "two_site_mpo_seconds": round(t_mpo, 3), "runs": runs}}
    if ref and "mixed" in runs:
        line["config"]["mixed_abs_delta_e_vs_fp64"] = abs(runs["mixed"]["final_energy"] - ref["final_energy"])
        line["config"]["mixed_speedup_vs_fp64"] = round(ref["sweep_seconds"] / runs["mixed"]["sweep_seconds"], 3)
    if "fp64_plain" in runs and ref and "mixed" in runs:
        line["config"]["mixed_speedup_vs_fp64_plain"] = round(runs["fp64_plain"]["sweep_seconds"] /
                                                              runs["mixed"]["sweep_seconds"], 3)
        line["config"]["fp64_speedup_vs_fp64_plain"] = round(runs["fp64_plain"]["sweep_seconds"] /
                                                             ref["sweep_seconds"], 3)
    print(json.dumps(line), flush=True)


def adjoint_bench(args, ctx):
    """Configuration 5: 16 orbitals, 16 electrons, SU2U1, S=0, M=512, converged two-site DMRG
    ground state; the integral gradients dE/dh and dE/d(pq|rs) (the 1- and 2-RDMs) from stored
    boundaries (qcmaquis_amd.adjoint), with every right boundary kept and with the kept ones capped
    at 25 % of their bytes; checked against central finite differences of <psi|H(h, eri)|psi> on 32
    random integral entries (step 1e-5), the derivative-MPO expectation values of a few entries
    (one chain growth each, the per-element measurement the adjoint replaces), the RDM traces and
    Euler's identity (sum theta dE/dtheta = E - E_core)."""
    from qcmaquis_amd import abi
    from qcmaquis_amd.adjoint import energy_gradient, rdms
    from qcmaquis_amd.dmrg import TwoSiteDMRG, right_normalize
    from qcmaquis_amd.qc import QCHamiltonian, make_integrals
    from qcmaquis_amd.rdm import eri_gradient, expval, h_gradient
    SU2 = abi.MQ_SYM_SU2U1
    L, N, M, seed = 16, 16, args.M or 512, 1005
    h, eri, ec = make_integrals(L, seed, 16)
    t0 = time.perf_counter()
    ham = QCHamiltonian(h, eri, ec, symmetry=SU2, gradient_layout=True)
    t_mpo = time.perf_counter() - t0
    d = TwoSiteDMRG(ham, N // 2, N // 2, M=M, ctx=ctx, tol=1e-10, max_iter=12)
    energies, prev, t_dmrg = [], None, time.perf_counter()
    for s in range(12):
        e, dt = d.sweep()
        energies.append(e)
        log("adjoint workload: sweep %d E %.12f (%.2f s)" % (s, e, dt))
        if prev is not None and abs(e - prev) < 1e-9:
            break
        prev = e
    t_dmrg = time.perf_counter() - t_dmrg
    runs = {}
    for name, cap in (("all_kept", None), ("capped_25pct", 0.25)):
        E, dh, deri, st = energy_gradient(ctx, ham, d.lay, d.data, mem_fraction=cap)
        runs[name] = {"seconds": round(st["seconds"], 3), "energy": E, "right_steps": st["right_steps"],
                      "regrown": st["regrown"], "all_right_MiB": round(st["all_right_bytes"] / 2 ** 20, 1),
                      "peak_kept_MiB": round(st["peak_kept_bytes"] / 2 ** 20, 1),
                      "peak_boundary_MiB": round(st["peak_boundary_bytes"] / 2 ** 20, 1),
                      "site_energy_spread": float(np.ptp(st["site_energies"]))}
        runs[name]["_g"] = (dh, deri)
        log("adjoint workload: %s %.3f s, E %.12f, %s" % (name, st["seconds"], E,
                                                        {k: v for k, v in runs[name].items() if k != "_g"}))
    dh, deri = runs["capped_25pct"].pop("_g")
    dh0, deri0 = runs["all_kept"].pop("_g")
    scale = max(np.max(np.abs(dh0)), np.max(np.abs(deri0)))
    cap_diff = max(np.max(np.abs(dh - dh0)), np.max(np.abs(deri - deri0))) / scale
    gamma, gam = rdms(dh, deri)
    euler = float(np.sum(dh * h) + np.sum(deri * eri)) - (runs["capped_25pct"]["energy"] - ec)
    # finite differences of the fixed normalised state's energy
    lay = list(d.lay)
    dat = [np.array(x) for x in d.data]
    right_normalize(lay, dat)
    hs = [(i, j) for i in range(L) for j in range(i + 1)]
    es = [(i, j, k, l) for i in range(L) for j in range(i + 1) for k in range(i + 1) for l in range(k + 1)
          if not (k == i and l > j)]
    rng = np.random.default_rng(5)
    picks = [("h", hs[i]) for i in rng.choice(len(hs), 12, replace=False)] + \
            [("eri", es[i]) for i in rng.choice(len(es), 20, replace=False)]
    step, worst, t_fd = 1e-5, 0.0, time.perf_counter()
    for kind, idx in picks:
        vals = []
        for sgn in (1.0, -1.0):
            hp, ep = h.copy(), eri.copy()
            (hp if kind == "h" else ep)[idx] += sgn * step
            vals.append(expval(ctx, lay, dat, QCHamiltonian(hp, ep, symmetry=SU2)))
        worst = max(worst, abs((vals[0] - vals[1]) / (2 * step) - (dh if kind == "h" else deri)[idx]))
    t_fd = time.perf_counter() - t_fd
    # the per-element measurement: one derivative-MPO chain growth per entry
    t1 = time.perf_counter()
    sample_h = hs[:2]
    sample_e = es[len(es) // 2:len(es) // 2 + 2]
    ref_h = h_gradient(ctx, lay, dat, sample_h, SU2)
    ref_e = eri_gradient(ctx, lay, dat, sample_e, SU2)
    t_elem = (time.perf_counter() - t1) / 4
    elem_diff = max(max(abs(a - dh[x]) for a, x in zip(ref_h, sample_h)),
                    max(abs(a - deri[x]) for a, x in zip(ref_e, sample_e)))
    n_params = len(hs) + len(es)
    line = {"metric": "integral gradients (1- and 2-RDM) from stored boundaries, configuration 5",
            "value": runs["capped_25pct"]["seconds"], "unit": "s (all dE/dh and dE/d(pq|rs), boundaries capped at 25%)",
            "n_gpus": 1, "higher_is_better": False, "dtype": "f64", "data": "synthetic (seeded rank-16 integrals)",
            "config": {"workload": f"configuration 5: {L} orbitals, {N} electrons, SU2U1, S=0, M={M}, converged "
                                   "two-site DMRG ground state",
                       "mpo_seconds": round(t_mpo, 3), "dmrg_seconds": round(t_dmrg, 3), "sweep_energies": energies,
                       "runs": runs, "capped_vs_all_kept_rel_diff": cap_diff,
                       "fd_check": {"entries": len(picks), "step": step, "max_abs_error": worst,
                                    "seconds": round(t_fd, 3)},
                       "per_element_measurement": {"entries": 4, "max_abs_diff": elem_diff,
                                                   "seconds_per_entry": round(t_elem, 4), "parameters": n_params,
                                                   "extrapolated_seconds_all": round(t_elem * n_params, 1)},
                       "euler_residual": euler, "trace_gamma": float(np.trace(gamma)),
                       "trace_Gamma_pprr": float(np.einsum("pprr->", gam))}}
    print(json.dumps(line), flush=True)


def batch_sweep_bench(args, ctx):
    """Configuration 4 as sweeps: `--instances` runs (sizes and integrals as in batch_bench), each
    a two-site DMRG from a right-normalised random MPS of bond dimension about M; one sweep of
    every run one after the other (TwoSiteDMRG.sweep), then one sweep of fresh copies swept
    concurrently (BatchTwoSiteDMRG: one batch plan and one device lock-step Davidson per step).
    FP64 Davidson, 8 iterations, tol 1e-8."""
    from qcmaquis_amd import abi
    from qcmaquis_amd.batch import BatchTwoSiteDMRG, scaled_bonds
    from qcmaquis_amd.dmrg import TwoSiteDMRG, right_normalize
    from qcmaquis_amd.model import MPSLayout
    from qcmaquis_amd.qc import QCHamiltonian, make_integrals
    rng = np.random.default_rng(1004)
    shapes = [(int(rng.integers(10, 21)), int(rng.choice([128, 256, 512]))) for _ in range(args.instances)]

    def make_runs():
        runs = []
        for i, (L, M) in enumerate(shapes):
            h, eri, ec = make_integrals(L, 1004 + 1 + i, 16)
            ham = QCHamiltonian(h, eri, ec)
            phys = ham.phys()
            bonds = scaled_bonds([phys] * L, (L // 2, L // 2), M)
            lay = [MPSLayout(bonds[k], phys, bonds[k + 1]) for k in range(L)]
            r = np.random.default_rng(1004 + 1 + i)
            data = [r.normal(size=x.n) for x in lay]
            right_normalize(lay, data)
            runs.append(TwoSiteDMRG(ham, L // 2, L // 2, M=M, ctx=ctx, tol=1e-8, max_iter=8, mps=(lay, data)))
        return runs

    t0 = time.perf_counter()
    runs = make_runs()
    log("%d runs built in %.1f s" % (len(runs), time.perf_counter() - t0))
    single = []
    t0 = time.perf_counter()
    for k, r in enumerate(runs):
        e, dt = r.sweep()
        single.append((e, list(r.iteration_results["Energy"])))
        log("run %d (L=%d M=%d): one sweep %.2f s, E=%.10f" % (k, r.L, r.M, dt, e))
    t_single = time.perf_counter() - t0
    del runs
    batch_runs = make_runs()
    b = BatchTwoSiteDMRG(batch_runs, abi.MQ_PREC_FP64)
    t0 = time.perf_counter()
    b.sweep()
    t_batch = time.perf_counter() - t0
    de = max(float(np.max(np.abs(np.array(s[1]) - np.array(r.iteration_results["Energy"]))))
             for s, r in zip(single, batch_runs))
    log("sweep of %d runs: one after the other %.2f s, concurrently %.2f s (plan %.2f, davidson %.2f, split and "
        "growth %.2f), max |dE| per step %.2e" % (len(batch_runs), t_single, t_batch, b.timing["plan"],
                                                  b.timing["davidson"], b.timing["accept"], de))
    line = {"metric": METRIC, "value": round(t_single / t_batch, 4), "unit": "speedup (one sweep, concurrent vs one "
            "after the other)", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(t_batch * 1e3, 1),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded integrals, random MPS)",
            "config": {"workload": "configuration 4 sweeps: %d independent runs (orbitals U{10..20}, half filling, "
                                   "TwoU1, M from {128, 256, 512}), one two-site sweep each" % len(batch_runs),
                       "seconds_one_after_the_other": round(t_single, 3), "seconds_concurrent": round(t_batch, 3),
                       "concurrent_timing_s": {k: round(v, 3) for k, v in b.timing.items()},
                       "max_abs_delta_e_per_step": de, "runs": [{"L": L, "M": M} for L, M in shapes]}}
    print(json.dumps(line), flush=True)


def batch_bench(args, ctx):
    """Configuration 4: independent instances (orbitals ~ U{10..20}, M from {128, 256, 512}, sizes
    from PCG64(1004), integrals of instance i from PCG64(1004 + 1 + i)), each at its middle
    two-site problem.  Times, with psi resident in HBM: every instance's own H*psi one after the
    other (sum over instances) and one H*psi of the batch plan over all instances (FP64 and FP32
    contraction); the batch's output is checked against the instances' own; and the Davidson of
    every instance (8 iterations, tol 1e-8) one after the other vs the lock-step Davidson over the
    batch (one batched H*psi per iteration)."""
    import ctypes as C

    import torch
    from qcmaquis_amd import abi
    from qcmaquis_amd.batch import config4_instance, lockstep_davidson
    from qcmaquis_amd.engine import HpsiPlan, HpsiPlanBatch
    rng = np.random.default_rng(1004)
    shapes = [(int(rng.integers(10, 21)), int(rng.choice([128, 256, 512]))) for _ in range(args.instances)]
    items, psis, recs = [], [], []
    sms = (C.c_float * 4)()
    bytes_plans = 0
    free0, total_hbm = torch.cuda.mem_get_info()
    t_all = time.perf_counter()
    for i, (L, M) in enumerate(shapes):
        t0 = time.perf_counter()
        item, psi0 = config4_instance(ctx, L, M, 1004 + 1 + i)
        tb = time.perf_counter() - t0
        plan = HpsiPlan(ctx, *item)
        st = plan.stats()
        x = torch.from_numpy(psi0).cuda()
        y = torch.empty_like(x)
        for _ in range(args.warmup):
            plan.apply_device(x.data_ptr(), y.data_ptr())
        ctx.sync()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            plan.apply_device(x.data_ptr(), y.data_ptr())
        ctx.sync()
        ms = (time.perf_counter() - t1) / args.steps * 1e3
        stage = np.zeros(4)
        for _ in range(3):
            ctx.check(ctx.lib.mq_hpsi_timed(plan.h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), 0, sms))
            stage += np.array(list(sms)) / 3
        t2 = time.perf_counter()
        e, _, dst = plan.davidson(psi0, tol=1e-8, max_iter=8)
        dav = time.perf_counter() - t2
        bytes_plans += 8 * (st["t_elements"] + st["s_elements"] + st["l_panel_elements"])
        recs.append({"L": L, "M": M, "psi": int(item[3].n), "flops": int(st["flops_total"]), "ms": ms,
                     "stage_ms": [round(v, 4) for v in stage], "n_terms": int(st["n_terms"]),
                     "n_lines": int(st["n_lines"]), "t_elements": int(st["t_elements"]),
                     "s_elements": int(st["s_elements"]), "union_elements": int(st["union_elements"]),
                     "n_t_tasks": int(st["n_t_tasks"]), "small_lines": int(st["small_lines"]),
                     "small_terms": int(st["small_terms"]), "davidson_s": dav, "e": e, "out": y.cpu().numpy()})
        items.append(item)
        psis.append(psi0)
        plan.close()
        del x, y
        free, _ = torch.cuda.mem_get_info()
        log("instance %d: L=%d M=%d psi %d, %.2f GFLOP, H*psi %.3f ms (stages %s), %d terms, davidson %.3f s "
            "(built in %.1f s), HBM used %.1f GB" % (i, L, M, item[3].n, st["flops_total"] / 1e9, ms,
                                                     recs[-1]["stage_ms"], st["n_terms"], dav, tb,
                                                     (free0 - free) / 1e9))
        if (free0 - free) + 1.3 * bytes_plans > 0.8 * free0:
            log("HBM: stopping at %d instances (boundaries + batch buffers would exceed 80%% of free HBM)" % (i + 1))
            break
    n = len(items)
    t0 = time.perf_counter()
    batch = HpsiPlanBatch(ctx, items)
    build = time.perf_counter() - t0
    bst = batch.stats()
    # the Davidson runs first, so that the timed H*psi launches are the last ones of the run (a
    # kernel trace of the whole run ends with them)
    t1 = time.perf_counter()
    e_ls, _, its, n_hpsi = lockstep_davidson(batch, psis, tol=1e-8, max_iter=8)
    lock_s = time.perf_counter() - t1
    # the same lock-step iteration on the device (mq_davidson_batch), FP64 and mixed
    dev = {}
    for name, prec in (("fp64", abi.MQ_PREC_FP64), ("mixed_fp32", abi.MQ_PREC_FP32)):
        t1 = time.perf_counter()
        e_d, _, its_d, st_d = batch.davidson(psis, tol=1e-8, max_iter=8, precision=prec)
        dev[name] = {"seconds": round(time.perf_counter() - t1, 3), "hpsi_seconds": round(st_d["hpsi_seconds"], 3),
                     "lockstep_iterations": int(st_d["iterations"]), "max_residual_norm": float(st_d["residual_norm"]),
                     "max_abs_delta_e_vs_one_after_the_other": float(max(abs(a - r["e"]) for a, r in zip(e_d, recs)))}
    x = torch.from_numpy(np.concatenate(psis)).cuda()
    y = torch.empty_like(x)
    timing = {}
    for name, prec in (("fp64", abi.MQ_PREC_FP64), ("mixed_fp32", abi.MQ_PREC_FP32)):
        for _ in range(args.warmup):
            batch.apply_device(x.data_ptr(), y.data_ptr(), prec)
        ctx.sync()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            batch.apply_device(x.data_ptr(), y.data_ptr(), prec)
        ctx.sync()
        timing[name] = (time.perf_counter() - t1) / args.steps * 1e3
        if name == "fp64":
            # the batch may split its GEMMs along K differently from the instances alone (the
            # split depends on the number of output tiles), so the sums can differ in rounding
            got = batch.split(y.cpu().numpy())
            err = max(float(np.linalg.norm(g - r["out"]) / np.linalg.norm(r["out"])) for g, r in zip(got, recs))
        bstage = np.zeros(4)
        for _ in range(3):
            ctx.check(ctx.lib.mq_hpsi_timed(batch.h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), prec, sms))
            bstage += np.array(list(sms)) / 3
        timing[name + "_stages"] = [round(v, 4) for v in bstage]
    flops = sum(r["flops"] for r in recs)
    assert flops == bst["flops_total"]
    seq_ms = sum(r["ms"] for r in recs)
    seq_dav = sum(r["davidson_s"] for r in recs)
    de = float(max(abs(a - r["e"]) for a, r in zip(e_ls, recs)))
    log("batch of %d: plan %.1f s, H*psi %.3f ms (sum of instances %.3f ms), FP32 %.3f ms; stages %s / %s; "
        "largest relative difference vs the instances %.2e; davidson lock-step device %s, host %.3f s, one after "
        "the other %.3f s, max |dE| %.2e" % (n, build, timing["fp64"], seq_ms, timing["mixed_fp32"],
                                             timing["fp64_stages"], timing["mixed_fp32_stages"], err, dev, lock_s,
                                             seq_dav, de))
    value = flops / (timing["fp64"] * 1e-3) / 1e12
    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(timing["fp64"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded integrals, random MPS)",
        "config": {
            "workload": "configuration 4: %d independent instances (orbitals U{10..20}, half filling, TwoU1, "
                        "M from {128, 256, 512}), middle two-site problem of each, one batched H*psi" % n,
            "step": "one H*psi over every instance (one batch plan), inputs resident in HBM",
            "instances": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items() if k != "out"}
                          for r in recs],
            "flops_per_step_reference_model": flops, "batch_plan_seconds": round(build, 2),
            "sum_of_instance_hpsi_ms": round(seq_ms, 4),
            "batch_speedup_vs_one_after_the_other": round(seq_ms / timing["fp64"], 3),
            "mixed_fp32": {"ms_per_step": round(timing["mixed_fp32"], 4),
                           "tflops": round(flops / (timing["mixed_fp32"] * 1e-3) / 1e12, 4)},
            "stage_ms_fp64": timing["fp64_stages"], "stage_ms_fp32": timing["mixed_fp32_stages"],
            "sum_of_instance_stage_ms": [round(sum(r["stage_ms"][k] for r in recs), 4) for k in range(4)],
            "max_rel_difference_vs_instances": err,
            "davidson": {"lockstep_device": dev, "lockstep_host_seconds": round(lock_s, 3),
                         "one_after_the_other_seconds": round(seq_dav, 3),
                         "batched_hpsi_calls": int(n_hpsi), "max_abs_delta_e": de,
                         "note": "8 iterations, tol 1e-8; lockstep_device = mq_davidson_batch (one batched H*psi "
                                 "and segmented vector launches per iteration, FP64 and FP32 contraction with FP64 "
                                 "refinement); lockstep_host = batch.py (host subspace algebra, host vectors); one "
                                 "after the other = each instance's device Davidson (mq_davidson, FP64)"},
            "hbm_total_gb": round(total_hbm / 1e9, 1),
            "seconds_building_instances": round(time.perf_counter() - t_all, 1)},
    }
    print(json.dumps(line), flush=True)


class ProfilerRegion:
    """roctxProfilerResume/Pause around the timed region: under `rocprofv3 --selected-regions` only
    the kernels of the timed H*psi calls are traced; without a profiler the calls do nothing."""

    def __init__(self):
        import ctypes as C
        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so"):
            try:
                self.lib = C.CDLL(os.path.join("/opt/rocm/lib", name))
                break
            except OSError:
                continue

    def __enter__(self):
        if self.lib is not None:
            self.lib.roctxProfilerResume(0)
        return self

    def __exit__(self, *exc):
        if self.lib is not None:
            self.lib.roctxProfilerPause(0)
        return False


def sweep_compare(ctx, L=24, M=512, seed=1002, warm_sweeps=2, n_sweeps=4, extra_sweeps=4, cpu_sweep=False):
    """Three two-site DMRG runs from one common start (a seeded random MPS after warm_sweeps FP64
    sweeps), Davidson 8 iterations and tol 1e-8 per step: FP64 with the grouped launches and mixed
    precision (FP32 contraction, FP64 refinement of every eigenvector), n_sweeps + extra_sweeps
    sweeps each, and FP64 with the reference's launch structure (one library dgemm per T task and
    per cohort, one gather per term), n_sweeps sweeps.  The common start keeps the comparison on
    precision: from a random MPS, runs whose Davidson results differ in the last digits can settle
    in different local minima of the sweep.  Reported per run: the wall time and final energy of
    every sweep; against the FP64 run, at sweep n_sweeps (all three runs) and at the last sweep
    (FP64 and mixed): |dE| of the final energy, the largest difference in truncated weight over
    the sweep's steps, and the final Davidson residual; speed-ups at sweep n_sweeps.  cpu_sweep: the
    CPU time of the FP64 run's last sweep, estimated as the H*psi flops of its Davidson iterations
    (the reference's flop model) over the CPU baseline's rate on the middle site problem after that
    sweep (the oracle on the host cores, cpu_baseline)."""
    from qcmaquis_amd import abi
    from qcmaquis_amd.dmrg import TwoSiteDMRG
    from qcmaquis_amd.qc import QCHamiltonian, make_integrals
    h, eri, ec = make_integrals(L, seed, 16)
    ham = QCHamiltonian(h, eri, ec)
    t0 = time.perf_counter()
    d = TwoSiteDMRG(ham, L // 2, L // 2, M=M, ctx=ctx, tol=1e-8, max_iter=8)
    warm_e = [d.sweep()[0] for _ in range(warm_sweeps)]
    start = ([l for l in d.lay], [x.copy() for x in d.data])
    warm_s = time.perf_counter() - t0
    del d
    log("sweeps: common start after %d FP64 sweeps (%.1f s), E %s" % (warm_sweeps, warm_s, warm_e))
    runs = {}
    for name, prec, count in (("fp64", abi.MQ_PREC_FP64, n_sweeps + extra_sweeps),
                              ("mixed", abi.MQ_PREC_FP32, n_sweeps + extra_sweeps),
                              ("fp64_plain", abi.MQ_PREC_FP64_PLAIN, n_sweeps)):
        ctx.set_option(abi.MQ_OPT_PLAIN_TABLES, int(prec == abi.MQ_PREC_FP64_PLAIN))
        d = TwoSiteDMRG(ham, L // 2, L // 2, M=M, ctx=ctx, tol=1e-8, max_iter=8, precision=prec, mps=start)
        rec = {"sweep_seconds": [], "sweep_energies": [], "truncated_weights": [], "residuals": []}
        for s in range(count):
            n0 = len(d.trunc)
            f0 = d.hpsi_flops
            e, dt = d.sweep()
            rec.setdefault("hpsi_flops", []).append(d.hpsi_flops - f0)
            rec["sweep_seconds"].append(round(dt, 3))
            rec["sweep_energies"].append(e)
            rec["truncated_weights"].append(d.trunc[n0:])
            rec["residuals"].append(d.last_residual)
        rec["final_residual_norm"] = d.last_residual
        rec["timing_s"] = {k: round(v, 3) for k, v in d.timing.items()}
        runs[name] = rec
        if name == "fp64" and cpu_sweep:
            try:
                cpu_est = cpu_sweep_estimate(ctx, d, ham, rec["hpsi_flops"][-1], rec["sweep_seconds"][-1])
            except Exception as e:  # the GPU measurement stands on its own
                log("cpu sweep estimate failed:", repr(e))
                cpu_est = None
        log("sweeps %s: %s s, E %s" % (name, rec["sweep_seconds"], ["%.10f" % x for x in rec["sweep_energies"]]))
        del d
    ctx.set_option(abi.MQ_OPT_PLAIN_TABLES, 0)
    ref = runs["fp64"]

    def versus(r, k):
        return {"abs_delta_e": abs(r["sweep_energies"][k] - ref["sweep_energies"][k]),
                "max_abs_delta_truncated_weight": float(np.max(np.abs(np.array(r["truncated_weights"][k]) -
                                                                      np.array(ref["truncated_weights"][k])))),
                "final_residual_norm": r["residuals"][k]}
    k4 = n_sweeps - 1
    cmp = {}
    for name in ("mixed", "fp64_plain"):
        r = runs[name]
        cmp[name] = {f"at_sweep_{n_sweeps}": versus(r, k4),
                     "sweep_seconds": r["sweep_seconds"][k4],
                     "speedup_vs_fp64_plain": round(runs["fp64_plain"]["sweep_seconds"][k4] / r["sweep_seconds"][k4], 3)}
    last = n_sweeps + extra_sweeps - 1
    cmp["mixed"][f"at_sweep_{last + 1}"] = versus(runs["mixed"], last)
    fp64_change = abs(ref["sweep_energies"][last] - ref["sweep_energies"][last - 1])
    out = {"runs": {k: {kk: vv for kk, vv in v.items() if kk not in ("truncated_weights", "residuals")}
                    for k, v in runs.items()},
           "vs_fp64": cmp, "fp64_final_residual_norm": ref["final_residual_norm"],
           "fp64_energy_change_in_last_sweep": fp64_change,
           "fp64_speedup_vs_fp64_plain": round(runs["fp64_plain"]["sweep_seconds"][k4] / ref["sweep_seconds"][k4], 3),
           "common_start": {"fp64_sweeps": warm_sweeps, "energies": warm_e, "seconds": round(warm_s, 3)},
           "cpu_fp64_sweep": cpu_est if cpu_sweep else None,
           "note": f"{L} orbitals, M={M}; from one common start (seeded random MPS after {warm_sweeps} FP64 sweeps) "
                   f"{n_sweeps + extra_sweeps} sweeps of FP64 and mixed and {n_sweeps} of the reference-style FP64, "
                   "Davidson 8 iterations / tol 1e-8 per two-site step; every run is swept for real (its own "
                   "energies, truncations and boundaries)"}
    log("sweep comparison: sweep %d %.2f s (fp64) / %.2f s (mixed) / %.2f s (fp64 plain); mixed |dE| %.2e after %d "
        "sweeps, %.2e after %d (FP64 moved %.2e in its last sweep)"
        % (n_sweeps, ref["sweep_seconds"][k4], runs["mixed"]["sweep_seconds"][k4], runs["fp64_plain"]["sweep_seconds"][k4],
           cmp["mixed"][f"at_sweep_{n_sweeps}"]["abs_delta_e"], n_sweeps,
           cmp["mixed"][f"at_sweep_{last + 1}"]["abs_delta_e"], last + 1, fp64_change))
    return out


def davidson_compare(plan, n, seed=3):
    """precision study on the benchmark site problem: Jacobi-Davidson from one start vector in
    FP64 and in mixed precision (FP32 contraction, FP64 refinement of the eigenvector), with the
    sweep budget (8 iterations, tol 1e-8) and converged (tol 1e-10, at most 60 iterations);
    energies, |dE| of mixed vs FP64, residual norms and wall times"""
    from qcmaquis_amd import abi
    # start vector of a sweep's later steps: close to the ground state (30 FP64 iterations from a
    # random vector), not the random vector itself
    t0 = time.perf_counter()
    e0, psi0, st0 = plan.davidson(np.random.default_rng(seed).normal(size=n), tol=1e-12, max_iter=30)
    out = {"start": {"fp64_iterations": int(st0["iterations"]), "energy": e0,
                     "residual_norm": float(st0["residual_norm"]), "seconds": round(time.perf_counter() - t0, 3)}}
    log("davidson start vector: E=%.10f, |r| %.2e after %d FP64 iterations" % (e0, st0["residual_norm"],
                                                                              st0["iterations"]))
    for name, tol, it in (("budget", 1e-8, 8), ("converged", 1e-10, 60)):
        rec = {}
        for prec_name, prec in (("fp64", abi.MQ_PREC_FP64), ("mixed", abi.MQ_PREC_FP32)):
            t0 = time.perf_counter()
            e, _, st = plan.davidson(psi0, tol=tol, max_iter=it, precision=prec)
            rec[prec_name] = {"energy": e, "iterations": int(st["iterations"]),
                              "refine_iterations": int(st["refine_iterations"]),
                              "residual_norm": float(st["residual_norm"]), "seconds": round(time.perf_counter() - t0, 3)}
        rec["abs_delta_e_mixed_vs_fp64"] = abs(rec["mixed"]["energy"] - rec["fp64"]["energy"])
        rec["tol"], rec["max_iter"] = tol, it
        out[name] = rec
        log("davidson %s: fp64 E=%.10f (%d it, |r| %.2e, %.2f s), mixed E=%.10f (%d it, %d FP64, |r| %.2e, %.2f s), "
            "|dE| %.2e" % (name, rec["fp64"]["energy"], rec["fp64"]["iterations"], rec["fp64"]["residual_norm"],
                           rec["fp64"]["seconds"], rec["mixed"]["energy"], rec["mixed"]["iterations"],
                           rec["mixed"]["refine_iterations"], rec["mixed"]["residual_norm"], rec["mixed"]["seconds"],
                           rec["abs_delta_e_mixed_vs_fp64"]))
    return out


def cpu_sweep_estimate(ctx, d, ham, sweep_flops, gpu_sweep_seconds):
    """CPU time of one configuration-2 FP64 sweep: the sweep's H*psi flops over the CPU baseline's
    rate on the middle two-site problem of the swept state (H*psi only: schedules, growth and the
    split are not counted, so the estimate is low)"""
    from qcmaquis_amd import abi
    from qcmaquis_amd.engine import HpsiPlan
    mid = d.L // 2 - 1
    for k in range(mid):
        d.step(k, True)
    mpo, Ld, Rd, lay2, _ = d.site_problem(mid)
    plan = HpsiPlan(ctx, mpo, Ld, Rd, lay2)
    flops = plan.stats()["flops_total"]
    plan.close()
    w = {"L": Ld, "R": Rd, "layout": lay2, "sym": getattr(ham, "sym", abi.MQ_SYM_TWOU1),
         "mpo_spec": ham.mpo(mid, True).to_spec()}
    cpu = cpu_baseline(w, flops, budget_s=16.0, what="the middle two-site problem of configuration 2")
    sec = sweep_flops / (cpu["value"] * 1e12)
    out = {"seconds_per_sweep_estimated": round(sec, 2), "sweep_hpsi_flops": int(sweep_flops),
           "rate_tflops": cpu["value"], "cores": cpu["cores"], "sample": cpu["sample"],
           "gpu_fp64_sweep_seconds": gpu_sweep_seconds, "gpu_speedup": round(sec / gpu_sweep_seconds, 1),
           "note": "the last FP64 sweep's Davidson H*psi flops (reference model) at the oracle's rate on the "
                   "middle two-site problem (%d flops per H*psi), H*psi only" % flops}
    log("cpu sweep estimate (configuration 2, H*psi only): %.1f s per sweep at %.4f TFLOP/s (GPU sweep %.2f s)"
        % (sec, cpu["value"], gpu_sweep_seconds))
    return out


def cpu_baseline(w, flops, budget_s=30.0, rate_guess_tflops=0.06, what="the benchmark site problem"):
    """the oracle (restated reference CPU path) on the same site problem, all host cores, in two
    threadings: OpenMP over the MPS blocks with one BLAS thread each, and one block at a time with
    every core in BLAS; the better one is the baseline.  When the whole H*psi would take more than
    about budget_s / 2 per threading at rate_guess_tflops, a bounded sample: the H*psi of every k-th
    input block (their flops counted in the same model, schedule excluded)"""
    import oracle
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")))
    Ls, Rs = w["L"].download(), w["R"].download()
    lay = w["layout"]
    psi = np.random.default_rng(1).normal(size=lay.n)
    est = flops / (rate_guess_tflops * 1e12)
    half = budget_s / 2
    variants = {}
    blas_name = "a triple loop"
    if est <= half:
        for name, omp, blas_threads in (("omp_blocks", cores, 1), ("blas_threads", 1, cores)):
            blas = oracle.use_blas(blas_threads)
            blas_name = "OpenBLAS" if blas else "a triple loop"
            t0 = time.perf_counter()
            oracle.hpsi(w["sym"], w["mpo_spec"], Ls, Rs, lay, psi, nthreads=omp)
            dt = time.perf_counter() - t0
            reps = 1
            while dt * (reps + 1) < half and reps < 5:
                t1 = time.perf_counter()
                oracle.hpsi(w["sym"], w["mpo_spec"], Ls, Rs, lay, psi, nthreads=omp)
                dt += time.perf_counter() - t1
                reps += 1
            per = dt / reps
            variants[name] = {"tflops": flops / per / 1e12, "seconds_per_hpsi": per, "hpsi": reps,
                              "openmp_threads": omp, "blas_threads": blas_threads}
        oracle.use_blas(1)
        best = max(variants, key=lambda k: variants[k]["tflops"])
        v = variants[best]
        return {"value": v["tflops"], "unit": "TFLOP/s", "cores": cores, "kind": "port",
                "sample": f"{v['hpsi']} full H*psi of {what} ({v['seconds_per_hpsi']:.3f} s each, "
                          f"schedule build excluded; {best}: {v['openmp_threads']} OpenMP threads over MPS blocks, "
                          f"{v['blas_threads']} {blas_name} threads)",
                "seconds_per_hpsi": v["seconds_per_hpsi"], "threadings": variants}
    nb = len(lay.left)
    stride = max(1, int(np.ceil(est / half)))
    mask = (np.arange(nb) % stride) == (stride // 2)
    for name, omp, blas_threads in (("omp_blocks", cores, 1), ("blas_threads", 1, cores)):
        blas = oracle.use_blas(blas_threads)
        blas_name = "OpenBLAS" if blas else "a triple loop"
        _, fl, sec = oracle.hpsi_sample(w["sym"], w["mpo_spec"], Ls, Rs, lay, psi, mask, nthreads=omp)
        variants[name] = {"tflops": fl / sec / 1e12, "seconds": sec, "openmp_threads": omp,
                          "blas_threads": blas_threads}
    oracle.use_blas(1)
    best = max(variants, key=lambda k: variants[k]["tflops"])
    rate = variants[best]["tflops"]
    return {"value": rate, "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "sample": f"H*psi of {int(mask.sum())} of the {nb} input blocks (every {stride}-th; {fl / flops:.1%} of "
                      f"the flops, schedule build excluded) in two threadings of {cores} cores, the better one "
                      f"({best}: {variants[best]['openmp_threads']} OpenMP threads over the blocks, "
                      f"{variants[best]['blas_threads']} {blas_name} threads, {variants[best]['seconds']:.1f} s)",
            "seconds_per_hpsi_extrapolated": flops / (rate * 1e12), "threadings": variants}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="qc3",
                    choices=["qc", "qc3", "qc3-sweep", "synthetic", "batch", "batch-sweep", "adjoint"])
    ap.add_argument("--plain-sweep", action="store_true", help="qc3-sweep: also sweep with the reference-style FP64")
    ap.add_argument("--sweep-runs", default="fp64,mixed", help="qc3-sweep: runs (fp64, mixed, fp64_plain)")
    ap.add_argument("--M", type=int, default=None, help="bond dimension of the qc / qc3 workloads")
    ap.add_argument("--regen-bonds", action="store_true", help="qc3: redo the warm-up half sweep")
    ap.add_argument("--instances", type=int, default=64, help="instances of the batch workload")
    ap.add_argument("--warm-sweeps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--dump-bonds", default=None, help="write the MPS bond indices of the workload (JSON)")
    ap.add_argument("--no-sweeps", action="store_true", help="skip the FP64 / mixed sweep comparison")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the reference-style FP64 H*psi (thousands of library launches: slow under counters)")
    ap.add_argument("--no-davidson", action="store_true", help="qc3: skip the Davidson precision study")
    ap.add_argument("--chunk-mb", type=int, default=None,
                    help="largest T + S of one plan chunk in MiB (0: one chunk; default: the library's rule)")
    ap.add_argument("--sblock-panel", type=int, default=None,
                    help="term-sum work order: blocks per panel (0: block by block; default: the library's rule)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="qc / qc3: schedule the benchmark plan's GEMM stages by the default tile rule instead of "
                         "the autotuner's timed pick (MQ_OPT_AUTOTUNE)")
    ap.add_argument("--no-overlap", action="store_true",
                    help="run the benchmark plan's chunks one after the other instead of overlapping each chunk's "
                         "output GEMMs with the next chunk's T-GEMMs and term sum (MQ_OPT_OVERLAP)")
    ap.add_argument("--sblock-xcd", type=int, default=None,
                    help="1: each XCD takes a contiguous range of the term-sum work items")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy inside the profiler region (HBM counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the site problem on a single MI355X "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)

    import torch
    from qcmaquis_amd import abi
    from qcmaquis_amd.engine import Context, HpsiPlan

    torch.cuda.init()
    if args.copy_calibration:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        with ProfilerRegion():
            b.copy_(a)
            torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8}), flush=True)
        return
    ctx = Context(0, abi.MQ_SYM_SU2U1 if args.workload in ("qc3", "qc3-sweep", "adjoint") else abi.MQ_SYM_TWOU1)
    if args.sblock_panel is not None:
        ctx.set_option(abi.MQ_OPT_SBLOCK_PANEL, args.sblock_panel)
    if args.sblock_xcd is not None:
        ctx.set_option(abi.MQ_OPT_SBLOCK_XCD, args.sblock_xcd)

    if args.workload == "adjoint":
        adjoint_bench(args, ctx)
        ctx.close()
        return
    if args.workload == "qc3-sweep":
        if args.chunk_mb is not None:
            ctx.set_option(abi.MQ_OPT_CHUNK_MB, args.chunk_mb)
        qc3_sweep_bench(args, ctx)
        ctx.close()
        return
    if args.workload in ("batch", "batch-sweep"):
        (batch_bench if args.workload == "batch" else batch_sweep_bench)(args, ctx)
        ctx.close()
        return
    t0 = time.perf_counter()
    if args.workload == "qc":
        w = qc_workload(ctx, M=args.M or 512, warm_sweeps=args.warm_sweeps, dump_bonds=args.dump_bonds)
    elif args.workload == "qc3":
        w = su2_workload(ctx, M=args.M or 2048, regen=args.regen_bonds)
    else:
        w = synthetic_workload(ctx)
    log("workload built in %.1f s: %s" % (time.perf_counter() - t0, w["desc"]))
    if args.chunk_mb is not None:
        ctx.set_option(abi.MQ_OPT_CHUNK_MB, args.chunk_mb)
    # the benchmark plan's GEMM stages take the autotuner's timed pick (plan creation, untimed);
    # the sweeps below keep the default rule (their stage keys rarely repeat)
    autotune = args.workload in ("qc", "qc3") and not args.no_autotune
    ctx.set_option(abi.MQ_OPT_AUTOTUNE, int(autotune))
    ctx.set_option(abi.MQ_OPT_OVERLAP, int(not args.no_overlap))
    plan = HpsiPlan(ctx, w["mpo"], w["L"], w["R"], w["layout"])
    ctx.set_option(abi.MQ_OPT_AUTOTUNE, 0)
    ctx.set_option(abi.MQ_OPT_OVERLAP, 1)
    st = plan.stats()
    log("plan:", {k: st[k] for k in ("n_t_tasks", "n_cohorts", "n_lines", "n_terms", "flops_total", "flops_gemm",
                                      "flops_gemm_exec", "bytes_compulsory", "t_elements", "s_elements",
                                      "s_line_elements", "gathered_elements", "n_sblocks", "union_elements",
                                      "sblock_macs", "n_chunks", "ts_buffer_elements", "build_seconds",
                                      "seconds_tables", "seconds_blocks")})
    n = w["layout"].n
    psi = torch.randn(n, dtype=torch.float64, device="cuda")
    out = torch.empty_like(psi)
    torch.cuda.synchronize()
    ctx.sync()
    for _ in range(args.warmup):
        plan.apply_device(psi.data_ptr(), out.data_ptr())
    ctx.sync()
    torch.cuda.synchronize()
    import ctypes as C
    stage = np.zeros(4)
    sms = (C.c_float * 4)()
    nrep = max(3, min(args.steps, 10))
    with ProfilerRegion():
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.apply_device(psi.data_ptr(), out.data_ptr())
        ctx.sync()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # per-stage HIP-event timing on the launch stream (separate pass, synchronous per step)
        for _ in range(nrep):
            ctx.check(ctx.lib.mq_hpsi_timed(plan.h, C.c_void_p(psi.data_ptr()), C.c_void_p(out.data_ptr()), 0, sms))
            stage += np.array(list(sms))
    stage /= nrep
    # mixed precision (FP32 operands, intermediates and matrix cores; psi' in FP64): same plan,
    # same timed protocol, error against the FP64 result of the same psi
    ref = out.clone()
    out32 = torch.empty_like(psi)
    for _ in range(args.warmup):
        plan.apply_device(psi.data_ptr(), out32.data_ptr(), abi.MQ_PREC_FP32)
    ctx.sync()
    torch.cuda.synchronize()
    stage32 = np.zeros(4)
    with ProfilerRegion():
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.apply_device(psi.data_ptr(), out32.data_ptr(), abi.MQ_PREC_FP32)
        ctx.sync()
        torch.cuda.synchronize()
        wall32 = time.perf_counter() - t0
        for _ in range(nrep):
            ctx.check(ctx.lib.mq_hpsi_timed(plan.h, C.c_void_p(psi.data_ptr()), C.c_void_p(out32.data_ptr()),
                                            abi.MQ_PREC_FP32, sms))
            stage32 += np.array(list(sms))
    stage32 /= nrep
    ms32 = wall32 / args.steps * 1e3
    err32 = float((out32 - ref).norm() / ref.norm())
    # split-BF16 GEMM products (MQ_PREC_BF16X3): the FP32 path with hi hi + hi lo + lo hi products on
    # the BF16 matrix cores (eager launches)
    out3 = torch.empty_like(psi)
    for _ in range(args.warmup):
        plan.apply_device(psi.data_ptr(), out3.data_ptr(), abi.MQ_PREC_BF16X3)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.apply_device(psi.data_ptr(), out3.data_ptr(), abi.MQ_PREC_BF16X3)
    ctx.sync()
    ms3 = (time.perf_counter() - t0) / args.steps * 1e3
    stage3 = np.zeros(4)
    for _ in range(nrep):
        ctx.check(ctx.lib.mq_hpsi_timed(plan.h, C.c_void_p(psi.data_ptr()), C.c_void_p(out3.data_ptr()),
                                        abi.MQ_PREC_BF16X3, sms))
        stage3 += np.array(list(sms))
    stage3 /= nrep
    err3 = float((out3 - ref).norm() / ref.norm())
    log("BF16x3 H*psi: %.3f ms, stages %s, relative error vs FP64 %.2e" % (ms3, np.round(stage3, 4).tolist(), err3))
    # the same H*psi as captured HIP graphs (psi copied into and out of plan-owned buffers)
    plan.use_graphs(True)
    graph_ms = {}
    outg = torch.empty_like(psi)
    for name, prec in (("fp64", abi.MQ_PREC_FP64), ("mixed_fp32", abi.MQ_PREC_FP32)):
        for _ in range(args.warmup):
            plan.apply_device(psi.data_ptr(), outg.data_ptr(), prec)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.apply_device(psi.data_ptr(), outg.data_ptr(), prec)
        ctx.sync()
        graph_ms[name] = round((time.perf_counter() - t0) / args.steps * 1e3, 4)
    plan.use_graphs(False)
    graph_ok = bool(torch.equal(outg, out32))  # last graph call was FP32 on the same psi
    # precision study of the Davidson on this site problem (the mixed path's energy against FP64)
    dav = davidson_compare(plan, n) if args.workload == "qc3" and not args.no_davidson else None
    # FP64 with the reference's launch structure: one library dgemm per T task and per output
    # block, one gather per term (MQ_PREC_FP64_PLAIN), on a plan of its own (made with the per-task
    # tables) and the same psi
    plain = None
    if not args.no_plain:
        plan.close()
        # one chunk: one dgemm per cohort over all its columns, as the reference issues them
        ctx.set_option(abi.MQ_OPT_PLAIN_TABLES, 1)
        ctx.set_option(abi.MQ_OPT_CHUNK_MB, 0)
        plan = HpsiPlan(ctx, w["mpo"], w["L"], w["R"], w["layout"])
        ctx.set_option(abi.MQ_OPT_PLAIN_TABLES, 0)
        ctx.set_option(abi.MQ_OPT_CHUNK_MB, abi.MQ_CHUNK_AUTO if args.chunk_mb is None else args.chunk_mb)
        outp = torch.empty_like(psi)
        plan.apply_device(psi.data_ptr(), outp.data_ptr(), abi.MQ_PREC_FP64_PLAIN)
        ctx.sync()
        nplain = max(2, min(args.steps, 5))
        stagep = np.zeros(4)
        t0 = time.perf_counter()
        for _ in range(nplain):
            plan.apply_device(psi.data_ptr(), outp.data_ptr(), abi.MQ_PREC_FP64_PLAIN)
        ctx.sync()
        msp = (time.perf_counter() - t0) / nplain * 1e3
        for _ in range(nplain):
            ctx.check(ctx.lib.mq_hpsi_timed(plan.h, C.c_void_p(psi.data_ptr()), C.c_void_p(outp.data_ptr()),
                                            abi.MQ_PREC_FP64_PLAIN, sms))
            stagep += np.array(list(sms))
        stagep /= nplain
        errp = float((outp - ref).norm() / ref.norm())
        log("FP64 plain H*psi: %.3f ms, stages %s, relative difference vs grouped FP64 %.2e" % (
            msp, np.round(stagep, 4).tolist(), errp))
        plain = {"ms_per_step": round(msp, 4), "tflops": round(st["flops_total"] / (msp * 1e-3) / 1e12, 4),
                 "grouped_fp64_speedup": round(msp / (wall / args.steps * 1e3), 3),
                 "mixed_grouped_speedup": round(msp / ms32, 3), "rel_difference_vs_grouped_fp64": errp,
                 "stage_ms": [round(x, 4) for x in stagep],
                 "launches": "one rocBLAS dgemm per executed T task, one S-assembly workgroup per line (%d) "
                             "gathering each term, one rocBLAS dgemm per cohort (%d) added into its output "
                             "block" % (st["n_lines"], st["n_cohorts"])}
    log("graph H*psi ms: %s (FP32 graph output equals eager: %s)" % (graph_ms, graph_ok))
    log("FP32 H*psi: %.3f ms, stages %s, relative error vs FP64 %.2e" % (ms32, np.round(stage32, 4).tolist(), err32))
    ms = wall / args.steps * 1e3
    log("timed %d H*psi: %.3f ms each" % (args.steps, ms))
    log("stage ms (set-up, T-GEMM, S-assembly, L*S GEMM):", np.round(stage, 4).tolist())
    value = st["flops_total"] / (ms * 1e-3) / 1e12
    gemm_ms = stage[1] + stage[3]
    gemm_tf = st["flops_gemm_exec"] / (gemm_ms * 1e-3) / 1e12
    # the term sum's algorithmic bytes: every T element read once and every S element written once
    # (FP64); as executed, every 16-line block reads the union of its T tiles (T elements read
    # union_elements / t_elements times)
    s_bytes = 8 * (st["t_elements"] + st["s_line_elements"])
    s_gbs = s_bytes / (stage[2] * 1e-3) / 1e9
    s_exec_gbs = 8 * (st["union_elements"] + st["s_line_elements"]) / (stage[2] * 1e-3) / 1e9
    rl_gemm = {"bound": "mfma", "kernel": GEMM_FAMILY,
               "achieved": round(gemm_tf, 4), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
               "frac": round(gemm_tf / PEAK_FP64_TFLOPS, 5), "traffic": None,
               "work_per_launch": "%d FP64 flops over the T-GEMM and L*S GEMM stages (%.4f ms)" % (
                   st["flops_gemm_exec"], gemm_ms)}
    rl_sasm = {"bound": "hbm", "kernel": "sblock_kernel<double>", "achieved": round(s_gbs, 2), "peak": PEAK_HBM_GBS,
               "unit": "GB/s", "frac": round(s_gbs / PEAK_HBM_GBS, 5), "traffic": None,
               "achieved_as_executed": round(s_exec_gbs, 2),
               "work_per_launch": "%d bytes: T read once (%d el) and S written once (%d el); as executed %d blocks "
                                  "of up to 64 lines read %d union T tile elements; %d terms, %d el in the per-term "
                                  "formulation" % (s_bytes, st["t_elements"], st["s_line_elements"], st["n_sblocks"],
                                                   st["union_elements"], st["n_terms"], st["gathered_elements"])}
    # the dominant kernel: the term sum is one launch per H*psi and the largest single share of its
    # time; the GEMM stages are spread over three tile-shape instances per stage and a reduce
    dominant, other = rl_sasm, rl_gemm
    # HBM bytes per launch (per H*psi for the GEMM family) from the committed rocprofv3 counter runs
    # of this default workload (FETCH_SIZE / WRITE_SIZE scaled by a device-copy calibration)
    try:
        prof = json.load(open(os.path.join(ROOT, "profiles", "traffic.json"))).get(args.workload, {})
        default_shape = args.workload != "qc" or args.warm_sweeps == 1
        for rl in (dominant, other):
            rec = prof.get(rl["kernel"])
            if rec and default_shape and not args.M:
                rl["traffic"] = round((rec["fetch_gb"] + rec["write_gb"]) * 1e9)
                rl["traffic_source"] = rec["source"]
    except (OSError, ValueError, KeyError):
        pass
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": {"qc": "synthetic (seeded integrals, random initial MPS)",
                 "qc3": "synthetic (seeded rank-80 integrals; MPS sectors from a warm-up half sweep at M=128, "
                        "scaled to M=2048, seeded random elements, normalised)"}.get(args.workload, "synthetic"),
        "config": dict(w["desc"], step="one H*psi (Davidson iteration) of the site problem, inputs resident in HBM",
                       flops_per_step_reference_model=st["flops_total"], gemm_flops_executed=st["flops_gemm_exec"],
                       stage_ms={"setup": round(stage[0], 4), "t_gemm": round(stage[1], 4),
                                 "s_assembly": round(stage[2], 4), "ls_gemm": round(stage[3], 4)},
                       hip_graph_ms_per_step=dict(graph_ms, fp32_output_equals_eager=graph_ok),
                       mixed_fp32={"ms_per_step": round(ms32, 4),
                                   "tflops": round(st["flops_total"] / (ms32 * 1e-3) / 1e12, 4),
                                   "speedup_vs_fp64": round(ms / ms32, 3), "rel_error_vs_fp64": err32,
                                   "stage_ms": [round(x, 4) for x in stage32]},
                       bf16x3={"ms_per_step": round(ms3, 4),
                               "tflops": round(st["flops_total"] / (ms3 * 1e-3) / 1e12, 4),
                               "speedup_vs_fp64": round(ms / ms3, 3), "rel_error_vs_fp64": err3,
                               "stage_ms": [round(x, 4) for x in stage3],
                               "note": "FP32 operands and term sum; GEMM products split into BF16 parts "
                                       "(hi hi + hi lo + lo hi) on the BF16 matrix cores, FP32 accumulation"},
                       fp64_plain=plain, davidson_precision=dav),
        "roofline": dominant,
        "roofline_other": other,
    }
    line["config"]["chunks_overlapped"] = bool(st["n_chunks"] > 1 and not args.no_overlap)
    if autotune:
        line["config"]["autotune"] = {"picks_t_ls": [st["gemm_policy_t"], st["gemm_policy_c"]],
                                      "seconds_tune": round(st["seconds_tune"], 3),
                                      "table": [{"key": k, "kind": kd, "pick": pk, "best_ms": round(b, 4),
                                                 "default_ms": round(dft, 4)} for k, kd, pk, b, dft in ctx.tuned()]}
    plan.close()
    if args.workload in ("qc", "qc3") and not args.no_sweeps:
        # real sweeps run at configuration 2 (24 orbitals, M = 512): a 40-orbital M = 2048 sweep takes
        # tens of minutes
        ctx2 = Context(0, abi.MQ_SYM_TWOU1) if args.workload == "qc3" else ctx
        line["config"]["sweeps_configuration_2"] = sweep_compare(ctx2, cpu_sweep=not args.no_cpu)
        if ctx2 is not ctx:
            ctx2.close()
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(w, st["flops_total"])
            log("cpu baseline:", line["cpu_baseline"])
        except Exception as e:  # the GPU measurement stands on its own
            log("cpu baseline failed:", repr(e))
            line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
