"""The product's plan builder (host logic of the HIP path) executed on the CPU by the table
emulator, against the CPU oracle: new boundary indices bit-exact, data and H*psi to 1e-12.
This covers the schedule / work-list construction and its bounds checks without a GPU; the GPU
parity tests cover the kernels that execute the same tables."""
import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.model import initial_boundary
from qcmaquis_amd.synthetic import TWOU1_SITE, fuse_index
from tests import emu
from tests.chains import make_chain, rel_err, same_index

CASES = [
    (abi.MQ_SYM_TWOU1, 1, dict(n_sites=5, total=(3, 2), max_m=3, D=6)),
    (abi.MQ_SYM_TWOU1, 2, dict(n_sites=5, total=(3, 2), max_m=6, D=9)),
    (abi.MQ_SYM_TWOU1, 3, dict(n_sites=4, total=(4, 4), max_m=5, D=6, phys=fuse_index(TWOU1_SITE, TWOU1_SITE))),
    (abi.MQ_SYM_SU2U1, 1, dict(n_sites=5, total=(4, 0), max_m=3, D=6)),
    (abi.MQ_SYM_SU2U1, 4, dict(n_sites=5, total=(4, 0), max_m=6, D=9)),
]


@pytest.mark.parametrize("sym,seed,kw", CASES)
def test_chain_emulated(oracle_lib, sym, seed, kw):
    ch = make_chain(seed, sym=sym, **kw)
    n = len(ch.mpos)
    Le = [initial_boundary(ch.bonds[0], ch.mpos[0].row_dim)]
    Lo = [initial_boundary(ch.bonds[0], ch.mpos[0].row_dim)]
    for k in range(n - 1):
        Le.append(emu.lstep(sym, ch.mpos[k], Le[-1], ch.layouts[k], ch.data[k], ch.layouts[k], ch.data[k]))
        Lo.append(oracle_lib.lstep(sym, ch.mpos[k], Lo[-1], ch.layouts[k], ch.data[k], ch.layouts[k], ch.data[k]))
        same_index(Le[-1], Lo[-1])
        assert rel_err(Le[-1].data, Lo[-1].data) < 1e-12
    Re = {n: initial_boundary(ch.bonds[n], ch.mpos[-1].col_dim)}
    Ro = {n: initial_boundary(ch.bonds[n], ch.mpos[-1].col_dim)}
    for k in range(n - 1, 0, -1):
        Re[k] = emu.rstep(sym, ch.mpos[k], Re[k + 1], ch.layouts[k], ch.data[k], ch.layouts[k], ch.data[k])
        Ro[k] = oracle_lib.rstep(sym, ch.mpos[k], Ro[k + 1], ch.layouts[k], ch.data[k], ch.layouts[k], ch.data[k])
        same_index(Re[k], Ro[k])
        assert rel_err(Re[k].data, Ro[k].data) < 1e-12
    rng = np.random.default_rng(seed)
    n_small = {}
    max_chunks = 0
    for k in range(n):
        psi = rng.normal(size=ch.layouts[k].n)
        b = oracle_lib.hpsi(sym, ch.mpos[k], Lo[k], Ro[k + 1], ch.layouts[k], psi)
        # every tile in 16-line blocks, a mix of blocks and small-tile lines, the automatic rule,
        # every tile term by term
        for small_max in (0, 6, -1, 256):
            a = emu.hpsi(sym, ch.mpos[k], Lo[k], Ro[k + 1], ch.layouts[k], psi, small_max)
            assert rel_err(a, b) < 1e-12, (k, small_max)
            n_small[small_max] = n_small.get(small_max, 0) + emu.last_small_lines
        # term-sum work items block by block, and element range by element range over panels of
        # two and of very many blocks (the block table's checks require every range exactly once)
        for panel in (0, 2, 1 << 20):
            emu.lib().emu_set_sblock_panel(panel)
            try:
                a = emu.hpsi(sym, ch.mpos[k], Lo[k], Ro[k + 1], ch.layouts[k], psi, 0)
            finally:
                emu.lib().emu_set_sblock_panel(16)
            assert rel_err(a, b) < 1e-12, (k, panel)
        # every schedule candidate of the GEMM autotuner
        for policy in range(7):
            emu.lib().emu_set_gemm_policy(policy)
            try:
                a = emu.hpsi(sym, ch.mpos[k], Lo[k], Ro[k + 1], ch.layouts[k], psi, -1)
            finally:
                emu.lib().emu_set_gemm_policy(-1)
            assert rel_err(a, b) < 1e-12, (k, "policy", policy)
        # output columns in chunks by right sector: one sector per chunk (budget 1 element), and
        # chunks of about a third of the intermediates, with the grouped and the reference-style
        # tables; chunks run on fresh NaN buffers and psi' starts as NaN (grouped), so a chunk
        # reading another chunk's intermediates or a column left unwritten shows
        chunks = []
        for budget in (1, "third"):
            if budget == "third":
                budget = max(1, emu.last_ts // 3)
            a = emu.hpsi(sym, ch.mpos[k], Lo[k], Ro[k + 1], ch.layouts[k], psi, -1, chunk_elems=budget)
            assert rel_err(a, b) < 1e-12, (k, budget)
            chunks.append(emu.last_chunks)
            # the reference-style tables of the same chunks
            a = emu.hpsi(sym, ch.mpos[k], Lo[k], Ro[k + 1], ch.layouts[k], psi, -1, chunk_elems=budget, plain=True)
            assert rel_err(a, b) < 1e-12, (k, budget, "plain")
        assert 1 <= chunks[1] <= chunks[0] <= len(ch.layouts[k].right)
        max_chunks = max(max_chunks, chunks[0])
    assert n_small[0] == 0 and 0 < n_small[6] < n_small[256]
    assert max_chunks >= 3


def test_nonsymmetric_left(oracle_lib):
    ch = make_chain(6, n_sites=4, max_m=4, D=6)
    Le = Lo = initial_boundary(ch.bonds[0], 1)
    for k in range(3):
        Le = emu.lstep(ch.sym, ch.mpos[k], Le, ch.layouts[k], ch.data[k], ch.layouts[k], ch.data[k], symmetric=False)
        Lo = oracle_lib.lstep(ch.sym, ch.mpos[k], Lo, ch.layouts[k], ch.data[k], ch.layouts[k], ch.data[k],
                              symmetric=False)
        same_index(Le, Lo)
        assert rel_err(Le.data, Lo.data) < 1e-12
