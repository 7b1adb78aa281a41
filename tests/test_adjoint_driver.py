"""The host logic of qcmaquis_amd.adjoint.energy_gradient on the CPU: its device calls (boundary
growth, H*psi plans, mq_hpsi_term_gradient) are replaced by the oracle's growth steps and the table
emulator's MPO-term adjoint, so the pass over the chain, the right-boundary store under a byte cap
(kept boundaries released behind the pass, the others regrown from the nearest kept one) and the
chain rule to the integrals are checked against determinant FCI without a GPU."""
import numpy as np
import pytest

from qcmaquis_amd import abi, adjoint
from qcmaquis_amd.model import initial_boundary
from qcmaquis_amd.qc import QCHamiltonian, make_integrals
from tests import emu
from tests.fci_rdm import fci_ground
from tests.test_qc_gradient import rdm_gradients, untruncated_ground

SU2 = abi.MQ_SYM_SU2U1


class FakeCtx:
    def set_option(self, option, value):
        pass


class FakeBoundary:
    alive = 0

    def __init__(self, spec):
        self.spec = spec
        FakeBoundary.alive += 1
        self.open = True

    def info(self):
        class I:
            pass
        i = I()
        i.total_elements = self.spec.data.size
        return i

    def close(self):
        if self.open:
            FakeBoundary.alive -= 1
        self.open = False

    @classmethod
    def initial(cls, ctx, index, aux_dim=1):
        return cls(initial_boundary(index, aux_dim))


class FakeMPO:
    def __init__(self, ctx, raw):
        self.spec = raw
        self.tospec = raw.to_spec()


class FakeMPS:
    def __init__(self, ctx, layout, data):
        self.layout, self.data = layout, np.array(data)

    def close(self):
        pass


class FakePlan:
    def __init__(self, ctx, mpo, left, right, layout):
        assert left.open and right.open
        self.mpo, self.left, self.right, self.layout = mpo, left, right, layout

    def close(self):
        pass


@pytest.fixture
def cpu_backend(monkeypatch, oracle_lib):
    def grow(fn):
        def step(ctx, mpo, b, bra, ket):
            assert b.open
            return FakeBoundary(fn(SU2, mpo.tospec, b.spec, bra.layout, bra.data, ket.layout, ket.data, nthreads=1))
        return step

    monkeypatch.setattr(adjoint, "DeviceBoundary", FakeBoundary)
    monkeypatch.setattr(adjoint, "DeviceMPO", FakeMPO)
    monkeypatch.setattr(adjoint, "DeviceMPS", FakeMPS)
    monkeypatch.setattr(adjoint, "HpsiPlan", FakePlan)
    monkeypatch.setattr(adjoint, "lstep", grow(oracle_lib.lstep))
    monkeypatch.setattr(adjoint, "rstep", grow(oracle_lib.rstep))
    monkeypatch.setattr(adjoint, "term_gradient", lambda plan, psi: emu.term_grad(
        SU2, plan.mpo.spec, plan.left.spec, plan.right.spec, plan.layout, psi))
    FakeBoundary.alive = 0
    return FakeCtx()


@pytest.mark.parametrize("cap", [None, 0.25])
def test_energy_gradient_driver_on_cpu(cpu_backend, oracle_lib, cap):
    L, N = 6, 6
    h, eri, _ = make_integrals(L, 77, rank=6)
    ham = QCHamiltonian(h, eri, 0.75, symmetry=SU2, gradient_layout=True)
    lay, data, e0 = untruncated_ground(ham, oracle_lib, L, N, 0)
    # hand the driver a state that is not normalised: it normalises a copy
    data = [d * 1.7 for d in data]
    E, dh, deri, st = adjoint.energy_gradient(cpu_backend, ham, lay, data, mem_fraction=cap)
    assert FakeBoundary.alive == 0  # every boundary the driver made is released
    e_fci, gamma, Gamma = fci_ground(h, eri, N // 2, N // 2, 0.75)
    assert abs(E - e_fci) < 1e-10 and np.max(np.abs(st["site_energies"] - e_fci)) < 1e-10
    rh, re = rdm_gradients(gamma, Gamma, L)
    assert np.max(np.abs(dh - rh)) < 1e-9 and np.max(np.abs(deri - re)) < 1e-9
    if cap is None:
        assert st["regrown"] == 0 and st["right_steps"] == L - 1
    else:
        assert st["peak_kept_bytes"] <= cap * st["all_right_bytes"]
        assert st["regrown"] > 0
    gamma2, _ = adjoint.rdms(dh, deri)
    assert np.allclose(gamma2, 0.5 * (gamma + gamma.T), rtol=0, atol=1e-9)
