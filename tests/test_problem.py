"""Host logic of the propagation driver and of the L-block restriction (no GPU).

Problem (pyprop/Problem.py:156-306): time bookkeeping for real and imaginary steps, the
Advance generator's yield schedule and stopping rule, renormalisation, and the
imaginary-time energy formula, driven here by a propagator stand-in that records its
calls and scales a host vector.

restrict_channels: for helium the field-free Hamiltonian is block-diagonal in the total
L, so the dense operator assembled from the restricted potentials must equal the L block
of the dense operator of the full system (SetupBigMatrix assembly, as test_oracle)."""
import math

import numpy as np
import pytest

from tests.test_oracle import dense_operator


class _Vec:
    """Host stand-in for Wavefunction: data, GetNorm, Normalize, Copy, Clear, InnerProduct."""

    def __init__(self, x):
        self.data = np.array(x, dtype=complex)

    def GetNorm(self):
        return math.sqrt(np.vdot(self.data, self.data).real)

    def Normalize(self):
        n = self.GetNorm()
        self.data /= n
        return n

    def Copy(self):
        return _Vec(self.data.copy())

    def Clear(self):
        self.data[:] = 0

    def InnerProduct(self, other):
        return complex(np.vdot(self.data, other.data))


class _Base:
    def __init__(self, energy):
        self.energy = energy

    def MultiplyHamiltonian(self, src, dst, t, dt):
        dst.data[:] = self.energy * src.data


class _Prop:
    """exp(-i dt E) on a host vector; records (t, dt) of every step."""

    def __init__(self, psi, energy):
        self.psi = psi
        self.calls = []
        self.BasePropagator = _Base(energy)
        self.E = energy
        self.setup = None

    def SetupStep(self, dt):
        self.setup = dt

    def AdvanceStep(self, t, dt):
        self.calls.append((t, dt))
        self.psi.data *= np.exp(-1j * complex(dt) * self.E)


def test_problem_real_time_bookkeeping():
    from pyprop_amd.problem import Problem
    psi = _Vec([1.0, 0.0])
    prop = _Prop(psi, 2.0)
    pr = Problem(psi, prop, timestep=0.01, duration=1.0, start_time=5.0)
    pr.SetupStep()
    assert prop.setup == 0.01 and isinstance(prop.setup, float)
    times = list(pr.Advance(10))
    assert len(prop.calls) == 100
    assert np.allclose(times, 5.0 + 0.1 * np.arange(1, 11))
    assert abs(pr.PropagatedTime - 6.0) < 1e-9
    assert np.allclose([c[0] for c in prop.calls], 5.0 + 0.01 * np.arange(100))
    # every step yields with yieldCount True; yieldEnd adds the end when it was not yielded
    pr.RestartPropagation(0.01, 0.0, 0.035)
    assert list(pr.Advance(True)) == pytest.approx([0.01, 0.02, 0.03, 0.04])
    pr.RestartPropagation(0.01, 0.0, 0.05)
    assert list(pr.Advance(2, yieldEnd=True)) == pytest.approx([0.02, 0.04, 0.05])
    assert pr.GetEnergy() == pytest.approx(2.0)


def test_problem_backward_time_and_renormalize():
    from pyprop_amd.problem import Problem
    psi = _Vec([3.0, 4.0])
    prop = _Prop(psi, 1.0)
    pr = Problem(psi, prop, timestep=-0.1, duration=1.0, start_time=2.0, renormalize=True)
    list(pr.Advance(5))
    assert len(prop.calls) == 10 and abs(pr.PropagatedTime - 1.0) < 1e-9
    assert abs(psi.GetNorm() - 1.0) < 1e-14


def test_problem_imaginary_time_energy():
    """dt = -i tau: PropagatedTime advances by tau; the norm decay over one unnormalised
    step gives E = -log(|psi|^2) / (2 tau) (Problem.py:281-306)."""
    from pyprop_amd.problem import Problem
    psi = _Vec([1.0, 1.0])
    prop = _Prop(psi, -0.75)
    pr = Problem(psi, prop, timestep=-0.05j, duration=0.5, renormalize=True)
    list(pr.Advance(True))
    assert len(prop.calls) == 10 and abs(pr.PropagatedTime - 0.5) < 1e-12
    assert all(c[1] == -0.05j for c in prop.calls)
    assert pr.GetEnergy() == pytest.approx(-0.75, abs=1e-12)
    assert prop.RenormalizeActive is True
    with pytest.raises(ValueError):
        Problem(psi, prop, timestep=0.1, duration=1.0).GetEnergyImTime()


@pytest.mark.parametrize("L", [0, 1])
def test_restrict_channels_is_the_L_block(L):
    from pyprop_amd.systems import helium_system, restrict_channels
    sysm = helium_system(lmax=1, L=(0, 1), order=4, n_breakpoints=8, rmax=10.0, n_inner=5, gamma=2.0)
    H, S = dense_operator(sysm, [1.0, 1.0, 0.0])
    idx = [c for c, ch in enumerate(sysm.channels) if ch.L == L]
    sub = restrict_channels(sysm, idx)
    assert all(not p.is_time_dependent for p in sub.potentials) and sub.pulse is None
    Hs, Ss = dense_operator(sub, [1.0] * len(sub.potentials))
    n = sysm.shape[1] * sysm.shape[2]
    rows = np.concatenate([np.arange(c * n, (c + 1) * n) for c in idx])
    assert np.array_equal(Hs, H[np.ix_(rows, rows)])
    assert np.allclose(Ss, S[np.ix_(rows, rows)], rtol=0, atol=0)
    # the field-free operator has no coupling out of the block
    others = np.setdiff1d(np.arange(H.shape[0]), rows)
    assert np.abs(H[np.ix_(rows, others)]).max() == 0
    # the preconditioner potentials follow the same channels
    assert all(p.shape[0] == len(idx) for p in sub.precond_potentials)
