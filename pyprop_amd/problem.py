"""The propagation driver: time bookkeeping around a propagator's AdvanceStep.

Problem mirrors pyprop/Problem.py:61-309 for the parts on the propagation path:
  SetupStep            Problem.py:113-140 (the propagator's SetupStep(dt))
  RestartPropagation   Problem.py:142-153
  AdvanceStep          Problem.py:156-171 (optional renormalisation; PropagatedTime += dt.real,
                       or += |dt| for purely imaginary dt)
  Advance              Problem.py:181-241 (generator yielding PropagatedTime yieldCount times,
                       the same stopping criteria for real and imaginary time)
  MultiplyHamiltonian  Problem.py:173-178
  GetEnergy            Problem.py:243-247
  GetEnergyExpectationValue  Problem.py:254-277 (normalises psi, <psi|H|psi> with the S metric)
  GetEnergyImTime      Problem.py:281-306 (-log(|psi|^2) / (2 |dt|) after one unnormalised step)

The reference builds the wavefunction and propagator from a config file (CreateWavefunction,
CreatePropagator); here the caller passes the objects, and timestep / duration / start_time
are the [Propagation] section's keys (ApplyConfigSection, Problem.py:309-316).  The
propagator is any object with SetupStep(dt), AdvanceStep(t, dt) and a BasePropagator whose
MultiplyHamiltonian(src, dst, t, dt) applies S^-1 H: CayleyPropagator, PamPropagator and
MixedCayleyPropagator of this package.
"""
from __future__ import annotations

import math
from typing import Iterator, Optional, Union

from .propagator import Wavefunction


class Problem:
    def __init__(self, psi: Wavefunction, propagator, timestep: Union[complex, float], duration: float,
                 start_time: float = 0.0, renormalize: bool = False):
        self.psi = psi
        self.Propagator = propagator
        self.TimeStep = complex(timestep)
        self.Duration = float(duration)
        self.StartTime = float(start_time)
        self.PropagatedTime = float(start_time)
        # the reference keeps this flag on the propagator (Propagator.RenormalizeActive)
        if not hasattr(propagator, "RenormalizeActive"):
            propagator.RenormalizeActive = bool(renormalize)
        elif renormalize:
            propagator.RenormalizeActive = True
        self.TempPsi: Optional[Wavefunction] = None
        self.IgnoreWarningRealEnergy = False

    # propagation ----------------------------------------------------------------------
    def _dt(self):
        """The step handed to the propagator: real when the timestep has no imaginary part."""
        return self.TimeStep.real if self.TimeStep.imag == 0 else self.TimeStep

    def SetupStep(self):
        """Problem.py:113-140 (the initial wavefunction is the caller's psi)."""
        if self.Propagator is not None:
            self.Propagator.SetupStep(self._dt())

    def RestartPropagation(self, timestep, startTime: float, propagationTime: float):
        """Problem.py:142-153."""
        self.TimeStep = complex(timestep)
        self.PropagatedTime = float(startTime)
        self.StartTime = float(startTime)
        self.Duration = float(propagationTime)
        if hasattr(self.Propagator, "RestartPropagation"):
            self.Propagator.RestartPropagation(self._dt(), startTime, propagationTime)

    def AdvanceStep(self):
        """One step from PropagatedTime; then the time bookkeeping of Problem.py:156-171."""
        self.Propagator.AdvanceStep(self.PropagatedTime, self._dt())
        if getattr(self.Propagator, "RenormalizeActive", False):
            self.psi.Normalize()
        if abs(self.TimeStep.real) < 1e-10:
            self.PropagatedTime += abs(self.TimeStep)
        else:
            self.PropagatedTime += self.TimeStep.real

    def Advance(self, yieldCount, duration: Optional[float] = None, yieldEnd: bool = False) -> Iterator[float]:
        """Generator advancing until the propagated time reaches the end; yields
        PropagatedTime `yieldCount` times (every step when yieldCount is a bool),
        Problem.py:181-241."""
        if duration is None:
            duration = self.Duration
        if isinstance(yieldCount, bool):
            yield_step = 1
        else:
            yield_step = max(1, int((duration / abs(self.TimeStep)) / yieldCount))
        if self.TimeStep.real == 0:
            def running():
                return (self.StartTime + self.Duration - self.PropagatedTime) > 0.5 * abs(self.TimeStep)
        else:
            end = self.StartTime + math.copysign(1.0, self.TimeStep.real) * self.Duration

            def running():
                return abs(self.PropagatedTime - end) > 0.5 * abs(self.TimeStep)
        index = 0
        prev = float("nan")
        while running():
            self.AdvanceStep()
            index += 1
            if index % yield_step == 0:
                yield self.PropagatedTime
                prev = self.PropagatedTime
        if yieldEnd and prev != self.PropagatedTime:
            yield self.PropagatedTime

    # observables ----------------------------------------------------------------------
    def MultiplyHamiltonian(self, srcPsi, dstPsi):
        """dst = S^-1 H(PropagatedTime) src (Problem.py:173-178)."""
        self.Propagator.BasePropagator.MultiplyHamiltonian(srcPsi, dstPsi, self.PropagatedTime, self._dt())

    def GetTempPsi(self) -> Wavefunction:
        if self.TempPsi is None:
            self.TempPsi = self.psi.Copy()
        return self.TempPsi

    def GetEnergy(self) -> float:
        """Problem.py:243-247."""
        if self.TimeStep.imag == 0:
            return self.GetEnergyExpectationValue()
        return self.GetEnergyImTime()

    def GetEnergyExpectationValue(self) -> float:
        """Normalise psi, then Re <psi|S (S^-1 H psi)> (Problem.py:254-277).  An energy with an
        imaginary part above 1e-10 raises a RuntimeWarning once (the reference logs a
        warning)."""
        import warnings
        self.psi.Normalize()
        tmp = self.GetTempPsi()
        tmp.Clear()
        self.MultiplyHamiltonian(self.psi, tmp)
        energy = self.psi.InnerProduct(tmp)
        if not self.IgnoreWarningRealEnergy and abs(energy.imag) > 1e-10:
            warnings.warn("Energy is not real (%s). Possible bug. Suppressing further warnings of this type" % energy,
                          RuntimeWarning)
            self.IgnoreWarningRealEnergy = True
        return energy.real

    def GetEnergyImTime(self) -> float:
        """Ground-state energy from the decay of the norm over one imaginary-time step
        without renormalisation (Problem.py:281-306)."""
        if self.TimeStep.imag == 0:
            raise ValueError("Can only find energy for imaginary time propagation")
        renorm = getattr(self.Propagator, "RenormalizeActive", False)
        self.psi.Normalize()
        self.Propagator.RenormalizeActive = False
        try:
            self.AdvanceStep()
        finally:
            self.Propagator.RenormalizeActive = renorm
        norm = self.psi.GetNorm()
        self.psi.data /= norm
        return -math.log(norm ** 2) / (2 * abs(self.TimeStep))
