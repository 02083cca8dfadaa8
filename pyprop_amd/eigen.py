"""Initial-state eigensolvers: pIRAM and GMRES shift-invert on the device.

PiramSolver mirrors pyprop/solver/PiramSolver.py:3-141 (Solve, GetEigenvalues,
GetEigenvector, SetEigenvector, error and convergence estimates) over pp_piram_*.
GMRESShiftInvertSolver mirrors pyprop/solver/GMRESShiftInvertSolver.py:1-132
(InverseIterations: dst = (H - shift S)^-1 S src, left-preconditioned GMRES with the
preconditioner set to Hamiltonian scaling 1 and overlap scaling -shift) over
pp_shift_invert_*.  Passed to PiramSolver as its matrix-vector product, as the reference's
`inverse_iterations = True` configuration does, the whole solve runs inside the library:
the operator is the library's own pp_shift_invert_apply, with no Python call per
matrix-vector product.

The Ritz values of the shift-inverted operator are theta = 1 / (E - shift); the wanted ones
have the smallest real part (functors.h:95-102), so a shift just above an eigenvalue finds
it first.  Energies: E = shift + 1 / theta (ShiftInvertEnergies).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _native as N
from . import device as D


def _fnptr(name: str) -> C.c_void_p:
    return C.cast(getattr(N.load(), name), C.c_void_p)


class HamiltonianOperator:
    """The Hamiltonian as a device operator for PiramSolver: mode 0 = H
    (MultiplyHamiltonianNoOverlap), 1 = S^-1 H (MultiplyHamiltonian), 2 = U^-T H U^-1
    (MultiplyHamiltonianBalancedOverlap); BasisPropagator.py:116-159.  Time scalings are
    those of `time`."""

    MODES = {"nooverlap": 0, "overlap": 1, "balanced": 2}

    def __init__(self, base, mode: str = "overlap", time: float = 0.0):
        base.SetupStep()
        self.base = base
        sc = np.ascontiguousarray(base.scalings(time), dtype=np.float64)
        h = C.c_void_p()
        N.check(N.load().pp_ham_op_create(base.ham, D.hptr(sc), self.MODES[mode], C.byref(h)))
        self.h = h

    def operator(self):
        return _fnptr("pp_ham_op_apply"), self.h

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_ham_op_destroy(self.h)
            self.h = None


class GMRESShiftInvertSolver:
    """dst = (H - shift S)^-1 S src (GMRESShiftInvertSolver.py:1-132)."""

    def __init__(self, psi, base, shift: float, preconditioner=None, krylov_basis_size: int = 20,
                 krylov_tolerance: float = 1e-10, krylov_double_orthogonalization: bool = True, time: float = 0.0,
                 factored=None, negate: bool = False):
        """preconditioner: a RadialTwoElectronPreconditioner (set up here with Hamiltonian
        scaling 1 and overlap scaling -shift, GMRESShiftInvertSolver.py:44-46) or None.
        factored: a FactoredHamiltonian to use for H (optional).
        negate: as a pIRAM operator, apply -(H - shift S)^-1 S (pp_shift_invert_set_negate):
        with a shift below the spectrum, pIRAM's wanted set is then the lowest energies;
        E = ShiftInvertEnergies(theta, shift, negate=True)."""
        self.psi = psi
        self.BasePropagator = base
        self.Shift = float(shift)
        self.Preconditioner = preconditioner
        base.SetupStep()
        if preconditioner is not None:
            preconditioner.SetHamiltonianScaling(1.0)
            preconditioner.SetOverlapScaling(-self.Shift)
            preconditioner.Setup(base)
        h = C.c_void_p()
        N.check(N.load().pp_shift_invert_create(base.ham, None if preconditioner is None else preconditioner.h,
                                                krylov_basis_size, krylov_tolerance,
                                                int(krylov_double_orthogonalization), self.Shift, C.byref(h)))
        self.h = h
        sc = np.ascontiguousarray(base.scalings(time), dtype=np.float64)
        N.check(N.load().pp_shift_invert_set_scalings(self.h, D.hptr(sc)))
        self.factored = factored
        if factored is not None:
            N.check(N.load().pp_shift_invert_set_factored(self.h, factored.h))
        self.negate = bool(negate)
        if negate:
            N.check(N.load().pp_shift_invert_set_negate(self.h, 1))

    def InverseIterations(self, srcPsi, destPsi, t=0.0, dt=0.0):
        """destPsi = (H - shift S)^-1 S srcPsi (GMRESShiftInvertSolver.py:100-126)."""
        src = getattr(srcPsi, "data", srcPsi)
        dst = getattr(destPsi, "data", destPsi)
        N.check(N.load().pp_shift_invert_solve(self.h, D.ptr(src), D.ptr(dst), D.stream_handle()))

    def operator(self):
        return _fnptr("pp_shift_invert_apply"), self.h

    def GetStatistics(self):
        solves, ops = C.c_int(), C.c_int()
        err = C.c_double()
        N.check(N.load().pp_shift_invert_stats(self.h, C.byref(solves), C.byref(ops), C.byref(err)))
        return {"InverseIterations": solves.value, "MultiplyHamiltonian": ops.value, "LastError": err.value}

    def PrintStatistics(self):
        print("Statistics for GMRESShiftInvertSolver")
        print(self.GetStatistics())

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_shift_invert_destroy(self.h)
            self.h = None


class PiramSolver:
    """Implicitly restarted Arnoldi (PiramSolver.py:3-141, piram.h) on the device.

    matrix_vector_func: an object with operator() -> (function address, context), such as
    HamiltonianOperator or GMRESShiftInvertSolver (inverse iterations)."""

    def __init__(self, psi, matrix_vector_func, krylov_basis_size: int = 15, krylov_eigenvalue_count: int = 1,
                 krylov_tolerance: float = 1e-10, krylov_max_iteration_count: int = 500,
                 krylov_use_random_start: bool = True, max_orthogonalization: int = 3):
        self.psi = psi
        self.MatrixVector = matrix_vector_func
        self.UseRandomStart = krylov_use_random_start
        self.EigenvalueCount = krylov_eigenvalue_count
        n = int(np.prod(psi.data.shape))
        h = C.c_void_p()
        N.check(N.load().pp_piram_create(n, krylov_basis_size, krylov_eigenvalue_count, krylov_tolerance,
                                         krylov_max_iteration_count, max_orthogonalization, C.byref(h)))
        self.h = h

    def Solve(self):
        """Run the Arnoldi iterations from psi (or the random start)."""
        fn, ctx = self.MatrixVector.operator()
        start = None if self.UseRandomStart else D.ptr(self.psi.data)
        N.check(N.load().pp_piram_solve(self.h, fn, ctx, start, D.stream_handle()))

    def _values(self):
        k = self.EigenvalueCount
        vals = np.zeros(2 * k)
        err = np.zeros(k)
        conv = np.zeros(k)
        cnt = C.c_int()
        N.check(N.load().pp_piram_eigenvalues(self.h, D.hptr(vals), D.hptr(err), D.hptr(conv), C.byref(cnt)))
        return vals[0::2] + 1j * vals[1::2], err, conv, cnt.value

    def GetEigenvalues(self):
        """Real parts of the wanted Ritz values (PiramSolver.py:116-121)."""
        return self._values()[0].real.copy()

    def GetComplexEigenvalues(self):
        return self._values()[0]

    def GetErrorEstimates(self):
        return self._values()[1]

    def GetConvergenceEstimates(self):
        """error bound - accuracy per wanted value; <= 0 means converged (piram.h:192-195)."""
        return self._values()[2]

    def GetConvergedEigenvalueCount(self):
        return self._values()[3]

    def GetEigenvector(self, index: int) -> np.ndarray:
        """Ritz vector `index` as a flat array, unit in the vector 2-norm (PiramSolver.py:123-134)."""
        out = D.torch.empty(self.psi.data.shape, dtype=D.torch.complex128, device=self.psi.data.device)
        N.check(N.load().pp_piram_eigenvector(self.h, index, D.ptr(out), D.stream_handle()))
        return out.cpu().numpy().ravel()

    def SetEigenvector(self, psi, eigenvectorIndex: int, normalize: bool = True):
        """psi <- the eigenvector, S-normalised when `normalize` (PiramSolver.py:136-141)."""
        N.check(N.load().pp_piram_eigenvector(self.h, eigenvectorIndex, D.ptr(psi.data), D.stream_handle()))
        if normalize:
            psi.Normalize()

    def GetStatistics(self):
        ops, rst, orth = C.c_int(), C.c_int(), C.c_int()
        N.check(N.load().pp_piram_stats(self.h, C.byref(ops), C.byref(rst), C.byref(orth)))
        return {"OperatorCount": ops.value, "RestartCount": rst.value, "OrthogonalizationCount": orth.value}

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_piram_destroy(self.h)
            self.h = None


def ShiftInvertEnergies(theta, shift: float, negate: bool = False):
    """Eigenvalues E of (H, S) from the Ritz values theta of (H - shift S)^-1 S (of its
    negative when negate)."""
    theta = np.asarray(theta)
    return shift - 1.0 / theta if negate else shift + 1.0 / theta


def ImaginaryTimePropagation(propagator, psi, timestep: float, steps: int, renormalize: bool = True, t0: float = 0.0):
    """Ground state by propagation in imaginary time, as the reference's FindGroundstate
    (examples/tensor/helium_stabilization/example.py:379-389) runs Problem.Advance with
    imtime: each step propagates with dt = -i timestep (exp(-timestep H) for the
    propagators here), then S-normalises psi (Problem.py:161-190).  The potentials are
    evaluated at the real time t0.  Returns the energy <psi|H|psi> after each step."""
    energies = []
    dt = -1j * float(timestep)
    base = propagator.BasePropagator
    for _ in range(steps):
        propagator.AdvanceStep(t0, dt)
        if renormalize:
            psi.Normalize()
        energies.append(base.GetEnergyExpectationValue(psi, t0))
    return np.array(energies)
