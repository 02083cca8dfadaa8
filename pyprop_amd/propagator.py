"""Host-side mirror of PyProp's basis propagation objects over the gfx950 library:

  OverlapMatrix / CombinedRepresentation   core/representation/{overlapmatrix,combinedrepresentation}.cpp
  Wavefunction                             core/wavefunction.{h,cpp}
  BasisPropagator                          pyprop/tensorpotential/BasisPropagator.py
  PamPropagator                            pyprop/propagator/PamPropagator.py
  CayleyPropagator                         pyprop/propagator/CayleyPropagator.py
  RadialTwoElectronPreconditioner          examples/tensor/helium_stabilization/preconditioner.py
  GmresWrapper / PampWrapper               core/krylov/{gmres,pamp}/*wrapper.cpp

Method names and argument meaning follow the reference; all vector work happens on the
device through the C ABI, and a propagation step crosses into Python once (the
reference crosses per matrix-vector product).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Sequence

import numpy as np

from . import _native as N
from . import device as D
from .bspline import BSplineBasis
from .systems import System, TensorPotentialData
from .tensorpotential import TensorPotential

OP_MULTIPLY, OP_SOLVE, OP_SQRT_MULTIPLY, OP_SQRT_SOLVE = 0, 1, 2, 3


def band_upper(S: np.ndarray, sd: int) -> np.ndarray:
    """Dense symmetric S -> BLAS upper band (n, sd+1), element (i, j) at [j][sd + i - j]."""
    n = S.shape[0]
    out = np.zeros((n, sd + 1))
    for j in range(n):
        for i in range(max(0, j - sd), j + 1):
            out[j, sd + i - j] = S[i, j]
    return out


class OverlapMatrix:
    """Banded B-spline overlap and its Cholesky factor (overlapmatrix.cpp:9-37)."""

    def __init__(self, S: np.ndarray, superdiagonals: int):
        self.n = S.shape[0]
        self.sd = superdiagonals
        self.S = S
        self.s_upper = band_upper(S, superdiagonals)
        h = C.c_void_p()
        N.check(N.load().pp_overlap_create(self.n, self.sd, D.hptr(self.s_upper), C.byref(h)))
        self.h = h

    def factor(self) -> np.ndarray:
        u = np.zeros_like(self.s_upper)
        N.check(N.load().pp_overlap_get_factor(self.h, D.hptr(u)))
        return u

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_overlap_destroy(self.h)
            self.h = None


class CombinedRepresentation:
    """Per-rank overlaps of a wavefunction (combinedrepresentation.cpp)."""

    def __init__(self, shape: Sequence[int], bases: Sequence[Optional[BSplineBasis]]):
        D.require_gpu()
        self.shape = tuple(int(s) for s in shape)
        self.bases = list(bases)
        self.overlaps = [None if b is None else OverlapMatrix(b.overlap_band(), b.order - 1) for b in bases]
        ext = np.array(self.shape, dtype=np.int32)
        arr = (C.c_void_p * len(self.shape))(*[None if o is None else o.h.value for o in self.overlaps])
        h = C.c_void_p()
        N.check(N.load().pp_repr_create(len(self.shape), D.hptr(ext), C.cast(arr, C.c_void_p), C.byref(h)))
        self.h = h
        self._work = None

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_repr_destroy(self.h)
            self.h = None

    def IsOrthogonalBasis(self, rank: int) -> bool:
        return self.overlaps[rank] is None

    def _op(self, op, conj, psi):
        data = getattr(psi, "data", psi)
        N.check(N.load().pp_repr_overlap(self.h, op, int(conj), D.ptr(data), D.stream_handle()))

    def MultiplyOverlap(self, psi):
        self._op(OP_MULTIPLY, False, psi)

    def SolveOverlap(self, psi):
        self._op(OP_SOLVE, False, psi)

    def MultiplySqrtOverlap(self, conj, psi):
        self._op(OP_SQRT_MULTIPLY, conj, psi)

    def SolveSqrtOverlap(self, conj, psi):
        self._op(OP_SQRT_SOLVE, conj, psi)

    def InnerProduct(self, a, b) -> complex:
        a = getattr(a, "data", a)
        b = getattr(b, "data", b)
        if self._work is None:
            self._work = D.zeros(self.shape)
        out = N.PPCplx()
        N.check(N.load().pp_repr_inner_product(self.h, D.ptr(a), D.ptr(b), D.ptr(self._work), C.byref(out),
                                               D.stream_handle()))
        return complex(out.re, out.im)


class Wavefunction:
    """complex128 C-ordered data on the device (core/wavefunction.h:10-116)."""

    def __init__(self, repr_: CombinedRepresentation, data=None):
        self.repr = repr_
        self.data = D.zeros(repr_.shape) if data is None else D.to_device(data)

    def GetData(self):
        return self.data

    def GetRepresentation(self):
        return self.repr

    def GetRank(self):
        return len(self.repr.shape)

    def Copy(self):
        return Wavefunction(self.repr, self.data.clone())

    def Clear(self):
        self.data.zero_()

    def InnerProduct(self, other) -> complex:
        return self.repr.InnerProduct(self, other)

    def GetNorm(self) -> float:
        """sqrt(<psi|S|psi>) (wavefunction.cpp:196-200)."""
        return math.sqrt(self.InnerProduct(self).real)

    def Normalize(self):
        n = self.GetNorm()
        self.data /= n
        return n

    def numpy(self) -> np.ndarray:
        return self.data.cpu().numpy()


class BasisPropagator:
    """Sum of device tensor potentials (BasisPropagator.py:1-168)."""

    def __init__(self, psi: Wavefunction, potentials: Sequence[TensorPotentialData], consolidate: bool = True,
                 generation: str = "factors"):
        """generation: how the device potentials are made, "factors" (device expansion of
        the separable factors) or "grid" (grid evaluation and B-spline projection on the
        device, as GeneratePotential does)."""
        self.psi = psi
        self.generation = generation
        # a potential with no basis pairs on some rank (a dipole coupling when only one total
        # L is present) adds nothing to H, as the reference's loop over zero pairs does; it
        # is left out instead of being given an empty device plan
        potentials = [p for p in potentials if all(g.pair_count > 0 for g in p.geometries)]
        self.PotentialList = [TensorPotential(p, psi.repr.shape, generation) for p in potentials]
        if consolidate:
            self.ConsolidatePotentials()
        self.ham = None

    def ConsolidatePotentials(self):
        """Merge potentials with equal storage, pairs and time dependence
        (BasisPropagator.py:68-107); merged data are summed in the reference layout."""
        pots = list(self.PotentialList)
        out = []
        while pots:
            cur = pots.pop(0)
            rest = []
            for other in pots:
                if cur.CanConsolidate(other):
                    d = cur.source
                    o = other.source
                    merged = TensorPotentialData(d.name + "+" + o.name, d.geometries, None, d.time_function,
                                                 np.concatenate([d.angular_coeffs, o.angular_coeffs]) if
                                                 d.angular_coeffs is not None and o.angular_coeffs is not None else None,
                                                 np.concatenate([d.radial_factors, o.radial_factors]) if
                                                 d.radial_factors is not None and o.radial_factors is not None else None,
                                                 d.grid_terms + o.grid_terms if d.grid_terms and o.grid_terms else None)
                    if merged.angular_coeffs is None:
                        merged._data = d.data + o.data
                    cur = TensorPotential(merged, cur.psi_shape, self.generation)
                else:
                    rest.append(other)
            out.append(cur)
            pots = rest
        self.PotentialList = out

    def SetupStep(self, dt=None):
        if self.ham is not None:
            return
        for p in self.PotentialList:
            p.SetupStep(dt)
        arr = (C.c_void_p * len(self.PotentialList))(*[p.plan.value for p in self.PotentialList])
        h = C.c_void_p()
        N.check(N.load().pp_ham_create(self.psi.repr.h, len(self.PotentialList), C.cast(arr, C.c_void_p),
                                       C.byref(h)))
        self.ham = h

    def __del__(self):
        if getattr(self, "ham", None) is not None and N._lib is not None:
            N._lib.pp_ham_destroy(self.ham)
            self.ham = None

    def scalings(self, t: float) -> np.ndarray:
        return np.array([p.time_scaling(t) for p in self.PotentialList], dtype=np.float64)

    def _mul(self, fn, src, dst, t):
        self.SetupStep()
        sc = self.scalings(t)
        N.check(fn(self.ham, D.hptr(sc), D.ptr(getattr(src, "data", src)), D.ptr(getattr(dst, "data", dst)),
                   D.stream_handle()))

    def MultiplyHamiltonianNoOverlap(self, srcPsi, destPsi, t, dt=None):
        self._mul(N.load().pp_ham_multiply_nooverlap, srcPsi, destPsi, t)

    def MultiplyHamiltonian(self, srcPsi, destPsi, t, dt=None):
        self._mul(N.load().pp_ham_multiply, srcPsi, destPsi, t)

    def MultiplyHamiltonianBalancedOverlap(self, srcPsi, destPsi, t, dt=None):
        self._mul(N.load().pp_ham_multiply_balanced, srcPsi, destPsi, t)

    def GetEnergyExpectationValue(self, psi: Wavefunction, t: float) -> float:
        """<psi|H|psi> / <psi|S|psi> (Problem.GetEnergyExpectationValue, Problem.py:257-279)."""
        tmp = Wavefunction(psi.repr)
        self.MultiplyHamiltonian(psi, tmp, t)
        e = psi.InnerProduct(tmp)
        return (e / psi.InnerProduct(psi)).real


class RadialTwoElectronPreconditioner:
    """Exact block-diagonal preconditioner (preconditioner.py:77-198, SuperLU variant)."""

    def __init__(self, psi: Wavefunction, system: System):
        self.psi = psi
        self.system = system
        self.HamiltonianScaling = 0j
        self.OverlapScaling = 1.0
        self.h = None

    def SetHamiltonianScaling(self, scalingH):
        self.HamiltonianScaling = complex(scalingH)

    def SetOverlapScaling(self, scalingS):
        self.OverlapScaling = complex(scalingS)

    def Setup(self, prop=None):
        sysm = self.system
        b1 = sysm.bases[1]
        b2 = sysm.bases[2]
        S1 = b1.overlap_band()
        S2 = b2.overlap_band()
        P = None
        g = sysm.precond_potentials[0].geometries
        for pot in sysm.precond_potentials:
            d = pot.data
            P = d.copy() if P is None else P + d
        P = np.ascontiguousarray(P, dtype=np.complex128)
        p1 = np.ascontiguousarray(g[1].pairs, dtype=np.int32)
        p2 = np.ascontiguousarray(g[2].pairs, dtype=np.int32)
        nc, n1, n2 = sysm.shape
        h = C.c_void_p()
        N.check(N.load().pp_precond_create(nc, n1, n2, b1.order - 1, D.hptr(np.ascontiguousarray(S1)),
                                           D.hptr(np.ascontiguousarray(S2)), D.hptr(P), len(p1), D.hptr(p1), len(p2),
                                           D.hptr(p2), N.cplx(self.HamiltonianScaling), N.cplx(self.OverlapScaling),
                                           C.byref(h)))
        self.h = h

    def Solve(self, psi):
        N.check(N.load().pp_precond_solve(self.h, D.ptr(getattr(psi, "data", psi)), D.stream_handle()))

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_precond_destroy(self.h)
            self.h = None


class CayleyPropagator:
    """(S + i dt/2 H) psi(t+dt) = (S - i dt/2 H) psi(t), H at t + |dt|/2, left
    preconditioned GMRES (CayleyPropagator.py:2-150)."""

    def __init__(self, psi: Wavefunction, base: BasisPropagator, preconditioner=None,
                 krylov_basis_size: int = 30, krylov_tolerance: float = 1e-10,
                 krylov_double_orthogonalization: bool = True):
        self.psi = psi
        self.BasePropagator = base
        self.Preconditioner = preconditioner
        self.basis_size = krylov_basis_size
        self.tol = krylov_tolerance
        self.dorth = krylov_double_orthogonalization
        self.h = None
        self.dt = None

    def SetupStep(self, dt):
        self.BasePropagator.SetupStep(dt)
        if self.Preconditioner is not None and self.Preconditioner.h is None:
            self.Preconditioner.SetHamiltonianScaling(1.0j * dt / 2.0)
            self.Preconditioner.SetOverlapScaling(1.0)
            self.Preconditioner.Setup(self)
        kron = getattr(self.Preconditioner, "kind", None) == "kron"
        if self.h is not None:  # a new step size: replace the solver
            N.check(N.load().pp_cayley_destroy(self.h))
            self.h = None
        h = C.c_void_p()
        N.check(N.load().pp_cayley_create(self.BasePropagator.ham,
                                          None if self.Preconditioner is None or kron else self.Preconditioner.h,
                                          self.basis_size, self.tol, int(self.dorth), C.byref(h)))
        self.h = h
        if kron:
            N.check(N.load().pp_cayley_set_kron(h, self.Preconditioner.h))
        if getattr(self, "factored", None) is not None:
            N.check(N.load().pp_cayley_set_factored(h, self.factored.h))
        self.dt = dt

    def SetFactoredHamiltonian(self, factored):
        """Multiply through a pyprop_amd.factored.FactoredHamiltonian (same operator, MFMA
        evaluation order) instead of the tensor potentials; None restores them."""
        self.factored = factored
        if self.h is not None:
            N.check(N.load().pp_cayley_set_factored(self.h, None if factored is None else factored.h))

    def AdvanceStep(self, t, dt):
        if self.h is None:
            self.SetupStep(dt)
        sc = self.BasePropagator.scalings(t + abs(dt / 2.0))
        N.check(N.load().pp_cayley_step(self.h, D.ptr(self.psi.data), float(dt), D.hptr(sc), D.stream_handle()))

    def GetErrorEstimate(self):
        e = C.c_double()
        n = C.c_int()
        N.check(N.load().pp_cayley_stats(self.h, C.byref(e), C.byref(n)))
        return e.value

    def GetOperatorCount(self):
        e = C.c_double()
        n = C.c_int()
        N.check(N.load().pp_cayley_stats(self.h, C.byref(e), C.byref(n)))
        return n.value

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_cayley_destroy(self.h)
            self.h = None


class PamPropagator:
    """pAMP step in the balanced-overlap form (PamPropagator.py:1-83)."""

    def __init__(self, psi: Wavefunction, base: BasisPropagator, krylov_basis_size: int = 20,
                 pade_order: int = 6, krylov_double_orthogonalization: bool = False,
                 krylov_exponentiation_method: int = 0):
        """krylov_exponentiation_method: 0 = Pade (ExponentiationPade), 1 = eigenvector
        exponential (ExponentiationEigenvector), pamp.h:43-47 / pampwrapper.cpp:56-60."""
        self.psi = psi
        self.BasePropagator = base
        self.basis_size = krylov_basis_size
        h = C.c_void_p()
        N.check(N.load().pp_pamp_create(int(np.prod(psi.repr.shape)), krylov_basis_size, pade_order,
                                        int(krylov_double_orthogonalization), C.byref(h)))
        self.h = h
        if krylov_exponentiation_method:
            N.check(N.load().pp_pamp_set_exponentiation(self.h, int(krylov_exponentiation_method)))

    def SetupStep(self, dt):
        self.BasePropagator.SetupStep(dt)

    def AdvanceStep(self, t, dt):
        self.BasePropagator.SetupStep(dt)
        sc = self.BasePropagator.scalings(t)
        N.check(N.load().pp_pamp_step_balanced(self.h, self.BasePropagator.ham, D.ptr(self.psi.data), N.cplx(dt),
                                               D.hptr(sc), D.stream_handle()))

    def GetPropagationErrorEstimate(self) -> complex:
        e = N.PPCplx()
        b = C.c_int()
        n = C.c_int()
        N.check(N.load().pp_pamp_stats(self.h, C.byref(e), C.byref(b), C.byref(n)))
        return complex(e.re, e.im)

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_pamp_destroy(self.h)
            self.h = None


class GmresWrapper:
    """GMRES on device vectors with a Python operator callback
    (krylov_GmresWrapper, gmreswrapper.cpp:84-110)."""

    def __init__(self, n: int, krylov_basis_size: int = 20, krylov_tolerance: float = 0.0,
                 krylov_double_orthogonalization: bool = True):
        h = C.c_void_p()
        N.check(N.load().pp_gmres_create(n, krylov_basis_size, krylov_tolerance,
                                         int(krylov_double_orthogonalization), C.byref(h)))
        self.h = h
        self.n = n
        self.m = krylov_basis_size

    def Solve(self, callback, rhs, sol, usePypropIntegration: bool = False):
        """sol = A^-1 rhs; callback(src_tensor, dst_tensor) computes dst = A src on the
        device (GmresWrapper::Solve, gmreswrapper.cpp:84-110).  Inner products are the
        plain vector ones; the S-metric integration functor (usePypropIntegration True) is
        not provided and raises."""
        if usePypropIntegration:
            raise NotImplementedError("GMRES with the S-metric integration functor is not provided")
        n = self.n

        def cb(ctx, pin, pout, stream):
            try:
                src = _tensor_view(pin, n)
                dst = _tensor_view(pout, n)
                callback(src, dst)
                return 0
            except Exception:
                return 1

        fn = N.OP_FN(cb)
        N.check(N.load().pp_gmres_solve(self.h, fn, None, D.ptr(rhs), D.ptr(sol), D.stream_handle()))

    def GetErrorEstimate(self) -> float:
        e = C.c_double()
        N.check(N.load().pp_gmres_stats(self.h, C.byref(e), None, None))
        return e.value

    def GetErrorEstimateList(self) -> np.ndarray:
        out = np.zeros(self.m)
        N.check(N.load().pp_gmres_error_list(self.h, D.hptr(out)))
        return out

    def GetOperatorCount(self) -> int:
        n = C.c_int()
        N.check(N.load().pp_gmres_stats(self.h, None, C.byref(n), None))
        return n.value

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_gmres_destroy(self.h)
            self.h = None


class PampWrapper:
    """pAMP on device vectors with a Python operator callback
    (krylov_PampWrapper, pampwrapper.cpp:18-108)."""

    def __init__(self, n: int, krylov_basis_size: int = 20, pade_order: int = 6,
                 krylov_double_orthogonalization: bool = False, krylov_exponentiation_method: int = 0):
        h = C.c_void_p()
        N.check(N.load().pp_pamp_create(n, krylov_basis_size, pade_order, int(krylov_double_orthogonalization),
                                        C.byref(h)))
        self.h = h
        self.n = n
        if krylov_exponentiation_method:
            N.check(N.load().pp_pamp_set_exponentiation(self.h, int(krylov_exponentiation_method)))

    def AdvanceStep(self, callback, psi, tempPsi, dt, t: float, usePypropIntegration: bool = False):
        """psi <- exp(-i dt A) psi, A applied by callback(srcPsi, destPsi, dt, t)
        (PampWrapper::AdvanceStep, pampwrapper.cpp:84-108).  As in ApplyOperator
        (pampwrapper.cpp:18-46) the callback sees psi and tempPsi pointed at the Krylov
        input and output vectors, the output zeroed first; psi and tempPsi are restored
        afterwards.  psi / tempPsi: Wavefunction objects of the same representation.
        The Krylov inner products are the plain vector ones (usePypropIntegration False,
        the balanced-overlap form PamPropagator.py:50-55 uses); the S-metric
        integration of the non-balanced form is not provided and raises."""
        if usePypropIntegration:
            raise NotImplementedError("pAMP with the S-metric integration functor is not provided; "
                                      "propagate in the balanced-overlap form instead")
        shape = tuple(psi.data.shape)
        n = self.n
        if int(np.prod(shape)) != n or tuple(tempPsi.data.shape) != shape:
            raise ValueError("psi and tempPsi must have %d elements of the same shape" % n)

        def cb(ctx, pin, pout, stream):
            old_in, old_out = psi.data, tempPsi.data
            try:
                src = _tensor_view(pin, n).view(shape)
                dst = _tensor_view(pout, n).view(shape)
                dst.zero_()
                psi.data, tempPsi.data = src, dst
                callback(psi, tempPsi, dt, t)
                return 0
            except Exception:
                return 1
            finally:
                psi.data, tempPsi.data = old_in, old_out

        fn = N.OP_FN(cb)
        N.check(N.load().pp_pamp_propagate(self.h, fn, None, D.ptr(psi.data), N.cplx(dt), D.stream_handle()))

    def GetPropagationErrorEstimate(self) -> complex:
        e = N.PPCplx()
        N.check(N.load().pp_pamp_stats(self.h, C.byref(e), None, None))
        return complex(e.re, e.im)

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_pamp_destroy(self.h)
            self.h = None


def _tensor_view(p, n):
    """A torch view of n complex128 device elements at address p (no copy)."""
    import torch
    addr = p if isinstance(p, int) else (p or 0)
    return _DevView.make(addr, n)


class _DevView:
    """Wrap a raw device pointer as a torch tensor via __cuda_array_interface__."""

    def __init__(self, addr, n):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<c16", "data": (addr, False), "version": 3,
                                         "strides": None}

    @staticmethod
    def make(addr, n):
        import torch
        return torch.as_tensor(_DevView(addr, n), device="cuda")
