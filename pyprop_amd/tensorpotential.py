"""Device tensor potentials: the host-side mirror of pyprop/tensorpotential/TensorPotential.py
over the gfx950 multiply plans of the native library.

    TensorPotential.MultiplyPotential(srcPsi, destPsi, t, dt)   dest += f(t) V src
    TensorPotential.CanConsolidate(other)                        TensorPotential.py:128-163
    TensorPotentialMultiply(storage_ids, V, scaling, src, dst, pairs)
                                     = core.TensorPotentialMultiply_<ids> (generator.py:1486-1493)
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _native as N
from . import device as D
from .systems import TensorPotentialData


def _pairs_args(geometries):
    rank = len(geometries)
    arr = (C.c_void_p * rank)()
    counts = np.zeros(rank, dtype=np.int32)
    keep = []
    for r, g in enumerate(geometries):
        pairs = getattr(g, "pairs", g)
        if pairs is None:
            arr[r] = None
            continue
        p = np.ascontiguousarray(pairs, dtype=np.int32)
        keep.append(p)
        arr[r] = p.ctypes.data
        counts[r] = p.shape[0]
    return arr, counts, keep


def _ids(storage_ids):
    return np.array([N.STORAGE[s] if isinstance(s, str) else int(s) for s in storage_ids], dtype=np.int32)


def TensorPotentialMultiply(storage_ids, potential, scaling, source, dest, pairs):
    """One-shot multiply with the semantics of the generated kernels: dest += V*scaling*src.
    `potential`, `source`, `dest`: device complex128 tensors; `pairs`: per-rank int32
    (P, 2) arrays (None for ranks without a pair list)."""
    lib = N.load()
    rank = len(storage_ids)
    ids = _ids(storage_ids)
    pext = np.array(potential.shape, dtype=np.int32)
    sext = np.array(source.shape, dtype=np.int32)
    dext = np.array(dest.shape, dtype=np.int32)
    arr, counts, keep = _pairs_args(list(pairs) + [None] * (rank - len(pairs)))
    N.check(lib.pp_tensor_potential_multiply(rank, D.hptr(ids), D.ptr(potential), D.hptr(pext), float(scaling),
                                             D.ptr(source), D.hptr(sext), D.ptr(dest), D.hptr(dext),
                                             C.cast(arr, C.c_void_p), D.hptr(counts), D.stream_handle()))
    return dest


class _CoreNamespace:
    """`core.TensorPotentialMultiply_<S0>_<S1>..._<Sn>(potential, scaling, source, dest, *pairs)`
    resolved by name as TensorPotential.SetupStep does (TensorPotential.py:83-87).  The
    trailing pair arguments are the Simp/Herm ranks' pairs, in rank order."""

    def __getattr__(self, name):
        prefix = "TensorPotentialMultiply_"
        if not name.startswith(prefix):
            raise AttributeError(name)
        ids = name[len(prefix):].split("_")
        for s in ids:
            if s not in N.STORAGE:
                raise AttributeError("no multiply function %s (storage %r)" % (name, s))

        def fn(potential, scaling, source, dest, *pair_args):
            it = iter(pair_args)
            pairs = [next(it) if s in ("Simp", "Herm") else None for s in ids]
            return TensorPotentialMultiply(ids, potential, scaling, source, dest, pairs)

        fn.__name__ = name
        return fn


core = _CoreNamespace()


class TensorPotential:
    """A tensor potential resident on the device (mirror of TensorPotential.py:3-163)."""

    def __init__(self, data: TensorPotentialData, psi_shape, generation: str = "factors"):
        """generation: "factors" expands the separable factors on the device; "grid"
        evaluates the potential on the quadrature grid and projects it on the device
        (GeneratePotential, TensorPotentialGenerator.py:80-274)."""
        if generation not in ("factors", "grid"):
            raise ValueError("generation must be 'factors' or 'grid'")
        self.generation = generation
        self.Name = data.name
        self.GeometryList = data.geometries
        self.IsTimeDependent = data.is_time_dependent
        self.TimeFunction = data.time_function
        self.OriginalTimeFunction = data.time_function
        self.DebugPotential = False
        self.source = data
        self.psi_shape = tuple(int(s) for s in psi_shape)
        self.plan = None
        self.stream_bytes = 0
        self.tuple_count = 0
        self.is_real = False

    # -- set-up --------------------------------------------------------------------
    def device_potential(self):
        """complex128 potential data in the reference layout on the device."""
        d = self.source
        if d.filename:
            # cached potential (TensorPotential.py:56-71): basis pairs checked against the file
            from .serialization import LoadTensorPotential
            return D.to_device(LoadTensorPotential(d.filename, d.dataset, self))
        if self.generation == "grid":
            from .potgen import generate_device_potential
            return generate_device_potential(d)
        if d._data is None and d.angular_coeffs is not None:
            import torch
            coeffs = D.to_device(np.ascontiguousarray(d.angular_coeffs, dtype=np.float64))
            radial = D.to_device(np.ascontiguousarray(d.radial_factors.reshape(d.radial_factors.shape[0], -1),
                                                      dtype=np.float64))
            out = torch.empty(d.shape, dtype=torch.complex128, device="cuda")
            T, P0 = d.angular_coeffs.shape
            N.check(N.load().pp_expand_potential(T, P0, radial.shape[1], D.ptr(coeffs), D.ptr(radial), D.ptr(out),
                                                 D.stream_handle()))
            return out
        return D.to_device(d.data)

    def SetupStep(self, timestep=None):
        """Create the device plan (TensorPotential.SetupStep, TensorPotential.py:80-87)."""
        if self.plan is not None:
            return
        lib = N.load()
        V = self.device_potential()
        rank = len(self.GeometryList)
        ids = _ids([g.storage_id for g in self.GeometryList])
        pext = np.array(V.shape, dtype=np.int32)
        sext = np.array(self.psi_shape, dtype=np.int32)
        arr, counts, keep = _pairs_args([g.pairs if g.storage_id in ("Simp", "Herm") else None
                                         for g in self.GeometryList])
        h = C.c_void_p()
        N.check(lib.pp_tp_create(rank, D.hptr(ids), D.ptr(V), D.hptr(pext), D.hptr(sext), C.cast(arr, C.c_void_p),
                                 D.hptr(counts), D.stream_handle(), C.byref(h)))
        self.plan = h
        del V
        sb = C.c_int64()
        tc = C.c_int64()
        ir = C.c_int()
        ik = C.c_int()
        N.check(lib.pp_tp_info(h, C.byref(sb), C.byref(tc), C.byref(ir), C.byref(ik)))
        self.stream_bytes = sb.value
        self.tuple_count = tc.value
        self.is_real = bool(ir.value)

    def __del__(self):
        if getattr(self, "plan", None) is not None and N._lib is not None:
            N._lib.pp_tp_destroy(self.plan)
            self.plan = None

    # -- multiply ------------------------------------------------------------------
    def time_scaling(self, t: float) -> float:
        return float(self.TimeFunction(t)) if self.IsTimeDependent else 1.0

    def MultiplyPotential(self, srcPsi, destPsi, t, timestep=None):
        """dest += f(t) V src (TensorPotential.py:92-112); srcPsi/destPsi are Wavefunctions
        or device tensors."""
        self.SetupStep()
        src = getattr(srcPsi, "data", srcPsi)
        dst = getattr(destPsi, "data", destPsi)
        N.check(N.load().pp_tp_multiply(self.plan, self.time_scaling(t), D.ptr(src), D.ptr(dst), D.stream_handle()))

    def CanConsolidate(self, other: "TensorPotential") -> bool:
        """TensorPotential.py:128-163."""
        if self.IsTimeDependent != other.IsTimeDependent:
            return False
        if self.IsTimeDependent and other.OriginalTimeFunction is not self.OriginalTimeFunction:
            return False
        if self.DebugPotential or other.DebugPotential:
            return False
        for a, b in zip(self.GeometryList, other.GeometryList):
            if a.storage_id != b.storage_id:
                return False
        for a, b in zip(self.GeometryList, other.GeometryList):
            if a.pairs.shape != b.pairs.shape or np.any(a.pairs != b.pairs):
                return False
        return True
