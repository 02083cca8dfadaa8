"""Batched propagation of many independent hydrogen-like atoms (uneven sizes) in one
persistent launch per pAMP step, with hipGraph-captured steps and an autotuned workgroup
size (pp_batch_* in the C ABI).

Each atom is the reduced-spherical x B-spline problem of hydrogen_system (H0 'Diag' x
BandNH, length-gauge dipole 'Simp' x BandNH, examples/tensor/reducedspherical/potential.cpp)
with its own radial size, lmax and pulse; every step is PamPropagator.AdvanceStep in the
balanced-overlap form (pyprop/propagator/PamPropagator.py:50-59).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from . import _native as N
from . import device as D
from .propagator import band_upper
from .systems import LaserSimpleLength, hydrogen_system, hydrogenic_1s


@dataclass
class AtomSpec:
    lmax: int
    n_bsplines: int
    pulse: LaserSimpleLength
    order: int = 7
    rmax: float = 80.0
    n_inner: int = 20
    gamma: float = 4.0


def random_specs(count: int, seed: int = 0, n_range=(80, 200), lmax_range=(4, 15), e0_range=(0.01, 0.1),
                 omega_range=(0.05, 0.3), cycles_range=(2, 8)) -> List[AtomSpec]:
    """Seeded draw of atom sizes and pulses (uniform in the given ranges; integer ranges
    inclusive)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        lmax = int(rng.integers(lmax_range[0], lmax_range[1] + 1))
        p = LaserSimpleLength(float(rng.uniform(*e0_range)), float(rng.uniform(*omega_range)),
                              float(rng.uniform(*cycles_range)))
        out.append(AtomSpec(lmax, n, p))
    return out


def atom_system(spec: AtomSpec):
    k = spec.order
    return hydrogen_system(lmax=spec.lmax, order=k, n_breakpoints=spec.n_bsplines - k + 4, rmax=spec.rmax,
                           n_inner=spec.n_inner, gamma=spec.gamma, pulse=spec.pulse)


def _band_dia(values, pairs, n, b):
    """BandNH potential values -> DIA [2b+1][n] (element (i, i+d-b) at [d][i])."""
    dia = np.zeros((2 * b + 1, n))
    keep = np.abs(pairs[:, 0] - pairs[:, 1]) <= b
    i = pairs[keep, 0]
    j = pairs[keep, 1]
    dia[j - i + b, i] = values[keep]
    return dia


def pack_atom(spec: AtomSpec):
    """Device-layout arrays of one atom: H0 DIA [nl][w][N], dipole R DIA [w][N], couplings
    [nl][2], overlap upper band [N][b+1], and the S-normalised 1s ground state in l = 0."""
    sysm = atom_system(spec)
    nl, n = sysm.shape
    basis = sysm.bases[1]
    b = basis.order - 1
    H0, dip = sysm.potentials
    g1 = H0.geometries[1]
    data0 = H0.data.real  # (nl, P2)
    h0 = np.stack([_band_dia(data0[l], g1.pairs, n, b) for l in range(nl)])
    gd0, gd1 = dip.geometries
    R = dip.radial_factors[0]
    r = _band_dia(R, gd1.pairs, n, b)
    cpl = np.zeros((nl, 2))
    for a, (l, lp) in enumerate(gd0.pairs):
        c = dip.angular_coeffs[0, a]
        if lp == l - 1:
            cpl[l, 0] = c
        elif lp == l + 1:
            cpl[l, 1] = c
        else:
            raise ValueError("dipole pair (%d, %d) is not l -> l +- 1" % (l, lp))
    su = band_upper(basis.overlap_band(), b)
    c1s, e1s = hydrogenic_1s(basis, 1.0, 0)
    psi = np.zeros((nl, n), complex)
    psi[0] = c1s
    return sysm, h0, r, cpl, su, psi


class BatchedAtoms:
    """A batch of independent atoms propagated by pAMP on the device."""

    def __init__(self, specs: Sequence[AtomSpec], krylov_basis_size: int = 20, pade_order: int = 6,
                 fp32_matvec: bool = False, packs=None):
        """packs: the pack_atom results of `specs` when already built (set-up is host work).

        fp32_matvec: banded H multiply in complex64 with FP32 accumulation.  Off by default:
        the kinetic band of an exponential radial grid spans ~6e6 a.u., so FP32 rounding of
        its entries and of the operand leaves an error of ~1e-7 of that in every product;
        with no refinement around the pAMP step it accumulates to ~3e-3 over 1000 steps,
        while the step, bound by latency, runs no faster (390 vs 387 steps/s measured on
        the 64-atom workload)."""
        packs = [pack_atom(s) for s in specs] if packs is None else packs
        self.packs = packs
        self.specs = list(specs)
        self.systems = [p[0] for p in packs]
        self.shapes = [p[0].shape for p in packs]
        bands = {s.bases[1].order - 1 for s in self.systems}
        if len(bands) != 1:
            raise ValueError("all atoms of a batch need the same B-spline order")
        band = bands.pop()
        nl = np.array([sh[0] for sh in self.shapes], dtype=np.int32)
        nn = np.array([sh[1] for sh in self.shapes], dtype=np.int32)
        h0 = np.ascontiguousarray(np.concatenate([p[1].ravel() for p in packs]))
        r = np.ascontiguousarray(np.concatenate([p[2].ravel() for p in packs]))
        cpl = np.ascontiguousarray(np.concatenate([p[3].ravel() for p in packs]))
        su = np.ascontiguousarray(np.concatenate([p[4].ravel() for p in packs]))
        pulse = np.ascontiguousarray(np.array([[s.pulse.amplitude, s.pulse.frequency, s.pulse.pulse_duration]
                                               for s in specs], dtype=np.float64))
        self.psi0 = [p[5] for p in packs]
        self.m = krylov_basis_size
        h = C.c_void_p()
        N.check(N.load().pp_batch_create(len(specs), D.hptr(nl), D.hptr(nn), band, D.hptr(h0), D.hptr(r),
                                         D.hptr(cpl), D.hptr(su), D.hptr(pulse), krylov_basis_size, pade_order,
                                         int(fp32_matvec), C.byref(h)))
        self.h = h
        self.set_state(self.psi0, np.zeros(len(specs)))

    @property
    def natoms(self):
        return len(self.specs)

    def elements(self) -> int:
        return int(sum(a * b for a, b in self.shapes))

    def set_state(self, psis, times):
        flat = np.ascontiguousarray(np.concatenate([np.asarray(p, complex).ravel() for p in psis]))
        t = np.ascontiguousarray(np.asarray(times, dtype=np.float64))
        N.check(N.load().pp_batch_set_state(self.h, D.hptr(flat), D.hptr(t)))

    def get_state(self):
        flat = np.zeros(self.elements(), complex)
        t = np.zeros(self.natoms)
        err = np.zeros(self.natoms, complex)
        N.check(N.load().pp_batch_get_state(self.h, D.hptr(flat), D.hptr(t), D.hptr(err)))
        out, o = [], 0
        for sh in self.shapes:
            k = sh[0] * sh[1]
            out.append(flat[o:o + k].reshape(sh))
            o += k
        return out, t, err

    def autotune(self, dt: float, trial_steps: int = 3):
        best = C.c_int()
        tim = np.zeros(5)
        N.check(N.load().pp_batch_autotune(self.h, float(dt), trial_steps, C.byref(best), D.hptr(tim),
                                           D.stream_handle()))
        return best.value, dict(zip((64, 128, 256, 512, 1024), tim.tolist()))

    def set_threads(self, threads: int):
        N.check(N.load().pp_batch_set_threads(self.h, threads))

    def step(self, dt: float, nsteps: int = 1):
        N.check(N.load().pp_batch_step(self.h, float(dt), nsteps, D.stream_handle()))

    def capture(self, dt: float, steps_per_graph: int):
        N.check(N.load().pp_batch_capture(self.h, float(dt), steps_per_graph, D.stream_handle()))

    def replay(self, replays: int):
        N.check(N.load().pp_batch_replay(self.h, replays, D.stream_handle()))

    def info(self):
        thr, grid = C.c_int(), C.c_int()
        el, lds = C.c_longlong(), C.c_longlong()
        N.check(N.load().pp_batch_info(self.h, C.byref(thr), C.byref(grid), C.byref(el), C.byref(lds)))
        return dict(threads=thr.value, grid=grid.value, elements=el.value, lds=lds.value)

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_batch_destroy(self.h)
            self.h = None
