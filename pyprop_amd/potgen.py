"""Potential generation on the device.

Mirrors TensorPotentialGenerator.GeneratePotential
(pyprop/tensorpotential/TensorPotentialGenerator.py:80-274): the potential is evaluated on
the quadrature grid of its B-spline ranks (the GridTerm list of a TensorPotentialData),
then represented in the basis one rank at a time, last rank first (:165-230), by
RepresentPotentialInBasisBSpline (core/tensorpotential/basis_bspline.h:10-47) with
BSplineGlobalOverlapIntegral entries (core/transform/bspline/bspline.cpp:372-466).  The
grid functions, both projections and the complex128 result stay on the device
(pp_potential_grid_*, pp_bspline_represent); the host builds only the per-rank
projection tables from the spline tables.

Rank 0 is the angular rank, whose pairs index the grid function's first axis (its
potential is already in the basis, GeometryInfo.UseGridRepresentation() is false); every
other rank is a B-spline rank.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Sequence

import numpy as np

from . import _native as N
from . import device as D


class BSplineProjection:
    """Projection of one B-spline rank onto a pair list (pp_bspline_projection)."""

    def __init__(self, basis, pairs: np.ndarray, differentiation: int = 0, hermitian: bool = False):
        B, D1, D2 = basis.tables()
        Bd = (B, D1, D2)[int(differentiation)]
        mask = basis.support_mask()
        Q = basis.grid.size
        first = np.argmax(mask, axis=1).astype(np.int32)
        count = mask.sum(axis=1).astype(np.int32)
        for i in range(mask.shape[0]):
            if count[i] and not mask[i, first[i]:first[i] + count[i]].all():
                raise ValueError("spline %d has a non-contiguous support on the grid" % i)
        pairs = np.ascontiguousarray(pairs, dtype=np.int32)
        w = np.ascontiguousarray(basis.grid_weights, dtype=np.float64)
        vals = np.ascontiguousarray(B * mask, dtype=np.float64)
        dvals = np.ascontiguousarray(Bd * mask, dtype=np.float64)
        h = C.c_void_p()
        N.check(N.load().pp_bspline_projection_create(B.shape[0], Q, int(basis.order), D.hptr(w), D.hptr(vals),
                                                      D.hptr(dvals), D.hptr(first), D.hptr(count),
                                                      int(differentiation), pairs.shape[0], D.hptr(pairs),
                                                      int(bool(hermitian)), C.byref(h)))
        self.h = h
        self.n_points = Q
        self.pair_count = pairs.shape[0]

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_bspline_projection_destroy(self.h)
            self.h = None


_cache: Dict[tuple, BSplineProjection] = {}


def projection_for(geometry, differentiation: int) -> BSplineProjection:
    """The projection of a B-spline geometry (cached per basis, storage, pairs and
    differentiation)."""
    if geometry.basis is None:
        raise ValueError("rank is not a B-spline rank")
    if geometry.storage_id not in ("Simp", "Herm", "Band", "BandNH"):
        raise ValueError("storage %r has no B-spline projection" % geometry.storage_id)
    b = geometry.basis
    # keyed on the basis's content (an id could be reused by a new basis after collection)
    key = (int(b.order), b.knots.tobytes(), b.grid.tobytes(), b.grid_weights.tobytes(), geometry.storage_id,
           geometry.pairs.tobytes(), int(differentiation))
    p = _cache.get(key)
    if p is None:
        p = BSplineProjection(geometry.basis, geometry.pairs, differentiation, geometry.storage_id == "Herm")
        _cache[key] = p
    return p


def represent(grid, projections: Sequence[BSplineProjection], out=None, accumulate: bool = False):
    """Represent a real device grid function (P0, Q1[, Q2]) in the pair bases of its
    radial ranks, last rank first; returns (or accumulates into) complex128
    (P0, P1[, P2])."""
    torch = D.torch
    lib = N.load()
    s = D.stream_handle()
    grid = grid.contiguous()
    P0 = grid.shape[0]
    if len(projections) == 1:
        (p1,) = projections
        if grid.shape[1] != p1.n_points:
            raise ValueError("grid does not match the projection")
        if out is None:
            out = torch.empty((P0, p1.pair_count), dtype=torch.complex128, device=grid.device)
        N.check(lib.pp_bspline_represent(p1.h, P0, 1, D.ptr(grid), D.ptr(out), 1, int(accumulate), s))
        return out
    if len(projections) != 2:
        raise ValueError("one or two B-spline ranks supported")
    p1, p2 = projections
    Q1, Q2 = grid.shape[1], grid.shape[2]
    if Q1 != p1.n_points or Q2 != p2.n_points:
        raise ValueError("grid does not match the projections")
    mid = torch.empty((P0, Q1, p2.pair_count), dtype=torch.float64, device=grid.device)
    N.check(lib.pp_bspline_represent(p2.h, P0 * Q1, 1, D.ptr(grid), D.ptr(mid), 0, 0, s))
    if out is None:
        out = torch.empty((P0, p1.pair_count, p2.pair_count), dtype=torch.complex128, device=grid.device)
    N.check(lib.pp_bspline_represent(p1.h, P0, p2.pair_count, D.ptr(mid), D.ptr(out), 1, int(accumulate), s))
    del mid
    return out


def grid_function(term, geometries):
    """A GridTerm evaluated on the device: real float64 (P0, Q1[, Q2])."""
    torch = D.torch
    lib = N.load()
    s = D.stream_handle()
    radial = geometries[1:]
    grids = [g.basis.grid for g in radial]
    P0 = term.coeffs.shape[1]
    coeffs = D.to_device(np.ascontiguousarray(term.coeffs, dtype=np.float64))
    n1 = grids[0].size
    n2 = grids[1].size if len(grids) == 2 else 1
    out = torch.empty((P0, n1, n2), dtype=torch.float64, device="cuda")
    if term.kind == "separable":
        T = term.coeffs.shape[0]
        f1 = D.to_device(np.ascontiguousarray(term.factors[0], dtype=np.float64).reshape(T, n1))
        f2 = (D.to_device(np.ascontiguousarray(term.factors[1], dtype=np.float64).reshape(T, n2)) if len(grids) == 2
              else torch.ones((T, 1), dtype=torch.float64, device="cuda"))
        N.check(lib.pp_potential_grid_separable(P0, T, D.ptr(coeffs), n1, D.ptr(f1), n2, D.ptr(f2), D.ptr(out), s))
    elif term.kind == "multipole":
        if len(grids) != 2:
            raise ValueError("a multipole grid needs two radial ranks")
        lams = np.asarray(term.lams, dtype=np.int32)
        if np.any(np.diff(lams) < 0) or lams.size != term.coeffs.shape[0]:
            raise ValueError("multipole orders must be ascending, one per coefficient row")
        dl = D.to_device(lams)
        r1 = D.to_device(np.ascontiguousarray(grids[0], dtype=np.float64))
        r2 = D.to_device(np.ascontiguousarray(grids[1], dtype=np.float64))
        N.check(lib.pp_potential_grid_multipole(P0, lams.size, D.ptr(dl), D.ptr(coeffs), n1, D.ptr(r1), n2,
                                                D.ptr(r2), D.ptr(out), s))
    else:
        raise ValueError("unknown grid term kind %r" % term.kind)
    return out if len(grids) == 2 else out.view(P0, n1)


def generate_device_potential(data):
    """The complex128 device potential of a TensorPotentialData from its grid terms
    (GeneratePotential), summed over the terms as ConsolidatePotentials sums potentials
    (BasisPropagator.py:68-107).  Each term is evaluated on the grid inside the projection
    of the last rank (pp_potential_represent_*): the grid itself is never stored."""
    if not data.grid_terms:
        raise ValueError("potential %r has no grid description" % data.name)
    torch = D.torch
    lib = N.load()
    s = D.stream_handle()
    radial = data.geometries[1:]
    if len(radial) not in (1, 2):
        raise ValueError("one or two B-spline ranks supported")
    P0 = data.geometries[0].pair_count
    out = torch.empty(data.shape, dtype=torch.complex128, device="cuda")
    work = None
    for k, term in enumerate(data.grid_terms):
        if len(term.differentiation) != len(radial):
            raise ValueError("one differentiation per radial rank")
        if term.coeffs.shape[1] != P0:
            raise ValueError("coefficients do not match the angular pairs")
        projs = [projection_for(g, d) for g, d in zip(radial, term.differentiation)]
        if len(radial) == 2 and work is None:
            work = torch.empty((P0, projs[0].n_points, projs[1].pair_count), dtype=torch.float64, device="cuda")
        coeffs = D.to_device(np.ascontiguousarray(term.coeffs, dtype=np.float64))
        T = term.coeffs.shape[0]
        if term.kind == "separable":
            fs = [D.to_device(np.ascontiguousarray(f, dtype=np.float64).reshape(T, g.basis.grid.size))
                  for f, g in zip(term.factors, radial)]
            if len(radial) == 1:
                N.check(lib.pp_potential_represent_separable(projs[0].h, None, P0, T, D.ptr(coeffs), D.ptr(fs[0]),
                                                             None, None, D.ptr(out), int(k > 0), s))
            else:
                N.check(lib.pp_potential_represent_separable(projs[0].h, projs[1].h, P0, T, D.ptr(coeffs),
                                                             D.ptr(fs[0]), D.ptr(fs[1]), D.ptr(work), D.ptr(out),
                                                             int(k > 0), s))
        elif term.kind == "multipole":
            if len(radial) != 2:
                raise ValueError("a multipole grid needs two radial ranks")
            lams = np.asarray(term.lams, dtype=np.int32)
            if np.any(np.diff(lams) < 0) or lams.size != T:
                raise ValueError("multipole orders must be ascending, one per coefficient row")
            dl = D.to_device(lams)
            r1 = D.to_device(np.ascontiguousarray(radial[0].basis.grid, dtype=np.float64))
            r2 = D.to_device(np.ascontiguousarray(radial[1].basis.grid, dtype=np.float64))
            N.check(lib.pp_potential_represent_multipole(projs[0].h, projs[1].h, P0, T, D.ptr(dl), D.ptr(coeffs),
                                                         D.ptr(r1), D.ptr(r2), D.ptr(work), D.ptr(out), int(k > 0),
                                                         s))
        else:
            raise ValueError("unknown grid term kind %r" % term.kind)
    return out


def generate_device_potential_staged(data):
    """The same potential with the grid stored between the steps (pp_potential_grid_*,
    then pp_bspline_represent per rank), GeneratePotential's steps 4 and 5 one by one."""
    if not data.grid_terms:
        raise ValueError("potential %r has no grid description" % data.name)
    out = None
    for term in data.grid_terms:
        if len(term.differentiation) != len(data.geometries) - 1:
            raise ValueError("one differentiation per radial rank")
        F = grid_function(term, data.geometries)
        projs = [projection_for(g, d) for g, d in zip(data.geometries[1:], term.differentiation)]
        out = represent(F, projs, out, accumulate=out is not None)
        del F
    return out
