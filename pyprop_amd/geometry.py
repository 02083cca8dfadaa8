"""Per-rank storage geometries of a tensor potential: the storage id the multiply
dispatches on and the (row, col) basis-pair list that indexes the potential data.

Reference: pyprop/tensorpotential/GeometryInfo.py:170-462 (common geometries),
pyprop/tensorpotential/basis/BSpline.py:8-109,200-222 (B-spline geometries),
pyprop/tensorpotential/basis/CoupledSphericalHarmonic.py:5-32,133-164 (coupled spherical),
pyprop/tensorpotential/basis/ReducedSphericalHarmonic.py:5-45,121-140 (reduced spherical).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import angular
from .bspline import (BSplineBasis, banded_hermitian_pairs, banded_nonhermitian_pairs,
                      banded_packed_pairs)

STORAGE_IDS = ("Simp", "Herm", "Diag", "Ident", "Band", "BandNH", "Dense")


@dataclass
class Geometry:
    """One rank of a tensor potential (GeometryInfo*)."""

    storage_id: str
    pairs: np.ndarray          # (P, 2) int32, row = dest index, col = source index
    use_grid: bool             # potential is evaluated on a grid and projected
    extent: int                # basis size of this rank
    basis: Optional[BSplineBasis] = None  # the B-spline basis of a B-spline rank

    @property
    def pair_count(self) -> int:
        return int(self.pairs.shape[0])

    def multiply_arguments(self):
        """Only pair-list storages pass their pairs to the multiply
        (GeometryInfo.py:197-202, 233-238; CoupledSphericalHarmonic.py:31-32)."""
        if self.storage_id in ("Simp", "Herm"):
            return [self.pairs]
        return []


def dense_pairs(n: int) -> np.ndarray:
    return np.array([(i, j) for i in range(n) for j in range(n)], dtype=np.int32).reshape(-1, 2)


def dense_hermitian_pairs(n: int) -> np.ndarray:
    return np.array([(i, j) for i in range(n) for j in range(i, n)], dtype=np.int32).reshape(-1, 2)


def bspline_geometry(basis: BSplineBasis, name: str) -> Geometry:
    """BasisfunctionBSpline.GetGeometryInfo (basis/BSpline.py:200-222)."""
    g = name.lower().strip()
    N = basis.number_of_bsplines
    k = basis.order
    if g == "identity":
        return Geometry("Ident", np.zeros((1, 2), np.int32), True, N, basis)
    if g == "banded-nonhermitian":
        return Geometry("BandNH", banded_nonhermitian_pairs(N, k - 1), True, N, basis)
    if g == "banded-packed":
        return Geometry("Simp", banded_packed_pairs(N, k), True, N, basis)
    if g == "banded-hermitian":
        return Geometry("Band", banded_hermitian_pairs(N, k), True, N, basis)
    if g in ("dense", "hermitian"):
        return Geometry("Simp", dense_pairs(N), True, N, basis)
    raise ValueError("geometry %r not supported by the B-spline basis" % name)


def coupled_geometry(channels, name: str) -> Geometry:
    """BasisfunctionCoupledSphericalHarmonic.GetGeometryInfo
    (basis/CoupledSphericalHarmonic.py:133-164); storage is always Simp."""
    g = name.lower().strip()
    n = len(channels)
    if g == "identity":
        return Geometry("Ident", np.zeros((1, 2), np.int32), False, n)
    if g == "dense":
        return Geometry("Simp", dense_pairs(n), False, n)
    return Geometry("Simp", angular.coupled_basis_pairs(channels, g), False, n)


def reduced_spherical_geometry(lmax: int, name: str) -> Geometry:
    """BasisfunctionReducedSphericalHarmonic.GetGeometryInfo
    (basis/ReducedSphericalHarmonic.py:121-140)."""
    g = name.lower().strip()
    n = lmax + 1
    if g == "identity":
        return Geometry("Ident", np.zeros((1, 2), np.int32), True, n)
    if g == "diagonal":
        return Geometry("Diag", np.stack([np.arange(n), np.arange(n)], 1).astype(np.int32), False, n)
    if g == "dense":
        return Geometry("Simp", dense_pairs(n), True, n)
    if g == "dipoleselectionrule":
        return Geometry("Simp", angular.reduced_dipole_pairs(lmax), True, n)
    raise ValueError("geometry %r not supported by the reduced spherical basis" % name)
