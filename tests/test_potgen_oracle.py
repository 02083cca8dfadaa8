"""Potential generation: the oracle's restatement of GeneratePotential's projection
(RepresentPotentialInBasisBSpline + BSplineGlobalOverlapIntegral, rank by rank) pinned by
closed forms, the hydrogen spectrum and separability, then used to check that the grid
description of every system potential (GridTerm lists) reproduces the potential the
separable factors give."""
import numpy as np
import pytest
import scipy.linalg

import oracle
from pyprop_amd.bspline import (BSplineBasis, banded_hermitian_pairs, banded_nonhermitian_pairs,
                                banded_packed_pairs, linear_breakpoints)
from pyprop_amd.geometry import dense_pairs
from pyprop_amd.systems import helium_system, hydrogen_system


def host_grid(term, geometries):
    """A GridTerm on the quadrature grid, in numpy (P0, Q1[, Q2])."""
    grids = [g.basis.grid for g in geometries[1:]]
    c = term.coeffs
    if term.kind == "separable":
        if len(grids) == 1:
            return np.einsum("ta,tq->aq", c, np.asarray(term.factors[0]).reshape(c.shape[0], -1))
        f1 = np.asarray(term.factors[0]).reshape(c.shape[0], -1)
        f2 = np.asarray(term.factors[1]).reshape(c.shape[0], -1)
        return np.einsum("ta,tq,tr->aqr", c, f1, f2)
    r1, r2 = grids
    rlo = np.minimum.outer(r1, r2)
    rhi = np.maximum.outer(r1, r2)
    G = np.stack([(rlo / rhi) ** lam / rhi for lam in term.lams])
    return np.einsum("ta,tqr->aqr", c, G)


def oracle_generate(data):
    """GeneratePotential on the oracle: grid, then the last rank, then the one before."""
    out = 0.0
    for term in data.grid_terms:
        F = host_grid(term, data.geometries)
        for r in reversed(range(1, len(data.geometries))):
            g = data.geometries[r]
            F = oracle.bspline_represent(g.basis, F, term.differentiation[r - 1], g.pairs, g.storage_id == "Herm",
                                         axis=r)
        out = out + F
    return out


def test_oracle_tables_match_host_basis():
    b = BSplineBasis(6, linear_breakpoints(0.0, 7.0, 11) ** 1.2, "zero")
    grid, gw, b0, b1, b2 = oracle.bspline_tables(b)
    B, D1, D2 = b.tables()
    assert np.allclose(grid, b.grid, rtol=0, atol=1e-14)
    assert np.allclose(gw, b.grid_weights, rtol=0, atol=1e-15)
    for x, y in ((b0, B), (b1, D1), (b2, D2)):
        assert np.abs(x - y).max() <= 1e-12 * np.abs(y).max()


def test_oracle_overlap_row_sums_closed_form():
    """F = 1, d = 0 on dense pairs gives S; in a vanilla basis sum_j B_j = 1, so
    sum_j S_ij = int B_i = (t_{i+k} - t_i) / k.  Symmetric, positive definite."""
    k = 5
    b = BSplineBasis(k, linear_breakpoints(0.0, 3.0, 7) ** 1.3, "vanilla")
    N = b.number_of_bsplines
    S = oracle.bspline_represent(b, np.ones((1, b.grid.size)), 0, dense_pairs(N), False, axis=1).reshape(N, N)
    t = b.knots
    assert np.allclose(S.sum(axis=1), [(t[i + k] - t[i]) / k for i in range(N)], rtol=1e-13)
    assert np.allclose(S, S.T, atol=1e-15)
    assert np.linalg.eigvalsh(S).min() > 0


def test_oracle_hydrogen_spectrum():
    """The hydrogen H0 grid terms (kinetic with d = 2, centrifugal + Coulomb) projected by
    the oracle on BandNH pairs: -1/(2 n^2) for the low levels of every l.  A wrong sign of
    the derivative reorder, a dropped term or a transposed pair fails this."""
    h = hydrogen_system()
    b = h.bases[1]
    N = b.number_of_bsplines
    H0 = oracle_generate(h.potentials[0])
    S = oracle.bspline_represent(b, np.ones((1, b.grid.size)), 0, dense_pairs(N), False, axis=1).reshape(N, N)
    pairs = banded_nonhermitian_pairs(N, 6)
    for l in range(4):
        M = np.zeros((N, N))
        for p, (i, j) in enumerate(pairs):
            if abs(i - j) <= 6:
                M[i, j] = H0[l, p]
        assert np.allclose(M, M.T, atol=1e-9 * np.abs(M).max())
        w = scipy.linalg.eigh(0.5 * (M + M.T), S, eigvals_only=True)
        ns = list(range(l + 1, 5))
        assert np.allclose(w[:len(ns)], [-0.5 / n ** 2 for n in ns], rtol=1e-7, atol=1e-9)


def test_oracle_two_ranks_separable_and_order_free(rng):
    """F = f1 (x) f2 projects to the outer product of the one-rank projections, and the
    rank order does not matter (a transposed pre/post index fails this)."""
    b1 = BSplineBasis(5, linear_breakpoints(0.0, 6.0, 9) ** 1.1, "zero")
    b2 = BSplineBasis(6, linear_breakpoints(0.0, 4.0, 8), "zero")
    p1 = banded_packed_pairs(b1.number_of_bsplines, 5)
    p2 = banded_nonhermitian_pairs(b2.number_of_bsplines, 5)
    f1 = rng.standard_normal(b1.grid.size)
    f2 = rng.standard_normal(b2.grid.size)
    F = np.stack([np.multiply.outer(f1, f2), 2.0 * np.multiply.outer(f1, f2)])
    V12 = oracle.bspline_represent(b1, oracle.bspline_represent(b2, F, 1, p2, False, axis=2), 2, p1, False, axis=1)
    V21 = oracle.bspline_represent(b2, oracle.bspline_represent(b1, F, 2, p1, False, axis=1), 1, p2, False, axis=2)
    v1 = oracle.bspline_represent(b1, f1[None], 2, p1, False, axis=1)[0]
    v2 = oracle.bspline_represent(b2, f2[None], 1, p2, False, axis=1)[0]
    ref = np.multiply.outer(v1, v2)
    scale = np.abs(ref).max()
    assert np.abs(V12[0] - ref).max() <= 1e-13 * scale
    assert np.abs(V12[1] - 2 * ref).max() <= 2e-13 * scale
    assert np.abs(V12 - V21).max() <= 1e-13 * scale


def test_oracle_herm_storage_halves_the_diagonal():
    b = BSplineBasis(5, linear_breakpoints(0.0, 5.0, 8), "zero")
    N = b.number_of_bsplines
    pairs = banded_hermitian_pairs(N, 5)
    F = np.ones((1, b.grid.size))
    full = oracle.bspline_represent(b, F, 0, pairs, False, axis=1)[0]
    half = oracle.bspline_represent(b, F, 0, pairs, True, axis=1)[0]
    diag = pairs[:, 0] == pairs[:, 1]
    assert np.allclose(half[diag], 0.5 * full[diag], rtol=1e-15)
    assert np.array_equal(half[~diag], full[~diag])


@pytest.mark.parametrize("which", ["hydrogen", "helium"])
def test_grid_terms_reproduce_the_factor_potentials(which):
    """Grid first, then projection (the reference's order) against projection first, then
    the sum over terms (the separable factors): equal to rounding for every potential,
    the preconditioner's included."""
    if which == "hydrogen":
        s = hydrogen_system(lmax=3, order=5, n_breakpoints=16, rmax=20.0, n_inner=6, gamma=2.0)
        pots = s.potentials
    else:
        s = helium_system(lmax=2, L=(0, 1, 2), order=5, n_breakpoints=12, rmax=15.0, n_inner=6, gamma=2.5)
        pots = s.potentials + s.precond_potentials
    for p in pots:
        V = oracle_generate(p)
        ref = p.data.real
        assert V.shape == ref.shape
        err = np.abs(V - ref).max() / np.abs(ref).max()
        assert err < 1e-13, (p.name, err)


@pytest.mark.parametrize("d", [1, 2])
def test_oracle_reorder_sign_by_parts(d):
    """With 'zero' end conditions every spline vanishes at both ends, so integration by
    parts gives int B_r B_c^(d) = (-1)^d int B_c B_r^(d) for d = 1 and (twice) for d = 2:
    the reference's reorder with factor (-1)^d must reproduce the direct integral
    int w B_row B_col^(d) for pairs below the diagonal too."""
    b = BSplineBasis(6, linear_breakpoints(0.0, 6.0, 10) ** 1.15, "zero")
    N = b.number_of_bsplines
    V = oracle.bspline_represent(b, np.ones((1, b.grid.size)), d, dense_pairs(N), False, axis=1).reshape(N, N)
    B, D1, D2 = b.tables()
    Bd = (B, D1, D2)[d]
    direct = np.einsum("q,rq,cq->rc", b.grid_weights, B, Bd)
    assert np.abs(V - direct).max() <= 1e-11 * np.abs(direct).max()
