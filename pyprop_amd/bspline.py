"""B-spline radial basis: breakpoints, knots, Gauss-Legendre quadrature, value tables
and the quadrature projection of grid functions onto B-spline pairs.

This is host-side set-up code (it runs once, before propagation); it produces the
overlap matrix S and the radial factors of every tensor potential.

Reference behaviour restated here:
  * breakpoint sequences      pyprop/bspline/breakpointsequences.py:7-38,82-101
  * continuity / knots        pyprop/bspline/BSpline.py:100-168
  * number of splines         pyprop/bspline/BSpline.py:33  (N = #knots - k)
  * Gauss-Legendre rule       pyprop/bspline/gausslegendrequadrature.py:18-42 (Golub-Welsch)
  * global grid and weights   pyprop/bspline/BSpline.py:190-242
  * value recursion           core/transform/bspline/bspline.cpp:26-76 (eps = 1e-15 guards)
  * first/second derivative   core/transform/bspline/bspline.cpp:82-161
  * pair integral             core/transform/bspline/bspline.cpp:372-466 (i<j reorder with (-1)^d)
  * point mapping             core/transform/bspline/bspline.h:173-176  ((x-1)(b-a)/2 + b)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

EPS = 1e-15  # core/transform/bspline/bspline.h:58


# ----------------------------------------------------------------------------------
# breakpoint sequences
# ----------------------------------------------------------------------------------

def linear_breakpoints(rmin: float, rmax: float, n: int) -> np.ndarray:
    """pyprop/bspline/breakpointsequences.py:31-38."""
    return np.array([rmin + (rmax - rmin) * i / float(n - 1) for i in range(n)])


def exponential_breakpoints(rmin: float, rmax: float, n: int, gamma: float) -> np.ndarray:
    """pyprop/bspline/breakpointsequences.py:7-28."""
    h = rmax - rmin
    return np.array([rmin + h * (math.exp(gamma * float(i - 1) / float(n - 1)) - 1.0)
                     / (math.exp(gamma) - 1.0) for i in range(1, n + 1)])


def exponential_linear_breakpoints(rmin: float, rpartition: float, rmax: float,
                                   n: int, gamma: float) -> np.ndarray:
    """Exponential points on [rmin, rpartition] followed by linear points with the last
    exponential spacing up to rmax (numpy.arange semantics, so the last point can lie
    just past rmax).  pyprop/bspline/breakpointsequences.py:82-101."""
    inner = exponential_breakpoints(rmin, rpartition, n, gamma)
    h = inner[-1] - inner[-2]
    outer = np.arange(rpartition + h, rmax + h, h)
    return np.concatenate([inner, outer])


def gauss_legendre(order: int):
    """Golub-Welsch nodes/weights on [-1, 1], ascending (gausslegendrequadrature.py:18-42)."""
    if order == 1:
        return np.array([0.0]), np.array([2.0])
    beta = np.array([1.0 / (2.0 * math.sqrt(1.0 - (2.0 * n) ** -2)) for n in range(1, order)])
    J = np.diag(beta, -1) + np.diag(beta, 1)
    E, V = np.linalg.eigh(J)
    idx = np.argsort(E)
    return E[idx], 2.0 * V[0, idx] ** 2


def scale_and_translate(x, a, b):
    """Map [-1,1] -> [a,b] exactly as core/transform/bspline/bspline.h:173-176."""
    return (x - 1.0) * (b - a) / 2.0 + b


# ----------------------------------------------------------------------------------
# basis
# ----------------------------------------------------------------------------------

@dataclass
class BSplineBasis:
    """An order-k B-spline basis on a breakpoint sequence with 'zero' or 'vanilla'
    continuity at the end points (pyprop/bspline/BSpline.py:100-135)."""

    order: int
    breakpoints: np.ndarray
    continuity: str = "zero"
    quad_order_additional: int = 0

    knots: np.ndarray = field(init=False)
    top_knot_map: np.ndarray = field(init=False)
    nodes: np.ndarray = field(init=False)
    weights: np.ndarray = field(init=False)
    grid: np.ndarray = field(init=False)          # global quadrature grid (Q,)
    grid_weights: np.ndarray = field(init=False)  # (Q,)
    interval_of_point: np.ndarray = field(init=False)

    def __post_init__(self):
        k = self.order
        bp = np.asarray(self.breakpoints, dtype=np.float64)
        self.breakpoints = bp
        nb = bp.size
        cont = np.full(nb, k - 1, dtype=int)
        if self.continuity == "zero":
            cont[0] = 1
            cont[-1] = 1
        elif self.continuity == "vanilla":
            cont[0] = 0
            cont[-1] = 0
        else:
            raise ValueError("unknown continuity %r" % self.continuity)
        self.continuity_seq = cont
        knots, top = [], []
        knot_index = -1
        for i in range(nb):
            mult = k - cont[i]
            knot_index += mult
            for _ in range(mult):
                knots.append(bp[i])
                top.append(knot_index)
        self.knots = np.array(knots)
        self.top_knot_map = np.array(top, dtype=int)
        self.number_of_bsplines = self.knots.size - k

        qorder = k + self.quad_order_additional
        self.nodes, self.weights = gauss_legendre(qorder)
        pts, wts, ivl = [], [], []
        for j in range(nb - 1):
            a, b = bp[j], bp[j + 1]
            for x, w in zip(self.nodes, self.weights):
                pts.append(scale_and_translate(x, a, b))
                wts.append((b - a) / 2.0 * w)
                ivl.append(j)
        self.grid = np.array(pts)
        self.grid_weights = np.array(wts)
        self.interval_of_point = np.array(ivl, dtype=int)
        self._tables = None

    # -- evaluation -------------------------------------------------------------
    def _levels(self, x: np.ndarray):
        """B(m, i, x) for m = 1..k, every i, following the recursion of
        core/transform/bspline/bspline.cpp:26-76 with the same eps guards."""
        t = self.knots
        k = self.order
        nk = t.size
        levels = {}
        B1 = np.zeros((nk - 1, x.size))
        for i in range(nk - 1):
            B1[i] = ((t[i] <= x) & (x < t[i + 1])).astype(np.float64)
        levels[1] = B1
        prev = B1
        for m in range(2, k + 1):
            cnt = nk - m
            cur = np.zeros((cnt, x.size))
            for i in range(cnt):
                b0 = prev[i]
                b1 = prev[i + 1]
                acc = np.zeros(x.size)
                m0 = b0 > EPS
                if m0.any():
                    acc[m0] += (x[m0] - t[i]) / (t[i + m - 1] - t[i]) * b0[m0]
                m1 = b1 > EPS
                if m1.any():
                    acc[m1] += (t[i + m] - x[m1]) / (t[i + m] - t[i + 1]) * b1[m1]
                cur[i] = acc
            levels[m] = cur
            prev = cur
        return levels

    def evaluate(self, x: np.ndarray):
        """Return (B, dB, d2B), each (N, len(x)), for the k-th order splines."""
        with np.errstate(divide="ignore", invalid="ignore"):
            return self._evaluate(x)

    def _evaluate(self, x: np.ndarray):
        x = np.atleast_1d(np.asarray(x, dtype=np.float64))
        t = self.knots
        k = self.order
        N = self.number_of_bsplines
        lv = self._levels(x)
        B = lv[k][:N].copy()
        D1 = np.zeros((N, x.size))
        D2 = np.zeros((N, x.size))
        Bkm1 = lv[k - 1]
        for i in range(N):
            b1 = Bkm1[i]
            b2 = Bkm1[i + 1]
            acc = np.zeros(x.size)
            m = b1 > EPS
            if m.any():
                acc[m] += (k - 1.0) / (t[i + k - 1] - t[i]) * b1[m]
            m = b2 > EPS
            if m.any():
                acc[m] -= (k - 1.0) / (t[i + k] - t[i + 1]) * b2[m]
            D1[i] = acc
        if k >= 3:
            Bkm2 = lv[k - 2]
            for i in range(N):
                B1 = Bkm2[i].copy()
                B2 = Bkm2[i + 1].copy()
                B3 = Bkm2[i + 2].copy()
                g1, g2, g3 = B1 > EPS, B2 > EPS, B3 > EPS
                if g1.any():
                    B1[g1] *= (k - 2.0) / (t[i + k - 2] - t[i])
                if g2.any():
                    B2[g2] *= (k - 2.0) / (t[i + k - 1] - t[i + 1])
                if g3.any():
                    B3[g3] *= (k - 2.0) / (t[i + k] - t[i + 2])
                acc = np.zeros(x.size)
                m = g1 | g2
                acc[m] += (k - 1.0) / (t[i + k - 1] - t[i]) * (B1[m] - B2[m])
                m = g2 | g3
                acc[m] -= (k - 1.0) / (t[i + k] - t[i + 1]) * (B2[m] - B3[m])
                D2[i] = acc
        return B, D1, D2

    def tables(self):
        """B, B', B'' on the global quadrature grid, each (N, Q)."""
        if self._tables is None:
            self._tables = self.evaluate(self.grid)
        return self._tables

    def support_mask(self) -> np.ndarray:
        """(N, Q) bool: quadrature points inside the knot span of each spline (the
        reference tabulates each spline only on its own non-degenerate knot intervals,
        bspline.cpp:169-201)."""
        N = self.number_of_bsplines
        k = self.order
        t = self.knots
        g = self.grid
        mask = np.zeros((N, g.size), dtype=bool)
        for i in range(N):
            mask[i] = (g > t[i]) & (g < t[i + k])
        return mask

    # -- projections -------------------------------------------------------------
    def projection_matrix(self, pairs: np.ndarray, differentiation: int = 0,
                          hermitian_half_diagonal: bool = False) -> np.ndarray:
        """P[q, p] such that V[p] = sum_q P[q, p] f(q) reproduces
        BSplineGlobalOverlapIntegral(f, d, row_p, col_p) (bspline.cpp:372-466):
        (i, j) = (min, max) of (row, col); factor (-1)^d when row > col; the sum runs
        over the common support and is zero when |i - j| >= k.  With
        hermitian_half_diagonal the diagonal pairs are halved, as
        RepresentPotentialInBasisBSpline does for "Herm" storage (basis_bspline.h:31-38)."""
        B, D1, D2 = self.tables()
        Bd = (B, D1, D2)[differentiation]
        mask = self.support_mask()
        w = self.grid_weights
        pairs = np.asarray(pairs)
        rows = pairs[:, 0]
        cols = pairs[:, 1]
        i = np.minimum(rows, cols)
        j = np.maximum(rows, cols)
        factor = np.where(rows > cols, (-1.0) ** differentiation, 1.0)
        valid = (j - i) < self.order
        common = mask[i] & mask[j]
        P = (w[None, :] * B[i] * Bd[j] * common).T  # (Q, P)
        P *= (factor * valid)[None, :]
        if hermitian_half_diagonal:
            P[:, rows == cols] *= 0.5
        return P

    def overlap_band(self) -> np.ndarray:
        """Overlap S_ij = sum_q w B_i B_j (BSplineOverlapIntegral) as a dense (N, N)
        symmetric matrix (zero outside |i-j| < k)."""
        N = self.number_of_bsplines
        pairs = np.array([(a, b) for a in range(N) for b in range(N)], dtype=np.int32)
        P = self.projection_matrix(pairs, 0)
        return P.sum(axis=0).reshape(N, N)

    def project(self, f: np.ndarray, pairs: np.ndarray, differentiation: int = 0,
                axis: int = -1, hermitian_half_diagonal: bool = False) -> np.ndarray:
        """Project the grid axis `axis` of f (length Q) onto the pair list."""
        P = self.projection_matrix(pairs, differentiation, hermitian_half_diagonal)
        f = np.moveaxis(np.asarray(f), axis, -1)
        out = f @ P
        return np.moveaxis(out, -1, axis)


def make_bspline(order: int, bpstype: str, xmin: float, xmax: float, xsize: int,
                 gamma: float = 2.0, xpartition: float | None = None,
                 continuity: str = "zero", quad_order_additional: int = 0) -> BSplineBasis:
    """Build a basis from the reference's config keys (pyprop/bspline/BSpline.py:46-97)."""
    if bpstype == "linear":
        bp = linear_breakpoints(xmin, xmax, xsize)
    elif bpstype == "exponential":
        bp = exponential_breakpoints(xmin, xmax, xsize, gamma)
    elif bpstype == "exponentiallinear":
        bp = exponential_linear_breakpoints(xmin, xpartition, xmax, xsize, gamma)
    else:
        raise ValueError("unsupported bpstype %r" % bpstype)
    return BSplineBasis(order, bp, continuity, quad_order_additional)


def banded_packed_pairs(N: int, k: int) -> np.ndarray:
    """GeometryInfoBSplineBanded pairs: row i, col j in [max(0,i-k+1), min(i+k,N))
    (pyprop/tensorpotential/basis/BSpline.py:30-48)."""
    out = []
    for i in range(N):
        for j in range(max(0, i - k + 1), min(i + k, N)):
            out.append((i, j))
    return np.array(out, dtype=np.int32)


def banded_hermitian_pairs(N: int, k: int) -> np.ndarray:
    """GeometryInfoBSplineBandedBlas pairs ("Band", zhbmv layout): index i*k + j holds
    (i, i+j), padding (0, N-1) beyond the edge (basis/BSpline.py:80-102)."""
    out = np.zeros((N * k, 2), dtype=np.int32)
    out[:, 1] = N - 1
    idx = 0
    for i in range(N):
        for j in range(k):
            if i + j < N:
                out[idx] = (i, i + j)
            else:
                out[idx] = (0, N - 1)
            idx += 1
    return out


def banded_nonhermitian_pairs(N: int, bands: int) -> np.ndarray:
    """GeometryInfoCommonBandedNonHermitian pairs ("BandNH", zgbmv layout): index
    j*(2b+1) + b + i - j holds (i, j) for |i-j| <= b, padding (0, N-1)
    (pyprop/tensorpotential/GeometryInfo.py:432-454)."""
    out = np.zeros((N * (2 * bands + 1), 2), dtype=np.int32)
    out[:, 1] = N - 1
    for i in range(N):
        for j in range(i - bands, i + bands + 1):
            if 0 <= j < N:
                out[j * (2 * bands + 1) + bands + i - j] = (i, j)
    return out
