"""Builders for the propagation workloads: the hydrogen atom in reduced spherical
harmonics x B-splines (rank 2) and helium in coupled spherical harmonics x B-splines x
B-splines (rank 3), with their tensor potentials in the reference's storage layout.

Host-side set-up.  Mirrors what TensorPotentialGenerator.GeneratePotential
(pyprop/tensorpotential/TensorPotentialGenerator.py:80-274) produces for the potential
classes of examples/tensor/helium_stabilization/{potential.cpp,coupledspherical.cpp,
coupledlength.cpp} and examples/tensor/reducedspherical/potential.cpp, followed by
BasisPropagator.ConsolidatePotentials (pyprop/tensorpotential/BasisPropagator.py:68-107).

Each potential is a sum of terms a_t(angular pair) x R_t(radial pairs), where R_t is the
B-spline projection of a grid function g_t(r1[, r2]) with the potential's per-rank
differentiation.  The projection follows BSplineGlobalOverlapIntegral; the reference
sums the terms on the quadrature grid before projecting, which agrees to rounding.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import angular
from .bspline import BSplineBasis, exponential_linear_breakpoints
from .geometry import Geometry, bspline_geometry, coupled_geometry, reduced_spherical_geometry


# ---------------------------------------------------------------------------------
# time functions
# ---------------------------------------------------------------------------------

@dataclass(frozen=True)
class LaserSimpleLength:
    """E0 sin^2(pi t / T) cos(omega t) on [0, T), else 0, with T = cycles*2*pi/omega
    (LaserFunctionSimpleLength, examples/tensor/reducedspherical/example.py:413-420)."""

    amplitude: float
    frequency: float
    cycles: float

    @property
    def pulse_duration(self) -> float:
        return self.cycles * 2.0 * math.pi / self.frequency

    def __call__(self, t: float) -> float:
        T = self.pulse_duration
        if 0 <= t < T:
            return self.amplitude * math.sin(t * math.pi / T) ** 2 * math.cos(t * self.frequency)
        return 0.0


# ---------------------------------------------------------------------------------
# containers
# ---------------------------------------------------------------------------------

@dataclass
class GridTerm:
    """One grid function of a potential, as GeneratePotential evaluates it on the
    quadrature grid before projecting (TensorPotentialGenerator.py:80-274), with the
    differentiation of each radial rank (the potential's `differentiation<r>` keys).

    kind "separable": F[a][q1][q2] = sum_t coeffs[t, a] factors[0][t, q1] factors[1][t, q2]
    (one radial rank: F[a][q] = sum_t coeffs[t, a] factors[0][t, q]);
    kind "multipole": F[a][q1][q2] = sum_t coeffs[t, a] r<^lams[t] / r>^(lams[t] + 1)."""

    kind: str
    coeffs: np.ndarray                 # (T, P0) real
    differentiation: tuple             # per radial rank
    factors: Optional[list] = None     # separable: per radial rank, (T, Q_r)
    lams: Optional[list] = None        # multipole: the lambda of each term


@dataclass
class TensorPotentialData:
    """Potential data in the reference storage layout plus its separable factors.

    The complex128 data array (shape = pair count of every rank) is expanded from the
    factors on first access: data[a, ...] = sum_t angular_coeffs[t, a] radial_factors[t, ...].
    """

    name: str
    geometries: list
    _data: Optional[np.ndarray] = None
    time_function: Optional[Callable] = None
    angular_coeffs: Optional[np.ndarray] = None   # (T, P0) real
    radial_factors: Optional[np.ndarray] = None   # (T, P1[, P2]) real
    grid_terms: Optional[list] = None             # GridTerm list: the same potential on the grid
    filename: Optional[str] = None                # load the data from this HDF5 file instead
    dataset: Optional[str] = None                 # (TensorPotential.py:56-71, group path in the file)

    @property
    def data(self) -> np.ndarray:
        if self._data is None:
            self._data = _expand(self.angular_coeffs, self.radial_factors)
        return self._data

    @property
    def shape(self):
        return tuple(g.pair_count for g in self.geometries)

    @property
    def storage_ids(self):
        return [g.storage_id for g in self.geometries]

    @property
    def pairs(self):
        return [g.pairs for g in self.geometries]

    @property
    def is_time_dependent(self) -> bool:
        return self.time_function is not None


@dataclass
class System:
    name: str
    shape: tuple
    bases: list                      # per rank: BSplineBasis or None (orthogonal)
    potentials: list                 # consolidated TensorPotentialData
    pulse: Optional[LaserSimpleLength] = None
    channels: list = field(default_factory=list)
    precond_potentials: list = field(default_factory=list)  # angular-diagonal H parts
    info: dict = field(default_factory=dict)

    @property
    def rank(self) -> int:
        return len(self.shape)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    def overlap(self, r: int) -> Optional[np.ndarray]:
        b = self.bases[r]
        return None if b is None else b.overlap_band()


# ---------------------------------------------------------------------------------
# radial projections
# ---------------------------------------------------------------------------------

def _proj(basis: BSplineBasis, geom: Geometry, d: int) -> np.ndarray:
    return basis.projection_matrix(geom.pairs, d, geom.storage_id == "Herm")


def radial_1d(basis, geom, g_values, d=0) -> np.ndarray:
    """sum_q P[q,p] g(q) for one radial rank."""
    return g_values @ _proj(basis, geom, d)


def radial_2d(b1, g1, b2, g2, G, d1=0, d2=0) -> np.ndarray:
    """P1^T G P2 for a grid function G(q1, q2)."""
    P1 = _proj(b1, g1, d1)
    P2 = _proj(b2, g2, d2)
    return P1.T @ (G @ P2)


def _expand(coeffs: np.ndarray, radial: np.ndarray) -> np.ndarray:
    """V[a, ...] = sum_t coeffs[t, a] radial[t, ...] (real), returned complex128."""
    T, P0 = coeffs.shape
    flat = np.ascontiguousarray(radial.reshape(T, -1))
    out = np.zeros((P0, flat.shape[1]), dtype=np.complex128)
    re = out.view(np.float64)[:, 0::2]
    chunk = 16
    for a0 in range(0, P0, chunk):
        re[a0:a0 + chunk] = coeffs[:, a0:a0 + chunk].T @ flat
    return out.reshape((P0,) + radial.shape[1:])


# ---------------------------------------------------------------------------------
# breakpoint helper
# ---------------------------------------------------------------------------------

def exponential_linear_for_count(rmin: float, rmax: float, n_inner: int, n_total: int,
                                 gamma: float) -> tuple:
    """Choose xpartition so that ExponentialLinearBreakpointSequence gives n_total
    breakpoints with the last one at rmax + 0.01 h (h = linear spacing)."""
    m = n_total - n_inner
    c = (math.exp(gamma) - math.exp(gamma * (n_inner - 2.0) / (n_inner - 1.0))) / (math.exp(gamma) - 1.0)
    rp = (rmax - rmin) / (1.0 + (m - 0.01) * c) + rmin
    bp = exponential_linear_breakpoints(rmin, rp, rmax, n_inner, gamma)
    if bp.size != n_total:
        raise RuntimeError("breakpoint count %d != %d" % (bp.size, n_total))
    return bp, rp


# ---------------------------------------------------------------------------------
# hydrogen, rank 2: (l, r)
# ---------------------------------------------------------------------------------

def hydrogen_system(lmax: int = 7, order: int = 7, n_breakpoints: int = 117,
                    rmax: float = 80.0, n_inner: int = 20, gamma: float = 4.0,
                    Z: float = 1.0, pulse: Optional[LaserSimpleLength] = None,
                    radial_storage: str = "banded-nonhermitian") -> System:
    """H atom: H = -1/2 d2/dr2 + l(l+1)/2r^2 - Z/r (angular 'Diagonal', storage Diag)
    plus the length-gauge z-dipole r cos(theta) E(t) (angular 'dipoleselectionrule',
    storage Simp), both over one B-spline radial rank (examples/tensor/reducedspherical
    config_radial.ini + potential.cpp:39-72,183-216,252-283)."""
    if pulse is None:
        pulse = LaserSimpleLength(0.05, 0.057, 4.0)
    bp, rp = exponential_linear_for_count(0.0, rmax, n_inner, n_breakpoints, gamma)
    basis = BSplineBasis(order, bp, "zero", 0)
    N = basis.number_of_bsplines
    nl = lmax + 1
    r = basis.grid
    gr = bspline_geometry(basis, radial_storage)

    gdiag = reduced_spherical_geometry(lmax, "diagonal")
    ls = np.arange(nl, dtype=np.float64)
    # three angular-diagonal terms: kinetic (d=2), centrifugal, Coulomb
    R_kin = radial_1d(basis, gr, np.full(r.size, -0.5), 2)
    R_inv2 = radial_1d(basis, gr, 1.0 / (2.0 * r * r), 0)
    R_coul = radial_1d(basis, gr, -Z / r, 0)
    coeffs = np.stack([np.ones(nl), ls * (ls + 1.0), np.ones(nl)])
    radial = np.stack([R_kin, R_inv2, R_coul])
    H0 = TensorPotentialData("RadialKineticEnergy+AngularKineticEnergy+CoulombPotential",
                             [gdiag, gr], None, None, coeffs, radial,
                             [GridTerm("separable", np.ones((1, nl)), (2,), [np.full((1, r.size), -0.5)]),
                              GridTerm("separable", coeffs[1:], (0,), [np.stack([1.0 / (2.0 * r * r), -Z / r])])])

    gdip = reduced_spherical_geometry(lmax, "dipoleselectionrule")
    cos_el = np.array([angular.reduced_cos_theta(int(a), int(b)) for a, b in gdip.pairs])
    R_r = radial_1d(basis, gr, r.copy(), 0)
    dip = TensorPotentialData("LaserPotentialLength", [gdip, gr], None, pulse,
                              cos_el[None, :], R_r[None, :],
                              [GridTerm("separable", cos_el[None, :], (0,), [r[None, :].copy()])])
    sysm = System("hydrogen", (nl, N), [None, basis], [H0, dip], pulse)
    sysm.info = dict(lmax=lmax, order=order, n_breakpoints=n_breakpoints, rmax=float(bp[-1]),
                     xpartition=rp, gamma=gamma, n_inner=n_inner, Z=Z)
    return sysm


# ---------------------------------------------------------------------------------
# helium, rank 3: (coupled channel, r1, r2)
# ---------------------------------------------------------------------------------

def helium_system(lmax: int = 5, L=(0, 1, 2), M: int = 0, order: int = 7,
                  n_breakpoints: int = 57, rmax: float = 40.0, n_inner: int = 20,
                  gamma: float = 4.0, Z: float = 2.0,
                  pulse: Optional[LaserSimpleLength] = None,
                  r12_rule: str = "selectionrule_r12_10",
                  radial_storage=("banded-packed", "banded-nonhermitian")) -> System:
    """Helium: H0 = T1 + T2 + l1(l1+1)/2r1^2 + l2(l2+1)/2r2^2 - Z/r1 - Z/r2 (angular
    'Diagonal'), the 1/r12 multipole expansion (angular 'SelectionRule_R12_<n>') and the
    length-gauge dipole (I1 r1 + I2 r2) E(t) ('SelectionRule_LinearPolarizedField'), all
    over (banded-packed r1) x (banded-nonhermitian r2) B-spline storage."""
    if pulse is None:
        pulse = LaserSimpleLength(0.05, 0.057, 4.0)
    bp, rp = exponential_linear_for_count(0.0, rmax, n_inner, n_breakpoints, gamma)
    basis = BSplineBasis(order, bp, "zero", 0)
    N = basis.number_of_bsplines
    channels = angular.default_coupled_index_iterator(lmax, L=list(L), M=[M])
    nc = len(channels)
    r = basis.grid
    g1 = bspline_geometry(basis, radial_storage[0])
    g2 = bspline_geometry(basis, radial_storage[1])

    def sep(f1, d1, f2, d2, ga=g1, gb=g2):
        a = radial_1d(basis, ga, f1, d1)
        b = radial_1d(basis, gb, f2, d2)
        return np.multiply.outer(a, b)

    one = np.ones(r.size)
    inv2 = 1.0 / (2.0 * r * r)
    # angular-diagonal part
    gdiag = coupled_geometry(channels, "diagonal")
    l1s = np.array([channels[a].l1 for a, _ in gdiag.pairs], dtype=np.float64)
    l2s = np.array([channels[a].l2 for a, _ in gdiag.pairs], dtype=np.float64)
    nd = gdiag.pair_count

    def h0_terms(ga, gb):
        radial = np.stack([
            sep(-0.5 * one, 2, one, 0, ga, gb),          # KineticEnergy1 (d1 = 2)
            sep(one, 0, -0.5 * one, 2, ga, gb),          # KineticEnergy2 (d2 = 2)
            sep(inv2, 0, one, 0, ga, gb),                # l1(l1+1)/2r1^2
            sep(one, 0, inv2, 0, ga, gb),                # l2(l2+1)/2r2^2
            sep(-Z / r, 0, one, 0, ga, gb) + sep(one, 0, -Z / r, 0, ga, gb),  # Coulomb
        ])
        coeffs = np.stack([np.ones(nd), np.ones(nd), l1s * (l1s + 1), l2s * (l2s + 1), np.ones(nd)])
        return coeffs, radial

    def h0_grid(nd_, l1v, l2v):
        c1 = np.ones((1, nd_))
        return [GridTerm("separable", c1, (2, 0), [-0.5 * one[None], one[None]]),
                GridTerm("separable", c1, (0, 2), [one[None], -0.5 * one[None]]),
                GridTerm("separable", np.stack([l1v * (l1v + 1), l2v * (l2v + 1), np.ones(nd_), np.ones(nd_)]),
                         (0, 0), [np.stack([inv2, one, -Z / r, one]), np.stack([one, inv2, one, -Z / r])])]

    c0, r0 = h0_terms(g1, g2)
    H0 = TensorPotentialData("RadialKineticEnergy1+RadialKineticEnergy2+AngularKineticEnergy+CoulombPotential",
                             [gdiag, g1, g2], None, None, c0, r0, h0_grid(nd, l1s, l2s))

    # 1/r12 multipoles: grid functions f_l(q1,q2) = r<^l / r>^(l+1)
    rmin_ = np.minimum.outer(r, r)
    rmax_ = np.maximum.outer(r, r)
    frac = rmin_ / rmax_

    def multipole_radial(ga, gb, lams):
        P1 = _proj(basis, ga, 0)
        P2 = _proj(basis, gb, 0)
        out = []
        for lam in lams:
            G = frac ** lam / rmax_
            out.append(P1.T @ (G @ P2))
        return np.stack(out)

    gr12 = coupled_geometry(channels, r12_rule)
    A = angular.r12_multipole_coefficients(channels, gr12.pairs)
    lams = sorted(A.keys())
    cR = np.stack([A[l] for l in lams])
    rR = multipole_radial(g1, g2, lams)
    R12 = TensorPotentialData("ElectronicCouplingPotential", [gr12, g1, g2],
                              None, None, cR, rR, [GridTerm("multipole", cR, (0, 0), lams=list(lams))])

    gdip = coupled_geometry(channels, "selectionrule_linearpolarizedfield")
    I1, I2 = angular.dipole_length_coefficients(channels, gdip.pairs)
    cD = np.stack([I1, I2])
    rD = np.stack([sep(r, 0, one, 0), sep(one, 0, r, 0)])
    dip = TensorPotentialData("DipolePotentialLength", [gdip, g1, g2],
                              None, pulse, cD, rD,
                              [GridTerm("separable", cD, (0, 0), [np.stack([r, one]), np.stack([one, r])])])

    # preconditioner potentials: angular-diagonal H0 + monopole (diagonal) part of 1/r12,
    # both as banded-packed x banded-packed (examples/tensor/helium_stabilization/config.ini:57-60,101-108)
    gp = bspline_geometry(basis, "banded-packed")
    cP0, rP0 = h0_terms(gp, gp)
    Amono = angular.r12_multipole_coefficients(channels, gdiag.pairs)
    lm = sorted(Amono.keys())
    cPm = np.stack([Amono[l] for l in lm])
    rPm = multipole_radial(gp, gp, lm)
    prec = [TensorPotentialData("PreconditionerH0", [gdiag, gp, gp], None, None, cP0, rP0, h0_grid(nd, l1s, l2s)),
            TensorPotentialData("ElectronicCouplingPotentialMonopoleTerm", [gdiag, gp, gp], None, None, cPm, rPm,
                                [GridTerm("multipole", cPm, (0, 0), lams=list(lm))])]

    sysm = System("helium", (nc, N, N), [None, basis, basis], [H0, R12, dip], pulse, channels, prec)
    sysm.info = dict(lmax=lmax, L=list(L), M=M, order=order, n_breakpoints=n_breakpoints,
                     rmax=float(bp[-1]), xpartition=rp, gamma=gamma, n_inner=n_inner, Z=Z,
                     r12_rule=r12_rule, multipoles=lams)
    return sysm


def _restrict_potential(p: TensorPotentialData, keep_map: dict, n_sub: int):
    """p with its rank-0 (angular) pairs restricted to the channels in keep_map
    (old index -> new index) and renumbered; None when no pair is left."""
    g0 = p.geometries[0]
    if g0.storage_id not in ("Simp", "Diag"):
        raise ValueError("channel restriction needs a Simp or Diag angular rank, not %s" % g0.storage_id)
    rows, cols = g0.pairs[:, 0], g0.pairs[:, 1]
    sel = np.array([int(r) in keep_map and int(c) in keep_map for r, c in zip(rows, cols)], dtype=bool)
    if not sel.any():
        return None
    new_pairs = np.array([(keep_map[int(r)], keep_map[int(c)]) for r, c in g0.pairs[sel]], dtype=np.int32).reshape(-1, 2)
    if g0.storage_id == "Diag":
        if not np.array_equal(new_pairs[:, 0], np.arange(n_sub)):
            raise ValueError("a Diag rank must keep every channel of the subset")
    g = Geometry(g0.storage_id, new_pairs, g0.use_grid, n_sub, g0.basis)
    terms = None
    if p.grid_terms:
        terms = [GridTerm(t.kind, np.ascontiguousarray(t.coeffs[:, sel]), t.differentiation, t.factors, t.lams)
                 for t in p.grid_terms]
    q = TensorPotentialData(p.name, [g] + list(p.geometries[1:]), None, p.time_function,
                            None if p.angular_coeffs is None else np.ascontiguousarray(p.angular_coeffs[:, sel]),
                            p.radial_factors, terms)
    if p.angular_coeffs is None:
        q._data = np.ascontiguousarray(p.data[sel])
    return q


def restrict_channels(system: System, channels, field_free: bool = True) -> System:
    """The same Hamiltonian on a subset of the rank-0 channels (for helium one total-L
    block): every potential keeps the angular pairs whose row and column both lie in the
    subset, renumbered in subset order; the radial data are shared, not recomputed.  With
    field_free the time-dependent potentials are left out (the field-free Hamiltonian whose
    bound states define ionisation, examples/tensor/helium_stabilization/example.py:567-570)."""
    channels = [int(c) for c in channels]
    if len(set(channels)) != len(channels) or not channels:
        raise ValueError("channels must be distinct and non-empty")
    keep_map = {c: i for i, c in enumerate(channels)}
    n_sub = len(channels)
    pots = []
    for p in system.potentials:
        if field_free and p.is_time_dependent:
            continue
        q = _restrict_potential(p, keep_map, n_sub)
        if q is not None:
            pots.append(q)
    prec = [q for q in (_restrict_potential(p, keep_map, n_sub) for p in system.precond_potentials) if q is not None]
    sub_channels = [system.channels[c] for c in channels] if system.channels else []
    info = dict(system.info)
    info["parent_channels"] = channels
    return System(system.name, (n_sub,) + tuple(system.shape[1:]), system.bases, pots,
                  None if field_free else system.pulse, sub_channels, prec, info)


def hydrogenic_1s(basis: BSplineBasis, Z: float, l: int = 0) -> np.ndarray:
    """Lowest generalized eigenvector of the 1-D radial Hamiltonian
    -1/2 d2/dr2 + l(l+1)/2r^2 - Z/r in the B-spline basis, S-normalised."""
    import scipy.linalg
    from .geometry import dense_pairs
    N = basis.number_of_bsplines
    g = Geometry("Simp", dense_pairs(N), True, N)
    r = basis.grid
    T = radial_1d(basis, g, np.full(r.size, -0.5), 2).reshape(N, N)
    V = radial_1d(basis, g, l * (l + 1) / (2 * r * r) - Z / r, 0).reshape(N, N)
    S = basis.overlap_band()
    H = T + V
    H = 0.5 * (H + H.T)
    w, v = scipy.linalg.eigh(H, S)
    c = v[:, 0]
    c /= math.sqrt(c @ S @ c)
    return c, w[0]


def helium_initial_state(system: System) -> np.ndarray:
    """Independent-particle 1s^2 state in the (l1=0, l2=0, L=0) channel, built from the
    He+ 1s radial eigenvector; S-normalised."""
    basis = system.bases[1]
    c, _ = hydrogenic_1s(basis, system.info["Z"])
    psi = np.zeros(system.shape, dtype=np.complex128)
    idx = [i for i, ch in enumerate(system.channels) if ch.l1 == 0 and ch.l2 == 0 and ch.L == 0]
    psi[idx[0]] = np.multiply.outer(c, c)
    return psi


def helium_smooth_start(system: System, seed: int = 17, amplitude: float = 0.05) -> np.ndarray:
    """The 1s^2 state plus, in every channel (l1, l2, L), a seeded complex multiple
    (`amplitude` times a standard complex normal) of the product of the lowest l1 state of
    charge 2 and the lowest l2 state of charge 1: a smooth start that populates all
    channels (unnormalised)."""
    rng = np.random.default_rng(seed)
    basis = system.bases[1]
    psi = helium_initial_state(system)
    lmax = max(max(ch.l1, ch.l2) for ch in system.channels)
    inner = [hydrogenic_1s(basis, 2.0, l)[0] for l in range(lmax + 1)]
    outer = [hydrogenic_1s(basis, 1.0, l)[0] for l in range(lmax + 1)]
    for c, ch in enumerate(system.channels):
        a = amplitude * complex(rng.standard_normal(), rng.standard_normal())
        psi[c] += a * np.multiply.outer(inner[ch.l1], outer[ch.l2])
    return psi
