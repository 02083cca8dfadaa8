"""Angular bases: coupled spherical harmonics |l1 l2 L M> (two electrons) and reduced
spherical harmonics |l m> (one electron), their channel enumeration, selection rules and
the angular matrix elements that multiply the radial tensors.

Reference behaviour restated here:
  * Clebsch-Gordan from the 3j symbol   core/representation/coupledspherical/coupledrange.h:17-32
  * channel enumeration                 pyprop/CoupledSphericalHarmonics.py:47-78
  * selection rules                     core/tensorpotential/coupledsphericalselectionrule.h:16-153,333-367
  * 1/r12 multipole angular factor      examples/tensor/helium_stabilization/coupledspherical.cpp:43-93,116-119
  * angular kinetic energy              examples/tensor/helium_stabilization/coupledspherical.cpp:150-210
  * length-gauge dipole factors I1, I2  examples/tensor/helium_stabilization/coupledlength.cpp:45-122
  * reduced-spherical dipole pairs      pyprop/tensorpotential/basis/ReducedSphericalHarmonic.py:20-38

The 3j symbols use the Racah closed form with exact integer factorials; GSL
(gsl_sf_coupling_3j_e), which the reference links, is not present, so agreement is to
floating-point rounding of the same closed form.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from fractions import Fraction
from functools import lru_cache

import numpy as np


@lru_cache(maxsize=None)
def wigner3j(j1: int, j2: int, j3: int, m1: int, m2: int, m3: int) -> float:
    """(j1 j2 j3; m1 m2 m3) for integer arguments via the Racah formula.  Returns 0 for
    violated triangle or projection rules (as GSL does)."""
    if m1 + m2 + m3 != 0:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    if j1 < 0 or j2 < 0 or j3 < 0:
        return 0.0
    f = math.factorial
    tri = Fraction(f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3), f(j1 + j2 + j3 + 1))
    pre = tri * f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j3 + m3) * f(j3 - m3)
    kmin = max(0, j2 - j3 - m1, j1 - j3 + m2)
    kmax = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    s = 0
    for k in range(kmin, kmax + 1):
        den = (f(k) * f(j1 + j2 - j3 - k) * f(j1 - m1 - k) * f(j2 + m2 - k)
               * f(j3 - j2 + m1 + k) * f(j3 - j1 - m2 + k))
        s += Fraction((-1) ** k, den)
    sign = -1 if (j1 - j2 - m3) % 2 else 1
    # sqrt of an exact rational times an exact rational sum
    val = float(s) * math.sqrt(float(pre))
    if s != 0:
        # better accuracy: sqrt(pre) * s computed as sqrt(pre * s^2) with sign
        val = math.copysign(math.sqrt(float(pre * s * s)), float(s))
    return sign * val


def clebsch_gordan(l1: int, l2: int, m1: int, m2: int, L: int, M: int) -> float:
    """<l1 m1 l2 m2 | L M> = (-1)^(M + l1 - l2) sqrt(2L+1) (l1 l2 L; m1 m2 -M)
    (coupledrange.h:22-31)."""
    v = wigner3j(l1, l2, L, m1, m2, -M)
    if v == 0.0:
        return 0.0
    sign = -1.0 if (M + l1 - l2) % 2 else 1.0
    return sign * math.sqrt(2 * L + 1.0) * v


@dataclass(frozen=True)
class CoupledIndex:
    l1: int
    l2: int
    L: int
    M: int


def default_coupled_index_iterator(lmax: int, L=None, M=(0,), full_tensor: bool = False):
    """DefaultCoupledIndexIterator.__iter__ (pyprop/CoupledSphericalHarmonics.py:47-78)."""
    if L is None:
        L = list(range(lmax + 1))
    if not hasattr(L, "__iter__"):
        L = [L]
    if not hasattr(M, "__iter__"):
        M = [M]
    Ls = sorted(set(int(x) for x in L))
    Ms = sorted(set(int(x) for x in M))
    out = []
    for curM in Ms:
        for curL in Ls:
            if abs(curM) > curL and not full_tensor:
                continue
            for l1 in range(lmax + 1):
                for l2 in range(lmax + 1):
                    if (abs(l1 - l2) <= curL <= l1 + l2) or full_tensor:
                        out.append(CoupledIndex(l1, l2, curL, curM))
    return out


# ---------------------------------------------------------------------------------
# selection rules (row-major over (left, right), coupledsphericalselectionrule.h:16-53)
# ---------------------------------------------------------------------------------

def _r12_sum(left: CoupledIndex, right: CoupledIndex, l3: int) -> float:
    """Inner m-sum of the multipole factor (coupledspherical.cpp:75-92)."""
    L, M = left.L, left.M
    l1p, l2p = left.l1, left.l2
    l1, l2 = right.l1, right.l2
    lstop = max(l1, l1p, l2, l2p)
    s = 0.0
    for m1p in range(-lstop, lstop + 1):
        m2p = M - m1p
        for m1 in range(-lstop, lstop + 1):
            m2 = M - m1
            m3 = m1 - m1p
            cur = clebsch_gordan(l1p, l2p, m1p, m2p, L, M)
            if cur == 0.0:
                continue
            cur *= clebsch_gordan(l1, l2, m1, m2, L, M)
            if cur == 0.0:
                continue
            cur *= clebsch_gordan(l1p, l1, -m1p, m1, l3, m3)
            cur *= clebsch_gordan(l2p, l2, -m2p, m2, l3, -m3)
            cur *= (-1.0) ** (m1p + m2p + m3)
            s += cur
    return s


def selection_rule_r12(left: CoupledIndex, right: CoupledIndex, cutoff: int | None) -> bool:
    """cutoff None -> CoupledSphericalSelectionRuleR12Old (selection rule 'r12');
    otherwise CoupledSphericalSelectionRuleR12(cutoff) ('r12_<cutoff>')."""
    nonzero = (left.L == right.L and left.M == right.M
               and left.L <= left.l1 + left.l2 and abs(left.l1 - left.l2) <= left.L
               and right.L <= right.l1 + right.l2 and abs(right.l1 - right.l2) <= right.L)
    if cutoff is None or not nonzero:
        return nonzero
    l1p, l2p, l1, l2 = left.l1, left.l2, right.l1, right.l2
    lo = max(abs(l1 - l1p), abs(l2 - l2p))
    hi = min(l1 + l1p, l2 + l2p, cutoff)
    total = 0.0
    for l3 in range(lo, hi + 1):
        c = clebsch_gordan(l1p, l1, 0, 0, l3, 0) * clebsch_gordan(l2p, l2, 0, 0, l3, 0)
        total += _r12_sum(left, right, l3) * c
    return abs(total) > 1e-14


def selection_rule_linear_polarized(left: CoupledIndex, right: CoupledIndex) -> bool:
    """coupledsphericalselectionrule.h:333-355."""
    if abs(left.L - right.L) == 1 and left.M == right.M:
        if abs(left.l1 - right.l1) == 1 and left.l2 == right.l2:
            return True
        if abs(left.l2 - right.l2) == 1 and left.l1 == right.l1:
            return True
    return False


def selection_rule_diagonal(left: CoupledIndex, right: CoupledIndex) -> bool:
    """coupledsphericalselectionrule.h:357-367."""
    return left == right


def coupled_basis_pairs(channels, rule: str) -> np.ndarray:
    """Angular pair list for a geometry name ("diagonal", "selectionrule_r12",
    "selectionrule_r12_<n>", "selectionrule_linearpolarizedfield") as
    basis/CoupledSphericalHarmonic.py:133-162 resolves it."""
    rule = rule.lower().strip()
    if rule == "diagonal":
        f = selection_rule_diagonal
    elif rule == "selectionrule_r12":
        f = lambda a, b: selection_rule_r12(a, b, None)  # noqa: E731
    elif rule.startswith("selectionrule_r12_"):
        n = int(rule[len("selectionrule_r12_"):])
        f = lambda a, b: selection_rule_r12(a, b, n)  # noqa: E731
    elif rule == "selectionrule_linearpolarizedfield":
        f = selection_rule_linear_polarized
    else:
        raise ValueError("unsupported coupled geometry %r" % rule)
    out = []
    for li, left in enumerate(channels):
        for ri, right in enumerate(channels):
            if f(left, right):
                out.append((li, ri))
    return np.array(out, dtype=np.int32).reshape(-1, 2)


# ---------------------------------------------------------------------------------
# angular matrix elements
# ---------------------------------------------------------------------------------

def r12_multipole_coefficients(channels, pairs) -> dict:
    """A[lambda][pair] such that the 1/r12 potential on the radial grid is
    sum_lambda A[lambda][pair] * r<^lambda / r>^(lambda+1)
    (coupledspherical.cpp:43-111; lambda runs without cutoff, 63-64)."""
    out: dict = {}
    for a, (li, ri) in enumerate(pairs):
        left = channels[li]
        right = channels[ri]
        if left.L != right.L or left.M != right.M:
            continue
        l1p, l2p, l1, l2 = left.l1, left.l2, right.l1, right.l2
        lo = max(abs(l1 - l1p), abs(l2 - l2p))
        hi = min(l1 + l1p, l2 + l2p)
        for l3 in range(lo, hi + 1):
            coef = math.sqrt((2.0 * l1p + 1.0) * (2.0 * l1 + 1.0) / (2.0 * l3 + 1.0))
            coef *= math.sqrt((2.0 * l2p + 1.0) * (2.0 * l2 + 1.0) / (2.0 * l3 + 1.0))
            coef *= 1.0 / (2 * l3 + 1.0)
            coef *= clebsch_gordan(l1p, l1, 0, 0, l3, 0)
            coef *= clebsch_gordan(l2p, l2, 0, 0, l3, 0)
            s = _r12_sum(left, right, l3) * coef
            arr = out.setdefault(l3, np.zeros(len(pairs)))
            arr[a] += s
    return out


def dipole_length_coefficients(channels, pairs):
    """(I1, I2) per pair so that the z-dipole potential is I1*r1 + I2*r2
    (coupledlength.cpp:55-107)."""
    def coeff(l, lp):
        return math.sqrt((2 * l + 1.0) * (2 * lp + 1.0) / (12 * math.pi))

    def cs(m):
        return 1.0 if m < 0 else (-1.0) ** m

    I1 = np.zeros(len(pairs))
    I2 = np.zeros(len(pairs))
    for a, (li, ri) in enumerate(pairs):
        left = channels[li]
        right = channels[ri]
        if abs(left.L - right.L) != 1 or left.M != right.M:
            continue
        L, M, l1, l2 = left.L, left.M, left.l1, left.l2
        Lp, l1p, l2p = right.L, right.l1, right.l2
        lstop = max(l1, l1p, l2, l2p)
        s1 = s2 = 0.0
        for m1 in range(-lstop, lstop + 1):
            m2 = M - m1
            m1p, m2p = m1, m2
            if abs(m1) > l1 or abs(m1p) > l1p or abs(m2) > l2 or abs(m2p) > l2p:
                continue
            base = clebsch_gordan(l1, l2, m1, m2, L, M) * clebsch_gordan(l1p, l2p, m1p, m2p, Lp, M)
            c1 = base * coeff(l1, l1p)
            c1 *= clebsch_gordan(l1, l1p, 0, 0, 1, 0) * clebsch_gordan(l1, l1p, -m1, m1p, 1, 0)
            c1 /= cs(m1)
            s1 += c1
            c2 = base * coeff(l2, l2p)
            c2 *= clebsch_gordan(l2, l2p, 0, 0, 1, 0) * clebsch_gordan(l2, l2p, -m2, m2p, 1, 0)
            c2 /= cs(m2)
            s2 += c2
        I1[a] = s1
        I2[a] = s2
    return I1, I2


# ---------------------------------------------------------------------------------
# reduced spherical harmonics |l m>, m fixed
# ---------------------------------------------------------------------------------

def reduced_dipole_pairs(lmax: int) -> np.ndarray:
    """'dipoleselectionrule' pairs: for l = 0..lmax, (l, l-1) then (l, l+1)
    (ReducedSphericalHarmonic.py:23-38)."""
    count = lmax + 1
    out = []
    for i in range(count):
        if i > 0:
            out.append((i, i - 1))
        if i < count - 1:
            out.append((i, i + 1))
    return np.array(out, dtype=np.int32)


def reduced_cos_theta(l: int, lp: int, m: int = 0) -> float:
    """<l m | cos(theta) | l' m> for normalised spherical harmonics: the closed form
    that the reference's Legendre quadrature projection (basis_reducedspherical.h)
    evaluates exactly."""
    if lp == l + 1:
        return math.sqrt(((l + 1.0) ** 2 - m * m) / ((2 * l + 1.0) * (2 * l + 3.0)))
    if lp == l - 1:
        return math.sqrt((l * l - m * m * 1.0) / ((2 * l - 1.0) * (2 * l + 1.0)))
    return 0.0
