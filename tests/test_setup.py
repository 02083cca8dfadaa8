"""Host set-up pins: B-spline basis, quadrature, angular couplings and the potentials they
generate, checked against closed forms and textbook results."""
import math

import numpy as np
import pytest
import scipy.linalg

from pyprop_amd import angular
from pyprop_amd.bspline import (BSplineBasis, banded_nonhermitian_pairs, banded_packed_pairs, gauss_legendre,
                                linear_breakpoints)
from pyprop_amd.geometry import dense_pairs, Geometry
from pyprop_amd.systems import (LaserSimpleLength, exponential_linear_for_count, helium_system, hydrogen_system,
                                radial_1d)


def test_gauss_legendre_vs_numpy():
    for q in (2, 5, 7, 12):
        x, w = gauss_legendre(q)
        xr, wr = np.polynomial.legendre.leggauss(q)
        assert np.allclose(x, xr, atol=1e-14)
        assert np.allclose(w, wr, atol=1e-14)


def test_knots_and_counts():
    b = BSplineBasis(7, linear_breakpoints(0, 10, 12), "zero")
    assert b.number_of_bsplines == 12 + 7 - 4  # N = n_bp + k - 4 for 'zero'
    assert np.sum(b.knots == 0.0) == 6 and np.sum(b.knots == 10.0) == 6
    bv = BSplineBasis(7, linear_breakpoints(0, 10, 12), "vanilla")
    assert bv.number_of_bsplines == 12 + 7 - 2


def test_partition_of_unity_and_overlap_row_sums():
    """Vanilla basis: sum_j B_j = 1 => sum_j S_ij = int B_i = (t_{i+k} - t_i)/k."""
    k = 5
    b = BSplineBasis(k, linear_breakpoints(0.0, 3.0, 7) ** 1.3, "vanilla")
    B, _, _ = b.tables()
    assert np.allclose(B.sum(axis=0), 1.0, atol=1e-13)
    S = b.overlap_band()
    t = b.knots
    ref = np.array([(t[i + k] - t[i]) / k for i in range(b.number_of_bsplines)])
    assert np.allclose(S.sum(axis=1), ref, rtol=1e-13)
    assert np.allclose(S, S.T)
    assert np.linalg.eigvalsh(S).min() > 0


def test_derivatives_by_finite_differences():
    b = BSplineBasis(6, linear_breakpoints(0.0, 5.0, 9), "zero")
    x = np.linspace(0.13, 4.87, 41)
    h = 1e-5
    B, D1, D2 = b.evaluate(x)
    Bp, _, _ = b.evaluate(x + h)
    Bm, _, _ = b.evaluate(x - h)
    assert np.allclose(D1, (Bp - Bm) / (2 * h), atol=1e-6)
    assert np.allclose(D2, (Bp - 2 * B + Bm) / h ** 2, atol=2e-3)


def test_kinetic_matrix_symmetric_zero_bc():
    b = BSplineBasis(7, linear_breakpoints(0.0, 8.0, 10), "zero")
    N = b.number_of_bsplines
    g = Geometry("Simp", dense_pairs(N), True, N)
    T = radial_1d(b, g, np.full(b.grid.size, -0.5), 2).reshape(N, N)
    assert np.allclose(T, T.T, atol=1e-11)


def test_hydrogen_spectrum():
    """-1/(2 n^2) for the low levels of every l from the (H0, S) generalized eigenproblem."""
    h = hydrogen_system()
    b = h.bases[1]
    N = b.number_of_bsplines
    S = b.overlap_band()
    H0 = h.potentials[0].data
    pairs = banded_nonhermitian_pairs(N, 6)
    for l in range(4):
        M = np.zeros((N, N))
        for p, (i, j) in enumerate(pairs):
            if abs(i - j) <= 6:
                M[i, j] = H0[l, p].real
        w = scipy.linalg.eigh(M, S, eigvals_only=True)
        ns = [n for n in range(l + 1, 5)]  # n <= 4: well inside the 80 a.u. box
        ref = np.array([-0.5 / n ** 2 for n in ns])
        assert np.allclose(w[:len(ns)], ref, rtol=1e-7, atol=1e-9)


def test_breakpoint_counts_match_configs():
    bp, rp = exponential_linear_for_count(0.0, 80.0, 20, 117, 4.0)
    assert bp.size == 117 and abs(bp[-1] - 80.0) < 0.01
    b = BSplineBasis(7, bp, "zero")
    assert b.number_of_bsplines == 120
    bp2, _ = exponential_linear_for_count(0.0, 40.0, 20, 57, 4.0)
    assert BSplineBasis(7, bp2, "zero").number_of_bsplines == 60


def test_pair_counts():
    assert len(banded_packed_pairs(60, 7)) == 738
    assert len(banded_nonhermitian_pairs(60, 6)) == 780
    assert len(banded_packed_pairs(120, 7)) == 1518


def test_wigner3j_against_sympy():
    from sympy.physics.wigner import wigner_3j
    rng = np.random.default_rng(0)
    for _ in range(200):
        j1, j2, j3 = (int(v) for v in rng.integers(0, 7, 3))
        m1 = int(rng.integers(-j1, j1 + 1))
        m2 = int(rng.integers(-j2, j2 + 1))
        assert abs(angular.wigner3j(j1, j2, j3, m1, m2, -m1 - m2) - float(wigner_3j(j1, j2, j3, m1, m2, -m1 - m2))) < 1e-14


def test_clebsch_gordan_orthogonality():
    l1, l2 = 2, 3
    for L in range(1, 6):
        for Lp in range(1, 6):
            s = sum(angular.clebsch_gordan(l1, l2, m1, -m1, L, 0) * angular.clebsch_gordan(l1, l2, m1, -m1, Lp, 0)
                    for m1 in range(-2, 3))
            assert abs(s - (1.0 if L == Lp else 0.0)) < 1e-13


def test_channel_and_pair_counts():
    ch = angular.default_coupled_index_iterator(5, L=[0, 1, 2], M=0)
    assert len(ch) == 42
    assert len(angular.coupled_basis_pairs(ch, "selectionrule_r12_10")) == 394
    assert len(angular.coupled_basis_pairs(ch, "selectionrule_linearpolarizedfield")) == 140
    assert len(angular.coupled_basis_pairs(ch, "diagonal")) == 42


def test_r12_monopole_of_s_waves_is_one():
    """For l1 = l2 = l1' = l2' = 0 (L = 0) the multipole sum is 1/r> (lambda = 0, factor 1)."""
    ch = angular.default_coupled_index_iterator(2, L=[0], M=0)
    pairs = angular.coupled_basis_pairs(ch, "diagonal")
    A = angular.r12_multipole_coefficients(ch, pairs)
    i0 = [i for i, c in enumerate(ch) if c.l1 == 0 and c.l2 == 0][0]
    assert abs(A[0][i0] - 1.0) < 1e-14
    # every diagonal element carries the monopole 1 (angular average of 1/r12 over a channel)
    assert np.allclose(A[0], 1.0, atol=1e-13)


def test_r12_coefficients_symmetric():
    ch = angular.default_coupled_index_iterator(3, L=[0, 1], M=0)
    pairs = angular.coupled_basis_pairs(ch, "selectionrule_r12")
    A = angular.r12_multipole_coefficients(ch, pairs)
    idx = {(int(a), int(b)): i for i, (a, b) in enumerate(pairs)}
    for lam, arr in A.items():
        for (a, b), i in idx.items():
            assert abs(arr[i] - arr[idx[(b, a)]]) < 1e-13


def test_dipole_coefficients_closed_form():
    """With l2 = l2' = 0 the coupled state |l1 0 L=l1 0> is the one-electron |l1 0>, so
    the coupled z1 element reduces to the one-electron <l1|cos|l1'>.  The reference's
    Coefficient(l, l') = sqrt((2l+1)(2l'+1)/(12 pi)) (coupledlength.cpp:113-116) makes its
    operator r Y_10 = sqrt(3/(4 pi)) r cos(theta); I2 vanishes."""
    ch = angular.default_coupled_index_iterator(3, L=[0, 1], M=0)
    pairs = angular.coupled_basis_pairs(ch, "selectionrule_linearpolarizedfield")
    I1, I2 = angular.dipole_length_coefficients(ch, pairs)
    for a, (li, ri) in enumerate(pairs):
        left, right = ch[li], ch[ri]
        if left.l2 == 0 and right.l2 == 0:
            ref = angular.reduced_cos_theta(left.l1, right.l1, 0)
            assert abs(abs(I1[a]) - math.sqrt(3.0 / (4.0 * math.pi)) * ref) < 1e-13
            assert abs(I2[a]) < 1e-14


def test_reduced_cos_theta_by_quadrature():
    from scipy.special import sph_harm_y
    x, w = np.polynomial.legendre.leggauss(30)
    th = np.arccos(x)
    for l in range(6):
        for lp in (l - 1, l + 1):
            if lp < 0:
                continue
            Yl = sph_harm_y(l, 0, th, 0.0).real
            Ylp = sph_harm_y(lp, 0, th, 0.0).real
            val = 2 * np.pi * np.sum(w * Yl * x * Ylp)
            assert abs(val - angular.reduced_cos_theta(l, lp)) < 1e-13


def test_laser_function():
    f = LaserSimpleLength(0.05, 0.057, 4.0)
    T = f.pulse_duration
    assert f(0.0) == 0.0 and f(T) == 0.0 and f(-1.0) == 0.0
    assert abs(f(T / 2) - 0.05 * math.cos(0.057 * T / 2)) < 1e-15


def test_helium_small_hamiltonian_bound_state():
    """Small helium basis: the lowest eigenvalue of H0 + R12 lies between the He ground
    state (-2.9037) and the independent-electron value plus the first-order repulsion."""
    from tests.test_oracle import dense_operator
    sysm = helium_system(lmax=1, L=(0,), order=6, n_breakpoints=12, rmax=12.0, n_inner=8, gamma=3.0)
    H, S = dense_operator(sysm, [1.0, 1.0, 0.0])
    w = scipy.linalg.eigh(H.real, S, eigvals_only=True)
    assert -2.91 < w[0] < -2.7


def test_hydrogen_bound_states_and_yield():
    """Field-free bound states of the hydrogen system: the lowest energies per l are
    -1/(2 n^2); the ground state has zero ionisation yield and a state with equal weight on
    the 1s state and a continuum (E > 0) eigenvector has yield 1/2."""
    import scipy.linalg
    from pyprop_amd.observables import bound_states, ionization_yield
    from pyprop_amd.systems import hydrogen_system
    sysm = hydrogen_system()
    bound = bound_states(sysm)
    for l, E, C in bound[:3]:
        for k in range(2):
            nq = l + 1 + k
            assert abs(E[k] + 0.5 / nq ** 2) < 1e-6
    nl, n = sysm.shape
    psi = np.zeros((nl, n), complex)
    psi[0] = bound[0][2][:, 0]
    assert abs(ionization_yield(psi, sysm, bound)) < 1e-12
    # a continuum eigenvector of l = 1
    from pyprop_amd.observables import _band_dense
    H0 = sysm.potentials[0]
    H = _band_dense(H0.data.real[1], H0.geometries[1].pairs, n)
    S = sysm.bases[1].overlap_band()
    w, v = scipy.linalg.eigh(0.5 * (H + H.T), S)
    kc = int(np.argmax(w > 0.5))
    psi[1] = v[:, kc]
    assert abs(ionization_yield(psi, sysm, bound) - 0.5) < 1e-10
