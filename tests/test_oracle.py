"""Pins of the CPU oracle (oracle/oracle.cpp) against what the mathematics and the
reference's own checks fix, independently of the oracle's code:

* rank-1 banded multiplies vs scipy's BLAS zgbmv / zhbmv (the routines the generated
  kernels call, generator.py:567-579, 709-723);
* every storage combination at rank 2 and 3 vs a dense matrix assembled from
  (V, pairs) exactly as the reference's SetupBigMatrix does
  (examples/tensor/helium_stabilization/twoelectron_test.py:34-61);
* the 6x6 known-answer matrix of examples/tensor/test/parallel_multiply2.py:351-381;
* overlap Cholesky vs LAPACK zpbtrf and BLAS ztbmv/ztbsv (overlapmatrix.cpp:35-149);
* Pade exponential vs scipy.linalg.expm incl. the TestPade matrices
  (core/krylov/pamp/pampwrapper.h:84-100);
* GMRES vs a dense solve, residual estimate vs true residual;
* pAMP vs dense expm of the balanced operator; Cayley vs the dense Cayley transform;
* exact block preconditioner vs numpy.linalg.solve.
"""
import numpy as np
import pytest
import scipy.linalg
import scipy.linalg.blas as blas
import scipy.linalg.lapack as lapack

import oracle
from tests.conftest import crandn

# ----------------------------------------------------------------------------- helpers


def pairs_for(storage, n, rng):
    """Reference pair layouts (GeometryInfo.py / basis/BSpline.py)."""
    from pyprop_amd.bspline import banded_hermitian_pairs, banded_nonhermitian_pairs
    from pyprop_amd.geometry import dense_hermitian_pairs, dense_pairs
    if storage == "Simp":
        allp = dense_pairs(n)
        sel = rng.random(len(allp)) < 0.5
        sel[0] = True
        return allp[sel]
    if storage == "Herm":
        return dense_hermitian_pairs(n)
    if storage == "Diag":
        return np.stack([np.arange(n), np.arange(n)], 1).astype(np.int32)
    if storage == "Dense":
        return dense_pairs(n)
    if storage == "BandNH":
        return banded_nonhermitian_pairs(n, 2)
    if storage == "Band":
        return banded_hermitian_pairs(n, 3)
    if storage == "Ident":
        return np.zeros((1, 2), np.int32)
    raise ValueError(storage)


def valid_mask(storage, pairs, n):
    """Potential slots that carry a matrix element (padding of the band layouts is unused)."""
    if storage == "BandNH":
        b = 2
        m = np.zeros(len(pairs), bool)
        for i in range(n):
            for j in range(max(0, i - b), min(n, i + b + 1)):
                m[j * (2 * b + 1) + b + i - j] = True
        return m
    if storage == "Band":
        k = 3
        m = np.zeros(len(pairs), bool)
        for i in range(n):
            for j in range(k):
                m[i * k + j] = i + j < n
        return m
    return np.ones(len(pairs), bool)


def big_matrix(storages, V, pairs, shape):
    """SetupBigMatrix (twoelectron_test.py:34-61): A[left, right] += V[i, j, ...] over the
    per-rank pair lists, for storages whose multiply is exactly the pair-list product."""
    N = int(np.prod(shape))
    A = np.zeros((N, N), complex)
    strides = [int(np.prod(shape[r + 1:])) for r in range(len(shape))]
    import itertools
    for idx in itertools.product(*[range(len(p)) for p in pairs]):
        left = sum(pairs[r][idx[r]][0] * strides[r] for r in range(len(shape)))
        right = sum(pairs[r][idx[r]][1] * strides[r] for r in range(len(shape)))
        A[left, right] += V[idx]
    return A


# ----------------------------------------------------------------------------- multiply


def test_bandnh_rank1_vs_zgbmv(rng):
    n, b = 11, 3
    V = crandn(rng, n * (2 * b + 1))
    x = crandn(rng, n)
    y0 = crandn(rng, n)
    y = y0.copy()
    oracle.tpmul(["BandNH"], V, 0.7, x, y, [None])
    a = V.reshape(n, 2 * b + 1).T  # (lda, n) Fortran band storage
    ref = blas.zgbmv(n, n, b, b, 0.7, np.asfortranarray(a), x, beta=1.0, y=y0.copy())
    assert np.allclose(y, ref, rtol=1e-14, atol=1e-14)


def test_band_rank1_vs_zhbmv(rng):
    n, k = 10, 4
    V = crandn(rng, n * k)
    x = crandn(rng, n)
    y0 = crandn(rng, n)
    y = y0.copy()
    oracle.tpmul(["Band"], V, -1.3, x, y, [None])
    a = V.reshape(n, k).T
    ref = blas.zhbmv(k - 1, -1.3, np.asfortranarray(a), x, beta=1.0, y=y0.copy(), lower=1)
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_known_answer_6x6():
    """parallel_multiply2.py:351-381: Simp storage of a 6x6 integer matrix vs dot(A, psi)."""
    A = np.array([[1, 2, 3, 0, 4, 0], [2, 5, 6, 7, 8, 9], [3, 6, 10, 11, 12, 0],
                  [0, 7, 11, 13, 14, 15], [4, 8, 12, 14, 16, 17], [0, 9, 0, 16, 17, 18]], float)
    pairs = np.array([(r, c) for r in range(6) for c in range(6) if A[r, c] != 0], np.int32)
    V = np.array([A[r, c] for r, c in pairs], complex)
    psi = np.random.default_rng(0).random(6) + 0j
    out = np.zeros(6, complex)
    oracle.tpmul(["Simp"], V, 1.0, psi, out, [pairs])
    assert np.allclose(out, A @ psi, rtol=1e-15, atol=1e-14)


PAIRLIST_STORAGES = ["Simp", "Diag", "Dense", "BandNH"]


@pytest.mark.parametrize("s0", PAIRLIST_STORAGES)
@pytest.mark.parametrize("s1", PAIRLIST_STORAGES)
def test_rank2_vs_setupbigmatrix(s0, s1, rng):
    shape = (4, 6)
    pairs = [pairs_for(s0, shape[0], rng), pairs_for(s1, shape[1], rng)]
    V = crandn(rng, (len(pairs[0]), len(pairs[1])))
    V *= np.multiply.outer(valid_mask(s0, pairs[0], shape[0]), valid_mask(s1, pairs[1], shape[1]))
    psi = crandn(rng, shape)
    out = np.zeros(shape, complex)
    oracle.tpmul([s0, s1], V, 0.37, psi, out, [p if s in ("Simp", "Herm") else None for s, p in zip([s0, s1], pairs)])
    A = big_matrix([s0, s1], V, pairs, shape)
    assert np.allclose(out.ravel(), 0.37 * A @ psi.ravel(), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("ids", [("Simp", "Simp", "BandNH"), ("Diag", "Simp", "BandNH"),
                                 ("Simp", "BandNH", "BandNH"), ("Dense", "Diag", "Simp")])
def test_rank3_vs_setupbigmatrix(ids, rng):
    shape = (3, 5, 6)
    pairs = [pairs_for(s, n, rng) for s, n in zip(ids, shape)]
    V = crandn(rng, tuple(len(p) for p in pairs))
    mask = np.ones(V.shape, bool)
    for r, (s, p, n) in enumerate(zip(ids, pairs, shape)):
        m = valid_mask(s, p, n)
        mask &= m.reshape([-1 if q == r else 1 for q in range(3)])
    V *= mask
    psi = crandn(rng, shape)
    out = np.zeros(shape, complex)
    oracle.tpmul(list(ids), V, 1.0, psi, out, [p if s in ("Simp", "Herm") else None for s, p in zip(ids, pairs)])
    A = big_matrix(ids, V, pairs, shape)
    assert np.allclose(out.ravel(), A @ psi.ravel(), rtol=1e-13, atol=1e-13)


def test_herm_rank1_closed_form(rng):
    """Herm adds V at (row, col) and conj(V) at (col, row); diagonal pairs get both."""
    n = 5
    pairs = pairs_for("Herm", n, rng)
    V = crandn(rng, len(pairs))
    A = np.zeros((n, n), complex)
    for (r, c), v in zip(pairs, V):
        A[r, c] += v
        A[c, r] += np.conj(v)
    x = crandn(rng, n)
    out = np.zeros(n, complex)
    oracle.tpmul(["Herm"], V, 1.0, x, out, [pairs])
    assert np.allclose(out, A @ x, rtol=1e-14)
    assert np.allclose(A, A.conj().T)


def test_ident_rank(rng):
    """Ident: potential independent of the rank -> identity on that rank."""
    shape = (4, 5)
    V = crandn(rng, (1, 5))  # Ident x Diag
    psi = crandn(rng, shape)
    out = np.zeros(shape, complex)
    oracle.tpmul(["Ident", "Diag"], V, 1.0, psi, out, [None, None])
    assert np.allclose(out, psi * V[0][None, :], rtol=1e-14)


def test_real_symmetric_outer_band_matches_symmetric_completion(rng):
    """Band at an outer rank with a real symmetric band and a real symmetric inner
    operator gives the symmetric matrix the band describes."""
    n0, n1, k = 5, 4, 2
    from pyprop_amd.bspline import banded_hermitian_pairs
    p0 = banded_hermitian_pairs(n0, k)
    a = rng.standard_normal(n0 * k)
    a *= np.array([i + j < n0 for i in range(n0) for j in range(k)])
    B = rng.standard_normal((n1, n1))
    B = B + B.T
    V = np.multiply.outer(a, np.array([B[i, i] for i in range(n1)])) + 0j  # Band x Diag
    psi = crandn(rng, (n0, n1))
    out = np.zeros((n0, n1), complex)
    oracle.tpmul(["Band", "Diag"], V, 1.0, psi, out, [None, None])
    A0 = np.zeros((n0, n0))
    for idx, (i, j) in enumerate(p0):
        if a[idx] != 0 or i == j:
            A0[i, j] = a[idx]
            A0[j, i] = a[idx]
    ref = A0 @ psi * np.diag(B)[None, :]
    assert np.allclose(out, ref, rtol=1e-13, atol=1e-13)


# ----------------------------------------------------------------------------- overlap


def _spd_band(n, sd, rng):
    S = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - sd), min(n, i + sd + 1)):
            S[i, j] = rng.random() * 0.1
    S = S + S.T + np.eye(n) * (2 * sd + 1)
    return S


def _upper(S, sd):
    from pyprop_amd.propagator import band_upper
    return band_upper(S, sd)


def test_overlap_factor_vs_zpbtrf(rng):
    n, sd = 12, 3
    S = _spd_band(n, sd, rng)
    su = _upper(S, sd)
    U = oracle.overlap_factor(su)
    ab = np.asfortranarray(su.T.astype(complex))  # (sd+1, n) upper band
    c, info = lapack.zpbtrf(ab, lower=0)
    assert info == 0
    assert np.allclose(U.T, c.real, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("op", [0, 1, 2, 3])
@pytest.mark.parametrize("conj", [False, True])
def test_overlap_ops_vs_blas(op, conj, rng):
    n, sd = 9, 2
    S = _spd_band(n, sd, rng)
    U = oracle.overlap_factor(_upper(S, sd))
    a = np.asfortranarray(U.T.astype(complex))
    x = crandn(rng, (3, n, 4))
    got = oracle.overlap_apply(U, op, conj, x.copy(), 1)
    Ud = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - sd), j + 1):
            Ud[i, j] = U[j, sd + i - j]
    assert np.allclose(Ud.T @ Ud, S, rtol=1e-13)
    for p in range(3):
        for q in range(4):
            v = x[p, :, q].copy()
            if op == 0:
                v = blas.ztbmv(sd, a, v, trans=0)
                v = blas.ztbmv(sd, a, v, trans=2)
                assert np.allclose(v, S @ x[p, :, q])
            elif op == 1:
                v = blas.ztbsv(sd, a, v, trans=2)
                v = blas.ztbsv(sd, a, v, trans=0)
                assert np.allclose(v, np.linalg.solve(S, x[p, :, q]))
            elif op == 2:
                v = blas.ztbmv(sd, a, v, trans=2 if conj else 0)
            else:
                v = blas.ztbsv(sd, a, v, trans=2 if conj else 0)
            assert np.allclose(got[p, :, q], v, rtol=1e-13, atol=1e-13)


# ----------------------------------------------------------------------------- Pade / pAMP


def test_pade_testpade_matrices():
    """TestPade inputs of core/krylov/pamp/pampwrapper.h:84-100, and their i-multiples."""
    A = np.array([[-49, 24], [24, 31]], complex)
    for M in (A, 1j * A, A / 40.0):
        got = oracle.pade_exp(M, 6)
        ref = scipy.linalg.expm(M)
        assert np.allclose(got, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def textbook_pade(A, p):
    """Diagonal [p/p] Pade approximant N(B)/D(B), c_j = (2p-j)! p! / ((2p)! j! (p-j)!),
    with the reference's scaling s = max(0, ceil(log2 max|A_ij|)) and s squarings."""
    import math
    c = [math.factorial(2 * p - j) * math.factorial(p) / (math.factorial(2 * p) * math.factorial(j)
                                                          * math.factorial(p - j)) for j in range(p + 1)]
    mx = np.abs(A).max()
    s = max(0, int(math.ceil(math.log2(mx)))) if mx > 0 else 0
    B = A / 2.0 ** s
    Nn = sum(c[j] * np.linalg.matrix_power(B, j) for j in range(p + 1))
    Dd = sum(c[j] * np.linalg.matrix_power(-B, j) for j in range(p + 1))
    E = np.linalg.solve(Dd, Nn)
    for _ in range(s):
        E = E @ E
    return E


@pytest.mark.parametrize("order", [3, 6, 9])
def test_pade_random_hermitian(order, rng):
    m = 12
    H = crandn(rng, (m, m))
    H = H + H.conj().T
    A = -1j * 0.3 * H
    got = oracle.pade_exp(A, order)
    # the even/odd Horner form with E = I + 2 (Q-P)^-1 P is the textbook approximant
    assert np.allclose(got, textbook_pade(A, order), rtol=1e-13, atol=1e-14)
    # and the approximant's truncation error at this scaling is small
    assert np.abs(got - scipy.linalg.expm(A)).max() < {3: 1e-3, 6: 1e-9, 9: 1e-13}[order]


def test_pamp_vs_expm(rng):
    n = 40
    H = crandn(rng, (n, n))
    H = (H + H.conj().T) / 2
    v = crandn(rng, n)
    dt = 0.05
    out, est, ops, bd = oracle.pamp_propagate(lambda x: H @ x, v, 20, dt)
    ref = scipy.linalg.expm(-1j * dt * H) @ v
    assert ops == 20
    assert np.linalg.norm(out - ref) < 1e-10 * np.linalg.norm(ref)
    assert abs(np.linalg.norm(out) - np.linalg.norm(v)) < 1e-10 * np.linalg.norm(v)


def test_pamp_happy_breakdown(rng):
    """An eigenvector spans a one-dimensional Krylov space."""
    n = 30
    H = np.diag(np.arange(n, dtype=float)) + 0j
    v = np.zeros(n, complex)
    v[3] = 1.0
    out, est, ops, bd = oracle.pamp_propagate(lambda x: H @ x, v, 10, 0.1)
    assert bd
    assert np.allclose(out, np.exp(-1j * 0.3) * v, atol=1e-14)


# ----------------------------------------------------------------------------- GMRES


def test_gmres_dense(rng):
    n = 50
    A = np.eye(n) * 4 + crandn(rng, (n, n)) * 0.2
    b = crandn(rng, n)
    x, err, ops, elist = oracle.gmres_solve(lambda v: A @ v, b, 40, 1e-11)
    assert np.linalg.norm(A @ x - b) < 1e-9
    # estimate is the true residual (absolute)
    assert abs(err - np.linalg.norm(A @ x - b)) < 1e-9
    nz = elist[elist > 0]
    assert np.all(np.diff(nz[1:]) <= 1e-12)  # monotone after the first (relative) entry


def test_gmres_tolerance_clamped(rng):
    n = 20
    A = np.eye(n) * 3 + crandn(rng, (n, n)) * 0.1
    b = crandn(rng, n)
    x, err, ops, _ = oracle.gmres_solve(lambda v: A @ v, b, 25, 0.0)
    assert err <= 1e-13 or ops == 25


# ----------------------------------------------------------------------------- systems


def small_helium():
    from pyprop_amd.systems import helium_system
    return helium_system(lmax=1, L=(0, 1), order=4, n_breakpoints=8, rmax=10.0, n_inner=5, gamma=2.0)


def oracle_system(sysm):
    s = oracle.System(sysm.shape)
    for p in sysm.potentials:
        s.add_potential(p.storage_ids, p.data, [g.pairs if g.storage_id in ("Simp", "Herm") else None
                                                for g in p.geometries])
    from pyprop_amd.propagator import band_upper
    for r, b in enumerate(sysm.bases):
        if b is not None:
            s.set_overlap(r, band_upper(b.overlap_band(), b.order - 1))
    return s


def dense_operator(sysm, scal):
    from pyprop_amd.geometry import dense_pairs
    N = sysm.size
    H = np.zeros((N, N), complex)
    for p, f in zip(sysm.potentials, scal):
        assert all(s in ("Simp", "BandNH", "Diag") for s in p.storage_ids)
        V = p.data.copy()
        for r, g in enumerate(p.geometries):
            if g.storage_id == "BandNH":
                n = g.extent
                b = (g.pair_count // n - 1) // 2
                m = valid_mask_band(n, b)
                V *= m.reshape([-1 if q == r else 1 for q in range(len(p.geometries))])
        H += f * big_matrix(p.storage_ids, V, [g.pairs for g in p.geometries], sysm.shape)
    mats = [np.eye(n) if b is None else b.overlap_band() for b, n in zip(sysm.bases, sysm.shape)]
    S = mats[0]
    for m in mats[1:]:
        S = np.kron(S, m)
    return H, S


def valid_mask_band(n, b):
    m = np.zeros(n * (2 * b + 1), bool)
    for i in range(n):
        for j in range(max(0, i - b), min(n, i + b + 1)):
            m[j * (2 * b + 1) + b + i - j] = True
    return m


def set_exact_preconditioner(s, sysm, dt):
    b1 = sysm.bases[1]
    P = sum(p.data for p in sysm.precond_potentials)
    g = sysm.precond_potentials[0].geometries
    s.set_preconditioner(b1.overlap_band(), b1.overlap_band(), b1.order - 1, P, g[1].pairs, g[2].pairs, 0.5j * dt)


@pytest.mark.parametrize("precond", [True, False])
def test_cayley_step_vs_dense(precond, rng):
    sysm = small_helium()
    s = oracle_system(sysm)
    scal = np.array([1.0, 1.0, 0.03])
    H, S = dense_operator(sysm, scal)
    psi = crandn(rng, sysm.shape)
    dt = 0.05
    if precond:
        set_exact_preconditioner(s, sysm, dt)
    m = 40 if precond else sysm.size
    out, err, ops = s.cayley_step(psi, dt, scal, m=m, tol=1e-12)
    A = S + 0.5j * dt * H
    B = S - 0.5j * dt * H
    ref = np.linalg.solve(A, B @ psi.ravel())
    assert err < 1e-11
    assert np.linalg.norm(out.ravel() - ref) < 1e-8 * np.linalg.norm(ref)
    # the Cayley transform is unitary in the S metric for Hermitian H
    n0 = np.vdot(psi.ravel(), S @ psi.ravel()).real
    n1 = np.vdot(out.ravel(), S @ out.ravel()).real
    assert abs(n1 - n0) < 1e-8 * n0


def test_hamiltonian_is_symmetric_and_energy_is_real(rng):
    sysm = small_helium()
    H, S = dense_operator(sysm, [1.0, 1.0, 0.05])
    assert np.allclose(H, H.T.conj(), atol=1e-10 * np.abs(H).max())
    s = oracle_system(sysm)
    psi = crandn(rng, sysm.shape)
    hp = s.multiply(psi, [1.0, 1.0, 0.05], mode=0)
    assert np.allclose(hp.ravel(), H @ psi.ravel(), rtol=1e-12, atol=1e-12)


def test_pamp_balanced_vs_dense(rng):
    from pyprop_amd.systems import hydrogen_system
    sysm = hydrogen_system(lmax=2, order=5, n_breakpoints=14, rmax=20.0, n_inner=6, gamma=2.0)
    s = oracle_system(sysm)
    scal = [1.0, 0.05]
    H, S = dense_operator(sysm, scal)
    psi = crandn(rng, sysm.shape)
    dt = 0.01
    out, est = s.pamp_step(psi, dt, scal, m=20)
    ref = scipy.linalg.expm(-1j * dt * np.linalg.solve(S, H)) @ psi.ravel()
    assert np.linalg.norm(out.ravel() - ref) < 1e-8 * np.linalg.norm(ref)


def test_preconditioner_exact_vs_dense(rng):
    sysm = small_helium()
    s = oracle_system(sysm)
    b1 = sysm.bases[1]
    P = sum(p.data for p in sysm.precond_potentials)
    g = sysm.precond_potentials[0].geometries
    dt = 0.05
    s.set_preconditioner(b1.overlap_band(), b1.overlap_band(), b1.order - 1, P, g[1].pairs, g[2].pairs, 0.5j * dt)
    x = crandn(rng, sysm.shape)
    got = s.precond_solve(x)
    S1 = b1.overlap_band()
    n1 = S1.shape[0]
    for c in range(sysm.shape[0]):
        M = np.kron(S1, S1).astype(complex)
        for q1, (i1, j1) in enumerate(g[1].pairs):
            for q2, (i2, j2) in enumerate(g[2].pairs):
                M[i1 * n1 + i2, j1 * n1 + j2] += 0.5j * dt * P[c, q1, q2]
        ref = np.linalg.solve(M, x[c].ravel())
        assert np.allclose(got[c].ravel(), ref, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("ids", [("Simp", "Simp", "BandNH"), ("Herm", "Band", "Diag"), ("Diag", "BandNH"),
                                 ("Band", "Simp", "Dense")], ids=lambda c: "_".join(c))
def test_threaded_multiply_is_bitwise_sequential(ids, rng):
    """The multiply split over host threads by rank-0 destination row sums every element in
    the sequential loop order, so it equals the one-thread result bit for bit."""
    shape = {2: (7, 23), 3: (9, 6, 11)}[len(ids)]
    pairs = [pairs_for(s, n, rng) for s, n in zip(ids, shape)]
    V = crandn(rng, tuple(len(p) for p in pairs))
    src = crandn(rng, shape)
    outs = []
    try:
        for nt in (1, 3, 8):
            oracle.set_threads(nt)
            dst = crandn(np.random.default_rng(5), shape)
            outs.append(oracle.tpmul(ids, V, 0.7, src, dst, pairs).copy())
    finally:
        oracle.set_threads(1)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_threaded_cayley_step_is_bitwise_sequential(rng):
    """Overlap passes (by fibre), preconditioner solve (by channel) and the multiply (by row)
    on host threads: a preconditioned Cayley step equals the one-thread step bit for bit."""
    sysm = small_helium()
    s = oracle_system(sysm)
    dt = 0.01
    set_exact_preconditioner(s, sysm, dt)
    psi = crandn(rng, sysm.shape)
    scal = np.array([1.0, 1.0, 0.04])
    res = []
    try:
        for nt in (1, 4):
            oracle.set_threads(nt)
            res.append((s.cayley_step(psi, dt, scal, m=30, tol=1e-12), s.overlap(1, False, psi),
                        s.precond_solve(psi), s.multiply(psi, scal, mode=2)))
    finally:
        oracle.set_threads(1)
    (a, b) = res
    assert np.array_equal(a[0][0], b[0][0]) and a[0][2] == b[0][2]
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


def test_eig_exp_normal_is_expm(rng):
    """EigenvectorExponential (pamp.h:567-608) forms V diag(exp(w)) V^H.  For the normal
    matrices pAMP meets in exact arithmetic (A = -i dt T, T Hermitian tridiagonal, the
    Hessenberg matrix of a Hermitian operator) that is exp(A)."""
    for m in (1, 2, 7, 20):
        d = rng.standard_normal(m)
        e = rng.standard_normal(m - 1) + 1j * rng.standard_normal(m - 1)
        T = np.diag(d).astype(complex) + np.diag(e, -1) + np.diag(e.conj(), 1)
        A = -0.3j * T
        assert np.allclose(oracle.eig_exp(A), scipy.linalg.expm(A), rtol=0, atol=1e-12)


def test_eig_exp_non_normal_matches_the_reference_formula(rng):
    """On a non-normal Hessenberg matrix the reference's V E V^H is not exp(A); the oracle
    must reproduce the formula itself: against numpy's eig (LAPACK zgeev, unit-norm
    eigenvectors; the formula does not depend on their phases)."""
    m = 12
    A = np.triu(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)), -1) * 0.4
    w, V = np.linalg.eig(A)
    ref = V @ np.diag(np.exp(w)) @ V.conj().T
    got = oracle.eig_exp(A)
    assert np.allclose(got, ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    assert not np.allclose(got, scipy.linalg.expm(A), atol=1e-3)


def test_pamp_eigenvector_method_vs_expm(rng):
    """pAMP with krylov_exponentiation_method = 1 on a Hermitian operator: the same
    propagation as the Pade method and as dense expm (the Hessenberg matrix is Hermitian up
    to rounding, so V E V^H = exp)."""
    n = 40
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = 0.5 * (X + X.conj().T) / np.sqrt(n)
    v = crandn(rng, n)
    dt = 0.05
    ref = scipy.linalg.expm(-1j * dt * H) @ v
    out_p, est_p, _, _ = oracle.pamp_propagate(lambda x: H @ x, v, 16, dt, method=0)
    out_e, est_e, _, _ = oracle.pamp_propagate(lambda x: H @ x, v, 16, dt, method=1)
    assert np.linalg.norm(out_e - ref) < 1e-10 * np.linalg.norm(ref)
    assert np.linalg.norm(out_e - out_p) < 1e-11 * np.linalg.norm(ref)
    assert abs(est_e - est_p) < 1e-10
