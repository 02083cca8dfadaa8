"""CPU pins of the learned-preconditioner oracle (oracle/mlp.py) and of the host set-up of
the Kronecker preconditioner (pyprop_amd.mixed.radial_eigenbasis).

The network has no reference counterpart, so its oracle is pinned by mathematics:
central finite differences of the loss for every parameter block (BF16 rounding off),
the closed-form derivative of GELU, BF16 round-to-nearest-even on hand-picked values,
and Adam against a hand-unrolled two-step update."""
import math

import numpy as np
import pytest

from oracle import mlp


def _tiny(rng, inp=12, hidden=8, nchan=3, rows=7):
    lay, total = mlp.layout(inp, hidden, nchan)
    flat = rng.standard_normal(total) * 0.4
    p = mlp.unpack(flat, inp, hidden, nchan)
    x = rng.standard_normal((rows, inp))
    t = rng.standard_normal((rows, inp))
    chan = rng.integers(0, nchan, rows)
    return p, x, t, chan, (inp, hidden, nchan)


def test_bf16_rounding_known_values():
    one = 1.0
    ulp = 2.0 ** -7  # BF16 spacing at 1
    vals = np.array([1.0, 1.0 + ulp / 2, 1.0 + 3 * ulp / 2, 1.0 + ulp / 4, -2.5, 3.0e-3])
    got = mlp.bf16(vals)
    # halfway 1 + ulp/2 ties to even (1.0); 1 + 1.5 ulp ties to even (1 + 2 ulp)
    assert got[0] == one
    assert got[1] == one
    assert got[2] == 1.0 + 2 * ulp
    assert got[3] == one
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_gelu_grad_closed_form():
    z = np.linspace(-5, 5, 41)
    h = 1e-6
    fd = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mlp.gelu_grad(z))) < 1e-8
    # GELU(0) = 0, GELU(z) -> z for large z, GELU(-z) -> 0
    assert mlp.gelu(np.array([0.0]))[0] == 0.0
    assert abs(mlp.gelu(np.array([8.0]))[0] - 8.0) < 1e-12


def test_mlp_gradient_vs_finite_differences(rng):
    p, x, t, chan, dims = _tiny(rng)
    sigma = 1.7
    loss, g = mlp.loss_and_grad(p, x, chan, t, sigma, round_bf16=False)
    h = 1e-6
    for name in mlp.NAMES:
        a = p[name]
        idx = [tuple(rng.integers(0, s) for s in a.shape) for _ in range(6)]
        for ix in idx:
            old = a[ix]
            a[ix] = old + h
            lp, _ = mlp.loss_and_grad(p, x, chan, t, sigma, round_bf16=False)
            a[ix] = old - h
            lm, _ = mlp.loss_and_grad(p, x, chan, t, sigma, round_bf16=False)
            a[ix] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[name][ix]) < 1e-6 * max(1.0, abs(fd)), (name, ix, fd, g[name][ix])


def test_mlp_embedding_gradient_only_for_present_channels(rng):
    p, x, t, chan, dims = _tiny(rng, nchan=5, rows=4)
    chan = np.array([0, 0, 2, 2])
    _, g = mlp.loss_and_grad(p, x, chan, t)
    assert np.all(g["emb"][[1, 3, 4]] == 0)
    assert np.any(g["emb"][0] != 0) and np.any(g["emb"][2] != 0)


def test_mlp_zero_output_layer_gives_bias_only(rng):
    p, x, t, chan, dims = _tiny(rng)
    p["W3"][:] = 0
    y, _ = mlp.forward(p, x, chan)
    assert np.allclose(y, p["b3"][None, :])


def test_adam_two_steps_by_hand():
    p = np.array([1.0, -2.0])
    m = np.zeros(2)
    v = np.zeros(2)
    g1 = np.array([0.5, -1.0])
    g2 = np.array([0.25, 2.0])
    lr, b1, b2, eps = 0.1, 0.9, 0.999, 1e-8
    mlp.adam_step(p, g1, m, v, 1, lr, b1, b2, eps)
    # step 1: m_hat = g1, v_hat = g1^2 -> p -= lr * sign(g1) (up to eps)
    assert np.allclose(p, [1.0 - 0.1, -2.0 + 0.1], atol=1e-7)
    mlp.adam_step(p, g2, m, v, 2, lr, b1, b2, eps)
    m2 = 0.9 * 0.1 * g1 + 0.1 * g2
    v2 = 0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2
    ref = np.array([0.9, -1.9]) - lr * (m2 / (1 - 0.81)) / (np.sqrt(v2 / (1 - 0.999 ** 2)) + eps)
    assert np.allclose(p, ref, rtol=1e-12)


def test_radial_eigenbasis_and_kronecker_identity():
    """E^T S E = I, E^T h_l E = diag(lam); the hydrogenic spectrum -Z^2/(2 n^2) for the
    lowest states; K_c built from (E, lam) equals S x S + a (h_l1 x S + S x h_l2)."""
    from pyprop_amd.mixed import radial_eigenbasis
    from pyprop_amd.systems import helium_system
    sysm = helium_system(lmax=2, L=(0, 1, 2), order=5, n_breakpoints=12, rmax=15.0, n_inner=6, gamma=2.5)
    E, lam, l1, l2, _ = radial_eigenbasis(sysm)
    S = sysm.bases[1].overlap_band()
    n = S.shape[0]
    Z = sysm.info["Z"]
    for l in range(E.shape[0]):
        assert np.allclose(E[l].T @ S @ E[l], np.eye(n), atol=1e-10)
    sysm2 = helium_system()
    E2, lam2, _, _, _ = radial_eigenbasis(sysm2)
    for l in range(3):
        for k in range(2):
            nq = l + 1 + k
            assert abs(lam2[l][k] + Z * Z / (2 * nq * nq)) < 1e-6
    # Kronecker inverse against the dense block
    a = 0.005j
    c = 5
    Ea, Eb = E[l1[c]], E[l2[c]]
    Sinv = np.linalg.inv(Ea @ Ea.T)  # S = E^-T E^-1
    h1 = Sinv @ Ea @ np.diag(lam[l1[c]]) @ Ea.T @ Sinv
    h2 = Sinv @ Eb @ np.diag(lam[l2[c]]) @ Eb.T @ Sinv
    K = np.kron(S, S) + a * (np.kron(h1, S) + np.kron(S, h2))
    rng = np.random.default_rng(3)
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Y = (Ea.T @ B @ Eb) / (1 + a * (lam[l1[c]][:, None] + lam[l2[c]][None, :]))
    X = Ea @ Y @ Eb.T
    assert np.linalg.norm(K @ X.ravel() - B.ravel()) < 1e-9 * np.linalg.norm(B)
