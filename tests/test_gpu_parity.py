"""GPU parity: the gfx950 path against the CPU oracle on identical seeded inputs.

Bar: FP64 tensor-potential multiply within 1e-12 relative (norm-wise) of the oracle;
propagated wavefunctions within 1e-10 relative (the Krylov solvers' own tolerances set
this bar; GMRES solves to an absolute residual of 1e-12 here).
"""
import sys

import numpy as np
import pytest

import oracle
from tests.conftest import crandn
from tests.test_oracle import big_matrix, pairs_for, valid_mask

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    from pyprop_amd import _native
    _native.load()
    return t


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


COMBOS = [
    ("BandNH",), ("Band",), ("Simp",), ("Herm",), ("Diag",), ("Dense",),
    ("Simp", "BandNH"), ("Diag", "BandNH"), ("Herm", "BandNH"), ("Band", "BandNH"), ("BandNH", "BandNH"),
    ("Simp", "Band"), ("Herm", "Band"), ("Dense", "Simp"), ("Ident", "BandNH"), ("Simp", "Herm"),
    ("Simp", "Simp", "BandNH"), ("Diag", "Simp", "BandNH"), ("Simp", "BandNH", "Band"), ("Herm", "Herm", "Diag"),
    ("Band", "Simp", "Dense"), ("Simp", "Simp", "Simp", "BandNH"),
]


@pytest.mark.parametrize("ids", COMBOS, ids=lambda c: "_".join(c))
@pytest.mark.parametrize("real", [False, True])
def test_tpmul_storages_vs_oracle(torch, ids, real, rng):
    from pyprop_amd.tensorpotential import TensorPotentialMultiply
    say("tpmul %s real=%s" % ("_".join(ids), real))
    shape = {1: (70,), 2: (5, 67), 3: (4, 5, 9), 4: (3, 3, 4, 8)}[len(ids)]
    pairs = [pairs_for(s, n, rng) for s, n in zip(ids, shape)]
    pshape = tuple(len(p) if s != "Ident" else 1 for s, p in zip(ids, pairs))
    V = crandn(rng, pshape)
    if real:
        V = V.real + 0j
    psi = crandn(rng, shape)
    dst0 = crandn(rng, shape)
    ref = dst0.copy()
    plist = [p if s in ("Simp", "Herm") else None for s, p in zip(ids, pairs)]
    oracle.tpmul(list(ids), V, 0.83, psi, ref, plist)
    out = dev(torch, dst0)
    TensorPotentialMultiply(list(ids), dev(torch, V), 0.83, dev(torch, psi), out, plist)
    got = out.cpu().numpy()
    assert rel(got, ref) < 1e-12


def test_tp_plan_repeated_and_f32(torch, rng):
    from pyprop_amd.systems import helium_system
    from pyprop_amd.tensorpotential import TensorPotential
    from pyprop_amd import _native as N, device as D
    say("plan reuse + fp32, small helium")
    sysm = helium_system(lmax=2, L=(0, 1, 2), order=5, n_breakpoints=12, rmax=15.0, n_inner=6, gamma=2.5)
    pot = sysm.potentials[1]
    tp = TensorPotential(pot, sysm.shape)
    tp.SetupStep()
    assert tp.is_real
    psi = crandn(rng, sysm.shape)
    ref = np.zeros(sysm.shape, complex)
    oracle.tpmul(pot.storage_ids, pot.data, 1.0, psi, ref, [g.pairs if g.storage_id in ("Simp", "Herm") else None
                                                             for g in pot.geometries])
    x = dev(torch, psi)
    for _ in range(3):
        y = torch.zeros_like(x)
        tp.MultiplyPotential(x, y, 0.0)
        assert rel(y.cpu().numpy(), ref) < 1e-12
    x32 = x.to(torch.complex64)
    y32 = torch.zeros_like(x32)
    N.check(N.load().pp_tp_multiply_f32(tp.plan, 1.0, D.ptr(x32), D.ptr(y32), D.stream_handle()))
    assert rel(y32.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("op", [0, 1, 2, 3])
@pytest.mark.parametrize("conj", [False, True])
@pytest.mark.parametrize("rank", [0, 1, 2])
def test_overlap_ops_vs_oracle(torch, op, conj, rank, rng):
    from pyprop_amd.bspline import BSplineBasis, linear_breakpoints
    from pyprop_amd.propagator import OverlapMatrix
    from pyprop_amd import _native as N, device as D
    import ctypes as C
    say("overlap op=%d conj=%s rank=%d" % (op, conj, rank))
    b = BSplineBasis(7, linear_breakpoints(0, 10, 15) ** 1.2, "zero")
    n = b.number_of_bsplines
    ov = OverlapMatrix(b.overlap_band(), 6)
    u_oracle = oracle.overlap_factor(ov.s_upper)
    assert np.allclose(ov.factor(), u_oracle, rtol=1e-13, atol=1e-15)
    shape = [3, 5, 7]
    shape[rank] = n
    x = crandn(rng, tuple(shape))
    ref = oracle.overlap_apply(u_oracle, op, conj, x.copy(), rank)
    t = dev(torch, x)
    ext = np.array(shape, dtype=np.int32)
    N.check(N.load().pp_overlap_apply(ov.h, op, int(conj), 3, D.hptr(ext), rank, D.ptr(t), D.stream_handle()))
    assert rel(t.cpu().numpy(), ref) < 1e-12


def _small_helium():
    from pyprop_amd.systems import helium_system
    return helium_system(lmax=2, L=(0, 1, 2), order=5, n_breakpoints=12, rmax=15.0, n_inner=6, gamma=2.5)


def _oracle_sys(sysm):
    from tests.test_oracle import oracle_system
    return oracle_system(sysm)


def _device(sysm, torch):
    from pyprop_amd.propagator import BasisPropagator, CombinedRepresentation, Wavefunction
    repr_ = CombinedRepresentation(sysm.shape, sysm.bases)
    psi = Wavefunction(repr_)
    base = BasisPropagator(psi, sysm.potentials)
    return repr_, psi, base


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_hamiltonian_modes_vs_oracle(torch, mode, rng):
    say("hamiltonian mode %d" % mode)
    sysm = _small_helium()
    s = _oracle_sys(sysm)
    repr_, psi, base = _device(sysm, torch)
    x = crandn(rng, sysm.shape)
    t = 123.0
    sc = base.scalings(t)
    ref = s.multiply(x, sc, mode=mode)
    src = dev(torch, x)
    dst = torch.zeros_like(src)
    [base.MultiplyHamiltonianNoOverlap, base.MultiplyHamiltonian, base.MultiplyHamiltonianBalancedOverlap][mode](
        src, dst, t)
    assert rel(dst.cpu().numpy(), ref) < 1e-12
    assert np.array_equal(src.cpu().numpy(), x)


def test_inner_product_and_norm(torch, rng):
    say("inner product")
    sysm = _small_helium()
    s = _oracle_sys(sysm)
    repr_, psi, base = _device(sysm, torch)
    a = crandn(rng, sysm.shape)
    b = crandn(rng, sysm.shape)
    ref = s.inner_product(a, b)
    got = repr_.InnerProduct(dev(torch, a), dev(torch, b))
    assert abs(got - ref) < 1e-12 * abs(ref)


def test_gmres_callback_vs_oracle(torch, rng):
    from pyprop_amd.propagator import GmresWrapper
    say("gmres with callback operator")
    sysm = _small_helium()
    s = _oracle_sys(sysm)
    repr_, psi, base = _device(sysm, torch)
    n = sysm.size
    shift = 3.0

    def op_cpu(v):
        return s.multiply(v.reshape(sysm.shape), [1, 1, 0], mode=1).ravel() + shift * v

    b = crandn(rng, n)
    ref, err, ops, elist = oracle.gmres_solve(op_cpu, b, 40, 1e-11, True)
    g = GmresWrapper(n, 40, 1e-11, True)

    def op_dev(src, dst):
        base.MultiplyHamiltonian(src.view(sysm.shape), dst.view(sysm.shape), 0.0)
        dst.add_(src, alpha=shift)

    x = torch.zeros(n, dtype=torch.complex128, device="cuda")
    g.Solve(op_dev, dev(torch, b), x)
    assert g.GetOperatorCount() == ops
    assert abs(g.GetErrorEstimate() - err) < 1e-9
    assert rel(x.cpu().numpy(), ref) < 1e-10


@pytest.mark.parametrize("m,dt,steps,tol", [(30, 0.002, 5, 1e-10), (20, 0.01, 1, 1e-7)])
def test_pamp_hydrogen_vs_oracle(torch, m, dt, steps, tol, rng):
    """The hydrogen radial spectrum reaches ~6e6 a.u. (exponential grid, order 7), so a
    20-vector Krylov space at dt = 0.01 resolves the propagator only to ~1e-8 per step
    and rounding differences grow from there; with m = 30, dt = 0.002 the expansion is
    converged to ~1e-12 (checked against dense expm in the oracle tests) and the two
    implementations must agree to 1e-10."""
    from pyprop_amd.propagator import BasisPropagator, CombinedRepresentation, PamPropagator, Wavefunction
    from pyprop_amd.systems import hydrogen_system, hydrogenic_1s
    say("pAMP hydrogen (8 x 120) m=%d dt=%g steps=%d" % (m, dt, steps))
    sysm = hydrogen_system()
    s = _oracle_sys(sysm)
    repr_ = CombinedRepresentation(sysm.shape, sysm.bases)
    x0 = np.zeros(sysm.shape, complex)
    for l, amp in ((0, 1.0), (1, 0.5j), (2, 0.25)):
        c, _ = hydrogenic_1s(sysm.bases[1], 1.0, l)
        x0[l] = amp * c
    psi = Wavefunction(repr_, x0)
    base = BasisPropagator(psi, sysm.potentials)
    prop = PamPropagator(psi, base, krylov_basis_size=m)
    ref = x0.copy()
    t = 100.0
    for k in range(steps):
        sc = base.scalings(t)
        ref, est = s.pamp_step(ref, dt, sc, m=m)
        prop.AdvanceStep(t, dt)
        t += dt
    got = psi.numpy()
    assert rel(got, ref) < tol
    assert abs(prop.GetPropagationErrorEstimate()) < 10 * abs(est) + 1e-12
    n0 = s.inner_product(x0, x0).real
    n1 = psi.InnerProduct(psi).real
    assert abs(n1 - n0) < 1e-8 * n0


def test_cayley_helium_small_vs_oracle(torch, rng):
    from pyprop_amd.propagator import CayleyPropagator, RadialTwoElectronPreconditioner
    from tests.test_oracle import set_exact_preconditioner
    say("Cayley small helium 3 steps with exact preconditioner")
    sysm = _small_helium()
    s = _oracle_sys(sysm)
    dt = 0.01
    set_exact_preconditioner(s, sysm, dt)
    repr_, psi, base = _device(sysm, torch)
    x0 = crandn(rng, sysm.shape)
    psi.data.copy_(dev(torch, x0))
    pc = RadialTwoElectronPreconditioner(psi, sysm)
    prop = CayleyPropagator(psi, base, pc, krylov_basis_size=30, krylov_tolerance=1e-12)
    ref = x0.copy()
    t = 50.0
    for k in range(3):
        sc = base.scalings(t + dt / 2)
        ref, err, ops = s.cayley_step(ref, dt, sc, m=30, tol=1e-12)
        prop.AdvanceStep(t, dt)
        assert prop.GetOperatorCount() == ops
        t += dt
    assert rel(psi.numpy(), ref) < 1e-10


def test_preconditioner_vs_oracle(torch, rng):
    from pyprop_amd.propagator import RadialTwoElectronPreconditioner
    from tests.test_oracle import set_exact_preconditioner
    say("block preconditioner solve")
    sysm = _small_helium()
    s = _oracle_sys(sysm)
    set_exact_preconditioner(s, sysm, 0.02)
    repr_, psi, base = _device(sysm, torch)
    pc = RadialTwoElectronPreconditioner(psi, sysm)
    pc.SetHamiltonianScaling(0.01j)
    pc.Setup()
    x = crandn(rng, sysm.shape)
    ref = s.precond_solve(x)
    t = dev(torch, x)
    pc.Solve(t)
    assert rel(t.cpu().numpy(), ref) < 1e-11


def test_helium_full_size_hamiltonian(torch, rng):
    """Full helium workload (42 x 60 x 60, H0 + R12 + dipole in Simp_Simp_BandNH): the
    fused gfx950 multiply against an independent evaluation of the same operator in its
    separable form sum_t (angular_t x radial_t) on the CPU (scipy.sparse)."""
    import scipy.sparse as sp
    from pyprop_amd.systems import helium_system
    say("full-size helium Hamiltonian (N = 151200)")
    sysm = helium_system()
    repr_, psi, base = _device(sysm, torch)
    nc, n1, n2 = sysm.shape
    x = crandn(rng, sysm.shape)
    t = 200.0
    sc = base.scalings(t)
    ref = np.zeros((nc, n1 * n2), complex)
    X = x.reshape(nc, n1 * n2)
    for p, f in zip(sysm.potentials, sc):
        g0, g1, g2 = p.geometries
        hb = (g2.pair_count // n2 - 1) // 2
        keep2 = np.abs(g2.pairs[:, 0] - g2.pairs[:, 1]) <= hb  # BandNH padding slots carry no element
        rows = (g1.pairs[:, 0][:, None] * n2 + g2.pairs[:, 0][None, :])[:, keep2].ravel()
        cols = (g1.pairs[:, 1][:, None] * n2 + g2.pairs[:, 1][None, :])[:, keep2].ravel()
        for c_t, R_t in zip(p.angular_coeffs, p.radial_factors):
            M = sp.csr_matrix((R_t[:, keep2].ravel(), (rows, cols)), shape=(n1 * n2, n1 * n2))
            Y = (M @ X.T).T
            A = sp.csr_matrix((c_t, (g0.pairs[:, 0], g0.pairs[:, 1])), shape=(nc, nc))
            ref += f * (A @ Y)
    src = dev(torch, x)
    dst = torch.zeros_like(src)
    base.MultiplyHamiltonianNoOverlap(src, dst, t)
    got = dst.cpu().numpy().reshape(nc, -1)
    assert rel(got, ref) < 1e-12
