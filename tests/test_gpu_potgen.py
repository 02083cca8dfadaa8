"""Potential generation on the GPU (pp_potential_grid_*, pp_bspline_represent): grid
evaluation and the rank-by-rank B-spline projection against the oracle's restatement of
GeneratePotential at small sizes, and against the separable-factor potentials at the
full helium size (different order of summation, equal to rounding)."""
import sys
import time

import numpy as np
import pytest

from tests.test_potgen_oracle import oracle_generate

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("which", ["hydrogen", "helium"])
def test_device_generation_vs_oracle(torch, which):
    from pyprop_amd.potgen import generate_device_potential, generate_device_potential_staged
    from pyprop_amd.systems import helium_system, hydrogen_system
    say("device potential generation vs oracle: %s" % which)
    if which == "hydrogen":
        s = hydrogen_system()
        pots = s.potentials
    else:
        s = helium_system(lmax=2, L=(0, 1, 2), order=5, n_breakpoints=12, rmax=15.0, n_inner=6, gamma=2.5)
        pots = s.potentials + s.precond_potentials
    for p in pots:
        ref = oracle_generate(p)
        for fn in (generate_device_potential, generate_device_potential_staged):
            V = fn(p).cpu().numpy()
            err = _rel(V.real, ref)
            say("  %-50s %-12s %s rel err %.2e" % (p.name[:50], "staged" if fn is generate_device_potential_staged else "fused", V.shape, err))
            assert V.shape == ref.shape
            assert np.all(V.imag == 0)
            assert err < 1e-13


def test_device_projection_one_rank_all_storages(torch, rng):
    """pp_bspline_represent alone on a random grid function, for every B-spline storage
    (Simp banded-packed, BandNH, Band, Herm = Band pairs with halved diagonal) and every
    differentiation, with post = 1 and post > 1 (a middle rank), plus accumulation."""
    from pyprop_amd import device as D
    from pyprop_amd.bspline import (BSplineBasis, banded_hermitian_pairs, banded_nonhermitian_pairs,
                                    banded_packed_pairs, linear_breakpoints)
    from pyprop_amd.potgen import BSplineProjection
    import oracle
    from pyprop_amd import _native as N
    say("device one-rank projection, all storages")
    b = BSplineBasis(7, linear_breakpoints(0.0, 12.0, 15) ** 1.2, "zero")
    n = b.number_of_bsplines
    cases = [("Simp", banded_packed_pairs(n, 7), False), ("BandNH", banded_nonhermitian_pairs(n, 6), False),
             ("Band", banded_hermitian_pairs(n, 7), False), ("Herm", banded_hermitian_pairs(n, 7), True)]
    for name, pairs, herm in cases:
        for d in (0, 1, 2):
            proj = BSplineProjection(b, pairs, d, herm)
            for pre, post in ((37, 1), (5, 23)):
                F = rng.standard_normal((pre, b.grid.size, post))
                ref = oracle.bspline_represent(b, F, d, pairs, herm, axis=1)
                src = torch.from_numpy(F).to("cuda")
                out = torch.empty((pre, pairs.shape[0], post), dtype=torch.float64, device="cuda")
                N.check(N.load().pp_bspline_represent(proj.h, pre, post, D.ptr(src), D.ptr(out), 0, 0,
                                                      D.stream_handle()))
                got = out.cpu().numpy()
                assert _rel(got, ref) < 1e-14, (name, d, pre, post)
                outc = torch.full((pre, pairs.shape[0], post), 1.0 + 2.0j, dtype=torch.complex128, device="cuda")
                N.check(N.load().pp_bspline_represent(proj.h, pre, post, D.ptr(src), D.ptr(outc), 1, 1,
                                                      D.stream_handle()))
                gc = outc.cpu().numpy()
                assert np.abs(gc.real - (ref + 1.0)).max() <= 1e-14 * max(1.0, np.abs(ref).max())
                assert np.all(gc.imag == 2.0)
    say("  ok")


def test_device_generation_full_helium_vs_factors(torch):
    """Full helium (42 x 60 x 60): every consolidated potential from the grid on the
    device against the device expansion of the separable factors; a Hamiltonian multiply
    with each."""
    from pyprop_amd.propagator import BasisPropagator, CombinedRepresentation, Wavefunction
    from pyprop_amd.potgen import generate_device_potential
    from pyprop_amd.tensorpotential import TensorPotential
    from pyprop_amd.systems import helium_system
    say("device potential generation, full helium")
    s = helium_system()
    for p in s.potentials:
        torch.cuda.synchronize()
        t0 = time.time()
        V = generate_device_potential(p)
        torch.cuda.synchronize()
        t1 = time.time()
        ref = TensorPotential(p, s.shape).device_potential()
        err = ((V - ref).abs().max() / ref.abs().max()).item()
        say("  %-60s %s: %.1f ms, rel err vs factors %.2e" % (p.name[:60], tuple(V.shape), 1e3 * (t1 - t0), err))
        assert err < 1e-12
        del V, ref
        torch.cuda.empty_cache()
    repr_ = CombinedRepresentation(s.shape, s.bases)
    psi = Wavefunction(repr_)
    x = torch.randn(s.shape, dtype=torch.complex128, device="cuda")
    outs = []
    for gen in ("factors", "grid"):
        base = BasisPropagator(psi, s.potentials, generation=gen)
        base.SetupStep()
        y = torch.empty_like(x)
        base.MultiplyHamiltonianNoOverlap(x, y, 100.0)
        outs.append(y)
        del base
        torch.cuda.empty_cache()
    err = ((outs[0] - outs[1]).abs().max() / outs[0].abs().max()).item()
    say("  H x, grid-generated vs factor potentials: rel %.2e" % err)
    assert err < 1e-12
