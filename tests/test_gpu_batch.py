"""GPU tests of batched pAMP (pyprop_amd.batch): every atom of a mixed-size batch against
the oracle's balanced pAMP step (FP64 multiply, converged Krylov parameters: 1e-10),
FP32-multiply error against FP64, graph replay against direct launches (bitwise), the
autotuner restoring the state, and the 64-atom random batch conserving the S-norm."""
import sys

import numpy as np
import pytest

from tests.test_oracle import oracle_system

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


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _small_specs():
    from pyprop_amd.batch import AtomSpec
    from pyprop_amd.systems import LaserSimpleLength
    return [AtomSpec(2, 30, LaserSimpleLength(0.05, 0.1, 2.0), rmax=30.0, n_inner=8),
            AtomSpec(4, 41, LaserSimpleLength(0.08, 0.2, 3.0), rmax=35.0, n_inner=8),
            AtomSpec(1, 25, LaserSimpleLength(0.02, 0.05, 2.0), rmax=25.0, n_inner=8)]


def _start(batch, rng):
    psis = []
    for sh in batch.shapes:
        x = rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
        psis.append(x / np.linalg.norm(x))
    times = np.array([0.3 * s.pulse.pulse_duration for s in batch.specs])
    batch.set_state(psis, times)
    return psis, times


def test_batch_fp64_vs_oracle(torch, rng):
    from pyprop_amd.batch import BatchedAtoms
    say("batched pAMP (FP64 multiply) vs oracle, 3 atoms, m=30 dt=0.002, 3 steps")
    specs = _small_specs()
    b = BatchedAtoms(specs, krylov_basis_size=30, fp32_matvec=False)
    psis, times = _start(b, rng)
    dt = 0.002
    b.step(dt, 3)
    got, t_end, err = b.get_state()
    for a, (sysm, spec) in enumerate(zip(b.systems, specs)):
        s = oracle_system(sysm)
        ref = psis[a].copy()
        t = times[a]
        for k in range(3):
            ref, est = s.pamp_step(ref, dt, [1.0, spec.pulse(t)], m=30)
            t += dt
        e = rel(got[a], ref)
        say("  atom %d %s: rel %.2e" % (a, sysm.shape, e))
        assert e < 1e-10
        assert abs(t_end[a] - t) < 1e-12


@pytest.mark.parametrize("start", ["ground", "random"])
def test_batch_fp32_multiply_error(torch, rng, start):
    """FP32-multiply error against the FP64 multiply after 10 steps: below 1e-6 from the
    1s ground state (the workload's start), below 1e-4 from random vectors, whose weight on
    the top of the radial spectrum (~6e6 a.u.) amplifies FP32 rounding."""
    from pyprop_amd.batch import BatchedAtoms
    say("batched pAMP FP32 multiply vs FP64, 10 steps, %s start" % start)
    specs = _small_specs()
    b64 = BatchedAtoms(specs, fp32_matvec=False)
    b32 = BatchedAtoms(specs, fp32_matvec=True)
    if start == "random":
        psis, times = _start(b64, rng)
    else:
        psis = b64.psi0
        times = np.array([0.3 * s.pulse.pulse_duration for s in specs])
        b64.set_state(psis, times)
    b32.set_state(psis, times)
    b64.step(0.01, 10)
    b32.step(0.01, 10)
    s64, _, _ = b64.get_state()
    s32, _, _ = b32.get_state()
    for a in range(len(specs)):
        e = rel(s32[a], s64[a])
        say("  atom %d: FP32-multiply error %.2e" % (a, e))
        assert e < (1e-6 if start == "ground" else 1e-4)


def test_batch_graph_replay_equals_direct(torch, rng):
    from pyprop_amd.batch import BatchedAtoms
    say("batched pAMP: hipGraph replay vs direct launches")
    specs = _small_specs()
    b1 = BatchedAtoms(specs)
    b2 = BatchedAtoms(specs)
    psis, times = _start(b1, rng)
    b2.set_state(psis, times)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        b1.step(0.01, 6)
        b2.capture(0.01, 3)
        b2.replay(1)
        torch.cuda.synchronize()
        b2.replay(1)
    torch.cuda.synchronize()
    s1, t1, _ = b1.get_state()
    s2, t2, _ = b2.get_state()
    for a in range(len(specs)):
        assert np.array_equal(s1[a], s2[a])
    assert np.array_equal(t1, t2)


def test_batch_autotune_restores_state(torch, rng):
    from pyprop_amd.batch import BatchedAtoms
    say("batched pAMP autotuner")
    b = BatchedAtoms(_small_specs())
    psis, times = _start(b, rng)
    before, tb, _ = b.get_state()
    best, tim = b.autotune(0.01, trial_steps=2)
    after, ta, _ = b.get_state()
    say("  best %d, ms/step %s" % (best, tim))
    assert best in (64, 128, 256, 512, 1024) and all(v > 0 for v in tim.values())
    assert b.info()["threads"] == best
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    assert np.array_equal(tb, ta)


def test_batch_64_random_atoms_conserve_norm(torch):
    from pyprop_amd.batch import BatchedAtoms, random_specs
    say("64 random atoms, 20 FP32-multiply steps: S-norm conservation")
    specs = random_specs(64, seed=2024)
    b = BatchedAtoms(specs)
    b.step(0.01, 20)
    got, t, err = b.get_state()
    for a, sysm in enumerate(b.systems):
        S = sysm.bases[1].overlap_band()
        nrm = np.einsum("li,ij,lj->", got[a].conj(), S, got[a]).real
        assert np.all(np.isfinite(got[a]))
        assert abs(nrm - 1.0) < 1e-5, (a, nrm)
    assert np.allclose(t, 0.2)


def test_batch_1000_steps_vs_oracle(torch):
    """The configured run length: 1000 steps (Krylov 20, dt = 0.01) from the 1s ground state
    at t = 0 through each atom's pulse, three atoms, steps captured 10 per graph.
    FP64 multiply (the batch default) against the oracle's pAMP on the same steps.  The
    m = 20 expansion is unconverged by design at the top of the radial spectrum, so rounding
    grows over the run; the bar is 100x what the oracle itself shows between two runs whose
    start differs by 1e-15 (relative), and at most 1e-6.
    FP32 multiply: its per-atom error against the oracle is reported and bounded by 1e-2
    (the reason FP64 is the default)."""
    from pyprop_amd.batch import BatchedAtoms
    say("batched pAMP, 1000 steps from the ground state vs oracle (FP64 and FP32 multiply)")
    specs = _small_specs()
    b64 = BatchedAtoms(specs)
    b32 = BatchedAtoms(specs, fp32_matvec=True, packs=b64.packs)
    steps, dt = 1000, 0.01
    stream = torch.cuda.Stream()  # graph capture needs a non-default stream
    with torch.cuda.stream(stream):
        for b in (b64, b32):
            b.capture(dt, 10)
            b.replay(steps // 10)
    torch.cuda.synchronize()
    s64, t64, _ = b64.get_state()
    s32, _, _ = b32.get_state()
    rng = np.random.default_rng(9)
    for a, (sysm, spec) in enumerate(zip(b64.systems, specs)):
        s = oracle_system(sysm)
        x0 = b64.psi0[a]
        pert = x0 + 1e-15 * np.linalg.norm(x0) * (rng.standard_normal(x0.shape) + 1j * rng.standard_normal(x0.shape)) \
            / np.sqrt(2 * x0.size)
        refs = []
        for start in (x0, pert):
            ref, t = start.copy(), 0.0
            for k in range(steps):
                ref, _ = s.pamp_step(ref, dt, [1.0, spec.pulse(t)], m=20)
                t += dt
            refs.append(ref)
        sens = rel(refs[1], refs[0])
        e64 = rel(s64[a], refs[0])
        e32 = rel(s32[a], refs[0])
        say("  atom %d %s: FP64 %.2e (oracle sensitivity %.2e), FP32 %.2e" % (a, sysm.shape, e64, sens, e32))
        assert abs(t64[a] - t) < 1e-9
        assert e64 < min(1e-6, max(100 * sens, 1e-12))
        assert e32 < 1e-2
