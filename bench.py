"""Helium rank-3 Cayley/Crank-Nicolson propagation on one MI355X.

Workload: helium in coupled spherical harmonics (L = 0..2, M = 0, l1, l2 <= 5: 42
channels) x 60 order-7 B-splines per electron on [0, 40] (N = 151,200 complex128);
H = H0 + 1/r12 (multipoles lambda <= 10) + length-gauge dipole E(t), stored as
Simp x banded-packed x banded-nonhermitian tensor potentials; dt = 0.01, E0 = 0.05,
omega = 0.057, sin^2 envelope of 4 cycles; start from the independent-particle 1s^2
state.  Data are synthetic (generated here), nothing is downloaded.

Three solvers of the same Cayley system are timed on the same steps:
  plain FP64      GMRES(30, 1e-10) with the exact block-diagonal preconditioner
  FP32 + Kron     complex64 flexible GMRES with the Kronecker preconditioner inside FP64
                  iterative refinement to |b - A x| <= 1e-10 |b|
  FP32 + learned  the same with the learned correction (trained here first, on pairs
                  from the plain solver under seeded random pulses)
`value` is the learned + FP32 rate; the errors are against the plain FP64 trajectory.

Prints one JSON line (see --help).  A "step" is one Cayley AdvanceStep.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def build_workload(torch):
    import numpy as np
    from pyprop_amd.propagator import (BasisPropagator, CayleyPropagator, CombinedRepresentation,
                                       RadialTwoElectronPreconditioner, Wavefunction)
    from pyprop_amd.systems import helium_initial_state, helium_system
    t0 = time.time()
    sysm = helium_system()
    log("helium set-up (host) %.1fs shape=%s" % (time.time() - t0, sysm.shape))
    repr_ = CombinedRepresentation(sysm.shape, sysm.bases)
    psi = Wavefunction(repr_, helium_initial_state(sysm))
    psi.Normalize()
    # the potentials are generated on the device from their grid functions (GeneratePotential)
    base = BasisPropagator(psi, sysm.potentials, generation="grid")
    t0 = time.time()
    base.SetupStep(0.01)
    torch.cuda.synchronize()
    log("device potentials (grid + projection) + plans %.1fs" % (time.time() - t0))
    return sysm, repr_, psi, base


def time_potgen(torch, sysm, reps=3):
    """Device generation (grid evaluation + two B-spline projections) of every helium
    potential: wall time per potential with HIP events on torch's current stream (the
    library's stream), after one generation that builds the projection tables."""
    from pyprop_amd.potgen import generate_device_potential
    out = {}
    for p in sysm.potentials:
        V = generate_device_potential(p)
        del V
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(reps):
            V = generate_device_potential(p)
            del V
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        P0, P1, P2 = p.shape
        Q1, Q2 = (g.basis.grid.size for g in p.geometries[1:])
        terms = len(p.grid_terms)
        # per grid term: grid write + read, intermediate (P0, Q1, P2) write + read, complex result
        # write (read + write when accumulating)
        nbytes = terms * (2 * 8 * P0 * Q1 * Q2 + 2 * 8 * P0 * Q1 * P2) + (2 * terms - 1) * 16 * P0 * P1 * P2
        out[p.name.split("+")[0]] = {"ms": ms, "shape": [P0, P1, P2], "grid_terms": terms,
                                     "bytes": nbytes, "GB_per_s": nbytes / ms / 1e6}
        log("device generation %-40s %s: %.2f ms, %.1f GB/s" % (p.name[:40], p.shape, ms, nbytes / ms / 1e6))
        torch.cuda.empty_cache()
    return out


def profiled_traffic():
    """HBM bytes per launch of the fused multiply from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = ROOT / "profiles" / "tp_dia_kernel_helium_fp64_matvec.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def time_matvec(torch, base, psi, reps=20):
    """Average duration of the fused Hamiltonian multiply (one kernel) with HIP events on
    the stream the library launches on (torch's current stream)."""
    src = psi.data.clone()
    dst = torch.zeros_like(src)
    for _ in range(3):
        base.MultiplyHamiltonianNoOverlap(src, dst, 100.0)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        base.MultiplyHamiltonianNoOverlap(src, dst, 100.0)
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps / 1e3


def time_stream_copy(torch, nbytes=1 << 30, reps=20):
    """Device-to-device copy rate (read + write bytes / time) of torch's copy kernel on a
    1 GiB buffer: the practical HBM ceiling that the multiply's rate is read against."""
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        dst.copy_(src)
    e1.record(st)
    e1.synchronize()
    t = e0.elapsed_time(e1) / reps / 1e3
    del src, dst
    torch.cuda.empty_cache()
    return 2 * nbytes / t / 1e9


def tuned_factored(base):
    """The factored operator with its radial kernel shape tuned for both precisions."""
    from pyprop_amd.factored import FactoredHamiltonian
    fh = FactoredHamiltonian(base)
    tune = {}
    for prec in (64, 32):
        best, tim = fh.tune(prec, trials=5)
        tune["fp%d" % prec] = {"best": fh.RADIAL_SHAPES[best],
                               "us": {k: round(1e3 * v, 1) for k, v in tim.items()}}
        tune["fp%d" % prec]["angular"] = fh.ANGULAR_VARIANTS[fh.angular_variant(prec)]
        tune["fp%d" % prec]["angular_us"] = {str(k): round(1e3 * v, 1) for k, v in fh.angular_timings.items()}
        log("factored radial tune FP%d: best %s  (%s); angular %s (%s)" % (
            prec, tune["fp%d" % prec]["best"], ", ".join("%s %.1f us" % (k, v) for k, v in tune["fp%d" % prec]["us"].items()),
            tune["fp%d" % prec]["angular"], ", ".join("%s: %.1f us" % (k, 1e3 * v) for k, v in fh.angular_timings.items())))
    return fh, tune


def time_factored(torch, fh, psi, dtype, reps=50):
    """Average duration of the factored multiply (angular + radial launch) in `dtype`."""
    src = psi.data.to(dtype).clone()
    dst = torch.zeros_like(src)
    for _ in range(3):
        fh.Multiply(src, dst, 100.0)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        fh.Multiply(src, dst, 100.0)
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps / 1e3


def factored_flops(fh, nc, n, precision):
    """(issued MFMA flops incl. padding, useful flops) of one factored multiply in the radial
    and angular variants chosen for `precision`.  Radial stage: band-tile variants with
    register reuse issue one 16x16x4 MFMA per (band-tile term, i1, jj, i2 tile, channel
    group, K step); the position variants ("pos<g>") one per (16 band terms, position,
    channel group, K step of the w x w window), channel groups rounded up to a multiple of
    g.  The separable terms run as two 1-D band passes on the vector units and are not
    counted.  Angular stage: two MFMAs (re, im) per (16 channels, 16 positions, K step of 4),
    each wave's share of the K steps rounded up to whole batches (12 steps, 4 in the
    all-rows form).  At helium these are the SQ_INSTS_MFMA counts the hardware reports
    (band tiles 1,487,808, position form 928,800, angular 259,200 per multiply)."""
    T, bw = fh.nterms, fh.bw
    w = 2 * bw + 1
    pairs1 = sum(min(n, i + bw + 1) - max(0, i - bw) for i in range(n))
    groups = math.ceil(nc / 8)
    name = fh.RADIAL_SHAPES[fh.radial_variant(precision)]
    if name.startswith("pos"):
        g = int(name[3])
        rad_mfma = math.ceil(fh.band_terms / 16) * n * n * math.ceil(groups / g) * g * ((w * w + 3) // 4)
    else:
        ks = (16 + 2 * bw + 3) // 4
        rad_mfma = fh.band_terms * pairs1 * ks * math.ceil(n / 16) * groups
    av = fh.ANGULAR_VARIANTS[fh.angular_variant(precision)]
    waves = int(av.replace("rows", ""))
    batch = 4 if av.startswith("rows") else 12
    nks = (T * nc + 3) // 4
    per = math.ceil(nks / waves)
    ang_steps = sum(math.ceil(max(0, min(nks, k0 + per) - k0) / batch) * batch
                    for k0 in range(0, waves * per, per))
    ang_mfma = math.ceil(n * n / 16) * math.ceil(nc / 16) * ang_steps * 2
    nnz_r = pairs1 * pairs1
    useful = T * nnz_r * 4 * nc + 2 * nc * T * nc * 2 * n * n
    return (rad_mfma + ang_mfma) * 2048, useful


def bench_batch(torch, steps=1000, per_graph=10, natoms=64, seed=2024, dt=0.01):
    """64 independent H atoms (N ~ U[80,200], lmax ~ U[4,15], seeded pulses): one
    persistent launch per pAMP step (Krylov 20, complex64 multiply, FP64 state), steps
    captured in a hipGraph, workgroup size from the autotuner.  The same steps with the FP64
    multiply give the error of the FP32 multiply."""
    import numpy as np
    from pyprop_amd.batch import BatchedAtoms, random_specs
    specs = random_specs(natoms, seed=seed)
    t0 = time.time()
    b32 = BatchedAtoms(specs, fp32_matvec=True)
    b64 = BatchedAtoms(specs, fp32_matvec=False, packs=b32.packs)
    log("batch: %d atoms set up (%.1fs), %d complex elements" % (natoms, time.time() - t0, b32.elements()))
    best, tim = b32.autotune(dt, trial_steps=3)
    b64.set_threads(best)
    stream = torch.cuda.Stream()
    res = {}
    for name, b in (("fp32", b32), ("fp64", b64)):
        with torch.cuda.stream(stream):
            b.capture(dt, per_graph)
            b.replay(1)  # warm-up
            torch.cuda.synchronize()
            b.set_state(b.psi0, np.zeros(natoms))
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            b.replay(steps // per_graph)
            torch.cuda.synchronize()
            res[name] = time.perf_counter() - t1
    s32, tt, _ = b32.get_state()
    s64, _, _ = b64.get_state()
    err = max(float(np.linalg.norm(x - y) / np.linalg.norm(y)) for x, y in zip(s32, s64))
    drift = 0.0
    for a, sysm in enumerate(b32.systems):
        S = sysm.bases[1].overlap_band()
        drift = max(drift, abs(float(np.einsum("li,ij,lj->", s32[a].conj(), S, s32[a]).real) - 1.0))
    sec = res["fp32"]
    out = {"atoms": natoms, "steps": steps, "krylov": 20, "dt": dt, "threads": best,
           "autotune_ms_per_step": tim, "elements": b32.elements(), "seconds": sec,
           "steps_per_s": steps / sec, "atom_steps_per_s": natoms * steps / sec,
           "matvecs_per_s": natoms * steps * 20 / sec, "fp64_multiply_steps_per_s": steps / res["fp64"],
           "fp32_vs_fp64_max_rel_error": err, "max_norm_drift": drift, "t_end": float(tt.max())}
    log("batch: %.1f steps/s (%.0f atom-steps/s, %.2e matvecs/s), threads %d, FP32 vs FP64 %.2e, norm drift %.1e" %
        (out["steps_per_s"], out["atom_steps_per_s"], out["matvecs_per_s"], best, err, drift))
    return out


def host_threads():
    """Host threads for the all-cores baseline: the CPU share given to this process (the
    OMP_NUM_THREADS convention when set, else the affinity mask)."""
    import os
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(sysm, n_ops_per_step, sample_frac=0.1, threads=1, reps=3):
    """The oracle (sequential C++ restatement of the reference loops) timed on the host:
    the tensor-potential multiply on `sample_frac` of every potential's angular pairs and
    the block-preconditioner solve of two channels, scaled to one Cayley step (n_ops
    operator applications, each one multiply + one preconditioner solve).  With threads > 1
    every thread runs the same kind of sample on its own slice of the angular pairs and its
    own two channels at the same time (the calls release the GIL), the reference's
    one-process-per-core model: the step time is the wall time per sample unit divided by
    the number of threads.  Each sample runs `reps` times; the median wall time counts."""
    import threading
    import numpy as np
    import oracle
    from pyprop_amd.systems import _expand
    rng = np.random.default_rng(3)
    nc, n1, n2 = sysm.shape
    psi = rng.standard_normal(sysm.shape) + 1j * rng.standard_normal(sysm.shape)

    def run_parallel(fns):
        ths = [threading.Thread(target=f) for f in fns]
        t0 = time.perf_counter()
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        return time.perf_counter() - t0

    # multiply: thread i takes pairs [i k, (i + 1) k) of every potential (wrapping around)
    t_mul = 0.0
    for p in sysm.potentials:
        P0 = p.geometries[0].pair_count
        k = max(1, int(P0 * sample_frac))
        jobs = []
        for i in range(threads):
            sel = (np.arange(k) + i * k) % P0
            V = _expand(p.angular_coeffs[:, sel], p.radial_factors)
            pairs0 = np.ascontiguousarray(p.geometries[0].pairs[sel])
            dst = np.zeros(sysm.shape, complex)
            jobs.append(lambda V=V, pairs0=pairs0, dst=dst, p=p: oracle.tpmul(
                p.storage_ids, V, 1.0, psi, dst, [pairs0, p.geometries[1].pairs, None]))
        wall = float(np.median([run_parallel(jobs) for _ in range(reps)]))
        t_mul += wall * P0 / (k * threads)
    # preconditioner: each thread factors (not timed) and solves (timed) two channels
    b1 = sysm.bases[1]
    S1 = b1.overlap_band()
    g = sysm.precond_potentials[0].geometries
    P = sum(_expand(x.angular_coeffs[:, :2], x.radial_factors) for x in sysm.precond_potentials)
    systems, xs = [], []
    for i in range(threads):
        s = oracle.System((2, n1, n2))
        s.set_preconditioner(S1, S1, b1.order - 1, P, g[1].pairs, g[2].pairs, 0.005j)
        systems.append(s)
        xs.append(psi[:2].copy())
    wall = run_parallel([lambda s=s, x=x: [s.precond_solve(x) for _ in range(reps)] for s, x in zip(systems, xs)])
    t_pc = wall / reps * nc / (2 * threads)
    step = (n_ops_per_step + 1) * (t_mul + t_pc)
    return dict(value=1.0 / step, unit="steps/s", cores=threads, kind="port",
                sample="oracle multiply on %d%% of each potential's angular pairs per thread, median of %d runs "
                       "(full multiply %.3fs), and 2-channel preconditioner solves per thread, %d runs (full solve "
                       "%.3fs), %d thread(s) at once, scaled to one step of %.1f operator applications + right-hand "
                       "side" % (int(sample_frac * 100), reps, t_mul, reps, t_pc, threads, n_ops_per_step))


def run_steps(torch, prop, psi, t0, dt, warmup, steps, stats=None):
    """W untimed steps then K timed steps from t0; returns (seconds, t_end)."""
    t = t0
    for _ in range(warmup):
        prop.AdvanceStep(t, dt)
        t += dt
    torch.cuda.synchronize()
    tstart = time.perf_counter()
    for _ in range(steps):
        prop.AdvanceStep(t, dt)
        if stats is not None:
            stats.append(prop.GetStats() if hasattr(prop, "GetStats") else {"ops": prop.GetOperatorCount()})
        t += dt
    torch.cuda.synchronize()
    return time.perf_counter() - tstart, t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-atom batched pAMP section")
    ap.add_argument("--profile-batch", type=int, default=0, help="only run this many batched steps (for rocprof)")
    ap.add_argument("--train-trajectories", type=int, default=64)
    ap.add_argument("--train-steps", type=int, default=24)
    ap.add_argument("--train-epochs", type=int, default=30)
    ap.add_argument("--profile-matvec", action="store_true", help="only run the matvec loop (for rocprof)")
    ap.add_argument("--profile-copy", action="store_true",
                    help="only copy 1 GiB device to device 20 times (HBM counter calibration for rocprof)")
    ap.add_argument("--profile-cayley", type=int, default=0, help="only run this many plain Cayley steps (for rocprof)")
    ap.add_argument("--profile-factored", action="store_true", help="only time the factored multiply (for rocprof)")
    ap.add_argument("--profile-potgen", action="store_true",
                    help="only time the device generation of the helium potentials (for rocprof)")
    ap.add_argument("--profile-learned", type=int, default=0,
                    help="only train the network on a few trajectories, then run this many mixed steps with the "
                         "learned preconditioner (for rocprof)")
    ap.add_argument("--profile-mixed", type=int, default=0,
                    help="only run this many mixed-precision steps with the Kronecker preconditioner (for rocprof)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: pyprop_amd has no multi-GPU path "
              "(the helium problem would shard by angular channel, one slice per GPU)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    assert torch.cuda.is_available(), "bench needs a GPU"
    from pyprop_amd.mixed import (KroneckerPreconditioner, LearnedPreconditioner, MixedCayleyPropagator,
                                  generate_training_pairs)
    from pyprop_amd.factored import FactoredHamiltonian
    from pyprop_amd.propagator import CayleyPropagator, RadialTwoElectronPreconditioner, Wavefunction

    if a.profile_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(20):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copied 20 x 1 GiB (read 1 GiB + write 1 GiB per copy kernel)")
        return
    if a.profile_batch:
        bench_batch(torch, steps=a.profile_batch, per_graph=min(10, a.profile_batch))
        return
    if a.profile_potgen:
        from pyprop_amd.systems import helium_system
        time_potgen(torch, helium_system())
        return
    sysm, repr_, psi, base = build_workload(torch)
    nnz = sum(int(np.prod(p.source.shape)) for p in base.PotentialList)
    N = sysm.size
    alg_bytes = 8 * nnz + 2 * 16 * N  # real float64 potential elements + src read + dst write
    dt = 0.01
    if a.profile_matvec:
        t = time_matvec(torch, base, psi, reps=50)
        log("matvec %.3f ms, %.1f GB/s" % (t * 1e3, alg_bytes / t / 1e9))
        return
    if a.profile_cayley:
        pc = RadialTwoElectronPreconditioner(psi, sysm)
        prop = CayleyPropagator(psi, base, pc, krylov_basis_size=30, krylov_tolerance=1e-10)
        prop.SetupStep(dt)
        sec, _ = run_steps(torch, prop, psi, 0.0, dt, 0, a.profile_cayley)
        log("%d Cayley steps %.2f ms/step, %d ops" % (a.profile_cayley, 1e3 * sec / a.profile_cayley,
                                                      prop.GetOperatorCount()))
        return
    if a.profile_factored:
        fh, _ = tuned_factored(base)
        for dtype in (torch.complex128, torch.complex64):
            t = time_factored(torch, fh, psi, dtype, reps=20)
            log("factored %s %.1f us" % (dtype, t * 1e6))
        return
    if a.profile_learned:
        pc = RadialTwoElectronPreconditioner(psi, sysm)
        plain = CayleyPropagator(psi, base, pc, krylov_basis_size=30, krylov_tolerance=1e-10)
        plain.SetupStep(dt)
        psi0 = psi.data.clone()
        kron = KroneckerPreconditioner(psi, sysm)
        learned = LearnedPreconditioner(kron, hidden=512, seed=1)
        x, tt, ch = generate_training_pairs(learned, base, plain, psi, dt, 4, 4, seed=2024)
        learned.fit(x, ch, tt, epochs=2, batch=256, lr=1e-3, seed=1)
        psi.data.copy_(psi0)
        prop = MixedCayleyPropagator(psi, base, kron, learned=learned, tolerance=1e-10,
                                     factored=tuned_factored(base)[0])
        prop.SetupStep(dt)
        sec, _ = run_steps(torch, prop, psi, 0.0, dt, 2, a.profile_learned)
        log("%d learned mixed steps %.2f ms/step, last %s" % (a.profile_learned, 1e3 * sec / a.profile_learned,
                                                              prop.GetStats()))
        return
    if a.profile_mixed:
        prop = MixedCayleyPropagator(psi, base, KroneckerPreconditioner(psi, sysm), tolerance=1e-10,
                                     factored=tuned_factored(base)[0])
        prop.SetupStep(dt)
        sec, _ = run_steps(torch, prop, psi, 0.0, dt, 2, a.profile_mixed)
        log("%d mixed steps %.2f ms/step, last %s" % (a.profile_mixed, 1e3 * sec / a.profile_mixed, prop.GetStats()))
        return

    t_mv = time_matvec(torch, base, psi)
    achieved = alg_bytes / t_mv / 1e9
    log("fused H multiply: %.3f ms  alg %.3f GB  -> %.1f GB/s (%.1f%% of 8 TB/s)" %
        (t_mv * 1e3, alg_bytes / 1e9, achieved, 100 * achieved / HBM_PEAK_GBS))
    copy_gbs = time_stream_copy(torch)
    log("device copy (1 GiB): %.1f GB/s; the multiply reaches %.1f%% of it" % (copy_gbs, 100 * achieved / copy_gbs))
    fh, fh_tune = tuned_factored(base)
    nc, n, _ = sysm.shape
    t_f64 = time_factored(torch, fh, psi, torch.complex128)
    t_f32 = time_factored(torch, fh, psi, torch.complex64)
    issued64, useful = factored_flops(fh, nc, n, 64)
    issued32, _ = factored_flops(fh, nc, n, 32)
    fact = {"fp64_ms": t_f64 * 1e3, "fp32_ms": t_f32 * 1e3, "fp64_matvecs_per_s": 1 / t_f64,
            "fp32_matvecs_per_s": 1 / t_f32, "fp64_mfma_flops_issued": issued64,
            "fp32_mfma_flops_issued": issued32, "useful_flops": useful,
            "fp64_mfma_util": issued64 / t_f64 / 78.6e12, "fp32_mfma_util": issued32 / t_f32 / 157.3e12,
            "radial_tune": fh_tune}
    log("factored multiply: FP64 %.1f us (%.1f%% of FP64 MFMA peak), FP32 %.1f us (%.1f%% of FP32 MFMA peak)" %
        (t_f64 * 1e6, 100 * fact["fp64_mfma_util"], t_f32 * 1e6, 100 * fact["fp32_mfma_util"]))
    psi0 = psi.data.clone()

    # 1. plain FP64 path: GMRES(30, 1e-10) with the exact block preconditioner
    t0 = time.time()
    pc = RadialTwoElectronPreconditioner(psi, sysm)
    plain = CayleyPropagator(psi, base, pc, krylov_basis_size=30, krylov_tolerance=1e-10)
    plain.SetupStep(dt)
    torch.cuda.synchronize()
    log("block preconditioner factorization + Cayley set-up %.1fs" % (time.time() - t0))
    norm0 = psi.GetNorm()
    ops = []
    sec_plain, t_end = run_steps(torch, plain, psi, 0.0, dt, a.warmup, a.steps, ops)
    plain_sps = a.steps / sec_plain
    ref = psi.data.clone()
    e_plain = base.GetEnergyExpectationValue(psi, t_end)
    n_plain = psi.GetNorm()
    ops_plain = float(np.mean([o["ops"] for o in ops]))
    log("plain FP64 Cayley: %.2f steps/s (%.2f ms/step), %.1f ops/step, |norm drift| %.2e, E = %.8f" %
        (plain_sps, 1e3 * sec_plain / a.steps, ops_plain, abs(n_plain - norm0), e_plain))
    # the same plain FP64 solver with the factored multiply in place of the tensor potentials
    psi.data.copy_(psi0)
    plain.SetFactoredHamiltonian(fh)
    sec_pf, _ = run_steps(torch, plain, psi, 0.0, dt, a.warmup, a.steps)
    plainf_sps = a.steps / sec_pf
    plainf_err = (torch.linalg.norm(psi.data - ref) / torch.linalg.norm(ref)).item()
    log("plain FP64 Cayley, factored multiply: %.2f steps/s (%.2f ms/step), err vs plain %.2e" %
        (plainf_sps, 1e3 * sec_pf / a.steps, plainf_err))

    # 2. learned correction: pairs from the plain solve under seeded random pulses, trained on the device
    psi.data.copy_(psi0)
    kron = KroneckerPreconditioner(psi, sysm)
    learned = LearnedPreconditioner(kron, hidden=512, seed=1)
    t0 = time.time()
    x, tt, ch = generate_training_pairs(learned, base, plain, psi, dt, a.train_trajectories, a.train_steps, seed=2024)
    torch.cuda.synchronize()
    t_gen = time.time() - t0
    t0 = time.time()
    loss = learned.fit(x, ch, tt, epochs=a.train_epochs, batch=256, lr=1e-3, seed=1)
    torch.cuda.synchronize()
    t_fit = time.time() - t0
    log("learned: %d training rows from %d plain steps (%.1fs), %d epochs (%.1fs), loss %.3e -> %.3e" %
        (x.shape[0], a.train_trajectories * a.train_steps, t_gen, a.train_epochs, t_fit,
         learned.loss_history[0], learned.loss_history[-1]))
    n_rows = int(x.shape[0])
    del x, tt, ch

    def mixed_run(with_learned):
        psi.data.copy_(psi0)
        prop = MixedCayleyPropagator(psi, base, kron, learned if with_learned else None, fallback=plain,
                                     tolerance=1e-10, factored=fh)
        prop.SetupStep(dt)
        st = []
        sec, t_e = run_steps(torch, prop, psi, 0.0, dt, a.warmup, a.steps, st)
        err = (torch.linalg.norm(psi.data - ref) / torch.linalg.norm(ref)).item()
        e = base.GetEnergyExpectationValue(psi, t_e)
        nn = psi.GetNorm()
        agg = {k: float(np.mean([s[k] for s in st])) for k in ("outer", "inner", "fp64_ops", "fp32_ops")}
        agg["fallbacks"] = int(sum(s["fallback"] for s in st))
        agg["max_rel_residual"] = float(max(s["rel_residual"] for s in st))
        return a.steps / sec, sec, err, abs(nn - n_plain), abs(e - e_plain), agg

    # what the network gains: residual left after one preconditioned direction on held-out
    # steps (the benchmark pulse, not seen in training), Kronecker alone vs with the network
    def first_direction_gain(nsteps=6):
        from pyprop_amd import _native as NN, device as DD
        psi.data.copy_(psi0)
        cur = psi.data.clone()
        h = torch.empty_like(cur)
        res = {"kron": [], "learned": []}
        tt = 0.0
        for _ in range(nsteps):
            sc = base.scalings(tt + dt / 2)
            NN.check(NN.load().pp_ham_multiply_nooverlap(base.ham, DD.hptr(sc), DD.ptr(cur), DD.ptr(h),
                                                         DD.stream_handle()))
            r = h * (-1j * dt)
            for name in ("kron", "learned"):
                r32 = r.to(torch.complex64)
                z = torch.empty_like(r32)
                if name == "learned":
                    learned.apply(r32, z, dt)
                else:
                    z.copy_(r32)
                    kron.SetHamiltonianScaling(0.5j * dt)
                    kron.Solve(z)
                z64 = z.to(torch.complex128)
                sz = z64.clone()
                repr_._op(0, False, sz)
                az = torch.empty_like(z64)
                NN.check(NN.load().pp_ham_multiply_nooverlap(base.ham, DD.hptr(sc), DD.ptr(z64), DD.ptr(az),
                                                             DD.stream_handle()))
                az = sz + 0.5j * dt * az
                al = torch.vdot(az.flatten(), r.flatten()) / torch.vdot(az.flatten(), az.flatten())
                res[name].append(float(torch.linalg.norm(r - al * az) / torch.linalg.norm(r)))
            plain.AdvanceStep(tt, dt)  # advance the reference state (psi aliases plain's state)
            cur.copy_(psi.data)
            tt += dt
        return {k: float(np.mean(v)) for k, v in res.items()}

    gain = first_direction_gain()
    log("first-direction residual on held-out steps: Kronecker %.3e, learned %.3e (%.1f%% lower)" %
        (gain["kron"], gain["learned"], 100 * (1 - gain["learned"] / gain["kron"])))

    kron_sps, kron_sec, kron_err, kron_dn, kron_de, kron_st = mixed_run(False)
    log("mixed FP32 + Kronecker: %.2f steps/s (%.2f ms/step), err vs plain %.2e, %s" %
        (kron_sps, 1e3 * kron_sec / a.steps, kron_err, kron_st))
    lrn_sps, lrn_sec, lrn_err, lrn_dn, lrn_de, lrn_st = mixed_run(True)
    log("mixed FP32 + learned: %.2f steps/s (%.2f ms/step), err vs plain %.2e, |dnorm| %.1e |dE| %.1e, %s" %
        (lrn_sps, 1e3 * lrn_sec / a.steps, lrn_err, lrn_dn, lrn_de, lrn_st))

    batch = None if a.no_batch else bench_batch(torch)
    potgen = time_potgen(torch, sysm)

    cpu = None
    cpu1 = None
    if not a.no_cpu_baseline:
        t0 = time.time()
        cpu1 = cpu_baseline(sysm, ops_plain, sample_frac=0.3)
        log("cpu baseline, 1 thread: %.4f steps/s (%.1fs)" % (cpu1["value"], time.time() - t0))
        nt = host_threads()
        t0 = time.time()
        cpu = cpu_baseline(sysm, ops_plain, threads=nt) if nt > 1 else cpu1
        log("cpu baseline, %d threads: %.4f steps/s (%.1fs)" % (nt, cpu["value"], time.time() - t0))

    out = {
        "metric": "CN steps/s, helium rank-3: learned+FP32 vs plain FP64 GPU at 1e-8 err; GB/s",
        "value": lrn_sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * lrn_sec / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 inner / c128 outer",
        "data": "synthetic (helium Hamiltonian generated in-process; independent-particle 1s^2 start; "
                "network trained in-process on plain-solve pairs from seeded random pulses)",
        "config": {"workload": "helium rank-3 Cayley-CN, 42 ch x 60 x 60 B-splines (N=151200), dt=0.01, "
                               "E0=0.05 w=0.057 4 cycles from t=0; learned+FP32: complex64 flexible GMRES(20) "
                               "with Kronecker + learned MLP (hidden 512, BF16 MFMA) preconditioner inside FP64 "
                               "refinement to |b-Ax| <= 1e-10 |b|, fallback to plain after 40 inner iterations; "
                               "all multiplies of the FP32 paths through the factored MFMA operator",
                   "plain_fp64_steps_per_s": plain_sps, "speedup_vs_plain_fp64": lrn_sps / plain_sps,
                   "plain_fp64_factored_steps_per_s": plainf_sps,
                   "speedup_vs_plain_fp64_factored": lrn_sps / plainf_sps,
                   "factored_multiply": fact,
                   "plain_fp64_ops_per_step": ops_plain,
                   "wavefunction_error_vs_plain_fp64": lrn_err, "norm_error": lrn_dn, "energy_error": lrn_de,
                   "learned_stats_per_step": lrn_st,
                   "first_direction_residual": gain,
                   "fp32_kron_steps_per_s": kron_sps, "fp32_kron_error": kron_err,
                   "fp32_kron_stats_per_step": kron_st,
                   "training": {"rows": n_rows, "plain_steps": a.train_trajectories * a.train_steps,
                                "epochs": a.train_epochs, "gen_s": t_gen, "fit_s": t_fit,
                                "loss_first": learned.loss_history[0], "loss_last": learned.loss_history[-1]},
                   "energy": e_plain, "norm_drift_plain": abs(n_plain - norm0),
                   "batched_h_atoms": batch,
                   "potential_generation": potgen,
                   "cpu_baseline_single_thread": cpu1},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(),
                     "traffic_source": "profiles/tp_dia_kernel_helium_fp64_matvec.md (rocprofv3 FETCH_SIZE + "
                                       "WRITE_SIZE, scaled by a 1 GiB copy)",
                     "kernel": "tp_dia_kernel (fused H0+R12+dipole FP64 multiply)", "ms": t_mv * 1e3,
                     "alg_bytes": alg_bytes, "copy_GBps": copy_gbs, "frac_of_copy": achieved / copy_gbs},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
