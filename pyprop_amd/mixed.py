"""Mixed-precision Cayley propagation with the Kronecker and learned preconditioners.

None of this has a counterpart in the reference; it solves the Cayley system of
pyprop/propagator/CayleyPropagator.py:71-150 faster:

  KroneckerPreconditioner   exact inverse of the separable part of each channel block of
                            S x S + i dt/2 H0 (a drop-in for RadialTwoElectronPreconditioner,
                            examples/tensor/helium_stabilization/preconditioner.py:101-198)
  LearnedPreconditioner     MLP correction on top of it, trained on the device from pairs
                            that the plain FP64 solve produces
  MixedCayleyPropagator     FP64 iterative refinement around complex64 flexible GMRES
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _native as N
from . import device as D
from .geometry import Geometry, dense_pairs
from .systems import LaserSimpleLength, System, radial_1d


def radial_eigenbasis(system: System):
    """Generalized eigenpairs of the one-electron radial Hamiltonians
    h_l = -1/2 d2/dr2 + l(l+1)/2r^2 - Z/r in the B-spline basis of the radial ranks:
    h_l E_l = S E_l diag(lam_l), E_l^T S E_l = I, for l = 0..lmax.
    Returns (E [nl][n][n], lam [nl][n], l1 [nchan], l2 [nchan], S [n][n])."""
    import scipy.linalg
    basis = system.bases[1]
    if system.bases[2] is not basis and system.bases[2].number_of_bsplines != basis.number_of_bsplines:
        raise ValueError("the Kronecker preconditioner needs the same radial basis for both electrons")
    n = basis.number_of_bsplines
    r = basis.grid
    g = Geometry("Simp", dense_pairs(n), True, n)
    Z = system.info["Z"]
    T = radial_1d(basis, g, np.full(r.size, -0.5), 2).reshape(n, n)
    Vc = radial_1d(basis, g, -Z / r, 0).reshape(n, n)
    Vl = radial_1d(basis, g, 1.0 / (2.0 * r * r), 0).reshape(n, n)
    S = basis.overlap_band()
    S = 0.5 * (S + S.T)
    l1 = np.array([ch.l1 for ch in system.channels], dtype=np.int32)
    l2 = np.array([ch.l2 for ch in system.channels], dtype=np.int32)
    nl = int(max(l1.max(), l2.max())) + 1
    E = np.zeros((nl, n, n))
    lam = np.zeros((nl, n))
    for l in range(nl):
        h = T + Vc + l * (l + 1) * Vl
        h = 0.5 * (h + h.T)
        w, v = scipy.linalg.eigh(h, S)
        E[l] = v
        lam[l] = w
    return E, lam, l1, l2, S


class KroneckerPreconditioner:
    """K_c = S x S + a (h_l1 x S + S x h_l2) per channel, inverted exactly through the
    radial eigenbases.  Duck-types the reference preconditioner interface
    (SetHamiltonianScaling / SetOverlapScaling / Setup / Solve, preconditioner.py:101-152);
    the overlap scaling must be 1."""

    kind = "kron"

    def __init__(self, psi, system: System):
        self.psi = psi
        self.system = system
        self.HamiltonianScaling = 0j
        self.OverlapScaling = 1.0
        self.h = None

    def SetHamiltonianScaling(self, scalingH):
        self.HamiltonianScaling = complex(scalingH)

    def SetOverlapScaling(self, scalingS):
        if complex(scalingS) != 1.0:
            raise ValueError("the Kronecker preconditioner supports overlap scaling 1 only")
        self.OverlapScaling = 1.0

    def Setup(self, prop=None):
        if self.h is not None:
            return
        E, lam, l1, l2, S = radial_eigenbasis(self.system)
        nc, n1, n2 = self.system.shape
        h = C.c_void_p()
        E = np.ascontiguousarray(E)
        lam = np.ascontiguousarray(lam)
        S = np.ascontiguousarray(S)
        N.check(N.load().pp_kron_create(nc, n1, E.shape[0], D.hptr(S), D.hptr(E), D.hptr(lam), D.hptr(l1),
                                        D.hptr(l2), C.byref(h)))
        self.h = h

    def Multiply(self, psi):
        '''psi <- K psi (complex128).'''
        self.Setup()
        t = getattr(psi, "data", psi)
        N.check(N.load().pp_kron_multiply(self.h, N.cplx(self.HamiltonianScaling), D.ptr(t), D.stream_handle()))

    def Solve(self, psi):
        self.Setup()
        t = getattr(psi, "data", psi)
        if t.dtype == D.torch.complex64:
            N.check(N.load().pp_kron_solve_f32(self.h, N.cplx(self.HamiltonianScaling), D.ptr(t), D.stream_handle()))
        else:
            N.check(N.load().pp_kron_solve(self.h, N.cplx(self.HamiltonianScaling), D.ptr(t), D.stream_handle()))

    def SolveInto(self, src, dst):
        """dst <- K^-1 src (complex64, distinct tensors)."""
        self.Setup()
        N.check(N.load().pp_kron_solve_f32_into(self.h, N.cplx(self.HamiltonianScaling), D.ptr(src), D.ptr(dst),
                                                D.stream_handle()))

    def SolveIntoWithOverlap(self, src, dst, sz):
        """dst <- K^-1 src and sz <- (S (x) S) dst, the latter formed in the same launch from the
        solve's middle product (one-launch form)."""
        self.Setup()
        N.check(N.load().pp_kron_solve_f32_overlap(self.h, N.cplx(self.HamiltonianScaling), D.ptr(src), D.ptr(dst),
                                                   D.ptr(sz), D.stream_handle()))

    FORMS = ("one-workgroup", "split", "fused")

    def SetForm(self, form: int):
        """The FP32 solve's kernel form (pp_kron_set_form): 0 one workgroup per channel,
        1 four per channel in two launches, 2 four per channel in one launch (default)."""
        self.Setup()
        N.check(N.load().pp_kron_set_form(self.h, int(form)))

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_kron_destroy(self.h)
            self.h = None


class LearnedPreconditioner:
    """BF16-MFMA MLP correction to the Kronecker preconditioner (see pp_learned_* in the
    C ABI).  Trained on the device by the library's own forward/backward kernels."""

    def __init__(self, kron: KroneckerPreconditioner, hidden: int = 512, seed: int = 1):
        kron.Setup()
        self.kron = kron
        nc, n1, n2 = kron.system.shape
        self.nchan, self.n, self.hidden = nc, n1, hidden
        self.inp = 2 * n1 * n1
        h = C.c_void_p()
        N.check(N.load().pp_learned_create(nc, n1, hidden, seed, C.byref(h)))
        self.h = h
        cnt = C.c_longlong()
        N.check(N.load().pp_learned_param_count(h, C.byref(cnt)))
        self.param_count = cnt.value
        self.sigma = 1.0
        self.loss_history = []

    # parameters -----------------------------------------------------------------------
    def layout(self):
        H, I, nc = self.hidden, self.inp, self.nchan
        sizes = [("W1", (H, I)), ("b1", (H,)), ("emb", (nc, H)), ("W2", (H, H)), ("b2", (H,)), ("W3", (I, H)),
                 ("b3", (I,))]
        out, o = {}, 0
        for name, shp in sizes:
            k = int(np.prod(shp))
            out[name] = (o, shp)
            o += k
        return out

    def get_params(self) -> np.ndarray:
        a = np.zeros(self.param_count, np.float32)
        N.check(N.load().pp_learned_get_params(self.h, D.hptr(a)))
        return a

    def set_params(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.size == self.param_count
        N.check(N.load().pp_learned_set_params(self.h, D.hptr(a)))

    def get_grads(self) -> np.ndarray:
        a = np.zeros(self.param_count, np.float32)
        N.check(N.load().pp_learned_get_grads(self.h, D.hptr(a)))
        return a

    def set_scale(self, sigma: float):
        self.sigma = float(sigma)
        N.check(N.load().pp_learned_set_scale(self.h, C.c_float(self.sigma)))

    # network --------------------------------------------------------------------------
    def _check_rows(self, x, chan, t=None):
        torch = D.torch
        if x.dtype != torch.float32 or x.dim() != 2 or x.shape[1] != self.inp or not x.is_contiguous():
            raise ValueError("x must be a contiguous float32 [rows][%d] device tensor" % self.inp)
        if chan.dtype != torch.int32 or chan.shape != (x.shape[0],):
            raise ValueError("chan must be an int32 [rows] device tensor")
        if t is not None and (t.dtype != torch.float32 or t.shape != x.shape or not t.is_contiguous()):
            raise ValueError("t must be a contiguous float32 tensor shaped like x")

    def forward(self, x, chan):
        self._check_rows(x, chan)
        y = D.torch.empty_like(x)
        N.check(N.load().pp_learned_forward(self.h, x.shape[0], D.ptr(x), D.ptr(chan), D.ptr(y), D.stream_handle()))
        return y

    def gradient(self, x, chan, t) -> float:
        self._check_rows(x, chan, t)
        loss = C.c_double()
        N.check(N.load().pp_learned_gradient(self.h, x.shape[0], D.ptr(x), D.ptr(chan), D.ptr(t), C.byref(loss),
                                             D.stream_handle()))
        return loss.value

    def train_step(self, x, chan, t, lr: float, want_loss: bool = False):
        self._check_rows(x, chan, t)
        loss = C.c_double()
        N.check(N.load().pp_learned_train_step(self.h, x.shape[0], D.ptr(x), D.ptr(chan), D.ptr(t), C.c_float(lr),
                                               C.byref(loss) if want_loss else None, D.stream_handle()))
        return loss.value if want_loss else None

    def apply(self, src, dst, dt: float):
        """dst = K^-1 (src + sigma |src| net(src_c / |src|)) (complex64 device tensors)."""
        N.check(N.load().pp_learned_apply(self.h, self.kron.h, float(dt), D.ptr(src), D.ptr(dst),
                                          D.stream_handle()))

    def prepare(self, r, d, dt: float):
        """Training rows from a solved system A d = r (complex128 psi-shaped)."""
        torch = D.torch
        x = torch.empty((self.nchan, self.inp), dtype=torch.float32, device="cuda")
        t = torch.empty_like(x)
        nrm = torch.empty(self.nchan, dtype=torch.float32, device="cuda")
        N.check(N.load().pp_learned_prepare(self.h, self.kron.h, float(dt), D.ptr(r), D.ptr(d), D.ptr(x), D.ptr(t),
                                            D.ptr(nrm), D.stream_handle()))
        return x, t, nrm

    # training -------------------------------------------------------------------------
    def fit(self, x, chan, t, epochs: int = 20, batch: int = 256, lr: float = 1e-3, seed: int = 0,
            log_every: int = 0):
        """Adam on sum |net(x) - t/sigma|^2 / rows over shuffled minibatches; sigma is set
        to the RMS of t.  Returns the final epoch's mean loss."""
        torch = D.torch
        rows = x.shape[0]
        sig = float(torch.sqrt(torch.mean(t.double() ** 2)).item())
        self.set_scale(max(sig, 1e-30))
        g = torch.Generator(device="cpu")
        g.manual_seed(seed)
        last = None
        for ep in range(epochs):
            perm = torch.randperm(rows, generator=g).to("cuda")
            tot, nb = 0.0, 0
            for b0 in range(0, rows - batch + 1, batch):
                idx = perm[b0:b0 + batch]
                xb = x.index_select(0, idx).contiguous()
                tb = t.index_select(0, idx).contiguous()
                cb = chan.index_select(0, idx).contiguous()
                want = (b0 + 2 * batch > rows)  # last batch of the epoch reports its loss
                l = self.train_step(xb, cb, tb, lr * (0.5 * (1 + math.cos(math.pi * ep / epochs))), want)
                if l is not None:
                    tot += l
                    nb += 1
            last = tot / max(nb, 1)
            self.loss_history.append(last)
            if log_every and (ep % log_every == 0 or ep == epochs - 1):
                print("[learned] epoch %d loss %.4e" % (ep, last), flush=True)
        return last

    def save(self, path):
        np.savez(path, params=self.get_params(), sigma=np.float64(self.sigma), hidden=self.hidden,
                 nchan=self.nchan, n=self.n)

    def load(self, path):
        z = np.load(path)
        assert int(z["hidden"]) == self.hidden and int(z["nchan"]) == self.nchan and int(z["n"]) == self.n
        self.set_params(z["params"])
        self.set_scale(float(z["sigma"]))

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_learned_destroy(self.h)
            self.h = None


@dataclass
class PulseDistribution:
    """Seeded random pulses E0 sin^2(pi t/T) cos(omega t): E0 in [e0_min, e0_max], omega in
    [w_min, w_max], `cycles` cycles; trajectories start at a random time in the pulse."""
    e0_min: float = 0.02
    e0_max: float = 0.1
    w_min: float = 0.05
    w_max: float = 0.2
    cycles: float = 4.0

    def draw(self, rng):
        p = LaserSimpleLength(float(rng.uniform(self.e0_min, self.e0_max)),
                              float(rng.uniform(self.w_min, self.w_max)), self.cycles)
        return p, float(rng.uniform(0.0, p.pulse_duration))


def scalings_for(base, pulse, t):
    """Per-potential time factors with `pulse` in place of every time-dependent factor."""
    return np.array([pulse(t) if p.IsTimeDependent else 1.0 for p in base.PotentialList], dtype=np.float64)


def generate_training_pairs(learned: LearnedPreconditioner, base, plain, psi0, dt: float, trajectories: int,
                            steps: int, seed: int = 0, pulses: PulseDistribution = None, rel_threshold=1e-4):
    """Propagate with the plain FP64 Cayley solver (exact block preconditioner) under seeded
    random pulses and keep, per step and channel block c with |r_c| > rel_threshold |r|, the
    training row of the solved correction system A d = r, r = b - A psi = -2 i (dt/2) H psi,
    d = psi(t+dt) - psi(t).
    Trajectories are chained (each starts from the end of the previous one).
    Returns device tensors x, t [rows][in] and chan [rows] (int32)."""
    torch = D.torch
    pulses = pulses or PulseDistribution()
    rng = np.random.default_rng(seed)
    psi = psi0.data.clone()
    h = torch.empty_like(psi)
    xs, ts, cs = [], [], []
    lib = N.load()
    for _ in range(trajectories):
        pulse, t0 = pulses.draw(rng)
        for k in range(steps):
            t = t0 + k * dt
            sc = scalings_for(base, pulse, t + abs(dt) / 2)
            N.check(lib.pp_ham_multiply_nooverlap(base.ham, D.hptr(sc), D.ptr(psi), D.ptr(h), D.stream_handle()))
            r = h * (-1j * dt)
            old = psi.clone()
            N.check(lib.pp_cayley_step(plain.h, D.ptr(psi), float(dt), D.hptr(sc), D.stream_handle()))
            d = psi - old
            x, tt, nrm = learned.prepare(r, d, dt)
            keep = nrm > rel_threshold
            idx = torch.nonzero(keep).flatten()
            xs.append(x.index_select(0, idx))
            ts.append(tt.index_select(0, idx))
            cs.append(idx.to(torch.int32))
    return torch.cat(xs).contiguous(), torch.cat(ts).contiguous(), torch.cat(cs).contiguous()


class MixedCayleyPropagator:
    """Cayley step by FP64 iterative refinement around complex64 flexible GMRES with the
    Kronecker preconditioner (and the learned correction when given).  Same AdvanceStep(t,
    dt) contract as CayleyPropagator (H at t + |dt|/2).  With `factored` (a
    pyprop_amd.factored.FactoredHamiltonian over the same potentials) every multiply uses
    the factored MFMA form."""

    def __init__(self, psi, base, kron: KroneckerPreconditioner, learned: Optional[LearnedPreconditioner] = None,
                 fallback=None, fallback_iterations: int = 40, krylov_basis_size: int = 20,
                 tolerance: float = 1e-10, inner_tolerance: float = 1e-5, max_outer: int = 8, factored=None,
                 device_mode: bool = False, learned_mode: int = 0, gram_schmidt_passes: int = 1,
                 graph: bool = False, deferred_finish: bool = True, overlap_from_solve: bool = True):
        """device_mode: after the first converged step, solve the inner least-squares problems
        on the device and synchronise once per step (pp_mixed_set_device_mode).
        learned_mode: the FGMRES directions the learned correction acts on: 0 the first of
        the first correction, 1 the first of every correction, 2 all
        (pp_mixed_set_learned_mode).
        gram_schmidt_passes: classical Gram-Schmidt passes per inner iteration (1 saves one
        launch per iteration; pp_mixed_set_gram_schmidt).
        graph: with device_mode, replay each device-mode step from a captured hipGraph
        (pp_mixed_set_graph; needs a factored operator whose angular forms are sparse).
        deferred_finish: let the next Kronecker solve finish a one-pass Gram-Schmidt that the
        host does not read first (one launch fewer; the same bits;
        pp_mixed_set_deferred_finish).
        overlap_from_solve: form the FP32 overlap product of each preconditioner output in its
        Kronecker solve, as F1 G F2^T, instead of the multiply's Cholesky-form passes
        (pp_mixed_set_overlap_from_solve)."""
        if graph and not device_mode:
            raise ValueError("graph mode needs device_mode")
        self.graph = graph
        self.learned_mode = learned_mode
        self.gram_schmidt_passes = gram_schmidt_passes
        self.deferred_finish = deferred_finish
        self.overlap_from_solve = overlap_from_solve
        self.psi = psi
        self.device_mode = device_mode
        self.factored = factored
        self.BasePropagator = base
        self.kron = kron
        self.learned = learned
        self.fallback = fallback
        self.fallback_iterations = fallback_iterations
        self.basis_size = krylov_basis_size
        self.tol = tolerance
        self.inner_tol = inner_tolerance
        self.max_outer = max_outer
        self.h = None

    def SetupStep(self, dt):
        self.BasePropagator.SetupStep(dt)
        self.kron.Setup()
        if self.fallback is not None and self.fallback.h is None:
            self.fallback.SetupStep(dt)
        if self.h is not None:  # a new step size: replace the solver
            N.check(N.load().pp_mixed_destroy(self.h))
            self.h = None
        h = C.c_void_p()
        N.check(N.load().pp_mixed_create(self.BasePropagator.ham, self.kron.h, self.basis_size, self.tol,
                                         self.inner_tol, self.max_outer, C.byref(h)))
        self.h = h
        if self.learned is not None:
            N.check(N.load().pp_mixed_set_learned(h, self.learned.h))
        if self.fallback is not None:
            N.check(N.load().pp_mixed_set_fallback(h, self.fallback.h, self.fallback_iterations))
        if self.factored is not None:
            N.check(N.load().pp_mixed_set_factored(h, self.factored.h))
        N.check(N.load().pp_mixed_set_device_mode(h, int(self.device_mode)))
        N.check(N.load().pp_mixed_set_learned_mode(h, int(self.learned_mode)))
        N.check(N.load().pp_mixed_set_gram_schmidt(h, int(self.gram_schmidt_passes)))
        N.check(N.load().pp_mixed_set_deferred_finish(h, int(self.deferred_finish)))
        N.check(N.load().pp_mixed_set_overlap_from_solve(h, int(self.overlap_from_solve)))
        N.check(N.load().pp_mixed_set_graph(h, int(self.graph)))

    def SetLearned(self, learned):
        self.learned = learned
        if self.h is not None:
            N.check(N.load().pp_mixed_set_learned(self.h, None if learned is None else learned.h))

    def AdvanceStep(self, t, dt, scalings=None):
        if self.h is None:
            self.SetupStep(dt)
        sc = self.BasePropagator.scalings(t + abs(dt / 2.0)) if scalings is None else scalings
        N.check(N.load().pp_mixed_step(self.h, D.ptr(self.psi.data), float(dt), D.hptr(sc), D.stream_handle()))

    def GetStats(self) -> dict:
        c = (C.c_int * 7)()
        r = (C.c_double * 2)()
        N.check(N.load().pp_mixed_stats(self.h, c, r))
        return dict(outer=c[0], inner=c[1], fp64_ops=c[2], fp32_ops=c[3], precond=c[4], learned=c[5],
                    fallback=c[6], residual=r[0], rel_residual=r[1])

    def GetGraphStats(self) -> dict:
        """Captures and replays of the device-mode step (graph mode)."""
        c, l = C.c_int(), C.c_int()
        N.check(N.load().pp_mixed_graph_stats(self.h, C.byref(c), C.byref(l)))
        return {"captures": c.value, "launches": l.value}

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_mixed_destroy(self.h)
            self.h = None
