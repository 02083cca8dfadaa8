"""Observables between propagation steps (the Problem driver's measurements,
pyprop/Problem.py:254-279 for the energy; ionisation by projection on the field-free bound
states as examples/tensor/helium_stabilization/example.py:557-604 does for helium).

Rank 2 (hydrogen-like (l, r)): the bound states come from a dense generalized eigensolve of
each l block on the host (120 x 120).  Rank 3 (helium (channel, r1, r2)): per total L, the
field-free Hamiltonian H0 + 1/r12 restricted to that L block is solved on the device by
pIRAM with GMRES shift-invert (the reference computes its bound-state files the same way,
pyprop/solver/PiramSolver.py + GMRESShiftInvertSolver.py), and the projections are S-metric
inner products on the device.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np


def _band_dense(values, pairs, n):
    M = np.zeros((n, n))
    b = int(np.max(np.abs(pairs[:, 0] - pairs[:, 1])))
    keep = np.abs(pairs[:, 0] - pairs[:, 1]) <= b
    M[pairs[keep, 0], pairs[keep, 1]] = values[keep]
    return M


def bound_states(system, emax: float = 0.0) -> List[Tuple[int, np.ndarray, np.ndarray]]:
    """Field-free bound states of a rank-2 (l, r) system: for every l the generalized
    eigenpairs H0_l c = E S c with E < emax, S-normalised.  Returns [(l, E[k], C[:, k])]."""
    import scipy.linalg
    if system.rank != 2:
        raise ValueError("bound_states supports (l, r) systems; use helium_bound_states for rank 3")
    H0 = system.potentials[0]
    g0, g1 = H0.geometries
    if g0.storage_id != "Diag":
        raise ValueError("the first potential must be the angular-diagonal H0")
    nl, n = system.shape
    S = system.bases[1].overlap_band()
    data = H0.data.real
    out = []
    for l in range(nl):
        H = _band_dense(data[l], g1.pairs, n)
        H = 0.5 * (H + H.T)
        w, v = scipy.linalg.eigh(H, S)
        keep = w < emax
        out.append((l, w[keep], v[:, keep]))
    return out


@dataclass
class BoundStateBlock:
    """The field-free bound states of one total-L block of a helium system.

    channels: indices of the block's channels in the full system (rank 0 of psi)
    energies: eigenvalues below the threshold, ascending
    vectors:  device complex128 tensor (k, len(channels), n1, n2), each S-normalised
    gram:     <phi_i|S|phi_j> (k x k), the identity up to the eigensolver's accuracy
    """
    L: int
    channels: np.ndarray
    energies: np.ndarray
    vectors: object
    gram: np.ndarray
    repr: object = None
    stats: Optional[dict] = None


def helium_bound_states(system, threshold: float = -2.0, count: int = 8, max_count: int = 60,
                        krylov_tolerance: float = 1e-10, max_restarts: int = 300, gmres_basis_size: int = 30,
                        gmres_tolerance: float = 1e-12, factored: bool = False) -> List[BoundStateBlock]:
    """Every field-free bound state below `threshold` (the He+ 1s threshold, -2.0 a.u., in
    FindIonizationProbability, example.py:557) of each total-L block of a rank-3 helium
    system, by pIRAM on the shift-inverted operator with GMRES and the exact block-diagonal
    preconditioner (PiramSolver / GMRESShiftInvertSolver as the reference computes its
    bound-state files).

    Per block, two stages, both with the shift below the block's spectrum, where H - sigma S
    and every channel block of the preconditioner are positive definite (the device block
    LU does not pivot):
      1. the lowest energy E_min, from one eigenvalue at sigma0 = -Z^2 - 0.5 (below every
         two-electron energy, each electron being bound by at most Z^2 / 2);
      2. sigma = E_min - 0.1 and the negated operator -(H - sigma S)^-1 S, whose Ritz values
         -1 / (E - sigma) put the lowest energies first in pIRAM's wanted set (smallest real
         part).  `count` states are requested and the request doubled (up to max_count)
         while every converged one still lies below the threshold.
    (A shift at the threshold itself would make the states below it the only negative Ritz
    values, but a request beyond their number then targets the cluster of high-energy
    values near 0+, which the restarted Arnoldi process cannot resolve.)"""
    from . import device as D
    from .eigen import GMRESShiftInvertSolver, PiramSolver, ShiftInvertEnergies
    from .propagator import BasisPropagator, CombinedRepresentation, RadialTwoElectronPreconditioner, Wavefunction
    from .systems import restrict_channels
    if system.rank != 3 or not system.channels:
        raise ValueError("helium_bound_states needs a rank-3 coupled-channel system")
    torch = D.torch
    Z = float(system.info.get("Z", 2.0))
    out = []
    for L in sorted(set(ch.L for ch in system.channels)):
        idx = np.array([c for c, ch in enumerate(system.channels) if ch.L == L], dtype=np.int64)
        sub = restrict_channels(system, idx)
        repr_ = CombinedRepresentation(sub.shape, sub.bases)
        psi = Wavefunction(repr_)
        base = BasisPropagator(psi, sub.potentials)
        base.SetupStep()
        fh = None
        if factored:
            from .factored import FactoredHamiltonian
            fh = FactoredHamiltonian(base)

        def run(shift, k):
            si = GMRESShiftInvertSolver(psi, base, shift, RadialTwoElectronPreconditioner(psi, sub),
                                        krylov_basis_size=gmres_basis_size, krylov_tolerance=gmres_tolerance,
                                        factored=fh, negate=True)
            solver = PiramSolver(psi, si, krylov_basis_size=min(128, max(2 * k + 10, 20)), krylov_eigenvalue_count=k,
                                 krylov_tolerance=krylov_tolerance, krylov_max_iteration_count=max_restarts)
            solver.Solve()
            theta = solver.GetComplexEigenvalues()
            conv = solver.GetConvergenceEstimates() <= 0
            if not np.all(conv):
                raise RuntimeError("L = %d: pIRAM did not converge %d states at shift %.4f (%s)" %
                                   (L, k, shift, solver.GetStatistics()))
            return si, solver, ShiftInvertEnergies(theta, shift, negate=True).real

        _, _, e0 = run(-Z * Z - 0.5, 1)
        shift = float(e0[0]) - 0.1
        k = max(1, count)
        while True:
            si, solver, energies = run(shift, k)
            if np.any(energies >= threshold) or 2 * k > max_count:
                break
            k *= 2
        if np.all(energies < threshold):
            raise RuntimeError("L = %d: all %d lowest states lie below the threshold; raise max_count" % (L, k))
        sel = np.nonzero(energies < threshold)[0]
        sel = sel[np.argsort(energies[sel])]
        vec = torch.empty((len(sel),) + tuple(sub.shape), dtype=torch.complex128, device=psi.data.device)
        tmp = Wavefunction(repr_)
        for j, i in enumerate(sel):
            solver.SetEigenvector(tmp, int(i), normalize=True)
            vec[j].copy_(tmp.data)
        gram = np.array([[repr_.InnerProduct(vec[i], vec[j]) for j in range(len(sel))] for i in range(len(sel))])
        st = dict(solver.GetStatistics(), **si.GetStatistics(), requested=k, shift=shift)
        out.append(BoundStateBlock(L, idx, energies[sel], vec, gram, repr_, st))
    return out


def _helium_yield(psi, system, blocks) -> float:
    """1 - |P psi|_S^2 / |psi|_S^2 with P the S-orthogonal projector on the bound states:
    per block, c_i = <phi_i|S|psi_L> and |P psi|^2 = c^H G^-1 c (G the Gram matrix), which is
    sum_i |c_i|^2 for S-orthonormal states (example.py:588-601)."""
    from . import device as D
    torch = D.torch
    if isinstance(psi, np.ndarray):
        data = D.to_device(np.ascontiguousarray(psi, dtype=np.complex128))
    else:
        data = getattr(psi, "data", psi)
    total = getattr(psi, "repr", None)
    if total is None:
        from .propagator import CombinedRepresentation
        total = CombinedRepresentation(system.shape, system.bases)
    norm2 = total.InnerProduct(data, data).real
    pb = 0.0
    for b in blocks:
        part = data.index_select(0, torch.as_tensor(b.channels, device=data.device)).contiguous()
        c = np.array([b.repr.InnerProduct(b.vectors[i], part) for i in range(b.vectors.shape[0])])
        if c.size:
            pb += float(np.real(np.vdot(c, np.linalg.solve(b.gram, c))))
    return 1.0 - pb / norm2


def ionization_yield(psi, system, bound=None) -> float:
    """1 - (probability in the field-free bound states) / <psi|S|psi>.

    Rank 2: psi a host array (l, r); bound from bound_states (all E < 0).
    Rank 3: psi a Wavefunction, device tensor or host array (channel, r1, r2); bound from
    helium_bound_states (all E below the -2.0 threshold), as FindIonizationProbability does
    (example.py:557-604; for an S-normalised psi its 1 - sum |<phi|psi>|^2)."""
    if system.rank == 3:
        return _helium_yield(psi, system, helium_bound_states(system) if bound is None else bound)
    bound = bound_states(system) if bound is None else bound
    S = system.bases[1].overlap_band()
    Spsi = psi @ S.T  # (nl, n): S applied on the radial rank
    norm2 = float(np.vdot(psi, Spsi).real)
    pb = 0.0
    for l, E, C in bound:
        proj = C.T @ Spsi[l]
        pb += float(np.sum(np.abs(proj) ** 2))
    return 1.0 - pb / norm2
