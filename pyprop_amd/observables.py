"""Observables between propagation steps (the Problem driver's measurements,
pyprop/Problem.py:254-279 for the energy; ionisation by projection on the field-free bound
states as examples/tensor/helium_stabilization/example.py:557-604 does for helium).

Rank 2 (hydrogen-like (l, r)): the bound states come from a dense generalized eigensolve of
each l block on the host (120 x 120).  Rank 3 (helium (channel, r1, r2)): per total L, the
field-free Hamiltonian H0 + 1/r12 restricted to that L block is solved on the device by
pIRAM with GMRES shift-invert at the ionisation threshold (the reference computes its
bound-state files the same way, pyprop/solver/PiramSolver.py + GMRESShiftInvertSolver.py),
and the projections are S-metric inner products on the device.
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
    system.

    Per block: pIRAM on the shift-inverted operator (H - sigma S)^-1 S with sigma =
    threshold, GMRES with the exact block-diagonal preconditioner.  Its Ritz values theta =
    1 / (E - sigma) are negative exactly for E < sigma, and pIRAM's wanted values are those
    with the smallest real part, so the wanted set fills with the states below the threshold
    first.  `count` states are requested; while every converged wanted value is still
    negative the request is doubled (up to max_count), so that none below the threshold is
    missed."""
    from . import device as D
    from .eigen import GMRESShiftInvertSolver, PiramSolver
    from .propagator import BasisPropagator, CombinedRepresentation, RadialTwoElectronPreconditioner, Wavefunction
    from .systems import restrict_channels
    if system.rank != 3 or not system.channels:
        raise ValueError("helium_bound_states needs a rank-3 coupled-channel system")
    torch = D.torch
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
        si = GMRESShiftInvertSolver(psi, base, threshold, RadialTwoElectronPreconditioner(psi, sub),
                                    krylov_basis_size=gmres_basis_size, krylov_tolerance=gmres_tolerance, factored=fh)
        k = max(1, count)
        while True:
            m = min(128, max(2 * k + 10, 20))
            solver = PiramSolver(psi, si, krylov_basis_size=m, krylov_eigenvalue_count=k,
                                 krylov_tolerance=krylov_tolerance, krylov_max_iteration_count=max_restarts)
            solver.Solve()
            theta = solver.GetComplexEigenvalues()
            conv = solver.GetConvergenceEstimates() <= 0
            below = theta.real < 0
            if not np.all(below) or 2 * k > max_count:
                break
            k *= 2
        if np.all(below):
            raise RuntimeError("L = %d: all %d wanted Ritz values lie below the threshold; raise max_count" % (L, k))
        if not np.all(conv[below]):
            raise RuntimeError("L = %d: pIRAM did not converge every state below the threshold (%s)" %
                               (L, solver.GetStatistics()))
        sel = np.nonzero(below)[0]
        energies = (threshold + 1.0 / theta[sel]).real
        order = np.argsort(energies)
        sel, energies = sel[order], energies[order]
        vec = torch.empty((len(sel),) + tuple(sub.shape), dtype=torch.complex128, device=psi.data.device)
        tmp = Wavefunction(repr_)
        for j, i in enumerate(sel):
            solver.SetEigenvector(tmp, int(i), normalize=True)
            vec[j].copy_(tmp.data)
        gram = np.array([[repr_.InnerProduct(vec[i], vec[j]) for j in range(len(sel))] for i in range(len(sel))])
        st = dict(solver.GetStatistics(), **si.GetStatistics(), requested=k)
        out.append(BoundStateBlock(L, idx, energies, vec, gram, repr_, st))
    return out


def _helium_yield(psi, system, blocks) -> float:
    """1 - |P psi|_S^2 / |psi|_S^2 with P the S-orthogonal projector on the bound states:
    per block, c_i = <phi_i|S|psi_L> and |P psi|^2 = c^H G^-1 c (G the Gram matrix), which is
    sum_i |c_i|^2 for S-orthonormal states (example.py:588-601)."""
    from . import device as D
    torch = D.torch
    data = getattr(psi, "data", None)
    if data is None:
        data = D.to_device(np.ascontiguousarray(psi))
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
