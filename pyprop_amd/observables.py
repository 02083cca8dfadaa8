"""Observables between propagation steps (the Problem driver's measurements,
pyprop/Problem.py:257-279 for the energy; ionisation by projection on the field-free bound
states as examples/tensor/helium_stabilization/example.py:557-604 does for helium).

Host-side evaluation on a copy of the state: these run once per output time, not per step.
"""
from __future__ import annotations

from typing import List, Tuple

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
        raise ValueError("bound_states supports (l, r) systems")
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


def ionization_yield(psi: np.ndarray, system, bound=None) -> float:
    """1 - sum over bound states |<phi|S|psi>|^2 / <psi|S|psi> for a rank-2 (l, r) state."""
    bound = bound_states(system) if bound is None else bound
    S = system.bases[1].overlap_band()
    Spsi = psi @ S.T  # (nl, n): S applied on the radial rank
    norm2 = float(np.vdot(psi, Spsi).real)
    pb = 0.0
    for l, E, C in bound:
        proj = C.T @ Spsi[l]
        pb += float(np.sum(np.abs(proj) ** 2))
    return 1.0 - pb / norm2
