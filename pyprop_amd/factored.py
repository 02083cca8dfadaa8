"""Factored form of a rank-3 (angular x radial x radial) Hamiltonian for the MFMA multiply.

Every potential the set-up builds is a sum of terms a_t(angular pair) x R_t(radial pairs)
(systems.py; for 1/r12 this is the multipole sum over lambda, the identity the reference's
CustomPotentialEvaluationR12 evaluates term by term,
examples/tensor/helium_stabilization/coupledspherical.cpp:60-111).  Hence
    H psi = sum_t f_p(t) (A_t (x) R_t) psi,
with A_t the dense nchan x nchan angular matrix of term t and R_t one radial operator
shared by all channels.  The device multiply (pp_fham_*) contracts the angular index first
on MFMA (W_t = f A_t X) and then applies every R_t to all channels at once (a banded
radial product, also on MFMA), instead of streaming one radial block per angular pair as
the reference storage does (generator.py:302-762).

R_t is stored diagonal-blocked: dia[t][i1][jj][d][i2] is the element
(row (i1, i2), column (j1, j2)) with j1 = i1 - bw + jj, j2 = i2 - bw + d.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _native as N
from . import device as D


def factor_potentials(potentials: Sequence, nchan: int, n: int):
    """(A [T][nchan][nchan], term_potential [T], dia [T][n][2bw+1][2bw+1][n], bw) from
    TensorPotentialData with (Simp angular) x (Simp banded r1) x (BandNH r2) geometries and
    separable factors (angular_coeffs, radial_factors).  term_potential[t] is the index of
    the potential (time scaling) term t belongs to."""
    A_list, tp_list, R_list = [], [], []
    bw = None
    for ip, p in enumerate(potentials):
        src = getattr(p, "source", p)
        if src.angular_coeffs is None or src.radial_factors is None:
            raise ValueError("potential %s has no separable factors" % src.name)
        g0, g1, g2 = src.geometries
        if g0.storage_id != "Simp" or g1.storage_id != "Simp" or g2.storage_id != "BandNH":
            raise ValueError("factored form needs Simp x Simp(banded) x BandNH storage")
        b2 = (g2.pair_count // n - 1) // 2
        p1 = g1.pairs
        b1 = int(np.max(np.abs(p1[:, 0] - p1[:, 1])))
        b = max(b1, b2)
        bw = b if bw is None else max(bw, b)
        for t in range(src.angular_coeffs.shape[0]):
            A = np.zeros((nchan, nchan))
            A[g0.pairs[:, 0], g0.pairs[:, 1]] = src.angular_coeffs[t]
            A_list.append(A)
            tp_list.append(ip)
            R_list.append((src.radial_factors[t], g1.pairs, g2.pairs, b2))
    w = 2 * bw + 1
    T = len(A_list)
    dia = np.zeros((T, n, w, w, n))
    for t, (R, p1, p2, b2) in enumerate(R_list):
        keep2 = np.abs(p2[:, 0] - p2[:, 1]) <= b2
        i2 = p2[keep2, 0]
        j2 = p2[keep2, 1]
        d = j2 - i2 + bw
        Rk = R[:, keep2]
        for q1 in range(p1.shape[0]):
            i1, j1 = int(p1[q1, 0]), int(p1[q1, 1])
            jj = j1 - i1 + bw
            dia[t, i1, jj, d, i2] = Rk[q1]
    return np.stack(A_list), np.array(tp_list, dtype=np.int32), dia, bw


class FactoredHamiltonian:
    """Device factored Hamiltonian over a BasisPropagator's potentials (same scalings)."""

    def __init__(self, base, overlap: bool = True):
        """overlap: also hold S1 (x) S2 of the radial ranks as a term, so that MultiplyOverlap
        (and the mixed-precision Cayley step) forms S + c H in one multiply."""
        shape = base.psi.repr.shape
        if len(shape) != 3 or shape[1] != shape[2]:
            raise ValueError("the factored multiply needs psi shaped (channels, n, n)")
        nchan, n, _ = shape
        A, tpot, dia, bw = factor_potentials(base.PotentialList, nchan, n)
        self.nterms = A.shape[0]
        self.angular_nnz = int(np.count_nonzero(A))  # non-zeros of all A_t (the sparse angular form's list)
        self.bw = bw
        self.shape = shape
        self.base = base
        A = np.ascontiguousarray(A)
        dia = np.ascontiguousarray(dia)
        h = C.c_void_p()
        N.check(N.load().pp_fham_create(nchan, n, self.nterms, bw, D.hptr(A), D.hptr(tpot), D.hptr(dia),
                                        len(base.PotentialList), C.byref(h)))
        self.h = h
        nb = C.c_int()
        N.check(N.load().pp_fham_band_terms(h, C.byref(nb)))
        self.band_terms = nb.value  # terms on the MFMA band tiles; the rest are separable
        self.has_overlap = False
        bases = base.psi.repr.bases
        if overlap and bases[1] is not None and bases[2] is not None:
            S1 = np.ascontiguousarray(bases[1].overlap_band(), dtype=np.float64)
            S2 = np.ascontiguousarray(bases[2].overlap_band(), dtype=np.float64)
            N.check(N.load().pp_fham_set_overlap(self.h, D.hptr(S1), D.hptr(S2)))
            self.has_overlap = True

    def Multiply(self, src, dst, t, addend=None, alpha=1.0):
        """dst = addend + alpha * H(t) src (complex128 or complex64 device tensors)."""
        sc = self.base.scalings(t)
        s = getattr(src, "data", src)
        d = getattr(dst, "data", dst)
        a = None if addend is None else getattr(addend, "data", addend)
        fn = N.load().pp_fham_multiply_f32 if s.dtype == D.torch.complex64 else N.load().pp_fham_multiply
        N.check(fn(self.h, D.hptr(sc), D.ptr(s), D.ptr(d), D.ptr(a), N.cplx(alpha), D.stream_handle()))

    def MultiplyOverlap(self, src, dst, t, alpha=1.0, beta=1.0, addend=None):
        """dst = addend + alpha * H(t) src + beta * (S (x) S) src (needs the overlap term)."""
        sc = self.base.scalings(t)
        s = getattr(src, "data", src)
        d = getattr(dst, "data", dst)
        a = None if addend is None else getattr(addend, "data", addend)
        fn = (N.load().pp_fham_multiply_overlap_f32 if s.dtype == D.torch.complex64
              else N.load().pp_fham_multiply_overlap)
        N.check(fn(self.h, D.hptr(sc), D.ptr(s), D.ptr(d), D.ptr(a), N.cplx(alpha), N.cplx(beta),
                   D.stream_handle()))

    # radial variants: "<terms>x<channel groups>" per wave on band tiles ("p": next band step
    # prefetched), "pos<groups>": each position's products of all band terms as one GEMM over
    # the w x w window, "pos<groups>w<positions>": the same with the window of a workgroup's
    # positions (consecutive along r2) copied into LDS once, "...s" the same with the window
    # swizzled instead of padded (FP32 only; a smaller window, three workgroups per CU) (the
    # position forms need the separable split; tune reports -1 ms for a variant that cannot
    # run)
    RADIAL_SHAPES = ("1x1", "2x2", "3x2", "3x3", "2x2p", "3x2p", "3x3p", "pos6", "pos3", "pos2",
                     "pos3w16", "pos2w16", "pos2w8", "pos1w16", "pos2w16s")

    def set_radial_variant(self, precision: int, variant: int):
        """Radial-stage kernel shape for the complex128 (precision 64) or complex64 (32)
        multiply: index into RADIAL_SHAPES."""
        N.check(N.load().pp_fham_set_radial_variant(self.h, precision, variant))

    def radial_variant(self, precision: int) -> int:
        v = C.c_int()
        N.check(N.load().pp_fham_get_radial_variant(self.h, precision, C.byref(v)))
        return v.value

    def tune(self, precision: int, trials: int = 5):
        """Time every radial variant on the device and keep the fastest, then the angular
        stage's wave count; returns (best radial variant, {shape: ms per multiply}) and leaves
        the angular timings in self.angular_timings."""
        best = C.c_int()
        tim = np.zeros(len(self.RADIAL_SHAPES) + len(self.ANGULAR_VARIANTS))
        N.check(N.load().pp_fham_tune(self.h, precision, trials, C.byref(best), D.hptr(tim), D.stream_handle()))
        nr = len(self.RADIAL_SHAPES)
        self.angular_timings = dict(zip(self.ANGULAR_VARIANTS, tim[nr:].tolist()))
        return best.value, dict(zip(self.RADIAL_SHAPES, tim[:nr].tolist()))

    ANGULAR_WAVES = (16, 8, 4)  # waves splitting K in the angular stage, one row block per workgroup
    # every angular variant: "<waves>" (one row block of 16 channels per workgroup),
    # "rows<waves>" (all row blocks of a position tile per workgroup, each Z element read once)
    # or "sparse<PV>" (only the non-zeros of the angular matrices, one wave per output channel
    # and 64 PV positions)
    ANGULAR_VARIANTS = ("16", "8", "4", "rows8", "rows16", "sparse1", "sparse2")

    def set_angular_variant(self, precision: int, variant: int):
        N.check(N.load().pp_fham_set_angular_variant(self.h, precision, variant))

    def angular_variant(self, precision: int) -> int:
        v = C.c_int()
        N.check(N.load().pp_fham_get_angular_variant(self.h, precision, C.byref(v)))
        return v.value

    def set_angular_waves(self, precision: int, waves: int):
        N.check(N.load().pp_fham_set_angular_waves(self.h, precision, waves))

    def angular_waves(self, precision: int) -> int:
        w = C.c_int()
        N.check(N.load().pp_fham_get_angular_waves(self.h, precision, C.byref(w)))
        return w.value

    def __del__(self):
        if getattr(self, "h", None) is not None and N._lib is not None:
            N._lib.pp_fham_destroy(self.h)
            self.h = None
