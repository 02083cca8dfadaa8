"""ORACLE -- test infrastructure only.

CPU restatement of PyProp's propagation hot path (oracle.cpp), used as the checker by
tests/, __graft_entry__.smoke() and as the CPU baseline leg of bench.py.  The product
package pyprop_amd never imports this module.

Each function of oracle.cpp cites the reference code it restates.  Its pins are listed
in tests/test_oracle.py: scipy's BLAS (zgbmv, zhbmv, ztbmv, ztbsv), LAPACK (zpbtrf,
zgels), scipy.linalg.expm, dense brute force assembled from (V, pairs) as the
reference's own SetupBigMatrix does, and the reference's 6x6 known-answer matrix.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
SRC = HERE / "oracle.cpp"

OP = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))

STORAGE = {"Simp": 0, "Herm": 1, "Diag": 2, "Ident": 3, "Band": 4, "BandNH": 5, "Dense": 6}


def build(force: bool = False) -> Path:
    """Compile oracle.cpp with g++ (recipe also in oracle/Makefile)."""
    LIB.parent.mkdir(exist_ok=True)
    if force or not LIB.exists() or LIB.stat().st_mtime < SRC.stat().st_mtime:
        tmp = LIB.with_suffix(".so.tmp")
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", str(SRC), "-o", str(tmp)]
        subprocess.run(cmd, check=True)
        os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        P = C.c_void_p
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_get_threads.restype = C.c_int
        L.oracle_tpmul.argtypes = [C.c_int, P, P, P, C.c_double, P, P, P, P]
        L.oracle_overlap_factor.argtypes = [C.c_int, C.c_int, P, P]
        L.oracle_overlap_apply.argtypes = [C.c_int, C.c_int, P, C.c_int, C.c_int, C.c_int, P, C.c_int, P]
        L.oracle_gmres_solve.argtypes = [C.c_int64, C.c_int, C.c_double, C.c_int, OP, P, P, P, P, P, P]
        L.oracle_pamp_propagate.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int, OP, P, P, C.c_double, C.c_double,
                                            P, P, P, C.c_int]
        L.oracle_pade_exp.argtypes = [C.c_int, P, C.c_int, P]
        L.oracle_eig_exp.argtypes = [C.c_int, P, P]
        L.oracle_bspline_tables.argtypes = [C.c_int, P, C.c_int, C.c_int, P, C.c_int, P, P, P, P, P, P, P]
        L.oracle_bspline_represent.argtypes = [C.c_int, P, C.c_int, C.c_int, P, C.c_int, P, P, C.c_int, C.c_int,
                                               P, C.c_int, C.c_int64, C.c_int64, P, P]
        L.oracle_sys_create.restype = C.c_void_p
        L.oracle_sys_create.argtypes = [C.c_int, P]
        L.oracle_sys_destroy.argtypes = [P]
        L.oracle_sys_add_potential.argtypes = [P, P, P, P, P, P]
        L.oracle_sys_set_overlap.argtypes = [P, C.c_int, C.c_int, C.c_int, P]
        L.oracle_sys_multiply.argtypes = [P, C.c_int, P, P, P]
        L.oracle_sys_overlap.argtypes = [P, C.c_int, C.c_int, P]
        L.oracle_sys_set_preconditioner.argtypes = [P, P, P, C.c_int, P, C.c_int, P, C.c_int, P,
                                                    C.c_double, C.c_double, C.c_double, C.c_double]
        L.oracle_sys_precond_solve.argtypes = [P, P]
        L.oracle_sys_cayley_step.argtypes = [P, P, C.c_double, P, C.c_int, C.c_double, C.c_int, P, P]
        L.oracle_sys_pamp_step.argtypes = [P, P, C.c_double, C.c_double, P, C.c_int, C.c_int, C.c_int, P, C.c_int]
        L.oracle_sys_inner_product.argtypes = [P, P, P, P]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """Host threads for the multiply, overlap passes and preconditioner solve (1 = the
    reference's sequential loops).  Results are bitwise the same for every count: each
    output element is summed by one thread in the sequential order."""
    lib().oracle_set_threads(int(n))


def get_threads() -> int:
    return int(lib().oracle_get_threads())


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _pairs_arg(pairs, rank):
    keep = []
    arr = (C.c_void_p * rank)()
    for r in range(rank):
        p = None if pairs is None else pairs[r]
        if p is None:
            arr[r] = None
        else:
            p = np.ascontiguousarray(p, dtype=np.int32)
            keep.append(p)
            arr[r] = p.ctypes.data
    return arr, keep


def tpmul(storage_ids, V, scaling, src, dst, pairs):
    """dst += V * scaling * src (in place on dst), reference loop order."""
    rank = len(storage_ids)
    ids = np.array([STORAGE[s] if isinstance(s, str) else int(s) for s in storage_ids], dtype=np.int32)
    V = np.ascontiguousarray(V, dtype=np.complex128)
    src = np.ascontiguousarray(src, dtype=np.complex128)
    assert dst.dtype == np.complex128 and dst.flags.c_contiguous
    pext = np.array(V.shape, dtype=np.int32)
    sext = np.array(src.shape, dtype=np.int32)
    parr, keep = _pairs_arg(pairs, rank)
    _check(lib().oracle_tpmul(rank, _ptr(ids), _ptr(V), _ptr(pext), float(scaling), _ptr(src), _ptr(sext),
                              _ptr(dst), C.cast(parr, C.c_void_p)))
    return dst


def overlap_factor(s_upper: np.ndarray) -> np.ndarray:
    s_upper = np.ascontiguousarray(s_upper, dtype=np.float64)
    n, ld = s_upper.shape
    u = np.empty_like(s_upper)
    _check(lib().oracle_overlap_factor(n, ld - 1, _ptr(s_upper), _ptr(u)))
    return u


def overlap_apply(u: np.ndarray, op: int, conj: bool, data: np.ndarray, rank: int) -> np.ndarray:
    n, ld = u.shape
    ext = np.array(data.shape, dtype=np.int32)
    _check(lib().oracle_overlap_apply(n, ld - 1, _ptr(np.ascontiguousarray(u)), op, int(conj), data.ndim,
                                      _ptr(ext), rank, _ptr(data)))
    return data


def _wrap_op(fn, n):
    def cb(ctx, pin, pout):
        try:
            a = np.ctypeslib.as_array(pin, shape=(2 * n,)).view(np.complex128)
            b = np.ctypeslib.as_array(pout, shape=(2 * n,)).view(np.complex128)
            b[:] = fn(a.copy())
            return 0
        except Exception:  # pragma: no cover - surfaced as an oracle error
            return 1
    return OP(cb)


def gmres_solve(apply, rhs, m, tol, dorth=True):
    n = rhs.size
    rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
    sol = np.zeros(n, dtype=np.complex128)
    err = np.zeros(1)
    ops = np.zeros(1, dtype=np.int32)
    elist = np.zeros(m)
    cb = _wrap_op(apply, n)
    _check(lib().oracle_gmres_solve(n, m, tol, int(dorth), cb, None, _ptr(rhs), _ptr(sol), _ptr(err), _ptr(ops),
                                    _ptr(elist)))
    return sol, float(err[0]), int(ops[0]), elist


def pamp_propagate(apply, v, m, dt, order=6, dorth=False, method=0):
    """method 0: Pade exponential of the Hessenberg matrix, 1: eigenvector exponential
    (krylov_exponentiation_method, pampwrapper.cpp:56-59)."""
    n = v.size
    v = np.ascontiguousarray(v, dtype=np.complex128).copy()
    est = np.zeros(2)
    ops = np.zeros(1, dtype=np.int32)
    bd = np.zeros(1, dtype=np.int32)
    cb = _wrap_op(apply, n)
    dt = complex(dt)
    _check(lib().oracle_pamp_propagate(n, m, order, int(dorth), cb, None, _ptr(v), dt.real, dt.imag, _ptr(est),
                                       _ptr(ops), _ptr(bd), int(method)))
    return v, complex(est[0], est[1]), int(ops[0]), bool(bd[0])


def eig_exp(A):
    """V diag(exp(w)) V^H for an upper Hessenberg A (unit-norm eigenvectors)."""
    m = A.shape[0]
    Af = np.asfortranarray(A, dtype=np.complex128)
    out = np.zeros((m, m), dtype=np.complex128, order="F")
    _check(lib().oracle_eig_exp(m, _ptr(Af), _ptr(out)))
    return np.asarray(out)


def pade_exp(A, order=6):
    m = A.shape[0]
    Af = np.asfortranarray(A, dtype=np.complex128)
    out = np.zeros((m, m), dtype=np.complex128, order="F")
    _check(lib().oracle_pade_exp(m, _ptr(Af), order, _ptr(out)))
    return np.asarray(out)


def _bspline_args(basis):
    knots = np.ascontiguousarray(basis.knots, dtype=np.float64)
    bp = np.ascontiguousarray(basis.breakpoints, dtype=np.float64)
    nodes = np.ascontiguousarray(basis.nodes, dtype=np.float64)
    weights = np.ascontiguousarray(basis.weights, dtype=np.float64)
    keep = (knots, bp, nodes, weights)
    return keep, (knots.size, _ptr(knots), int(basis.order), bp.size, _ptr(bp), nodes.size, _ptr(nodes),
                  _ptr(weights))


def bspline_tables(basis):
    """(grid, weights, B, B', B'') of the oracle's own B-spline evaluation, from the
    basis's knots, breakpoints and reference-interval Gauss rule."""
    keep, args = _bspline_args(basis)
    Q = (basis.breakpoints.size - 1) * basis.nodes.size
    N = basis.knots.size - basis.order
    grid, gw = np.zeros(Q), np.zeros(Q)
    b0, b1, b2 = (np.zeros((N, Q)) for _ in range(3))
    _check(lib().oracle_bspline_tables(*args, _ptr(grid), _ptr(gw), _ptr(b0), _ptr(b1), _ptr(b2)))
    return grid, gw, b0, b1, b2


def bspline_represent(basis, src, differentiation, pairs, hermitian, axis):
    """Project the grid axis `axis` of the real array src onto the pair list (one rank of
    RepresentPotentialInBasisBSpline)."""
    keep, args = _bspline_args(basis)
    src = np.ascontiguousarray(src, dtype=np.float64)
    pairs = np.ascontiguousarray(pairs, dtype=np.int32)
    pre = int(np.prod(src.shape[:axis]))
    post = int(np.prod(src.shape[axis + 1:]))
    out = np.zeros(src.shape[:axis] + (pairs.shape[0],) + src.shape[axis + 1:])
    _check(lib().oracle_bspline_represent(*args, int(differentiation), pairs.shape[0], _ptr(pairs), int(hermitian),
                                          pre, post, _ptr(src), _ptr(out)))
    return out


class System:
    """Hamiltonian + overlaps + optional exact block preconditioner on the CPU."""

    def __init__(self, shape):
        self.shape = tuple(int(s) for s in shape)
        ext = np.array(self.shape, dtype=np.int32)
        self.h = lib().oracle_sys_create(len(self.shape), _ptr(ext))
        self.nterms = 0
        self._keep = []

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_sys_destroy(self.h)
            self.h = None

    def add_potential(self, storage_ids, V, pairs):
        rank = len(storage_ids)
        ids = np.array([STORAGE[s] if isinstance(s, str) else int(s) for s in storage_ids], dtype=np.int32)
        V = np.ascontiguousarray(V, dtype=np.complex128)
        pext = np.array(V.shape, dtype=np.int32)
        parr, keep = _pairs_arg(pairs, rank)
        counts = np.array([0 if (pairs is None or pairs[r] is None) else len(pairs[r]) for r in range(rank)],
                          dtype=np.int32)
        _check(lib().oracle_sys_add_potential(self.h, _ptr(ids), _ptr(V), _ptr(pext), C.cast(parr, C.c_void_p),
                                              _ptr(counts)))
        self.nterms += 1

    def set_overlap(self, rank, s_upper):
        s_upper = np.ascontiguousarray(s_upper, dtype=np.float64)
        n, ld = s_upper.shape
        _check(lib().oracle_sys_set_overlap(self.h, rank, n, ld - 1, _ptr(s_upper)))

    def set_preconditioner(self, s1, s2, band, P, pairs1, pairs2, scaling_h, scaling_s=1.0):
        s1 = np.ascontiguousarray(s1, dtype=np.float64)
        s2 = np.ascontiguousarray(s2, dtype=np.float64)
        P = np.ascontiguousarray(P, dtype=np.complex128)
        p1 = np.ascontiguousarray(pairs1, dtype=np.int32)
        p2 = np.ascontiguousarray(pairs2, dtype=np.int32)
        sh, ss = complex(scaling_h), complex(scaling_s)
        _check(lib().oracle_sys_set_preconditioner(self.h, _ptr(s1), _ptr(s2), band, _ptr(P), len(p1), _ptr(p1),
                                                   len(p2), _ptr(p2), sh.real, sh.imag, ss.real, ss.imag))

    def multiply(self, src, scalings=None, mode=0):
        src = np.ascontiguousarray(src, dtype=np.complex128)
        dst = np.zeros_like(src)
        sc = np.ones(self.nterms) if scalings is None else np.asarray(scalings, dtype=np.float64)
        _check(lib().oracle_sys_multiply(self.h, mode, _ptr(sc), _ptr(src), _ptr(dst)))
        return dst

    def overlap(self, op, conj, psi):
        psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
        _check(lib().oracle_sys_overlap(self.h, op, int(conj), _ptr(psi)))
        return psi

    def precond_solve(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
        _check(lib().oracle_sys_precond_solve(self.h, _ptr(psi)))
        return psi

    def cayley_step(self, psi, dt, scalings, m=30, tol=1e-10, dorth=True):
        psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
        sc = np.asarray(scalings, dtype=np.float64)
        err = np.zeros(1)
        ops = np.zeros(1, dtype=np.int32)
        _check(lib().oracle_sys_cayley_step(self.h, _ptr(psi), dt, _ptr(sc), m, tol, int(dorth), _ptr(err),
                                            _ptr(ops)))
        return psi, float(err[0]), int(ops[0])

    def pamp_step(self, psi, dt, scalings, m=20, order=6, dorth=False, method=0):
        psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
        sc = np.asarray(scalings, dtype=np.float64)
        est = np.zeros(2)
        dt = complex(dt)
        _check(lib().oracle_sys_pamp_step(self.h, _ptr(psi), dt.real, dt.imag, _ptr(sc), m, order, int(dorth),
                                          _ptr(est), int(method)))
        return psi, complex(est[0], est[1])

    def inner_product(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        b = np.ascontiguousarray(b, dtype=np.complex128)
        out = np.zeros(2)
        _check(lib().oracle_sys_inner_product(self.h, _ptr(a), _ptr(b), _ptr(out)))
        return complex(out[0], out[1])
