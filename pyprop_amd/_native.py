"""ctypes binding of the gfx950 library libpyprop_amd.so (C ABI in include/pyprop_amd.h).

There is no fallback: if the library is missing or no GPU is visible, the product path
raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_lib" / "libpyprop_amd.so"

OP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)


class PPCplx(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class NativeError(RuntimeError):
    pass


STORAGE = {"Simp": 0, "Herm": 1, "Diag": 2, "Ident": 3, "Band": 4, "BandNH": 5, "Dense": 6}
OVERLAP_OPS = {"multiply": 0, "solve": 1, "sqrt_multiply": 2, "sqrt_solve": 3}

# (name, restype, argtypes) for every exported symbol of include/pyprop_amd.h
P = C.c_void_p
I = C.c_int
D = C.c_double
I64 = C.c_int64
SIGNATURES = [
    ("pp_last_error", C.c_char_p, []),
    ("pp_version", C.c_char_p, []),
    ("pp_stream_synchronize", I, [P]),
    ("pp_tensor_potential_multiply", I, [I, P, P, P, D, P, P, P, P, P, P, P]),
    ("pp_tp_create", I, [I, P, P, P, P, P, P, P, P]),
    ("pp_tp_multiply", I, [P, D, P, P, P]),
    ("pp_tp_multiply_f32", I, [P, C.c_float, P, P, P]),
    ("pp_tp_info", I, [P, P, P, P, P]),
    ("pp_tp_destroy", I, [P]),
    ("pp_expand_potential", I, [I, I64, I64, P, P, P, P]),
    ("pp_bspline_projection_create", I, [I, I, I, P, P, P, P, P, I, I, P, I, P]),
    ("pp_bspline_projection_destroy", I, [P]),
    ("pp_bspline_projection_info", I, [P, P, P, P]),
    ("pp_bspline_represent", I, [P, I64, I64, P, P, I, I, P]),
    ("pp_potential_grid_multipole", I, [I64, I, P, P, I, P, I, P, P, P]),
    ("pp_potential_grid_separable", I, [I64, I, P, I, P, I, P, P, P]),
    ("pp_potential_represent_multipole", I, [P, P, I64, I, P, P, P, P, P, P, I, P]),
    ("pp_potential_represent_separable", I, [P, P, I64, I, P, P, P, P, P, I, P]),
    ("pp_overlap_create", I, [I, I, P, P]),
    ("pp_overlap_get_factor", I, [P, P]),
    ("pp_overlap_destroy", I, [P]),
    ("pp_overlap_apply", I, [P, I, I, I, P, I, P, P]),
    ("pp_repr_create", I, [I, P, P, P]),
    ("pp_repr_destroy", I, [P]),
    ("pp_repr_overlap", I, [P, I, I, P, P]),
    ("pp_repr_inner_product", I, [P, P, P, P, P, P]),
    ("pp_repr_inner_product_noalloc", I, [P, P, P, P, P]),
    ("pp_ham_create", I, [P, I, P, P]),
    ("pp_ham_destroy", I, [P]),
    ("pp_ham_multiply_nooverlap", I, [P, P, P, P, P]),
    ("pp_ham_multiply", I, [P, P, P, P, P]),
    ("pp_ham_multiply_balanced", I, [P, P, P, P, P]),
    ("pp_vec_dot", I, [I64, P, P, P, P]),
    ("pp_vec_norm", I, [I64, P, P, P]),
    ("pp_gmres_create", I, [I64, I, D, I, P]),
    ("pp_gmres_destroy", I, [P]),
    ("pp_gmres_solve", I, [P, OP_FN, P, P, P, P]),
    ("pp_gmres_stats", I, [P, P, P, P]),
    ("pp_gmres_error_list", I, [P, P]),
    ("pp_pamp_create", I, [I64, I, I, I, P]),
    ("pp_pamp_destroy", I, [P]),
    ("pp_pamp_set_exponentiation", I, [P, I]),
    ("pp_pamp_propagate", I, [P, OP_FN, P, P, PPCplx, P]),
    ("pp_pamp_stats", I, [P, P, P, P]),
    ("pp_precond_create", I, [I, I, I, I, P, P, P, I, P, I, P, PPCplx, PPCplx, P]),
    ("pp_precond_destroy", I, [P]),
    ("pp_precond_solve", I, [P, P, P]),
    ("pp_cayley_create", I, [P, P, I, D, I, P]),
    ("pp_cayley_destroy", I, [P]),
    ("pp_cayley_step", I, [P, P, D, P, P]),
    ("pp_cayley_stats", I, [P, P, P]),
    ("pp_pamp_step_balanced", I, [P, P, P, PPCplx, P, P]),
    ("pp_kron_create", I, [I, I, I, P, P, P, P, P, P]),
    ("pp_kron_multiply", I, [P, PPCplx, P, P]),
    ("pp_kron_destroy", I, [P]),
    ("pp_kron_solve", I, [P, PPCplx, P, P]),
    ("pp_kron_solve_f32", I, [P, PPCplx, P, P]),
    ("pp_kron_solve_f32_into", I, [P, PPCplx, P, P, P]),
    ("pp_kron_solve_f32_overlap", I, [P, PPCplx, P, P, P, P]),
    ("pp_kron_set_form", I, [P, I]),
    ("pp_cayley_set_kron", I, [P, P]),
    ("pp_learned_create", I, [I, I, I, C.c_ulonglong, P]),
    ("pp_learned_destroy", I, [P]),
    ("pp_learned_param_count", I, [P, P]),
    ("pp_learned_get_params", I, [P, P]),
    ("pp_learned_set_params", I, [P, P]),
    ("pp_learned_get_grads", I, [P, P]),
    ("pp_learned_set_scale", I, [P, C.c_float]),
    ("pp_learned_forward", I, [P, I, P, P, P, P]),
    ("pp_learned_gradient", I, [P, I, P, P, P, P, P]),
    ("pp_learned_train_step", I, [P, I, P, P, P, C.c_float, P, P]),
    ("pp_learned_prepare", I, [P, P, D, P, P, P, P, P, P]),
    ("pp_learned_apply", I, [P, P, D, P, P, P]),
    ("pp_gemm_bf16", I, [I, I, I, P, I64, I64, P, I64, I64, P, P]),
    ("pp_mixed_create", I, [P, P, I, D, D, I, P]),
    ("pp_mixed_destroy", I, [P]),
    ("pp_mixed_set_learned", I, [P, P]),
    ("pp_mixed_set_learned_mode", I, [P, I]),
    ("pp_mixed_set_fallback", I, [P, P, I]),
    ("pp_mixed_step", I, [P, P, D, P, P]),
    ("pp_mixed_stats", I, [P, P, P]),
    ("pp_mixed_set_factored", I, [P, P]),
    ("pp_mixed_set_device_mode", I, [P, I]),
    ("pp_mixed_set_gram_schmidt", I, [P, I]),
    ("pp_mixed_set_deferred_finish", I, [P, I]),
    ("pp_mixed_set_overlap_from_solve", I, [P, I]),
    ("pp_mixed_set_graph", I, [P, I]),
    ("pp_mixed_graph_stats", I, [P, P, P]),
    ("pp_cayley_set_factored", I, [P, P]),
    ("pp_fham_create", I, [I, I, I, I, P, P, P, I, P]),
    ("pp_fham_destroy", I, [P]),
    ("pp_fham_multiply", I, [P, P, P, P, P, PPCplx, P]),
    ("pp_fham_multiply_f32", I, [P, P, P, P, P, PPCplx, P]),
    ("pp_fham_set_radial_variant", I, [P, I, I]),
    ("pp_fham_get_radial_variant", I, [P, I, P]),
    ("pp_fham_set_angular_waves", I, [P, I, I]),
    ("pp_fham_get_angular_waves", I, [P, I, P]),
    ("pp_fham_set_angular_variant", I, [P, I, I]),
    ("pp_fham_band_terms", I, [P, P]),
    ("pp_fham_get_angular_variant", I, [P, I, P]),
    ("pp_fham_tune", I, [P, I, I, P, P, P]),
    ("pp_fham_set_overlap", I, [P, P, P]),
    ("pp_piram_create", I, [I64, I, I, D, I, I, P]),
    ("pp_piram_destroy", I, [P]),
    ("pp_piram_solve", I, [P, P, P, P, P]),
    ("pp_piram_eigenvalues", I, [P, P, P, P, P]),
    ("pp_piram_eigenvector", I, [P, I, P, P]),
    ("pp_piram_stats", I, [P, P, P, P]),
    ("pp_shift_invert_create", I, [P, P, I, D, I, D, P]),
    ("pp_shift_invert_destroy", I, [P]),
    ("pp_shift_invert_set_scalings", I, [P, P]),
    ("pp_shift_invert_set_factored", I, [P, P]),
    ("pp_shift_invert_solve", I, [P, P, P, P]),
    ("pp_shift_invert_apply", I, [P, P, P, P]),
    ("pp_shift_invert_set_negate", I, [P, I]),
    ("pp_shift_invert_stats", I, [P, P, P, P]),
    ("pp_ham_op_create", I, [P, P, I, P]),
    ("pp_ham_op_destroy", I, [P]),
    ("pp_ham_op_apply", I, [P, P, P, P]),
    ("pp_fham_multiply_overlap", I, [P, P, P, P, P, PPCplx, PPCplx, P]),
    ("pp_fham_multiply_overlap_f32", I, [P, P, P, P, P, PPCplx, PPCplx, P]),
    ("pp_batch_create", I, [I, P, P, I, P, P, P, P, P, I, I, I, P]),
    ("pp_batch_destroy", I, [P]),
    ("pp_batch_set_state", I, [P, P, P]),
    ("pp_batch_get_state", I, [P, P, P, P]),
    ("pp_batch_autotune", I, [P, D, I, P, P, P]),
    ("pp_batch_set_threads", I, [P, I]),
    ("pp_batch_step", I, [P, D, I, P]),
    ("pp_batch_capture", I, [P, D, I, P]),
    ("pp_batch_replay", I, [P, I, P]),
    ("pp_batch_info", I, [P, P, P, P, P]),
]

_lib = None


def load(path: Path | None = None):
    """Load (once) and return the native library; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise NativeError("%s is missing: build it with `python -m pyprop_amd.build` "
                          "(there is no CPU fallback)" % p)
    lib = C.CDLL(str(p))
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().pp_last_error().decode()
        raise NativeError("pyprop_amd error %d: %s" % (rc, msg))


def cplx(z) -> PPCplx:
    z = complex(z)
    return PPCplx(z.real, z.imag)
