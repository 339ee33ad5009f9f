"""ctypes binding of libshk.so (the C ABI declared in include/shk.h).

Loading fails loudly when the library is missing: there is no CPU fallback
anywhere in the package.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libshk.so")

SHK_OK = 0
SHK_F64, SHK_F32 = 0, 1
SHK_ROWSPARSE, SHK_ROWSYMSPARSE, SHK_ARRAYROWSPARSE = 0, 1, 2
SHK_PREC_IDENTITY, SHK_PREC_JACOBI = 0, 1


class ShkIteration(ctypes.Structure):
    """Mirror of shk_iteration (Seldon's Iteration<double>)."""
    _fields_ = [("tolerance", ctypes.c_double), ("facteur_reste", ctypes.c_double),
                ("max_iter", ctypes.c_int), ("nb_iter", ctypes.c_int),
                ("error_code", ctypes.c_int), ("fail_convergence", ctypes.c_int),
                ("print_level", ctypes.c_int), ("init_guess_null", ctypes.c_int),
                ("parameter_restart", ctypes.c_int), ("hist_cap", ctypes.c_int),
                ("hist_len", ctypes.c_int), ("hist", ctypes.c_void_p),
                ("hist_it", ctypes.c_void_p)]


class ShkIrParams(ctypes.Structure):
    _fields_ = [("tolerance", ctypes.c_double), ("max_refinements", ctypes.c_int),
                ("inner_restart", ctypes.c_int), ("inner_max_iter", ctypes.c_int),
                ("inner_tolerance", ctypes.c_double), ("inner_method", ctypes.c_int),
                ("refinements", ctypes.c_int), ("total_inner_iter", ctypes.c_int),
                ("final_relres", ctypes.c_double), ("relres_hist", ctypes.c_void_p)]


# (name, restype, argtypes); every int-returning function is a status call
_P, _I, _I64, _D, _U = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_size_t
_PI, _PD = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
_PP = ctypes.POINTER(ctypes.c_void_p)
_PI64 = ctypes.POINTER(ctypes.c_int64)

SIGNATURES = [
    ("shk_last_error", ctypes.c_char_p, []),
    ("shk_version", ctypes.c_char_p, []),
    ("shk_ctx_create", _I, [_I, _PP]),
    ("shk_ctx_destroy", _I, [_P]),
    ("shk_ctx_set_stream", _I, [_P, _P]),
    ("shk_ctx_synchronize", _I, [_P]),
    ("shk_ctx_device_info", _I, [_P, _PI, _PI, ctypes.c_char_p, _I]),
    ("shk_malloc", _I, [_P, _U, _PP]),
    ("shk_free", _I, [_P, _P]),
    ("shk_memcpy_h2d", _I, [_P, _P, _P, _U]),
    ("shk_memcpy_d2h", _I, [_P, _P, _P, _U]),
    ("shk_memcpy_d2d", _I, [_P, _P, _P, _U]),
    ("shk_csr_create", _I, [_P, _I, _I, _I, _I, _I64, _P, _P, _P, _I, _PP]),
    ("shk_csr_create_layout", _I, [_P, _I, _I, _I, _I, _I64, _P, _P, _P, _I, _I, _PP]),
    ("shk_csr_layout", _I, [_P, _PI]),
    ("shk_csr_destroy", _I, [_P]),
    ("shk_csr_set_values", _I, [_P, _P, _P, _I]),
    ("shk_csr_info", _I, [_P, _PI, _PI, _PI64, _PI64, _PI, _PI]),
    ("shk_csr_spmv_bytes", _D, [_P, _I]),
    ("shk_spmv", _I, [_P, _P, _D, _P, _D, _P]),
    ("shk_spmv_out", _I, [_P, _P, _D, _P, _D, _P, _P]),
    ("shk_mlt", _I, [_P, _P, _P, _P]),
    ("shk_spmv_trans", _I, [_P, _P, _D, _P, _D, _P]),
    ("shk_csr_write", _I, [_P, _P, ctypes.c_char_p]),
    ("shk_sor", _I, [_P, _P, _P, _P, _D, _I, _I, _I]),
    ("shk_csr_read", _I, [_P, ctypes.c_char_p, _I, _I, _I, _PP]),
    ("shk_vec_write", _I, [_P, _I, ctypes.c_int64, _P, ctypes.c_char_p, _I]),
    ("shk_vec_read", _I, [_P, _I, ctypes.c_char_p, _I, ctypes.c_int64, _P, ctypes.POINTER(ctypes.c_int64)]),
    ("shk_csr_write_hb", _I, [_P, _P, ctypes.c_char_p]),
    ("shk_csr_read_hb", _I, [_P, ctypes.c_char_p, _I, _I, _PP]),
    ("shk_coo_to_csr", _I, [_P, _I, ctypes.c_int64, _P, _P, _P, _I, _I, _I, _I, ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_int), _P, _P, _P]),
    ("shk_mlt_trans", _I, [_P, _P, _P, _P]),
    ("shk_mlt_dot", _I, [_P, _P, _P, _P, _PD]),
    ("shk_dot", _I, [_P, _I, _I64, _P, _P, _PD]),
    ("shk_dot_device", _I, [_P, _I, _I64, _P, _P, _P]),
    ("shk_nrm2", _I, [_P, _I, _I64, _P, _PD]),
    ("shk_axpy", _I, [_P, _I, _I64, _D, _P, _P]),
    ("shk_scal", _I, [_P, _I, _I64, _D, _P]),
    ("shk_copy", _I, [_P, _I, _I64, _P, _P]),
    ("shk_fill", _I, [_P, _I, _I64, _D, _P]),
    ("shk_axpy_dot", _I, [_P, _I, _I64, _D, _P, _P, _P, _PD]),
    ("shk_convert", _I, [_P, _I64, _I, _P, _I, _P]),
    ("shk_jacobi", _I, [_P, _P, _P, _P]),
    ("shk_cg", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_cg_noalloc", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_bicgstab", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_cgs", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_bicg", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_tfqmr", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_qmr", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_cgne", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_lsqr", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_bicgcr", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_qmrsym", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_minres", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_symmlq", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_qcgs", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_gcr", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_bicgstabl", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_gmres", _I, [_P, _P, _P, _P, _I, ctypes.POINTER(ShkIteration), _PI]),
    ("shk_solve_ir", _I, [_P, _P, _P, _P, _P, _I, ctypes.POINTER(ShkIrParams), _PI]),
    ("shk_sddmm", _I, [_P, _P, _P, _P, _P]),
    ("shk_cg_batched", _I, [_P, _P, _I, _P, _P, _P, _D, _I, _I, _I, _I, _P, _P]),
    ("shk_cg_batched_ir", _I, [_P, _P, _P, _I, _P, _P, _P, _D, _I, _D, _I, _I, _I, _P, _P]),
    ("shk_time_spmv", _I, [_P, _P, _I, _P, _P, _I, _PD]),
]

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: the HIP library must be built "
                           "(python -c 'import __graft_entry__ as g; g.build()'); "
                           "there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return [name for name, _, _ in SIGNATURES]


class ShkError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"shk status {status}: {msg}")
        self.status = status


def check(status):
    if status != SHK_OK:
        raise ShkError(status, load().shk_last_error().decode(errors="replace"))
    return status
