"""ctypes binding of liblfr.so (the C ABI declared in include/lfr.h).

The library is loaded from ``lfric_core_amd/_lib/liblfr.so``.  There is no
fallback: if the library is missing or a HIP device is absent every call
raises ``LfrError``.
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "liblfr.so"

LFR_OK = 0
ERRORS = {1: "LFR_ERR_ARG", 2: "LFR_ERR_HIP", 3: "LFR_ERR_NOMEM", 4: "LFR_ERR_NOT_CONVERGED",
          5: "LFR_ERR_NAN", 6: "LFR_ERR_SINGULAR"}


class LfrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_D = C.c_double

# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "lfr_init": [_I, C.POINTER(_P)],
    "lfr_finalise": [_P],
    "lfr_last_error": [_P],
    "lfr_set_stream": [_P, _P],
    "lfr_get_stream": [_P],
    "lfr_synchronize": [_P],
    "lfr_malloc": [_P, C.c_size_t, C.POINTER(_P)],
    "lfr_free": [_P, _P],
    "lfr_memcpy_h2d": [_P, _P, _P, C.c_size_t],
    "lfr_memcpy_d2h": [_P, _P, _P, C.c_size_t],
    "lfr_space_create": [_P, _P, _I, _I, _I, _I, C.POINTER(_P)],
    "lfr_space_destroy": [_P],
    "lfr_plan_create": [_P, _P, _P, _P, _I, _I, _I, C.POINTER(_P)],
    "lfr_plan_destroy": [_P],
    "lfr_plan_info": [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I64)],
    "lfr_matrix_vector": [_P, _I, _P, _P, _P, _P, _P, _P, _I],
    "lfr_dg_matrix_vector": [_P, _I, _P, _P, _P, _P, _P],
    "lfr_dg_inc_matrix_vector": [_P, _I, _P, _P, _P, _P, _P],
    "lfr_transpose_matrix_vector": [_P, _I, _P, _P, _P, _P, _P, _P, _I],
    "lfr_matrix_vector_ndata": [_P, _I, _P, _P, _P, _P, _P, _P, _I, _I],
    "lfr_compute_mass_matrix_w0": [_P, _I, _P, _P, _P, _P, _P, _P],
    "lfr_compute_mass_matrix": [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lfr_compute_derham_matrices": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lfr_intergrid_create": [_P, _P, _P, _P, _I, _I, C.POINTER(_P)],
    "lfr_intergrid_destroy": [_P],
    "lfr_restrict_scalar": [_P, _I, _P, _P, _P],
    "lfr_prolong_scalar": [_P, _I, _P, _P, _P],
    "lfr_restrict_w2": [_P, _I, _P, _P, _P],
    "lfr_prolong_w2": [_P, _I, _P, _P, _P],
    "lfr_tri_solve": [_P, _I, _P, _P, _P, _P, _P, _P],
    "lfr_operator_tri_solve": [_P, _I, _P, _P, _P, _P],
    "lfr_lma_pack": [_P, _I, _P, _P, _P, C.POINTER(_P)],
    "lfr_lma_packed_destroy": [_P],
    "lfr_lma_packed_info": [_P, C.POINTER(C.c_longlong)],
    "lfr_matrix_vector_packed": [_P, _P, _P, _P, _I],
    "lfr_lma_outer": [_P, _I, _P, _P, _P, _P, _P, _D, _I],
    "lfr_vec_setval": [_P, _I, _I64, _P, _D],
    "lfr_vec_copy": [_P, _I, _I64, _P, _P],
    "lfr_vec_axpy": [_P, _I, _I64, _D, _P, _P],
    "lfr_vec_aypx": [_P, _I, _I64, _D, _P, _P],
    "lfr_vec_axpby": [_P, _I, _I64, _D, _P, _D, _P, _P],
    "lfr_vec_scale": [_P, _I, _I64, _D, _P],
    "lfr_vec_divide": [_P, _I, _I64, _P, _P, _P],
    "lfr_vec_multiply": [_P, _I, _I64, _P, _P, _P],
    "lfr_vec_convert": [_P, _I, _P, _I, _P, _I64],
    "lfr_vec_dot": [_P, _I, _I64, _P, _P, C.POINTER(_D)],
    "lfr_vec_dot_dev": [_P, _I, _I64, _P, _P, _P],
    "lfr_vec_dot_dev_noalloc": [_P, _I, _I64, _P, _P, _P],
    "lfr_tune_set": [_P, C.c_char_p, _I],
    "lfr_tune_get": [_P, C.c_char_p, C.POINTER(_I)],
    "lfr_tune_count": [_P, C.POINTER(_I)],
    "lfr_tune_entry": [_P, _I, C.c_char_p, _I, C.POINTER(_I)],
    "lfr_tune_clear": [_P],
    "lfr_tune_save": [_P, C.c_char_p],
    "lfr_tune_load": [_P, C.c_char_p],
    # solver entry points take structure pointers; lfric_core_amd.solvers
    # installs the typed signatures before calling them
    "lfr_columnwise_op_appinv": [_P, _I, _P, _P, _P, _I, _I, _I, _P, _P, _I],
    "lfr_mg_create": [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _D, _I, _I, _I, C.POINTER(_P)],
    "lfr_mg_destroy": [_P],
    "lfr_mg_block_sizes": [_P, _P],
    "lfr_mg_solve": [_P, _P, _P, _I, _P],
    "lfr_space_set_boundary": [_P, _P],
    "lfr_enforce_bc": [_P, _I, _P, _P],
    "lfr_cma_create": [_P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P, _P, _I, _P, _P, C.POINTER(_P)],
    "lfr_cma_destroy": [_P],
    "lfr_columnwise_op_asm": [_P, _I, _P, _P, _P, _I],
    "lfr_columnwise_op_app": [_P, _I, _P, _P, _P, _P, _P, _P],
    "lfr_multi_create": [_P, _I, _I, _P, _P, _P, _P, _P, _I, _I, C.POINTER(_P)],
    "lfr_multi_destroy": [_P],
    "lfr_multi_info": [_P, _P, _P, _P, _P],
    "lfr_matrix_vector_multi": [_P, _P, _P, _P, _I, _I],
    "lfr_cg": None,
    "lfr_refined_cg": None,
    "lfr_refined_cg_multi": None,
    "lfr_gmres": None,
    # n-field vectors, abstract operators, preconditioners and the generic
    # solvers: typed in lfric_core_amd.solvers
    "lfr_vector_size": None,
    "lfr_vector_length": None,
    "lfr_vector_dot": None,
    "lfr_vector_norm": None,
    "lfr_vector_field_norm": None,
    "lfr_linop_lma": None,
    "lfr_linop_block": None,
    "lfr_linop_callback": None,
    "lfr_linop_destroy": [_P],
    "lfr_linop_apply": [_P, _P, _P, _P],
    "lfr_precon_null": None,
    "lfr_precon_diagonal": None,
    "lfr_precon_columnwise": None,
    "lfr_precon_callback": None,
    "lfr_precon_destroy": [_P],
    "lfr_precon_apply": [_P, _P, _P, _P],
    "lfr_iterative_solve": None,
    "lfr_iterative_solve_bytes": [_P, C.POINTER(C.c_double)],
    "lfr_mm_diagonal": [_P, _I, _P, _P, _P, _P],
    "lfr_cg_tape_create": None,
    "lfr_cg_tape_destroy": [_P],
    "lfr_cg_tape_info": [_P, C.POINTER(_I64), C.POINTER(_I)],
    "lfr_cg_tape_forward": [_P, _P, _P, _P],
    "lfr_cg_tape_reverse": [_P, _P, _P, _P, _P, _P, _I],
}
_RESTYPE = {"lfr_last_error": C.c_char_p, "lfr_get_stream": _P}

_lib = None
_lock = threading.Lock()


def lib():
    """Load liblfr.so once; raise if it is not built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not LIB_PATH.exists():
                    raise LfrError(2, f"{LIB_PATH} is not built: run "
                                      "`make -C lfric_core_amd/csrc` or __graft_entry__.build()")
                L = C.CDLL(str(LIB_PATH))
                for name, args in SIGNATURES.items():
                    if args is None or not hasattr(L, name):
                        continue
                    fn = getattr(L, name)
                    fn.argtypes = args
                    fn.restype = _RESTYPE.get(name, _I)
                _lib = L
    return _lib


def exported_symbols() -> list[str]:
    """Names in SIGNATURES that the loaded library exports."""
    L = lib()
    return [n for n in SIGNATURES if hasattr(L, n)]


def check(rc: int, ctx=None) -> None:
    if rc != LFR_OK:
        msg = lib().lfr_last_error(ctx).decode() if ctx else "(no context)"
        raise LfrError(rc, msg)
