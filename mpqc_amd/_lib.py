"""ctypes binding of ``libmpqc_ccsd.so`` (the C-ABI in ``include/mpqc_ccsd.h``).

The library is the only compute path: there is no CPU fallback.  If it is not
built, loading raises ``ImportError`` naming the build step.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmpqc_ccsd.so")

OK = 0
ERR_ARG, ERR_HIP, ERR_STATE, ERR_NONFINITE, ERR_DIIS, ERR_NOMEM, ERR_NOCONV = 1, 2, 3, 4, 5, 6, 7
PREC_F64, PREC_F32, PREC_BF16X3, PREC_F16X3, PREC_BF16X6 = 0, 1, 2, 3, 4
LAYOUT_DENSE, LAYOUT_TA_TILES = 0, 1
MEM_HOST, MEM_DEVICE = 0, 1
INT_FIA, INT_FIJ, INT_FAB, INT_GIJAB, INT_GIJKL, INT_GIAJB, INT_GIABC, INT_GIJKA, INT_GABCD = range(9)
INT_XAI, INT_XIJ, INT_XAB = 9, 10, 11
DF_INTEGRAL_IDS = {"Xai": INT_XAI, "Xij": INT_XIJ, "Xab": INT_XAB}
INTEGRAL_IDS = {"Fia": INT_FIA, "Fij": INT_FIJ, "Fab": INT_FAB, "Gijab": INT_GIJAB,
                "Gijkl": INT_GIJKL, "Giajb": INT_GIAJB, "Giabc": INT_GIABC,
                "Gijka": INT_GIJKA, "Gabcd": INT_GABCD}

# every symbol declared in include/mpqc_ccsd.h
EXPORTS = [
    "mpqc_ccsd_version", "mpqc_ccsd_ctx_create", "mpqc_ccsd_ctx_destroy", "mpqc_ccsd_last_error",
    "mpqc_ccsd_set_option", "mpqc_ccsd_set_integral", "mpqc_ccsd_set_df_basis",
    "mpqc_ccsd_transform_ao", "mpqc_ccsd_transform_df", "mpqc_ccsd_get_integral",
    "mpqc_ccsd_triples", "mpqc_ccsd_eom_ee", "mpqc_ccsd_eom_stats", "mpqc_ccsd_set_direct_ao",
    "mpqc_ccsd_set_orbital_energies",
    "mpqc_ccsd_solver_init", "mpqc_ccsd_set_amplitudes", "mpqc_ccsd_get_amplitudes",
    "mpqc_ccsd_get_residuals", "mpqc_ccsd_get_dressed_fock", "mpqc_ccsd_init_guess",
    "mpqc_ccsd_residual", "mpqc_ccsd_energy_error", "mpqc_ccsd_update", "mpqc_ccsd_solve",
    "mpqc_ccsd_get_trace", "mpqc_ccsd_pp_ladder", "mpqc_ccsd_gemm", "mpqc_ccsd_gemm_noalloc",
    "mpqc_ccsd_get_kernel_times", "mpqc_ccsd_jacobian_transpose", "mpqc_ccsd_lambda_solve",
    "mpqc_ccsd_get_lambda", "mpqc_ccsd_energy_gradient", "mpqc_ccsd_get_integral_gradient",
    "mpqc_ccsd_batch_create", "mpqc_ccsd_batch_destroy", "mpqc_ccsd_batch_last_error",
    "mpqc_ccsd_batch_set_integral", "mpqc_ccsd_batch_set_amplitudes", "mpqc_ccsd_batch_residual",
    "mpqc_ccsd_batch_get_residuals", "mpqc_ccsd_batch_tune", "mpqc_ccsd_batch_time",
    "mpqc_ccsd_batch_time_stages", "mpqc_ccsd_batch_get_stages",
]

_lib = None


class CCSDError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


def load():
    """Loads the library (once) and declares the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`"
                          " or `python mpqc_amd/build.py`")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    D = ctypes.c_double
    I64 = ctypes.c_int64
    pI64 = ctypes.POINTER(ctypes.c_int64)
    pD = ctypes.POINTER(ctypes.c_double)
    pI = ctypes.POINTER(ctypes.c_int)
    sig = {
        "mpqc_ccsd_version": (ctypes.c_char_p, []),
        "mpqc_ccsd_ctx_create": (I, [I, I, pI64, I, pI64, I, I, ctypes.POINTER(P)]),
        "mpqc_ccsd_ctx_destroy": (I, [P]),
        "mpqc_ccsd_last_error": (ctypes.c_char_p, [P]),
        "mpqc_ccsd_set_option": (I, [P, ctypes.c_char_p, D]),
        "mpqc_ccsd_set_integral": (I, [P, I, P, I, I]),
        "mpqc_ccsd_set_orbital_energies": (I, [P, pD, pD]),
        "mpqc_ccsd_set_df_basis": (I, [P, I, pI64, I]),
        "mpqc_ccsd_transform_ao": (I, [P, I, P, P, P, P, I]),
        "mpqc_ccsd_get_integral": (I, [P, I, P, I, I]),
        "mpqc_ccsd_triples": (I, [P, pD, pD]),
        "mpqc_ccsd_eom_ee": (I, [P, I, I, D, I, P, P]),
        "mpqc_ccsd_eom_stats": (I, [P, P, P]),
        "mpqc_ccsd_set_direct_ao": (I, [P, I, P, P, P, I]),
        "mpqc_ccsd_transform_df": (I, [P, I, P, P, P, P, P, I]),
        "mpqc_ccsd_solver_init": (I, [P, I, I, D, I, I]),
        "mpqc_ccsd_set_amplitudes": (I, [P, P, P, I, I]),
        "mpqc_ccsd_get_amplitudes": (I, [P, P, P, I, I]),
        "mpqc_ccsd_get_residuals": (I, [P, P, P, I, I]),
        "mpqc_ccsd_get_dressed_fock": (I, [P, P, P]),
        "mpqc_ccsd_init_guess": (I, [P]),
        "mpqc_ccsd_residual": (I, [P]),
        "mpqc_ccsd_energy_error": (I, [P, pD, pD]),
        "mpqc_ccsd_update": (I, [P]),
        "mpqc_ccsd_solve": (I, [P, I, D, I, pD, pI]),
        "mpqc_ccsd_get_trace": (I, [P, I, pI, P, P, P, P, P]),
        "mpqc_ccsd_pp_ladder": (I, [I, I, P, P, P, D, D, I, P]),
        "mpqc_ccsd_get_kernel_times": (I, [P, I, pI, P, P, P, P]),
        "mpqc_ccsd_jacobian_transpose": (I, [P, P, P, P, P, I, I]),
        "mpqc_ccsd_lambda_solve": (I, [P, I, D, I, pI]),
        "mpqc_ccsd_get_lambda": (I, [P, P, P, I, I]),
        "mpqc_ccsd_energy_gradient": (I, [P]),
        "mpqc_ccsd_get_integral_gradient": (I, [P, I, P, I, I]),
        "mpqc_ccsd_batch_create": (I, [I, pI, pI, I, ctypes.POINTER(P)]),
        "mpqc_ccsd_batch_destroy": (I, [P]),
        "mpqc_ccsd_batch_last_error": (ctypes.c_char_p, [P]),
        "mpqc_ccsd_batch_set_integral": (I, [P, I, I, P, I, I]),
        "mpqc_ccsd_batch_set_amplitudes": (I, [P, I, P, P, I, I]),
        "mpqc_ccsd_batch_residual": (I, [P]),
        "mpqc_ccsd_batch_get_residuals": (I, [P, I, P, P, I, I]),
        "mpqc_ccsd_batch_tune": (I, [P, I, I, pD]),
        "mpqc_ccsd_batch_time": (I, [P, I, pD]),
        "mpqc_ccsd_batch_time_stages": (I, [P, I]),
        "mpqc_ccsd_batch_get_stages": (I, [P, I, pI, P, P, P, P, P, P]),
        "mpqc_ccsd_gemm": (I, [I, I, I, I, I, D, P, I64, I64, P, I64, I64, D, P, I64, I64, I, I, P]),
        "mpqc_ccsd_gemm_noalloc": (I, [I, I, I, I, I, D, P, I64, I64, P, I64, I64, D, P, I64, I64, I, I, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(status, ctx=None):
    if status != OK:
        lib = load()
        msg = lib.mpqc_ccsd_last_error(ctx).decode(errors="replace")
        raise CCSDError(status, msg)
