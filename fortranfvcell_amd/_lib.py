"""ctypes binding of the C ABI in include/fvcell.h (libfvcell.so).

The library is loaded from this package directory only; there is no
fallback: a missing or unloadable library raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfvcell.so")
FORTRAN_LIB_PATH = os.path.join(HERE, "libfvcell_fortran.a")

FVC_OK, FVC_MAXIT, FVC_BREAKDOWN, FVC_NONFINITE, FVC_UNSUPPORTED, FVC_DEVICE_ERROR, FVC_BAD_ARGUMENT = range(7)
NOALLOC_WORKSPACE_BYTES = 32 << 20   # FVC_NOALLOC_WORKSPACE_BYTES
STATUS_NAMES = {0: "converged", 1: "maxit", 2: "breakdown", 3: "nonfinite", 4: "unsupported",
                5: "device_error", 6: "bad_argument"}
PCG, BICGSTAB, GMRES = 1, 2, 3
CSRD, CSR, CSRC, ELLPACK = 1, 2, 3, 4
X0_AS_OPENMP, X0_AS_SEQUENTIAL, X0_ZERO, X0_GIVEN = 0, 1, 2, 3
STOP_REFERENCE, STOP_RELRES = 0, 1
ENGINE_AUTO, ENGINE_KERNELS, ENGINE_PERSISTENT = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
REC_FUSED, REC_REFERENCE = 0, 1
FLAG_COL16, FLAG_TILES = 1, 2
ORTH_MGS, ORTH_CGS2, ORTH_CGS2_LAG = 0, 1, 2

# Every symbol include/fvcell.h declares.
EXPORTS = [
    "fvc_default_options", "fvc_init", "fvc_last_error", "fvc_solver", "fvc_solve", "fvc_keep_matrix",
    "fvc_solve_dev", "fvc_solve_dev_noalloc",
    "fvc_solve_batch", "fvc_solve_batch_dev", "fvc_solve_batch_csrd", "fvc_solve_batch_csrd_dev",
    "fvc_ell_transpose_dev", "fvc_grad_k_dev",
    "fvc_assemble_heat_dev", "fvc_gform_dev", "fvc_simple_update_dev", "fvc_assemble_pc_dev",
    "fvc_assemble_mom_dev",
    "fvc_matvec", "fvc_matvec_dev", "fvc_dot", "fvc_dot_dev", "fvc_getres", "fvc_getres_dev",
    "fvc_malloc", "fvc_free", "fvc_memcpy_htod", "fvc_memcpy_dtoh", "fvc_memcpy_dtod", "fvc_memset_dev",
    "fvc_synchronize", "fvc_device_info", "fvc_event_create", "fvc_event_destroy",
    "fvc_event_record", "fvc_event_elapsed_ms",
]


class Options(C.Structure):
    _fields_ = [("x0_rule", C.c_int32), ("gmres_k", C.c_int32), ("stop_rule", C.c_int32),
                ("true_residual", C.c_int32), ("engine", C.c_int32), ("block_size", C.c_int32),
                ("grid_blocks", C.c_int32), ("precision", C.c_int32), ("batch_work_exp10", C.c_int32),
                ("barrier", C.c_int32), ("recurrence", C.c_int32), ("refresh_every", C.c_int32),
                ("refresh_drop_e3", C.c_int32), ("row_order", C.c_int32), ("lane_exchange", C.c_int32),
                ("gmres_orth", C.c_int32), ("batch_rows_per_wave", C.c_int32),
                ("batch_launch_systems", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("inner_iterations", C.c_int32),
                ("engine", C.c_int32), ("energy", C.c_double), ("conv", C.c_double),
                ("last", C.c_double), ("true_relres", C.c_double), ("seconds", C.c_double),
                ("grid_blocks", C.c_int32), ("block_size", C.c_int32), ("flags", C.c_int32),
                ("reserved_i", C.c_int32), ("reserved", C.c_double * 3)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class FvcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"fvcell error {code} ({STATUS_NAMES.get(code, '?')}): {msg}")
        self.code = code


_lib = None


def load(path: str = LIB_PATH):
    """Load libfvcell.so; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(path)
    P, I, D, SZ = C.c_void_p, C.c_int32, C.c_double, C.c_size_t
    L.fvc_default_options.argtypes = [C.POINTER(Options)]
    L.fvc_default_options.restype = None
    L.fvc_init.argtypes = [I]
    L.fvc_last_error.restype = C.c_char_p
    L.fvc_last_error.argtypes = []
    L.fvc_solver.argtypes = [P, P, P, P, P, P, P, I, I, I, D, I, I, I, I, I, C.POINTER(Stats)]
    L.fvc_solve.argtypes = [I, I, I, I, I, P, P, P, P, P, P, P, D, I, C.POINTER(Options), C.POINTER(Stats)]
    L.fvc_keep_matrix.argtypes = [I]
    L.fvc_solve_dev.argtypes = [I, I, I, I, I, P, P, P, P, P, P, P, D, I, C.POINTER(Options),
                                C.POINTER(Stats), P]
    L.fvc_solve_dev_noalloc.argtypes = [I, I, I, I, I, P, P, P, P, P, P, P, D, I, C.POINTER(Options),
                                        C.POINTER(Stats), P]
    L.fvc_solve_batch.argtypes = [I, I, P, I, P, P, P, P, P, D, I, C.POINTER(Options), P]
    L.fvc_solve_batch_dev.argtypes = [I, I, P, I, P, P, P, P, P, D, I, C.POINTER(Options), P, P]
    L.fvc_solve_batch_csrd.argtypes = [I, I, P, P, P, P, P, P, P, P, D, I, C.POINTER(Options), P]
    L.fvc_solve_batch_csrd_dev.argtypes = [I, I, P, P, P, P, P, P, P, P, D, I, C.POINTER(Options), P, P]
    L.fvc_ell_transpose_dev.argtypes = [I, I, P, P, P, P, P]
    L.fvc_grad_k_dev.argtypes = [I, I, P, P, P, P, P, P, P, P, P, P]
    L.fvc_assemble_heat_dev.argtypes = [I, I, P, P, P, P, P, P, P, P, P, P, P, P, P]
    L.fvc_gform_dev.argtypes = [I, I, I, P, P, P, P, P, P, P, P]
    L.fvc_assemble_pc_dev.argtypes = [I, I, I, I, I, I] + [P] * 16
    L.fvc_assemble_mom_dev.argtypes = [I, I, I, I, I, I, D, D, I, I, I] + [P] * 20
    L.fvc_simple_update_dev.argtypes = [I, I, P, P, P, P, P, P, P, D, D, I, P]
    L.fvc_matvec.argtypes = [I, I, I, I, P, P, P, P, P, P, P]
    L.fvc_matvec_dev.argtypes = [I, I, I, I, P, P, P, P, P, P, P, P]
    L.fvc_dot.argtypes = [P, P, I, C.POINTER(D)]
    L.fvc_dot_dev.argtypes = [P, P, I, C.POINTER(D), P]
    L.fvc_getres.argtypes = [P, P, P, I, I, I]
    L.fvc_getres_dev.argtypes = [P, P, P, I, I, I, P]
    L.fvc_malloc.argtypes = [C.POINTER(P), SZ]
    L.fvc_free.argtypes = [P]
    L.fvc_memcpy_htod.argtypes = [P, P, SZ]
    L.fvc_memcpy_dtoh.argtypes = [P, P, SZ]
    L.fvc_memset_dev.argtypes = [P, C.c_int, SZ]
    L.fvc_memcpy_dtod.argtypes = [P, P, SZ, P]
    L.fvc_synchronize.argtypes = []
    L.fvc_device_info.argtypes = [C.POINTER(I), C.c_char_p, I]
    L.fvc_event_create.argtypes = [C.POINTER(P)]
    L.fvc_event_destroy.argtypes = [P]
    L.fvc_event_record.argtypes = [P, P]
    L.fvc_event_elapsed_ms.argtypes = [P, P, C.POINTER(C.c_float)]
    for name in EXPORTS:
        f = getattr(L, name)
        if f.restype is C.c_int and name not in ("fvc_last_error",):
            f.restype = C.c_int
    _lib = L
    return L


def check(rc: int, allow=(FVC_OK,)) -> int:
    if rc not in allow:
        msg = load().fvc_last_error().decode(errors="replace")
        raise FvcError(rc, msg)
    return rc


def default_options(**kw) -> Options:
    o = Options()
    load().fvc_default_options(C.byref(o))
    names = {f for f, _ in Options._fields_ if f != "reserved"}
    for k, v in kw.items():
        if k not in names:
            raise TypeError(f"unknown solver option {k!r}")
        setattr(o, k, v)
    return o
