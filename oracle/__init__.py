"""The oracle: a CPU restatement of fortranFVCell's solver path (fvc_oracle.c)
and, where it was built, the reference's own Fortran compiled from its
sources (oracle/_ref/libfvcref.so).

TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py use this package as the checker; the product
(fortranfvcell_amd) never imports it.

Array arguments follow the reference: pointers to element 1, 1-based index
values, float64 values, int32 indices.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import sys
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libfvcref.so")

# matvec kinds of fvc_oracle.c
MV_CSRD, MV_CSR, MV_CSRC, MV_CSRC_SYM, MV_CSRD_SYM, MV_ELLPACK, MV_CSRD_ILU2 = 1, 2, 3, 4, 5, 6, 7
DOT_SEQ, DOT_L4 = 0, 1
# matvec kinds of ref_bridge.f90
REF_KIND = {MV_CSRD: 1, MV_CSR: 2, MV_CSRC: 3, MV_CSRC_SYM: 4, MV_CSRD_SYM: 5, MV_ELLPACK: 6,
            MV_CSRD_ILU2: 7}


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("inner", C.c_int32), ("status", C.c_int32),
                ("pad", C.c_int32), ("energy", C.c_double), ("conv", C.c_double),
                ("last", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


_lib = None
_ref = None


def build(ref: bool | None = None) -> None:
    """Compile the C restatement; when the reference's sources exist, also the
    reference's solver path (oracle/_ref/libfvcref.so), the reference program
    linked with its own solver (oracle/_ref/libmvf_ref.so) and, once the
    drop-in is built, the same program linked against the drop-in
    (oracle/_ref/libmvf_dropin.so)."""
    targets = ["oracle"]
    if ref or (ref is None and os.path.isdir("/root/reference/src")):
        targets += ["ref", "refprog"]
        pkg = os.path.join(os.path.dirname(HERE), "fortranfvcell_amd")
        if all(os.path.exists(os.path.join(pkg, f)) for f in ("libfvcell.so", "libfvcell_fortran.a")):
            targets.append("dropin")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


DROPIN_PATH = os.path.join(HERE, "_ref", "libmvf_dropin.so")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run oracle.build())")
        L = C.CDLL(LIB_PATH)
        D, I, P = C.c_double, C.c_int32, C.c_void_p
        L.orc_dot.restype = D
        L.orc_dot.argtypes = [P, P, I]
        L.orc_dot_L4.restype = D
        L.orc_dot_L4.argtypes = [P, P, I]
        L.orc_matvec.argtypes = [I, I, P, P, P, P, P, P, P]
        L.orc_pcg.argtypes = [I, I, P, P, P, P, P, P, P, P, D, I, I, I, C.POINTER(Stats)]
        L.orc_pbicgstab.argtypes = [I, I, P, P, P, P, P, P, P, P, D, I, I, I, C.POINTER(Stats)]
        L.orc_gmres.argtypes = [I, I, P, P, P, P, P, P, P, P, I, D, I, C.POINTER(Stats)]
        L.orc_solver.argtypes = [I, I, I, I, I, P, P, P, P, P, P, P, D, I, C.POINTER(Stats)]
        L.orc_getres.argtypes = [P, P, P, I, I, I]
        _lib = L
    return _lib


def have_ref() -> bool:
    return os.path.exists(REF_PATH)


def ref(pool_words: int = 1 << 27, omp: int = 0, nthreads: int = 1):
    """The reference library (None when it was not built)."""
    global _ref
    if _ref is None:
        if not have_ref():
            return None
        L = C.CDLL(REF_PATH)
        D, I, P = C.c_double, C.c_int32, C.c_void_p
        L.ref_init.argtypes = [C.c_int64, I, I]
        L.ref_open_log.argtypes = [C.c_char_p, I]
        L.ref_close_log.argtypes = []
        L.ref_solver.argtypes = [P, P, P, P, P, P, P, I, I, I, D, I, I, I, I]
        L.ref_matvec.argtypes = [I, I, P, P, P, P, P, P, P]
        L.ref_dot.restype = D
        L.ref_dot.argtypes = [P, P, I]
        L.ref_dot_ompl4.restype = D
        L.ref_dot_ompl4.argtypes = [P, P, I]
        L.ref_gmres.argtypes = [I, I, I, P, P, P, P, P, P, P, P, I, D, I, I]
        L.ref_pcg.argtypes = [I, I, I, I, P, P, P, P, P, P, P, P, D, I, I, I]
        L.ref_pbicgstab.argtypes = [I, I, I, I, P, P, P, P, P, P, P, P, D, I, I, I]
        L.ref_wtime.restype = D
        L.ref_wtime.argtypes = []
        # the reference's init_malloc prints a banner on unit 6; keep it off
        # the caller's stdout (a benchmark's stdout is one JSON line)
        sys.stdout.flush()
        saved, quiet = os.dup(1), os.open(os.devnull, os.O_WRONLY)
        try:
            os.dup2(quiet, 1)
            L.ref_init(pool_words, omp, nthreads)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
            os.close(quiet)
        dev_null = b"/dev/null"
        L.ref_open_log(dev_null, len(dev_null))   # unit 13, the solver log
        _ref = L
    return _ref


def ref_set_omp(omp: int, nthreads: int) -> None:
    r = ref()
    r.ref_init(0, omp, nthreads)


def ref_open_log(path: str) -> None:
    b = path.encode()
    ref().ref_open_log(b, len(b))


# the solver's unit-13 line (src/Iterativos.f:91-93,216-218,429-431)
LOG_RE = re.compile(r"^(.*?)it\s*(\d+) energy norm\s+(\S+) tol\s+(\S+)\s*$")


def ref_logged(fn):
    """Run fn() (a call into the reference library) with unit 13 pointed at a
    temporary file; return (fn's result, label, iterations, energy) parsed
    from the line the reference's solver wrote there."""
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        log = os.path.join(td, "solver.log")
        ref_open_log(log)
        out = fn()
        ref().ref_close_log()
        ref_open_log("/dev/null")
        line = open(log).read().strip().splitlines()[-1]
    m = LOG_RE.match(line)
    return out, m.group(1).strip(), int(m.group(2)), float(m.group(3).replace("D", "E"))


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _dummy(a):
    return a if a is not None else np.zeros(1)


def kind_of(fmt: dict, unsym: bool = True) -> int:
    m = fmt["matriz"]
    if m == 4:
        return MV_ELLPACK
    if m == 2:
        return MV_CSR
    if m == 1:
        return MV_CSRD_SYM if fmt.get("sym") else MV_CSRD
    if m == 3:
        return MV_CSRC_SYM if fmt.get("sym") else MV_CSRC
    raise ValueError(m)


def _vals(fmt):
    """(ad, au, al) as the reference's solver receives them."""
    ad = _f64(fmt["ad"])
    al = _f64(fmt.get("al"))
    au = _f64(fmt.get("au"))
    if au is None:
        au = al          # ELLPACK and unsymmetric CSRD: au == al (src/Datastruct.f:77,175)
    return ad, _dummy(au), _dummy(al)


def matvec(fmt: dict, x: np.ndarray, kind: int | None = None, use_ref: bool = False) -> np.ndarray:
    kind = kind_of(fmt) if kind is None else kind
    ad, au, al = _vals(fmt)
    ia, ja = _i32(fmt["iax"]), _i32(fmt["ja"])
    x = _f64(x)
    y = np.zeros(fmt["neq"])
    if use_ref:
        ref().ref_matvec(REF_KIND[kind], fmt["neq"], _p(ad), _p(au), _p(al), _p(ia), _p(ja), _p(x), _p(y))
    else:
        lib().orc_matvec(kind, fmt["neq"], _p(ad), _p(au), _p(al), _p(ia), _p(ja), _p(x), _p(y))
    return y


def dot(a, b, kind: int = DOT_SEQ, use_ref: bool = False) -> float:
    a, b = _f64(a), _f64(b)
    if use_ref:
        f = ref().ref_dot_ompl4 if kind == DOT_L4 else ref().ref_dot
    else:
        f = lib().orc_dot_L4 if kind == DOT_L4 else lib().orc_dot
    return f(_p(a), _p(b), len(a))


def solver(fmt: dict, code: int, unsym: bool = True, omp: bool = True, tol: float = 1e-10,
           maxit: int = 50000, x0: np.ndarray | None = None, b: np.ndarray | None = None):
    """orc_solver: the src/Solver.f dispatch.  Returns (x, b_after, stats)."""
    ad, au, al = _vals(fmt)
    ia, ja = _i32(fmt["iax"]), _i32(fmt["ja"])
    x = np.zeros(fmt["neq"]) if x0 is None else _f64(x0).copy()
    bb = _f64(fmt["b"] if b is None else b).copy()
    st = Stats()
    lib().orc_solver(code, fmt["matriz"], int(unsym), int(omp), fmt["neq"], _p(ad), _p(au), _p(al),
                     _p(x), _p(bb), _p(ia), _p(ja), tol, maxit, C.byref(st))
    return x, bb, st.as_dict()


def ref_solver(fmt: dict, code: int, unsym: bool = True, omp: bool = False, nthreads: int = 1,
               tol: float = 1e-10, maxit: int = 50000, x0=None, b=None, nlit: int = 2):
    """The reference's own solver() (oracle/_ref).  It stops the process on
    non-convergence, so call it only on systems that converge."""
    r = ref()
    r.ref_init(0, int(omp), nthreads)
    ad, au, al = _vals(fmt)
    ia, ja = _i32(fmt["iax"]), _i32(fmt["ja"])
    x = np.zeros(fmt["neq"]) if x0 is None else _f64(x0).copy()
    bb = _f64(fmt["b"] if b is None else b).copy()
    r.ref_solver(_p(ad), _p(au), _p(al), _p(x), _p(bb), _p(ia), _p(ja), fmt["neq"], fmt["nad"],
                 int(unsym), tol, maxit, code, fmt["matriz"], nlit)
    return x, bb


def krylov(method: str, fmt: dict, tol: float = 1e-10, maxit: int = 50000, x0=None, b=None,
           dot_kind: int = DOT_SEQ, init: bool = True, k: int = 30, kind: int | None = None,
           use_ref: bool = False, omp: bool = False, nthreads: int = 1):
    """Direct pcg / pbicgstab / gmres with an explicit matvec and dot."""
    kind = kind_of(fmt) if kind is None else kind
    ad, au, al = _vals(fmt)
    ia, ja = _i32(fmt["iax"]), _i32(fmt["ja"])
    m = ad.copy()
    x = np.zeros(fmt["neq"]) if x0 is None else _f64(x0).copy()
    bb = _f64(fmt["b"] if b is None else b).copy()
    st = Stats()
    n = fmt["neq"]
    if use_ref:
        r = ref()
        r.ref_init(0, int(omp), nthreads)
        args = (_p(ad), _p(au), _p(al), _p(ia), _p(ja), _p(m), _p(bb), _p(x))
        if method == "pcg":
            r.ref_pcg(REF_KIND[kind], int(omp), n, fmt["nad"], *args, tol, maxit, int(init), 2)
        elif method == "bicgstab":
            r.ref_pbicgstab(REF_KIND[kind], int(omp), n, fmt["nad"], *args, tol, maxit, int(init), 2)
        elif method == "gmres":
            r.ref_gmres(REF_KIND[kind], n, fmt["nad"], *args, k, tol, maxit, 2)
        return x, bb, None
    L = lib()
    args = (_p(ad), _p(au), _p(al), _p(ia), _p(ja), _p(m), _p(bb), _p(x))
    if method == "pcg":
        L.orc_pcg(kind, n, *args, tol, maxit, dot_kind, int(init), C.byref(st))
    elif method == "bicgstab":
        L.orc_pbicgstab(kind, n, *args, tol, maxit, dot_kind, int(init), C.byref(st))
    elif method == "gmres":
        L.orc_gmres(kind, n, *args, k, tol, maxit, C.byref(st))
    else:
        raise ValueError(method)
    return x, bb, st.as_dict()
