"""fortranfvcell_amd -- fortranFVCell's linear-solver path on MI355X (gfx950).

The reference's solver object (src/Solver.f: ``solver`` and ``getRes``, with
src/Iterativos*.f and src/Matvec*.f underneath) re-built as hand-written HIP
kernels behind a C ABI (include/fvcell.h, libfvcell.so).  This module is the
Python side of that ABI: it mirrors the reference's call surface with numpy
arrays in the reference's layouts, and gives device-resident buffers for
callers that keep systems in HBM.

No CPU fallback exists: every call goes to the GPU library, and a missing
library or GPU raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (BICGSTAB, CSR, CSRC, CSRD, ELLPACK, ENGINE_AUTO, ENGINE_KERNELS,
                   ENGINE_PERSISTENT, FLAG_COL16, FLAG_TILES, GMRES, ORTH_CGS2, ORTH_CGS2_LAG, ORTH_MGS, PCG,
                   PREC_FP64, PREC_MIXED,
                   REC_FUSED, REC_REFERENCE,
                   STOP_REFERENCE,
                   STOP_RELRES, X0_AS_OPENMP,
                   X0_AS_SEQUENTIAL, X0_GIVEN, X0_ZERO, FvcError, Options, Stats, check,
                   default_options, load)

__all__ = ["solver", "solve", "keep_matrix", "solve_dev", "matvec", "matvec_dev", "dot", "dot_dev", "getres",
           "DeviceArray", "DeviceSystem", "Options", "Stats", "FvcError", "default_options", "load",
           "PCG", "BICGSTAB", "GMRES", "CSRD", "CSR", "CSRC", "ELLPACK"]

_NONFATAL = (_lib.FVC_OK, _lib.FVC_MAXIT, _lib.FVC_BREAKDOWN, _lib.FVC_NONFINITE)


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return C.c_void_p(a.ptr)
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays must be contiguous")
    return a.ctypes.data_as(C.c_void_p)


def _f64(a, name):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError(f"{name} must be a contiguous float64 array (real*8)")
    return a


def _i32(a, name):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a
    if a.dtype != np.int32 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError(f"{name} must be a contiguous int32 array (integer)")
    return a


def init(device: int = 0) -> None:
    check(load().fvc_init(device))


def device_info() -> dict:
    cus = C.c_int32(0)
    name = C.create_string_buffer(256)
    check(load().fvc_device_info(C.byref(cus), name, 256))
    return {"compute_units": cus.value, "name": name.value.decode()}


# ----------------------------------------------------------- the drop-in --

def solver(ad, au, al, x, b, iax, ja, neq, nad, unsym, tol, maxit, code, matriz, nlit=2,
           omp_solver=True, raise_on_failure=True) -> dict:
    """``solver(ad,au,al,x,b,iax,ja,neq,nad,unsym,tol,maxit,code,matriz,nlit)``
    of src/Solver.f:28-29 on the GPU.  ``x`` (float64, in/out) is updated in
    place; ``b`` is left intact (the reference overwrites it).
    ``omp_solver`` is the reference's openmpSolver flag: it selects which
    paths warm-start from ``x`` (src/Solver.f:94-260).  Returns the stats;
    non-convergence raises FvcError like the reference stops, unless
    ``raise_on_failure`` is False."""
    L = load()
    st = Stats()
    au = al if au is None else au
    rc = L.fvc_solver(_ptr(_f64(ad, "ad")), _ptr(_f64(au, "au")), _ptr(_f64(al, "al")),
                      _ptr(_f64(x, "x")), _ptr(_f64(b, "b")), _ptr(_i32(iax, "iax")),
                      _ptr(_i32(ja, "ja")), int(neq), int(nad), int(bool(unsym)), float(tol),
                      int(maxit), int(code), int(matriz), int(nlit), int(bool(omp_solver)),
                      C.byref(st))
    check(rc, (_lib.FVC_OK,) if raise_on_failure else _NONFATAL)
    return st.as_dict()


def keep_matrix(keep: bool = True) -> None:
    """The next host-array ``solver`` / ``solve`` call on this device reuses
    the device copy of A uploaded by the previous one, if the format, sizes
    and array addresses match; the caller guarantees the contents are
    unchanged (the two pressure-correction solves of a SIMPLE step,
    src/Simple.f:590,621).  One-shot."""
    check(load().fvc_keep_matrix(int(bool(keep))))


def _fmt_arrays(fmt: dict):
    ad = np.ascontiguousarray(fmt["ad"], dtype=np.float64)
    al = fmt.get("al")
    al = None if al is None else np.ascontiguousarray(al, dtype=np.float64)
    au = fmt.get("au")
    au = al if au is None else np.ascontiguousarray(au, dtype=np.float64)
    iax = np.ascontiguousarray(fmt["iax"], dtype=np.int32)
    ja = np.ascontiguousarray(fmt["ja"], dtype=np.int32)
    return ad, au, al, iax, ja


def solve(fmt: dict, code: int, x0=None, b=None, tol: float = 1e-10, maxit: int = 50000,
          unsym: bool | None = None, options: Options | None = None, raise_on_failure=False, **opt):
    """Solve a system given in one of the reference layouts (a dict as built by
    ``problems.to_ell/to_csrd/to_csr/to_csrc``).  Returns (x, stats)."""
    L = load()
    ad, au, al, iax, ja = _fmt_arrays(fmt)
    if unsym is None:
        unsym = not fmt.get("sym", False)
    neq = int(fmt["neq"])
    x = np.zeros(neq) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    bb = np.ascontiguousarray(fmt["b"] if b is None else b, dtype=np.float64)
    o = options if options is not None else default_options(**opt)
    st = Stats()
    rc = L.fvc_solve(int(code), int(fmt["matriz"]), int(unsym), neq, int(fmt["nad"]), _ptr(ad),
                     _ptr(au if au is not None else ad), _ptr(al if al is not None else ad), _ptr(x),
                     _ptr(bb), _ptr(iax), _ptr(ja), float(tol), int(maxit), C.byref(o), C.byref(st))
    check(rc, (_lib.FVC_OK,) if raise_on_failure else _NONFATAL)
    return x, st.as_dict()


def matvec(fmt: dict, x, unsym: bool | None = None) -> np.ndarray:
    """y = A x with the reference's matvec_* for this layout (src/Matvec.f)."""
    ad, au, al, iax, ja = _fmt_arrays(fmt)
    if unsym is None:
        unsym = not fmt.get("sym", False)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(int(fmt["neq"]))
    check(load().fvc_matvec(int(fmt["matriz"]), int(unsym), int(fmt["neq"]), int(fmt["nad"]),
                            _ptr(ad), _ptr(au if au is not None else ad),
                            _ptr(al if al is not None else ad), _ptr(iax), _ptr(ja), _ptr(x), _ptr(y)))
    return y


def dot(a, b) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    r = C.c_double(0.0)
    check(load().fvc_dot(_ptr(a), _ptr(b), len(a), C.byref(r)))
    return r.value


def getres(u, x, num, neq, numel, flag) -> None:
    """getRes (src/Solver.f:368-391) on host arrays, in place."""
    check(load().fvc_getres(_ptr(_f64(u, "u")), _ptr(_f64(x, "x")), _ptr(_i32(num, "num")),
                            int(neq), int(numel), int(bool(flag))))


# ------------------------------------------------------- device residency --

class DeviceArray:
    """A buffer in HBM allocated through the library (fvc_malloc)."""

    def __init__(self, nbytes: int, dtype, shape):
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape)
        self.nbytes = int(nbytes)
        p = C.c_void_p(0)
        check(load().fvc_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes, a.dtype, a.shape)
        if a.nbytes:
            check(load().fvc_memcpy_htod(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes))
        return d

    @classmethod
    def empty(cls, shape, dtype=np.float64) -> "DeviceArray":
        shape = (shape,) if np.isscalar(shape) else tuple(shape)
        return cls(int(np.prod(shape)) * np.dtype(dtype).itemsize, dtype, shape)

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(load().fvc_memcpy_htod(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes))

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(load().fvc_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            load().fvc_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceSystem:
    """A system resident in HBM in the reference layout: the arrays the
    reference's solver receives, uploaded once."""

    def __init__(self, fmt: dict, unsym: bool | None = None):
        ad, au, al, iax, ja = _fmt_arrays(fmt)
        self.matriz = int(fmt["matriz"])
        self.neq = int(fmt["neq"])
        self.nad = int(fmt["nad"])
        self.unsym = (not fmt.get("sym", False)) if unsym is None else unsym
        self.ad = DeviceArray.from_host(ad)
        self.al = DeviceArray.from_host(al) if al is not None else None
        self.au = DeviceArray.from_host(au) if (au is not None and au is not al) else self.al
        self.iax = DeviceArray.from_host(iax)
        # ELLPACK's iax is {w, neq+1}: handed over from host memory, so that a
        # solve or SpMV on resident data needs no device-to-host read
        self._iax_arg = iax if self.matriz == ELLPACK else self.iax
        self.ja = DeviceArray.from_host(ja)
        self.b = DeviceArray.from_host(np.ascontiguousarray(fmt["b"], dtype=np.float64))
        self.x = DeviceArray.empty(self.neq)

    @classmethod
    def resident_ell(cls, ad: "DeviceArray", al: "DeviceArray", b: "DeviceArray", ja,
                     unsym: bool = True) -> "DeviceSystem":
        """An ELLPACK system whose values are already in HBM (assembled on the
        device): ad (neq), al (neq, w), b (neq) DeviceArrays, taken as they
        are; ja (neq, w) the 1-based int32 columns, uploaded once."""
        ja = np.ascontiguousarray(ja, dtype=np.int32)
        neq, w = al.shape
        if ad.shape != (neq,) or b.shape != (neq,) or ja.shape != (neq, w):
            raise ValueError("ad (neq), al (neq, w), b (neq) and ja (neq, w) expected")
        self = cls.__new__(cls)
        self.matriz, self.neq, self.unsym = ELLPACK, int(neq), bool(unsym)
        self.nad = int(neq * w)
        self.ad, self.al, self.au, self.b = ad, al, al, b
        iax = np.array([w, neq + 1], dtype=np.int32)
        self.iax = DeviceArray.from_host(iax)
        self._iax_arg = iax
        self.ja = DeviceArray.from_host(ja)
        self.x = DeviceArray.empty(self.neq)
        return self

    def solve(self, code: int, tol: float = 1e-10, maxit: int = 50000, x0=None, options=None,
              stream=None, noalloc: bool = False, **opt) -> dict:
        """fvc_solve_dev on the resident arrays (fvc_solve_dev_noalloc with
        noalloc: the workspace is the library's fixed per-device array)."""
        if x0 is not None:
            self.x.upload(np.asarray(x0, dtype=np.float64))
        o = options if options is not None else default_options(**opt)
        st = Stats()
        fn = load().fvc_solve_dev_noalloc if noalloc else load().fvc_solve_dev
        rc = fn(code, self.matriz, int(self.unsym), self.neq, self.nad,
                                  _ptr(self.ad), _ptr(self.au or self.ad), _ptr(self.al or self.ad),
                                  _ptr(self.x), _ptr(self.b), _ptr(self._iax_arg), _ptr(self.ja),
                                  float(tol), int(maxit), C.byref(o), C.byref(st), stream)
        check(rc, _NONFATAL)
        return st.as_dict()

    def matvec(self, x: DeviceArray, y: DeviceArray, stream=None) -> None:
        check(load().fvc_matvec_dev(self.matriz, int(self.unsym), self.neq, self.nad, _ptr(self.ad),
                                    _ptr(self.au or self.ad), _ptr(self.al or self.ad), _ptr(self._iax_arg),
                                    _ptr(self.ja), _ptr(x), _ptr(y), stream))


def matvec_dev(system: DeviceSystem, x: DeviceArray, y: DeviceArray, stream=None) -> None:
    system.matvec(x, y, stream)


def solve_dev(system: DeviceSystem, code: int, **kw) -> dict:
    return system.solve(code, **kw)


def dot_dev(a: DeviceArray, b: DeviceArray, n: int | None = None, stream=None) -> float:
    r = C.c_double(0.0)
    n = a.shape[0] if n is None else n
    check(load().fvc_dot_dev(_ptr(a), _ptr(b), int(n), C.byref(r), stream))
    return r.value


# ---------------------------------------------------------------- batches --

def concat_ell(fmts: list) -> dict:
    """Concatenate ELLPACK systems of one width (each in the reference layout,
    local 1-based ja) into the batch layout of fvc_solve_batch."""
    w = int(fmts[0]["iax"][0])
    if any(int(f["iax"][0]) != w for f in fmts):
        raise ValueError("all systems of a batch must share the ELLPACK width")
    sizes = np.array([int(f["neq"]) for f in fmts], dtype=np.int64)
    row_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    return dict(w=w, nsys=len(fmts), row_offsets=row_off,
                ad=np.ascontiguousarray(np.concatenate([f["ad"] for f in fmts]), dtype=np.float64),
                al=np.ascontiguousarray(np.concatenate([f["al"] for f in fmts]), dtype=np.float64),
                ja=np.ascontiguousarray(np.concatenate([f["ja"] for f in fmts]), dtype=np.int32),
                b=np.ascontiguousarray(np.concatenate([f["b"] for f in fmts]), dtype=np.float64))


def concat_csrd(fmts: list) -> dict:
    """Concatenate CSRD systems (the reference's matriz 1 layout with unsym:
    ad, then each row's off-diagonals, 1-based ia and ja local to the system)
    into the batch layout of fvc_solve_batch_csrd."""
    if any(int(f["matriz"]) != CSRD or f.get("sym", False) for f in fmts):
        raise ValueError("concat_csrd takes nonsymmetric CSRD systems (problems.to_csrd)")
    sizes = np.array([int(f["neq"]) for f in fmts], dtype=np.int64)
    nvals = np.array([len(f["ja"]) for f in fmts], dtype=np.int64)
    row_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    val_off = np.concatenate([[0], np.cumsum(nvals)]).astype(np.int32)
    return dict(matriz=CSRD, nsys=len(fmts), row_offsets=row_off, val_offsets=val_off,
                ad=np.ascontiguousarray(np.concatenate([f["ad"] for f in fmts]), dtype=np.float64),
                al=np.ascontiguousarray(np.concatenate([f["al"] for f in fmts]), dtype=np.float64),
                ia=np.ascontiguousarray(np.concatenate([f["iax"] for f in fmts]), dtype=np.int32),
                ja=np.ascontiguousarray(np.concatenate([f["ja"] for f in fmts]), dtype=np.int32),
                b=np.ascontiguousarray(np.concatenate([f["b"] for f in fmts]), dtype=np.float64))


def _batch_stats(sts, n):
    return [sts[i].as_dict() for i in range(n)]


def solve_batch(batch: dict, tol: float = 1e-10, maxit: int = 50000, x0=None, options=None, **opt):
    """Jacobi-PCG on every system of a batch (concat_ell or concat_csrd) in
    one persistent launch.  Returns (x concatenated, list of per-system
    stats)."""
    n = int(batch["row_offsets"][-1])
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    o = options if options is not None else default_options(**opt)
    sts = (Stats * batch["nsys"])()
    if int(batch.get("matriz", ELLPACK)) == CSRD:
        rc = load().fvc_solve_batch_csrd(PCG, int(batch["nsys"]), _ptr(batch["row_offsets"]),
                                         _ptr(batch["val_offsets"]), _ptr(batch["ad"]), _ptr(batch["al"]), _ptr(x),
                                         _ptr(batch["b"]), _ptr(batch["ia"]), _ptr(batch["ja"]), float(tol),
                                         int(maxit), C.byref(o), sts)
    else:
        rc = load().fvc_solve_batch(PCG, int(batch["nsys"]), _ptr(batch["row_offsets"]), int(batch["w"]),
                                    _ptr(batch["ad"]), _ptr(batch["al"]), _ptr(x), _ptr(batch["b"]),
                                    _ptr(batch["ja"]), float(tol), int(maxit), C.byref(o), sts)
    check(rc, _NONFATAL)
    return x, _batch_stats(sts, batch["nsys"])


class DeviceBatch:
    """A batch of ELLPACK (concat_ell) or CSRD (concat_csrd) systems resident
    in HBM."""

    def __init__(self, batch: dict):
        self.nsys = int(batch["nsys"])
        self.matriz = int(batch.get("matriz", ELLPACK))
        self.w = int(batch.get("w", 0))
        self.n = int(batch["row_offsets"][-1])
        self.row_offsets = DeviceArray.from_host(batch["row_offsets"])
        self.ad = DeviceArray.from_host(batch["ad"])
        self.al = DeviceArray.from_host(batch["al"])
        self.ja = DeviceArray.from_host(batch["ja"])
        self.b = DeviceArray.from_host(batch["b"])
        self.x = DeviceArray.empty(self.n)
        if self.matriz == CSRD:
            self.val_offsets = DeviceArray.from_host(batch["val_offsets"])
            self.ia = DeviceArray.from_host(batch["ia"])

    def solve(self, tol: float = 1e-10, maxit: int = 50000, options=None, stream=None, **opt):
        o = options if options is not None else default_options(**opt)
        sts = (Stats * self.nsys)()
        if self.matriz == CSRD:
            rc = load().fvc_solve_batch_csrd_dev(PCG, self.nsys, _ptr(self.row_offsets), _ptr(self.val_offsets),
                                                 _ptr(self.ad), _ptr(self.al), _ptr(self.x), _ptr(self.b),
                                                 _ptr(self.ia), _ptr(self.ja), float(tol), int(maxit), C.byref(o),
                                                 sts, stream)
        else:
            rc = load().fvc_solve_batch_dev(PCG, self.nsys, _ptr(self.row_offsets), self.w, _ptr(self.ad),
                                            _ptr(self.al), _ptr(self.x), _ptr(self.b), _ptr(self.ja),
                                            float(tol), int(maxit), C.byref(o), sts, stream)
        check(rc, _NONFATAL)
        return _batch_stats(sts, self.nsys)
