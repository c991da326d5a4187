"""The reference's own MatrixMarket export (exportCoo -> csrToCoo ->
writeCoo, src/ExportCoo.f:23-62, src/CsrToCoo.f:29-179), called from the
reference program's objects compiled from its sources
(oracle/_ref/libmvf_dropin.so).

TEST INFRASTRUCTURE ONLY: it pins fortranfvcell_amd.mtx, the project's
writer, byte for byte.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import post

_POOL_WORDS = 1 << 24    # the reference's memory pool, in 4-byte words (64 MiB)
_pool_ready = False


def have_dropin() -> bool:
    return post.have_dropin()


def _lib():
    """The reference program's objects, with their memory pool set up once
    (src/Malloc.f:init_malloc; exportCoo allocates its COO arrays there)."""
    global _pool_ready
    L = post._dropin()
    if not _pool_ready:
        C.c_int64.in_dll(L, "_QMmallocEmaxmem").value = _POOL_WORDS
        # init_malloc prints a banner on unit 6: keep it off the caller's stdout
        saved, quiet = os.dup(1), os.open(os.devnull, os.O_WRONLY)
        try:
            os.dup2(quiet, 1)
            L.init_malloc_()
        finally:
            os.dup2(saved, 1)
            os.close(saved)
            os.close(quiet)
        _pool_ready = True
    return L


def ref_export_coo(fmt: dict, code: int, unsym: bool, workdir: str, it: int = 0) -> tuple:
    """exportCoo on the arrays of fmt (CSRD: code 1, CSRC: code 3) with the
    file prefix 'sys' in workdir.  The reference names the files
    name('sys', it, 52) = 'syscoo_<it>' (src/Filenames.f:147-150) plus '.mtx'
    and '_b.mtx' (src/CsrToCoo.f:135-178).  Returns the two paths."""
    L = _lib()
    ad = np.ascontiguousarray(fmt["ad"], dtype=np.float64)
    al = np.ascontiguousarray(fmt["al"], dtype=np.float64)
    au = al if fmt.get("au") is None else np.ascontiguousarray(fmt["au"], dtype=np.float64)
    b = np.ascontiguousarray(fmt["b"], dtype=np.float64)
    iax = np.ascontiguousarray(fmt["iax"], dtype=np.int32)
    ja = np.ascontiguousarray(fmt["ja"], dtype=np.int32)
    prename = b"sys".ljust(80)                      # character*80, blank-padded
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    ri = lambda v: C.byref(C.c_int32(v))
    L.exportcoo_.restype = None
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        # trailing argument: the length of the character dummy
        L.exportcoo_(p(iax), p(ja), p(al), p(ad), p(au), p(b), ri(int(fmt["neq"])), ri(int(fmt["nad"])),
                     ri(1 if unsym else 0), C.c_char_p(prename), ri(77), ri(it), ri(code), C.c_size_t(80))
    finally:
        os.chdir(cwd)
    base = os.path.join(workdir, f"syscoo_{it}")
    return base + ".mtx", base + "_b.mtx"
