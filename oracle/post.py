"""The oracle for the post-solve step of a SIMPLE iteration: the pressure-
correction gradient of gform (src/Pform.f:466-676) for the cells that take
cell_si_p_i_ggl (types 63/65/73/75, src/Celllib.f:127-157), iws = 2: its
Green-Gauss reconstruction (src/CellibPressure.f:611-650), and simpleUpdate
(src/Simple.f:871-898).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The restatement is vectorised over cells; every cell's arithmetic is the
reference's, operation by operation, so that results are bit-identical.
Arrays follow the reference's Fortran layout, held as C arrays with the
first Fortran index fastest: x(2, nnode) -> (nnode, 2), ix(nen+1, numel) ->
(numel, nen+1), nelcon(nen, numel) -> (numel, nen), pedge(nen+1, numel) ->
(numel, nen+1), grad(2, numel) -> (numel, 2); indices 1-based, a boundary
neighbour negative.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DROPIN_PATH = os.path.join(HERE, "_ref", "libmvf_dropin.so")

# local edge -> (first node, second node), 0-based (setcell, src/Celllib.f:541-583)
SN = {3: ((0, 1), (1, 2), (2, 0)), 4: ((0, 1), (1, 2), (2, 3), (3, 0))}


def _centroid(P, nsh):
    """cellGeom2D's centre: the vertex sum, then times 1/nshared
    (src/Celllib.f:725-733).  P: (..., nsh, 2)."""
    s = np.zeros(P.shape[:-2] + (2,))
    for l in range(nsh):
        s = s + P[..., l, :]
    return (1.0 / nsh) * s


def _cross_norm(a, b):
    """|a x b| for in-plane vectors: vet (src/Matvec.f:304-316) with the third
    components zero, then sqrt(dot_local(c, c, 3)) (src/Matvec.f:404-421)."""
    c1 = a[..., 1] * 0.0 - 0.0 * b[..., 1]
    c2 = 0.0 * b[..., 0] - a[..., 0] * 0.0
    c3 = a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]
    return np.sqrt(((0.0 + c1 * c1) + c2 * c2) + c3 * c3)


def _area(E, nsh):
    """areacell (src/Celllib.f:509-522): areatriacell (:282-297) or
    areaquadcell (:309-358), the mean of the two diagonal splits."""
    if nsh == 3:
        return 0.5 * _cross_norm(E[..., 0, :], E[..., 1, :])
    a = 0.5 * _cross_norm(E[..., 0, :], E[..., 1, :])
    a = a + 0.5 * _cross_norm(E[..., 2, :], E[..., 3, :])
    a = a + 0.5 * _cross_norm(E[..., 0, :], E[..., 3, :])
    a = a + 0.5 * _cross_norm(E[..., 1, :], E[..., 2, :])
    return 0.5 * a


def gform_p(x, ix, nelcon, pedge, u, grad):
    """New gradient and flux limiter of every cell (grad is the old gradient,
    used by the extrapolating boundary condition)."""
    numel, nsh = nelcon.shape
    sn = SN[nsh]
    X = x[ix[:, :nsh] - 1]                              # (numel, nsh, 2)
    xc = _centroid(X, nsh)
    # the central cell's edges: unit normals and centre-to-edge distances
    # (src/Celllib.f:737-808)
    n = np.empty((numel, nsh, 2))
    ca = np.empty((numel, nsh))
    for l, (a, b) in enumerate(sn):
        e = X[:, b] - X[:, a]
        meta = np.sqrt((0.0 + e[:, 0] * e[:, 0]) + e[:, 1] * e[:, 1])
        eu0, eu1 = e[:, 0] / meta, e[:, 1] / meta
        n[:, l, 0], n[:, l, 1] = eu1, -eu0
        xm = 0.5 * (X[:, b] + X[:, a])
        df = xm - xc
        ca[:, l] = df[:, 0] * n[:, l, 0] + df[:, 1] * n[:, l, 1]
    # neighbour centres, or ghost centres across boundary edges, and the
    # values there (src/CellibPressure.f:432-463, iws = 2)
    inner = nelcon > 0
    nb = np.where(inner, nelcon - 1, 0)
    xcn = _centroid(x[ix[nb, :nsh] - 1], nsh)           # (numel, nsh, 2)
    gama = 2.0 * ca
    ghost = xc[:, None, :] + gama[:, :, None] * n
    xcg = np.where(inner[:, :, None], xcn, ghost)
    uc = u[:, None]
    extrap = uc + gama * (grad[:, None, 0] * n[:, :, 0] + grad[:, None, 1] * n[:, :, 1])
    pe = pedge[:, :nsh]
    bc = np.where(pe == 3, 0.0, np.where(pe == 6, np.broadcast_to(uc, pe.shape), extrap))
    ug = np.where(inner, u[nb], bc)
    # Green-Gauss over the polygon of the neighbour centres
    # (src/CellibPressure.f:611-650)
    E = np.empty((numel, nsh, 2))
    um = np.empty((numel, nsh))
    for l, (a, b) in enumerate(sn):
        E[:, l] = xcg[:, b] - xcg[:, a]
        um[:, l] = 0.5 * (ug[:, b] + ug[:, a])
    area = _area(E, nsh)
    g0 = np.zeros(numel)
    g1 = np.zeros(numel)
    for l in range(nsh):
        g0 = g0 + um[:, l] * E[:, l, 1]
        g1 = g1 + um[:, l] * (-E[:, l, 0])
    return np.stack([g0 / area, g1 / area], axis=1), np.ones(numel)


def simple_update(w, pre, u1, u2, prec, gradpc, d, under_pc, simple_r):
    """simpleUpdate (src/Simple.f:871-898): w = u* - d grad p'; p += underPc p'
    (SIMPLE only).  Returns new (w, pre); w and d are (numel, 2)."""
    w = np.array(w, dtype=np.float64, copy=True)
    pre = np.array(pre, dtype=np.float64, copy=True)
    w[:, 0] = u1 - d[:, 0] * gradpc[:, 0]
    w[:, 1] = u2 - d[:, 1] * gradpc[:, 1]
    if not simple_r:
        pre = pre + under_pc * prec
    return w, pre


# ------------------------------------------------ the reference's own code --

def have_dropin() -> bool:
    return os.path.exists(DROPIN_PATH)


def _dropin():
    L = C.CDLL(DROPIN_PATH, mode=C.RTLD_LOCAL)
    return L


class _OmpControl(C.Structure):
    # COMMON /omp_control/ (include/openmp.fi:16-18)
    _fields_ = [("num_threads", C.c_int32), ("openmpCell", C.c_int32), ("openmpSolver", C.c_int32),
                ("nThreadsCell", C.c_int32), ("nThreadsSolver", C.c_int32), ("nThreadsDot", C.c_int32),
                ("nThreadsMatVec", C.c_int32)]


def ref_gform_p(x, ix, nelcon, pedge, u, grad, cell_type):
    """The reference's gform on the same arrays (sequential loop: openmpCell
    false), ndf = 1, iws = iws1 = 2, lib = 2: returns (grad, fluxl)."""
    L = _dropin()
    ctl = _OmpControl.in_dll(L, "omp_control_")
    ctl.openmpCell = 0
    numel, nsh = nelcon.shape
    nst = 3
    f64 = lambda a: np.asfortranarray(a, dtype=np.float64)
    i32 = lambda a: np.asfortranarray(a, dtype=np.int32)
    # C (numel, k) arrays are Fortran (k, numel): pass their transposes' memory
    xa = np.ascontiguousarray(x, dtype=np.float64)
    ixa = np.ascontiguousarray(ix, dtype=np.int32)
    nca = np.ascontiguousarray(nelcon, dtype=np.int32)
    pea = np.ascontiguousarray(pedge, dtype=np.int32)
    ua = np.ascontiguousarray(u, dtype=np.float64)
    g = np.ascontiguousarray(grad, dtype=np.float64).copy()
    fluxl = np.zeros(numel)
    sedge = np.zeros((numel, nsh + 1, nst))
    e = np.ones((1, 10))
    ls = np.zeros((numel, nsh, 2))
    ie = np.array([cell_type], np.int32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    ri = lambda v: C.byref(C.c_int32(v))
    L.gform_.restype = None
    L.gform_(p(ua), p(g), p(fluxl), p(xa), p(sedge), p(e), p(ls), p(ie), p(nca), p(pea), p(ixa),
             ri(numel), ri(2), ri(nsh), ri(nsh), ri(1), ri(nst), ri(2), ri(2), ri(2))
    del f64, i32
    return g, fluxl


def ref_simple_update(w, pre, u1, u2, prec, gradpc, d, under_u, under_pc, simple_r):
    """The reference's simpleUpdate on the same arrays: returns (w, pre)."""
    L = _dropin()
    numel = len(u1)
    wa = np.ascontiguousarray(w, dtype=np.float64).copy()
    pa = np.ascontiguousarray(pre, dtype=np.float64).copy()
    p = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.c_void_p)
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (u1, u2, prec, gradpc, d)]
    L.simpleupdate_.restype = None
    L.simpleupdate_(wa.ctypes.data_as(C.c_void_p), pa.ctypes.data_as(C.c_void_p),
                    *[k.ctypes.data_as(C.c_void_p) for k in keep],
                    C.byref(C.c_int32(numel)), C.byref(C.c_int32(2)),
                    C.byref(C.c_double(under_u)), C.byref(C.c_double(under_pc)),
                    C.byref(C.c_int32(1 if simple_r else 0)))
    del p
    return wa, pa
