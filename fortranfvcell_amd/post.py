"""The SIMPLE steps around the solves, on the device: the assembly of the
momentum systems (pform with cell_si_eb_i_ggl, src/Pform.f:65-436;
src/Simple.f:425-431,456-462) and of the pressure-correction system (pform
with cell_si_p_i_ggl; src/Simple.f:583-588), the gradient of gform
(src/Pform.f:466-676; src/Simple.f:601,635) and simpleUpdate
(src/Simple.f:871-898, called at :642), over the C ABI
(fvc_assemble_mom_dev, fvc_assemble_pc_dev, fvc_gform_dev,
fvc_simple_update_dev).

``DeviceMesh`` keeps the mesh arrays the gradient reads resident in HBM, in
the reference's layout (x(2,nnode), ix(nen+1,numel), nelcon(nen,numel),
pedge(nen+1,numel)); the fields are DeviceArrays, so p', its gradient and the
corrected velocities never leave the GPU between the solve and the update.
"""
from __future__ import annotations

import numpy as np

from . import DeviceArray, _ptr, check, load


class DeviceMesh:
    def __init__(self, x, ix, nelcon, pedge):
        x = np.ascontiguousarray(x, dtype=np.float64)
        ix = np.ascontiguousarray(ix, dtype=np.int32)
        nelcon = np.ascontiguousarray(nelcon, dtype=np.int32)
        pedge = np.ascontiguousarray(pedge, dtype=np.int32)
        numel, nen = nelcon.shape
        if x.ndim != 2 or x.shape[1] != 2:
            raise ValueError("x must be (nnode, 2), the reference's x(2,nnode)")
        if ix.shape != (numel, nen + 1) or pedge.shape != (numel, nen + 1):
            raise ValueError("ix and pedge must be (numel, nen+1)")
        if nen not in (3, 4):
            raise ValueError("triangles and quads only")
        nodes = ix[:, :nen]
        if nodes.min() < 1 or nodes.max() > len(x):
            raise ValueError("ix holds node numbers outside 1..nnode")
        if nelcon.max() > numel or np.any(nelcon == 0):
            raise ValueError("nelcon holds neighbours outside 1..numel (walls are negative)")
        self.numel, self.nnode, self.nen = numel, len(x), nen
        self.x = DeviceArray.from_host(x)
        self.ix = DeviceArray.from_host(ix)
        self.nelcon = DeviceArray.from_host(nelcon)
        self.pedge = DeviceArray.from_host(pedge)

    def assemble_pc(self, num: DeviceArray, e: DeviceArray, sedge: DeviceArray, w: DeviceArray,
                    d: DeviceArray, gradp: DeviceArray, p: DeviceArray, ad: DeviceArray, al: DeviceArray,
                    b: DeviceArray, rescell: DeviceArray, int_vel: int = 5, stream=None) -> None:
        """The pressure-correction system into ad (neq), al (neq, nen), b
        (neq) and rescell (numel).  num (numel) int32 1-based equations; e
        (nmat, 10) material properties (density in column 2); sedge (numel,
        nen+1, 3) prescribed edge values; w, d, gradp (numel, 2); p (numel)."""
        neq = ad.shape[0]
        if al.shape != (neq, self.nen) or b.shape != (neq,) or rescell.shape != (self.numel,):
            raise ValueError("ad (neq), al (neq, nen), b (neq) and rescell (numel) expected")
        if num.shape != (self.numel,) or num.dtype != np.int32:
            raise ValueError("num must be int32 (numel,)")
        if e.ndim != 2 or e.shape[1] != 10 or sedge.shape != (self.numel, self.nen + 1, 3):
            raise ValueError("e must be (nmat, 10) and sedge (numel, nen+1, 3)")
        for f, shape in ((w, (self.numel, 2)), (d, (self.numel, 2)), (gradp, (self.numel, 2)), (p, (self.numel,))):
            if f.shape != shape:
                raise ValueError(f"field of shape {f.shape}, expected {shape}")
        check(load().fvc_assemble_pc_dev(self.numel, self.nnode, self.nen, neq, e.shape[0], int(int_vel),
                                         _ptr(self.x), _ptr(self.ix), _ptr(self.nelcon), _ptr(self.pedge),
                                         _ptr(num), _ptr(e), _ptr(sedge), _ptr(w), _ptr(d), _ptr(gradp),
                                         _ptr(p), _ptr(ad), _ptr(al), _ptr(b), _ptr(rescell), stream))

    def assemble_mom(self, comp: int, num: DeviceArray, e: DeviceArray, sedge: DeviceArray, u: DeviceArray,
                     pres: DeviceArray, w: DeviceArray, grad1: DeviceArray, grad2: DeviceArray,
                     gradp: DeviceArray, d: DeviceArray, ad: DeviceArray, al: DeviceArray, b: DeviceArray,
                     rescell: DeviceArray, im: DeviceArray, dt: float, under_u: float = 1.0,
                     s_pressure: bool = True, simple_c: bool = False, bs: bool = False, stream=None) -> None:
        """Momentum component comp (1 or 2) into ad (neq), al (neq, nen), b
        (neq), rescell (numel), im[:, comp-1] and im[:, comp+1] of im (numel,
        4) and d[:, comp-1] of d (numel, 2).  u, pres (numel); w, grad1 (of
        u), grad2 (of the other component), gradp (numel, 2); e (nmat, 10)
        with viscosity, density, ..., column 6 the transient coefficient and
        column 7 the limiter code 1..7."""
        neq = ad.shape[0]
        if al.shape != (neq, self.nen) or b.shape != (neq,) or rescell.shape != (self.numel,):
            raise ValueError("ad (neq), al (neq, nen), b (neq) and rescell (numel) expected")
        if num.shape != (self.numel,) or num.dtype != np.int32:
            raise ValueError("num must be int32 (numel,)")
        if e.ndim != 2 or e.shape[1] != 10 or sedge.shape != (self.numel, self.nen + 1, 3):
            raise ValueError("e must be (nmat, 10) and sedge (numel, nen+1, 3)")
        for f, shape in ((u, (self.numel,)), (pres, (self.numel,)), (w, (self.numel, 2)),
                         (grad1, (self.numel, 2)), (grad2, (self.numel, 2)), (gradp, (self.numel, 2)),
                         (d, (self.numel, 2)), (im, (self.numel, 4))):
            if f.shape != shape:
                raise ValueError(f"field of shape {f.shape}, expected {shape}")
        check(load().fvc_assemble_mom_dev(self.numel, self.nnode, self.nen, neq, e.shape[0], int(comp), float(dt),
                                          float(under_u), int(bool(s_pressure)), int(bool(simple_c)), int(bool(bs)),
                                          _ptr(self.x), _ptr(self.ix), _ptr(self.nelcon), _ptr(self.pedge),
                                          _ptr(num), _ptr(e), _ptr(sedge), _ptr(u), _ptr(pres), _ptr(w),
                                          _ptr(grad1), _ptr(grad2), _ptr(gradp), _ptr(d), _ptr(ad), _ptr(al),
                                          _ptr(b), _ptr(rescell), _ptr(im), stream))

    def gform(self, u: DeviceArray, grad: DeviceArray, fluxl: DeviceArray, stream=None) -> None:
        """grad (numel, 2): the old gradient on entry, the new one on return
        (the ghost values of extrapolating walls use the old one)."""
        check(load().fvc_gform_dev(self.numel, self.nnode, self.nen, _ptr(self.x), _ptr(self.ix),
                                   _ptr(self.nelcon), _ptr(self.pedge), _ptr(u), _ptr(grad),
                                   _ptr(fluxl), stream))


def ell_pattern(num, nelcon, neq: int | None = None) -> np.ndarray:
    """The ELLPACK column array of the pressure-correction system, (neq, nen)
    int32: ellPackJa (src/EllPack.f:19-42) on id(j, num(c)) = num(nelcon(j, c)),
    a wall slot pointing at equation 1."""
    num = np.asarray(num, dtype=np.int32)
    nelcon = np.asarray(nelcon, dtype=np.int32)
    neq = int(num.max()) if neq is None else int(neq)
    inner = nelcon > 0
    nb = np.where(inner, nelcon - 1, 0)
    ja = np.ones((neq, nelcon.shape[1]), dtype=np.int32)
    own = num > 0
    ja[num[own] - 1] = np.where(inner[own] & (num[nb[own]] > 0), num[nb[own]], 1)
    return ja


def simple_update(numel: int, w: DeviceArray, pre: DeviceArray, u1: DeviceArray, u2: DeviceArray,
                  prec: DeviceArray, gradpc: DeviceArray, d: DeviceArray, under_u: float,
                  under_pc: float, simple_r: bool, stream=None) -> None:
    """w = (u1, u2) - d * gradPc; pre += underPc * preC unless SIMPLEC."""
    check(load().fvc_simple_update_dev(int(numel), 2, _ptr(w), _ptr(pre), _ptr(u1), _ptr(u2), _ptr(prec),
                                       _ptr(gradpc), _ptr(d), float(under_u), float(under_pc),
                                       int(bool(simple_r)), stream))
