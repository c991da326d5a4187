"""The post-solve step of a SIMPLE iteration on the device: the pressure-
correction gradient of gform (src/Pform.f:466-676; src/Simple.f:601,635) and
simpleUpdate (src/Simple.f:871-898, called at :642), over the C ABI
(fvc_gform_dev, fvc_simple_update_dev).

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

    def gform(self, u: DeviceArray, grad: DeviceArray, fluxl: DeviceArray, stream=None) -> None:
        """grad (numel, 2): the old gradient on entry, the new one on return
        (the ghost values of extrapolating walls use the old one)."""
        check(load().fvc_gform_dev(self.numel, self.nnode, self.nen, _ptr(self.x), _ptr(self.ix),
                                   _ptr(self.nelcon), _ptr(self.pedge), _ptr(u), _ptr(grad),
                                   _ptr(fluxl), stream))


def simple_update(numel: int, w: DeviceArray, pre: DeviceArray, u1: DeviceArray, u2: DeviceArray,
                  prec: DeviceArray, gradpc: DeviceArray, d: DeviceArray, under_u: float,
                  under_pc: float, simple_r: bool, stream=None) -> None:
    """w = (u1, u2) - d * gradPc; pre += underPc * preC unless SIMPLEC."""
    check(load().fvc_simple_update_dev(int(numel), 2, _ptr(w), _ptr(pre), _ptr(u1), _ptr(u2), _ptr(prec),
                                       _ptr(gradpc), _ptr(d), float(under_u), float(under_pc),
                                       int(bool(simple_r)), stream))
