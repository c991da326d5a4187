"""Differentiable finite-volume solve: T = A(k)^-1 (b(k) + q) with a
hand-written backward.

Forward: assemble the heat operator A(k), b(k) on the GPU
(fvc_assemble_heat_dev, the reference's coefficient forms, the same
arithmetic as problems.HeatOperator.assemble), solve with Jacobi-BiCGSTAB on
the GPU.
Backward: the adjoint solve A^T lam = dL/dT, again Jacobi-BiCGSTAB on the GPU,
on the same ELLPACK structure with transposed values (fvc_ell_transpose_dev),
then the face-flux gradient kernel (fvc_grad_k_dev) gives dL/dk; dL/dq = lam.

``FVSolve`` wraps this as a torch.autograd.Function on CPU float64 tensors;
the arithmetic runs through the C ABI (libfvcell.so) on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import (BICGSTAB, ELLPACK, STOP_RELRES, X0_GIVEN, X0_ZERO, DeviceArray, _ptr, check,
               default_options, load)
from . import problems as P


class HeatSolver:
    """Device-resident forward and adjoint solves of one HeatOperator."""

    def __init__(self, op: P.HeatOperator, tol: float = 1e-13, maxit: int = 200000):
        self.op = op
        mesh = op.mesh
        self.neq, self.w = mesh.neq, mesh.w
        self.tol, self.maxit = tol, maxit
        inner = mesh.nbr >= 0
        self.nad = int(inner.sum())
        ja = np.where(inner, mesh.nbr + 1, 1).astype(np.int32)
        self.iax = np.array([self.w, self.neq + 1], np.int32)   # host: read without a sync
        self.ja = DeviceArray.from_host(ja.ravel())
        self.back = DeviceArray.from_host(P.ell_back_slots(mesh.nbr).ravel())
        self.geo = DeviceArray.from_host(np.ascontiguousarray(op.geo.ravel()))
        self.alpha = DeviceArray.from_host(np.ascontiguousarray(np.where(inner, op.alpha, 0.0).ravel()))
        self.wall_g = DeviceArray.from_host(np.ascontiguousarray(op.wall_g.ravel()))
        self.wall_T = DeviceArray.from_host(np.ascontiguousarray(op.wall_T.ravel()))
        self.conv = DeviceArray.from_host(np.ascontiguousarray(op.conv.ravel()))
        self.conv_f = DeviceArray.from_host(np.ascontiguousarray(op.convF))
        self.q = DeviceArray.empty(self.neq)
        n, nw = self.neq, self.neq * self.w
        self.ad = DeviceArray.empty(n)
        self.al = DeviceArray.empty(nw)
        self.alT = DeviceArray.empty(nw)
        self.b = DeviceArray.empty(n)
        self.T = DeviceArray.empty(n)
        self.g = DeviceArray.empty(n)
        self.lam = DeviceArray.empty(n)
        self.k = DeviceArray.empty(n)
        self.gk = DeviceArray.empty(n)
        self.last_forward = None
        self.last_adjoint = None

    def _bicgstab(self, al: DeviceArray, x: DeviceArray, b: DeviceArray, warm: bool = False,
                  strict: bool = True) -> dict:
        from . import _NONFATAL, Stats
        o = default_options(stop_rule=STOP_RELRES, x0_rule=X0_GIVEN if warm else X0_ZERO)
        st = Stats()
        rc = load().fvc_solve_dev(BICGSTAB, ELLPACK, 1, self.neq, self.nad, _ptr(self.ad), _ptr(al),
                                  _ptr(al), _ptr(x), _ptr(b), _ptr(self.iax), _ptr(self.ja), self.tol,
                                  self.maxit, C.byref(o), C.byref(st), None)
        check(rc, (0,) if strict else _NONFATAL)
        return st.as_dict()

    def load(self, k: np.ndarray, q: np.ndarray | None = None) -> None:
        """Assemble A(k), b(k) + q on the GPU (fvc_assemble_heat_dev): only k
        (and q) cross PCIe; the geometry and convection data stay resident."""
        self.k.upload(np.asarray(k, dtype=np.float64))
        qp = None
        if q is not None:
            self.q.upload(np.asarray(q, dtype=np.float64))
            qp = _ptr(self.q)
        check(load().fvc_assemble_heat_dev(self.neq, self.w, _ptr(self.ja), _ptr(self.k), qp, _ptr(self.geo),
                                           _ptr(self.alpha), _ptr(self.conv), _ptr(self.conv_f),
                                           _ptr(self.wall_g), _ptr(self.wall_T), _ptr(self.ad), _ptr(self.al),
                                           _ptr(self.b), None))

    def forward(self, k: np.ndarray, q: np.ndarray | None = None) -> np.ndarray:
        self.load(k, q)
        self.last_forward = self._bicgstab(self.al, self.T, self.b)
        return self.T.get()

    def solve_rhs(self, r: np.ndarray, restarts: int = 8) -> np.ndarray:
        """A x = r with the loaded matrix (x is returned; T is left alone).
        BiCGSTAB's shadow residual is the initial residual; for a very
        localised r it can become orthogonal to the iterates (breakdown).
        The solve then restarts from its last finite iterate, whose residual
        is spread out."""
        self.g.upload(np.asarray(r, dtype=np.float64))
        st = self._bicgstab(self.al, self.lam, self.g, strict=False)
        for _ in range(restarts):
            if st["status"] == 0:
                break
            if st["status"] not in (2, 3):
                raise RuntimeError(f"correction solve did not converge: {st}")
            st = self._bicgstab(self.al, self.lam, self.g, warm=True, strict=False)
        if st["status"] != 0:
            raise RuntimeError(f"correction solve did not converge: {st}")
        return self.lam.get()

    def adjoint_at(self, k: np.ndarray, q: np.ndarray | None, T: np.ndarray, gT: np.ndarray):
        """(dL/dk, dL/dq) at the forward state (k, q, T) given dL/dT: A(k) is
        assembled again and T uploaded first, so the result does not depend
        on which forward solve ran last on this solver."""
        self.load(k, q)
        self.T.upload(np.asarray(T, dtype=np.float64))
        return self.adjoint(gT)

    def adjoint(self, gT: np.ndarray):
        """(dL/dk, dL/dq) for the last forward solve, given dL/dT."""
        L = load()
        check(L.fvc_ell_transpose_dev(self.neq, self.w, _ptr(self.ja), _ptr(self.back), _ptr(self.al),
                                      _ptr(self.alT), None))
        self.g.upload(np.asarray(gT, dtype=np.float64))
        self.last_adjoint = self._bicgstab(self.alT, self.lam, self.g)
        check(L.fvc_grad_k_dev(self.neq, self.w, _ptr(self.ja), _ptr(self.k), _ptr(self.T), _ptr(self.lam),
                               _ptr(self.geo), _ptr(self.alpha), _ptr(self.wall_g), _ptr(self.wall_T),
                               _ptr(self.gk), None))
        lam = self.lam.get()
        return self.gk.get(), lam


def make_fvsolve():
    """torch.autograd.Function class (torch imported on first use)."""
    import torch

    class FVSolve(torch.autograd.Function):
        """T = FVSolve.apply(k, q, solver): differentiable in k and q.  Each
        application keeps its own k, q and T, so one solver may be applied
        several times (time steps, several loss terms) before a backward."""

        @staticmethod
        def forward(ctx, k, q, solver: HeatSolver):
            T = torch.from_numpy(solver.forward(k.detach().cpu().numpy(), q.detach().cpu().numpy()).copy())
            ctx.solver = solver
            ctx.save_for_backward(k, q, T)
            return T

        @staticmethod
        def backward(ctx, gT):
            k, q, T = ctx.saved_tensors
            gk, gq = ctx.solver.adjoint_at(k.detach().cpu().numpy(), q.detach().cpu().numpy(),
                                           T.detach().cpu().numpy(), gT.detach().cpu().numpy())
            return torch.from_numpy(gk), torch.from_numpy(gq), None

    return FVSolve


def loss_and_grad(solver: HeatSolver, k: np.ndarray, subset: np.ndarray, q: np.ndarray | None = None):
    """L = mean(T[subset]^2), with dL/dk and dL/dq from one adjoint solve."""
    T = solver.forward(k, q)
    Ls = float(np.mean(T[subset] ** 2))
    gT = np.zeros_like(T)
    gT[subset] = 2.0 * T[subset] / len(subset)
    gk, gq = solver.adjoint(gT)
    return Ls, gk, gq, T
