"""The finite-difference reference of the adjoint tests (tests/heat_fd.py),
checked on the CPU: the face-by-face right-hand side equals the assembled
difference of the two systems, and LU-based central differences agree with
an exact sparse-LU adjoint gradient."""
import numpy as np
import scipy.sparse.linalg as sla

from fortranfvcell_amd import problems as P
from heat_fd import fd_gradient_lu, perturbation_rhs


def test_perturbation_rhs_matches_assembled_difference():
    op, k, subset = P.heat_problem(24, subset_frac=0.05)
    s0 = op.assemble(k)
    T = np.linspace(0.0, 1.0, s0.neq) ** 2
    A0 = P.scipy_csr(s0)
    for c in (0, 23, 300, 575):
        for dk in (0.3 * k[c], -0.2 * k[c]):
            kk = k.copy()
            kk[c] += dk
            s = op.assemble(kk)
            want = (s.b - s0.b) - (P.scipy_csr(s) @ T - A0 @ T)
            got = perturbation_rhs(op, k, T, c, dk)
            assert np.allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())


def test_lu_central_differences_match_exact_adjoint():
    op, k, subset = P.heat_problem(40, subset_frac=0.02)
    s = op.assemble(k)
    A = P.scipy_csr(s).tocsc()
    lu = sla.splu(A)
    T = lu.solve(s.b)
    T = T + lu.solve(s.b - A @ T)
    g = np.zeros(s.neq)
    g[subset] = 2.0 * T[subset] / len(subset)
    lam = lu.solve(g, trans="T")
    lam = lam + lu.solve(g - A.T @ lam, trans="T")
    gk = op.grad_k(k, T, lam)
    cells = np.random.default_rng(3).choice(s.neq, 12, replace=False)
    fd = fd_gradient_lu(op, k, T, subset, cells)
    assert np.all(np.abs(fd - gk[cells]) <= 1e-6 * np.abs(gk[cells])), (fd, gk[cells])
