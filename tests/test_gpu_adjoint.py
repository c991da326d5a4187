"""Differentiable solve: adjoint BiCGSTAB on A^T plus the face-flux gradient
kernel, against central finite differences (relative error <= 1e-5), an FP64
CPU adjoint built on the reference's own solver (oracle/_ref), and through
torch.autograd."""
import time

import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle
from fortranfvcell_amd import problems as P
from heat_fd import loss_difference, perturbation_rhs

pytestmark = pytest.mark.gpu

FD_REL_STEP = 1e-4   # central-difference step, h = 1e-4 * k_c


def _loss_grad(solver, k, subset):
    from fortranfvcell_amd.adjoint import loss_and_grad
    return loss_and_grad(solver, k, subset)


def _fd_correction(op, solver, k, T, subset, c):
    """Central difference of L = mean(T[S]^2) in k_c, each side evaluated as
    T(k +- h e_c) = T + d+-,  d+- = A(k +- h e_c)^-1 r+-  (r from perturbation_rhs):
    the exact difference of the two solutions (for the solution T of A(k) T = b),
    solved for directly so that it does not cancel against T.  The loss
    difference is formed as mean((d+ - d-) (2 T + d+ + d-)) over S, equal to
    mean((T + d+)^2) - mean((T + d-)^2) but free of the cancellation against L
    itself."""
    h = FD_REL_STEP * k[c]
    d = []
    for sgn in (+1.0, -1.0):
        kk = k.copy()
        kk[c] += sgn * h
        solver.load(kk)
        d.append(solver.solve_rhs(perturbation_rhs(op, k, T, c, sgn * h))[subset])
    return loss_difference(T, subset, d[0], d[1], h)


def test_adjoint_vs_direct_fd_small(fvc):
    from fortranfvcell_amd.adjoint import HeatSolver
    op, k, subset = P.heat_problem(48, subset_frac=0.05)
    solver = HeatSolver(op, tol=1e-14)
    L0, gk, gq, T = _loss_grad(solver, k, subset)
    rng = np.random.default_rng(1)
    for c in rng.choice(op.mesh.neq, 8, replace=False):
        h = FD_REL_STEP * k[c]
        Ls = []
        for sgn in (1.0, -1.0):
            kk = k.copy()
            kk[c] += sgn * h
            s = op.assemble(kk)
            Tn = sla.spsolve(P.scipy_csr(s).tocsc(), s.b)
            Ls.append(np.mean(Tn[subset] ** 2))
        fd = (Ls[0] - Ls[1]) / (2 * h)
        assert abs(gk[c] - fd) <= 1e-5 * abs(fd) + 1e-14, (c, gk[c], fd)
        # the correction-solve form used at full size agrees with the direct form
        fdc = _fd_correction(op, solver, k, T, subset, c)
        assert abs(fdc - fd) <= 1e-5 * abs(fd) + 1e-14, (c, fdc, fd)


def test_adjoint_vs_reference_cpu_adjoint(fvc, orc):
    from fortranfvcell_amd.adjoint import HeatSolver
    op, k, subset = P.heat_problem(40, subset_frac=0.05)
    solver = HeatSolver(op, tol=1e-14)
    L0, gk, lam, T = _loss_grad(solver, k, subset)
    s = op.assemble(k)
    g = np.zeros(s.neq)
    g[subset] = 2 * T[subset] / len(subset)
    # the reference's BiCGSTAB (src/Iterativos.f:140) on the explicit A^T
    sT = P.transpose_system(s)
    fT = P.to_ell(sT)
    if oracle.have_ref():
        lam_ref, _, _ = orc.krylov("bicgstab", fT, b=g, tol=1e-15, use_ref=True)
    else:
        lam_ref, _, _ = orc.krylov("bicgstab", fT, b=g, tol=1e-15)
    lam_direct = sla.spsolve(P.scipy_csr(s).T.tocsc(), g)
    assert np.linalg.norm(lam_ref - lam_direct) <= 1e-8 * np.linalg.norm(lam_direct)
    assert np.linalg.norm(lam - lam_ref) <= 1e-8 * np.linalg.norm(lam_ref)
    gk_ref = op.grad_k(k, T, lam_ref)
    assert np.linalg.norm(gk - gk_ref) <= 1e-8 * np.linalg.norm(gk_ref)


def test_torch_autograd_function(fvc):
    torch = pytest.importorskip("torch")
    from fortranfvcell_amd.adjoint import HeatSolver, make_fvsolve
    FVSolve = make_fvsolve()
    op, k, subset = P.heat_problem(32, subset_frac=0.05)
    solver = HeatSolver(op, tol=1e-14)
    kt = torch.tensor(k, dtype=torch.float64, requires_grad=True)
    qt = torch.zeros(op.mesh.neq, dtype=torch.float64, requires_grad=True)
    T = FVSolve.apply(kt, qt, solver)
    idx = torch.from_numpy(subset)
    loss = (T[idx] ** 2).mean()
    loss.backward()
    L0, gk, gq, _ = _loss_grad(solver, k, subset)
    assert abs(loss.item() - L0) <= 1e-14 * abs(L0)
    # torch's dL/dT rounds differently from numpy's 2 T / |S|: same solve to ~tol
    assert np.allclose(kt.grad.numpy(), gk, rtol=1e-8, atol=1e-12 * np.abs(gk).max())
    assert np.allclose(qt.grad.numpy(), gq, rtol=1e-8, atol=1e-12 * np.abs(gq).max())
    # dL/dq against central differences on a few cells (L is quadratic in q)
    for c in (0, 17, 500):
        e = np.zeros(op.mesh.neq)
        e[c] = 1e-3
        Lp = np.mean(solver.forward(k, e)[subset] ** 2)
        Lm = np.mean(solver.forward(k, -e)[subset] ** 2)
        fd = (Lp - Lm) / 2e-3
        assert abs(gq[c] - fd) <= 1e-6 * abs(fd) + 1e-15


def test_torch_two_applications_before_backward(fvc):
    """One solver applied to two conductivities (two time levels, two loss
    terms) before one backward: each application's gradient is taken at its
    own k, q and T, not at the solver's last forward state."""
    torch = pytest.importorskip("torch")
    from fortranfvcell_amd.adjoint import HeatSolver, make_fvsolve
    FVSolve = make_fvsolve()
    op, k1, subset = P.heat_problem(32, subset_frac=0.05)
    k2 = k1 * np.exp(0.3 * np.sin(np.arange(op.mesh.neq)))
    q2 = 1e-2 * np.cos(np.arange(op.mesh.neq))
    solver = HeatSolver(op, tol=1e-14)
    kt1 = torch.tensor(k1, dtype=torch.float64, requires_grad=True)
    kt2 = torch.tensor(k2, dtype=torch.float64, requires_grad=True)
    qt1 = torch.zeros(op.mesh.neq, dtype=torch.float64, requires_grad=True)
    qt2 = torch.tensor(q2, dtype=torch.float64, requires_grad=True)
    idx = torch.from_numpy(subset)
    T1 = FVSolve.apply(kt1, qt1, solver)
    T2 = FVSolve.apply(kt2, qt2, solver)
    ((T1[idx] ** 2).mean() + (T2[idx] ** 2).mean()).backward()
    from fortranfvcell_amd.adjoint import loss_and_grad
    _, gk1, gq1, _ = loss_and_grad(solver, k1, subset)
    _, gk2, gq2, _ = loss_and_grad(solver, k2, subset, q2)
    assert np.abs(gk1 - gk2).max() > 1e-3 * np.abs(gk1).max(), "the two applications must differ"
    for g, ref in ((kt1.grad, gk1), (kt2.grad, gk2), (qt1.grad, gq1), (qt2.grad, gq2)):
        assert np.allclose(g.numpy(), ref, rtol=1e-8, atol=1e-12 * np.abs(ref).max())


def test_adjoint_full_config_512(fvc):
    """The differentiation config: 512x512 jittered quads, lognormal k, 1%
    subset; 32 random cells against central finite differences, and all cells
    against a sparse-LU FP64 adjoint."""
    from fortranfvcell_amd.adjoint import HeatSolver
    op, k, subset = P.heat_problem(512)
    solver = HeatSolver(op, tol=1e-15)
    t0 = time.perf_counter()
    L0, gk, gq, T = _loss_grad(solver, k, subset)
    print(f"forward {solver.last_forward['iterations']} it {solver.last_forward['seconds'] * 1e3:.1f} ms, "
          f"adjoint {solver.last_adjoint['iterations']} it {solver.last_adjoint['seconds'] * 1e3:.1f} ms, "
          f"wall {time.perf_counter() - t0:.2f} s", flush=True)
    rng = np.random.default_rng(2026)
    cells = rng.choice(op.mesh.neq, 32, replace=False)
    worst = 0.0
    gmax = np.abs(gk).max()
    rows = []
    for c in cells:
        fd = _fd_correction(op, solver, k, T, subset, c)
        err = abs(gk[c] - fd) / max(abs(fd), 1e-300)
        rows.append((int(c), gk[c], fd, err))
        worst = max(worst, err)
    for r in rows:
        print(f"  cell {r[0]:7d} adjoint {r[1]: .6e} fd {r[2]: .6e} rel {r[3]:.1e} |g|/max {abs(r[1]) / gmax:.1e}")
    print(f"32 cells: worst relative error vs central FD {worst:.2e}", flush=True)
    for c, g, fd, err in rows:
        assert err <= 1e-5, (c, g, fd, err)
    # every cell against an FP64 CPU adjoint by sparse LU (A^T lam = dL/dT,
    # refined twice), with T solved the same way
    s = op.assemble(k)
    A = P.scipy_csr(s).tocsc()
    lu = sla.splu(A)
    Tl = lu.solve(s.b)
    for _ in range(2):
        Tl = Tl + lu.solve(s.b - A @ Tl)
    g = np.zeros(s.neq)
    g[subset] = 2.0 * Tl[subset] / len(subset)
    lam = lu.solve(g, trans="T")
    for _ in range(2):
        lam = lam + lu.solve(g - A.T @ lam, trans="T")
    gk_lu = op.grad_k(k, Tl, lam)
    err_all = np.linalg.norm(gk - gk_lu) / np.linalg.norm(gk_lu)
    print(f"all {s.neq} cells vs sparse-LU adjoint: relative 2-norm error {err_all:.2e}", flush=True)
    assert err_all <= 1e-8


def test_adjoint_peclet100(fvc):
    """The differentiation config with config 2's velocity field scaled to
    its largest cell Peclet number, 100, at 128x128 cells: the largest mesh
    on which Jacobi-BiCGSTAB solves this steady problem.  Measured on MI355X
    (tol 1e-13, true residual): 128^2 forward 57,813 iterations, adjoint
    1,976; 256^2 the forward goes non-finite after 1,523 iterations (the
    adjoint converges in 6,067) and GMRES(30) reaches 9e-8 in 4,000 cycles;
    512^2 the forward breaks down after 984 iterations, the adjoint after
    10,823, GMRES(30) 1.4e-2 in 4,000 cycles.  So the 512^2 configuration runs
    at Peclet 10 (test_adjoint_full_config_512) and the Peclet-100 field is
    checked here: the adjoint against a sparse-LU FP64 adjoint on every cell
    and the gradient against LU central differences on 12 cells."""
    from fortranfvcell_amd.adjoint import HeatSolver
    from heat_fd import fd_gradient_lu
    op, k, subset = P.heat_problem(128, peclet_max=100.0)
    solver = HeatSolver(op, tol=1e-13)
    L0, gk, gq, T = _loss_grad(solver, k, subset)
    print(f"Pe 100, 128^2: forward {solver.last_forward['iterations']} it, adjoint "
          f"{solver.last_adjoint['iterations']} it", flush=True)
    s = op.assemble(k)
    A = P.scipy_csr(s).tocsc()
    lu = sla.splu(A)
    Tl = lu.solve(s.b)
    for _ in range(2):
        Tl = Tl + lu.solve(s.b - A @ Tl)
    assert np.linalg.norm(T - Tl) <= 1e-9 * np.linalg.norm(Tl)
    g = np.zeros(s.neq)
    g[subset] = 2.0 * Tl[subset] / len(subset)
    lam = lu.solve(g, trans="T")
    for _ in range(2):
        lam = lam + lu.solve(g - A.T @ lam, trans="T")
    assert np.linalg.norm(gq - lam) <= 1e-8 * np.linalg.norm(lam)
    gk_lu = op.grad_k(k, Tl, lam)
    err_all = np.linalg.norm(gk - gk_lu) / np.linalg.norm(gk_lu)
    print(f"  all cells vs sparse-LU adjoint: {err_all:.2e}", flush=True)
    assert err_all <= 1e-8
    cells = np.sort(np.random.default_rng(100).choice(s.neq, 12, replace=False))
    fd = fd_gradient_lu(op, k, Tl, subset, cells)
    rel = np.abs(gk[cells] - fd) / np.abs(fd)
    print(f"  12 cells vs LU central differences: worst {rel.max():.1e}", flush=True)
    assert np.all(rel <= 1e-5)


def test_gpu_assembly_is_bit_identical_to_host(fvc):
    """fvc_assemble_heat_dev against problems.HeatOperator.assemble: the same
    coefficient forms in the same order, so equal to the last bit."""
    from fortranfvcell_amd.adjoint import HeatSolver
    op, k, subset = P.heat_problem(64, subset_frac=0.05)
    solver = HeatSolver(op)
    q = np.sin(np.arange(op.mesh.neq) * 0.01)
    for qq in (None, q):
        solver.load(k, qq)
        s = op.assemble(k, qq)
        inner = s.nbr >= 0
        assert np.array_equal(solver.ad.get(), s.ad)
        assert np.array_equal(solver.al.get(), np.where(inner, s.off, 0.0).ravel())
        assert np.array_equal(solver.b.get(), s.b)
