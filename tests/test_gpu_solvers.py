"""GPU Krylov solves through the C ABI against the oracle.

Bar (stated by the path's contract): solutions agree with the FP64 oracle to
||x - x_ref|| / ||x_ref|| <= 1e-8 at relative residual ~1e-10.  Iteration
counts agree within a small band: the GPU's dot products sum in a fixed
tree rather than left to right, which moves the last bits of every scalar.
The two GPU engines (persistent kernel, kernel per phase) share the row order and
reduction trees and must agree bit for bit.
"""
import numpy as np
import scipy.sparse.linalg as sla
import pytest

import oracle
from fortranfvcell_amd import problems as P

pytestmark = pytest.mark.gpu

X_TOL = 1e-8


def _rel(x, y):
    return np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)


def _iters_close(a, b):
    return abs(a - b) <= max(2, int(0.03 * b))


def tol_for_relres(orc, s, f, code, unsym, target=1e-10):
    """Largest tol in a decade ladder for which the oracle's solution has
    ||b - A x|| / ||b|| <= target (the stopping rules are not the residual)."""
    A = P.scipy_csr(s)
    for tol in (1e-10, 3e-11, 1e-11, 3e-12, 1e-12, 3e-13):
        xo, _, so = orc.solver(f, code, unsym=unsym, omp=True, tol=tol, maxit=20000)
        assert so["status"] == 0
        if np.linalg.norm(s.b - A @ xo) <= target * np.linalg.norm(s.b):
            return tol, xo, so
    raise AssertionError("oracle never reached the target residual")


CASES = [
    # (name, builder, code, layout)
    ("pcg_ell_poisson32", lambda: P.config_poisson(32), 1, "ell"),
    ("pcg_csrd_poisson32", lambda: P.config_poisson(32), 1, "csrd"),
    ("pcg_csrdsym_lognormal24", lambda: P.config_pressure_lognormal(24), 1, "csrd_sym"),
    ("pcg_ell3_tri16", lambda: P.config_pressure_lognormal(16, triangles=True), 1, "ell"),
    ("bicg_ell_advdiff32", lambda: P.config_advdiff(32, peclet_max=60.0), 2, "ell"),
    ("bicg_csrd_advdiff32", lambda: P.config_advdiff(32, peclet_max=60.0), 2, "csrd"),
    ("bicg_csrc_advdiff24", lambda: P.config_advdiff(24, peclet_max=60.0), 2, "csrc"),
    ("gmres_csrd_advdiff24", lambda: P.config_advdiff(24, peclet_max=60.0), 3, "csrd"),
    ("gmres_csrc_advdiff24", lambda: P.config_advdiff(24, peclet_max=60.0), 3, "csrc"),
]

LAYOUT = {"ell": P.to_ell, "csrd": P.to_csrd, "csr": P.to_csr,
          "csrd_sym": lambda s: P.to_csrd(s, sym=True), "csrc": lambda s: P.to_csrc(s, unsym=True),
          "csrc_sym": lambda s: P.to_csrc(s, unsym=False)}


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_solver_parity_vs_oracle(fvc, orc, case):
    name, build, code, lay = case
    s = build()
    f = LAYOUT[lay](s)
    unsym = not lay.endswith("sym")
    print(f"solve {name} neq={s.neq}", flush=True)
    # the reference path with openmpSolver = .true., at the tolerance where the
    # reference's own stopping rule reaches a true relative residual of 1e-10
    tol, xo, so = tol_for_relres(orc, s, f, code, unsym)
    x0 = np.zeros(s.neq)
    stg = fvc.solver(f["ad"], f.get("au"), f["al"] if f.get("al") is not None else f["ad"], x0,
                     np.ascontiguousarray(f["b"]), f["iax"], f["ja"], s.neq, f["nad"], unsym,
                     tol, 20000, code, f["matriz"], 2, omp_solver=True)
    assert stg["status"] == 0
    assert _rel(x0, xo) <= X_TOL, (name, _rel(x0, xo))
    assert _iters_close(stg["iterations"], so["iterations"]), (stg["iterations"], so["iterations"])
    assert abs(stg["energy"] - so["energy"]) <= 1e-8 * abs(so["energy"])


KRYLOV2 = [c for c in CASES if c[2] in (1, 2)]


@pytest.mark.parametrize("rec", [0, 1], ids=["fused", "reference_order"])
@pytest.mark.parametrize("case", KRYLOV2, ids=[c[0] for c in KRYLOV2])
def test_recurrences_vs_oracle(fvc, orc, case, rec):
    """PCG and BiCGSTAB in both recurrences (one/two grid reductions per
    iteration, or the reference's expressions in its order) against the
    oracle at the same tolerance."""
    name, build, code, lay = case
    s = build()
    f = LAYOUT[lay](s)
    unsym = not lay.endswith("sym")
    print(f"recurrence {rec} {name} neq={s.neq}", flush=True)
    tol, xo, so = tol_for_relres(orc, s, f, code, unsym)
    x, st = fvc.solve(f, code, tol=tol, maxit=20000, recurrence=rec)
    assert st["status"] == 0
    assert _rel(x, xo) <= X_TOL, (name, rec, _rel(x, xo))
    assert _iters_close(st["iterations"], so["iterations"]), (st["iterations"], so["iterations"])
    assert abs(st["energy"] - so["energy"]) <= 1e-8 * abs(so["energy"])


# GMRES has a single recurrence
ENGINE_CASES = [(code, lay, rec) for code in (1, 2, 3) for lay in ("ell", "csrd") for rec in (0, 1)
                if not (code == 3 and rec == 1)]


@pytest.mark.parametrize("code,lay,rec", ENGINE_CASES,
                         ids=[f"{c}-{l}-{'fused' if r == 0 else 'reference_order'}" for c, l, r in ENGINE_CASES])
def test_engines_bit_identical(fvc, code, lay, rec):
    s = P.config_poisson(40) if code == 1 else P.config_advdiff(40, peclet_max=60.0)
    f = LAYOUT[lay](s)
    print(f"engines code={code} {lay} recurrence={rec}", flush=True)
    xp, sp = fvc.solve(f, code, engine=fvc.ENGINE_PERSISTENT, gmres_k=20, recurrence=rec)
    xk, sk = fvc.solve(f, code, engine=fvc.ENGINE_KERNELS, gmres_k=20, recurrence=rec)
    assert sp["status"] == 0 and sk["status"] == 0
    assert sp["engine"] == fvc.ENGINE_PERSISTENT and sk["engine"] == fvc.ENGINE_KERNELS
    assert np.array_equal(xp, xk)
    assert sp["iterations"] == sk["iterations"]
    assert sp["energy"] == sk["energy"]


ROW_CASES = [(1, "ell", 0), (2, "ell", 0), (3, "csrd", 0), (1, "csrd_sym", 0), (2, "csrc", 0), (1, "ell", 1)]


@pytest.mark.parametrize("G", [64, 60], ids=["8_groups", "1_group"])
@pytest.mark.parametrize("code,lay,prec", ROW_CASES,
                         ids=[f"{c}-{l}-{'mixed' if p else 'fp64'}" for c, l, p in ROW_CASES])
def test_row_orders(fvc, orc, code, lay, prec, G):
    """Rows to blocks as one contiguous chunk per block (row_order 1) or as
    XCD tiles (row_order 2: 64 blocks make 8 groups of 8 sweeping an eighth
    of the rows each, 60 blocks one group sweeping all of them).  With
    64-thread blocks every block holds several rows per thread, so the two
    orders give the rows to different threads.  Both reach the oracle's
    solution, and under tiles the two engines agree bit for bit."""
    s = P.config_poisson(96) if code == 1 else P.config_advdiff(96, peclet_max=60.0)
    f = LAYOUT[lay](s)
    unsym = not lay.endswith("sym")
    print(f"row orders code={code} {lay} precision={prec} G={G} neq={s.neq}", flush=True)
    opts = dict(grid_blocks=G, block_size=64)
    if prec:
        A = P.scipy_csr(s)
        xd = sla.spsolve(A.tocsc(), s.b)
        runs = [fvc.solve(f, code, tol=1e-10, precision=fvc.PREC_MIXED, row_order=ro, engine=e, **opts)
                for ro, e in ((1, fvc.ENGINE_AUTO), (2, fvc.ENGINE_PERSISTENT), (2, fvc.ENGINE_KERNELS))]
        for x, st in runs:
            assert st["status"] == 0 and st["grid_blocks"] == G
            assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)
    else:
        tol, xo, so = tol_for_relres(orc, s, f, code, unsym)
        runs = [fvc.solve(f, code, tol=tol, maxit=20000, row_order=ro, engine=e, **opts)
                for ro, e in ((1, fvc.ENGINE_AUTO), (2, fvc.ENGINE_PERSISTENT), (2, fvc.ENGINE_KERNELS))]
        for x, st in runs:
            assert st["status"] == 0 and st["grid_blocks"] == G
            assert _rel(x, xo) <= X_TOL, _rel(x, xo)
            assert _iters_close(st["iterations"], so["iterations"]), (st["iterations"], so["iterations"])
    (xp, sp), (xk, sk) = runs[1], runs[2]
    assert np.array_equal(xp, xk) and sp["iterations"] == sk["iterations"]


def test_repeat_solves_bit_identical(fvc):
    s = P.config_poisson(48)
    f = P.to_ell(s)
    x1, s1 = fvc.solve(f, 1)
    x2, s2 = fvc.solve(f, 1)
    assert np.array_equal(x1, x2) and s1["iterations"] == s2["iterations"]


@pytest.mark.parametrize("k", [5, 30])
def test_gmres_ellpack_any_k(fvc, orc, k):
    # the reference has no ELLPACK branch in solver(); gmres itself takes any
    # matvec and k (src/Iterativos.f:271): checked against the oracle's gmres
    s = P.config_advdiff(32, peclet_max=60.0)
    f = P.to_ell(s)
    A = P.scipy_csr(s)
    for tol in (1e-10, 3e-11, 1e-11, 3e-12, 1e-12):
        xo, _, so = orc.krylov("gmres", f, k=k, tol=tol)
        if np.linalg.norm(s.b - A @ xo) <= 1e-10 * np.linalg.norm(s.b):
            break
    xg, sg = fvc.solve(f, 3, gmres_k=k, tol=tol)
    assert sg["status"] == 0
    assert _rel(xg, xo) <= X_TOL
    assert _iters_close(sg["iterations"], so["iterations"])
    assert abs(sg["inner_iterations"] - so["inner"]) <= max(3, int(0.03 * so["inner"]))


def test_x0_rules_follow_the_reference(fvc, orc):
    # ELLPACK PCG warm-starts from x on both reference paths; CSRD PCG does
    # only with openmpSolver; GMRES always starts from zero.
    s = P.config_poisson(24)
    xs = np.linspace(-1.0, 1.0, s.neq)
    for lay, omp in (("ell", True), ("ell", False), ("csrd", True), ("csrd", False)):
        f = LAYOUT[lay](s)
        xo, _, so = orc.solver(f, 1, omp=omp, x0=xs)
        x0 = xs.copy()
        st = fvc.solver(f["ad"], None, f["al"], x0, np.ascontiguousarray(f["b"]), f["iax"], f["ja"],
                        s.neq, f["nad"], True, 1e-10, 20000, 1, f["matriz"], 2, omp_solver=omp)
        assert _rel(x0, xo) <= X_TOL
        assert _iters_close(st["iterations"], so["iterations"]), (lay, omp)


def test_true_residual_and_relres_rule(fvc):
    s = P.config_advdiff(48, peclet_max=80.0)
    f = P.to_ell(s)
    x, st = fvc.solve(f, 2, tol=1e-10, stop_rule=fvc.STOP_RELRES)
    assert st["status"] == 0
    r = np.linalg.norm(s.b - P.scipy_csr(s) @ x) / np.linalg.norm(s.b)
    assert r <= 2e-10
    assert abs(st["true_relres"] - r) <= 1e-3 * r + 1e-16


@pytest.mark.parametrize("rec", [0, 1], ids=["fused", "reference_order"])
def test_maxit_counts(fvc, rec):
    # maxit stops after exactly maxit iterations in both recurrences; the
    # x returned is that iteration's (checked against a longer run's prefix
    # through the residual decrease)
    s = P.config_advdiff(24, peclet_max=60.0)
    f = P.to_ell(s)
    for code in (1, 2):
        g = P.to_ell(P.config_poisson(24)) if code == 1 else f
        for engine in (fvc.ENGINE_PERSISTENT, fvc.ENGINE_KERNELS):
            x3, st3 = fvc.solve(g, code, maxit=3, recurrence=rec, engine=engine)
            assert st3["status"] == 1 and st3["iterations"] == 3, (code, engine, st3)
            x3r, _ = fvc.solve(g, code, maxit=3, recurrence=1 - rec, engine=engine)
            assert _rel(x3, x3r) <= 1e-10, (code, engine)


@pytest.mark.parametrize("code", [1, 2])
def test_relres_restarts_stop_at_the_fp64_floor(fvc, code):
    # a true relative residual of 1e-17 is below what FP64 can reach: the
    # restarts stop once one no longer halves it, instead of running all eight
    s = P.config_poisson(48) if code == 1 else P.config_advdiff(48, peclet_max=60.0)
    f = P.to_ell(s)
    _, st13 = fvc.solve(f, code, tol=1e-13, stop_rule=fvc.STOP_RELRES, x0_rule=fvc.X0_ZERO)
    cap = 4 * st13["iterations"] + 50
    x, st = fvc.solve(f, code, tol=1e-17, stop_rule=fvc.STOP_RELRES, x0_rule=fvc.X0_ZERO, maxit=cap)
    # converged on the recursive residual, or (BiCGSTAB's can stall) at the cap
    assert st["status"] in (0, 1, 2, 3)
    assert st["iterations"] <= cap, (st["iterations"], st13["iterations"])
    r = np.linalg.norm(s.b - P.scipy_csr(s) @ x) / np.linalg.norm(s.b)
    assert r <= 1e-13 and abs(st["true_relres"] - r) <= 0.5 * r + 1e-16


def test_status_codes(fvc):
    s = P.config_poisson(16)
    f = P.to_ell(s)
    _, st = fvc.solve(f, 1, maxit=2)
    assert st["status"] == 1 and st["iterations"] == 2
    _, st = fvc.solve(f, 1, b=np.zeros(s.neq))
    assert st["status"] in (2, 3)
    bad = s.b.copy()
    bad[5] = np.nan
    _, st = fvc.solve(f, 2, b=bad)
    assert st["status"] == 3
    with pytest.raises(fvc.FvcError):
        fvc.solve(dict(f, matriz=7), 1, raise_on_failure=True)
    with pytest.raises(fvc.FvcError):
        fvc.solve(P.to_csr(s), 3, gmres_k=1000, raise_on_failure=True)
    with pytest.raises(fvc.FvcError, match="row_order"):
        fvc.solve(f, 1, row_order=3)


def test_device_resident_solve_matches_host(fvc):
    s = P.config_advdiff(40, peclet_max=60.0)
    f = P.to_ell(s)
    xh, sh = fvc.solve(f, 2)
    ds = fvc.DeviceSystem(f)
    sd = ds.solve(2, x0=np.zeros(s.neq))
    assert np.array_equal(ds.x.get(), xh)
    assert sd["iterations"] == sh["iterations"]


def test_pcg_csr_full_storage(fvc, orc):
    # matriz 2 (full CSR): the reference's PCG call is malformed
    # (src/Solver.f:142-143); here Jacobi uses the true diagonal, checked
    # against the oracle's pcg with the CSR matvec and m = diag(A).
    s = P.config_poisson(24)
    f = P.to_csr(s)
    ad, au, al = oracle._vals(f)
    x = np.zeros(s.neq)
    b = s.b.copy()
    st = oracle.Stats()
    import ctypes as C
    m = np.ascontiguousarray(s.ad)
    orc.lib().orc_pcg(orc.MV_CSR, s.neq, oracle._p(ad), oracle._p(au), oracle._p(al),
                      oracle._p(oracle._i32(f["iax"])), oracle._p(oracle._i32(f["ja"])), oracle._p(m),
                      oracle._p(b), oracle._p(x), 1e-10, 20000, 0, 1, C.byref(st))
    xg, sg = fvc.solve(f, 1)
    assert sg["status"] == 0 and _rel(xg, x) <= X_TOL


@pytest.mark.parametrize("code,prec", [(1, 0), (2, 0), (1, 1)], ids=["pcg", "bicgstab", "pcg_mixed"])
def test_more_partials_than_loaded_ahead(fvc, code, prec):
    """A kernel-per-phase launch loads each thread's terms of the previous
    phase's partial sums ahead, up to four; with more blocks than four times
    the block size it sums them in a loop instead.  That path must converge
    like the other and repeat bit for bit."""
    s = P.config_poisson(384) if code == 1 else P.config_advdiff(384, peclet_max=60.0)
    f = P.to_ell(s)
    print(f"many blocks code={code} precision={prec} neq={s.neq}", flush=True)
    opts = dict(tol=1e-10, maxit=20000, stop_rule=fvc.STOP_RELRES, x0_rule=fvc.X0_ZERO, precision=prec,
                engine=fvc.ENGINE_KERNELS, block_size=256)
    xa, sa = fvc.solve(f, code, grid_blocks=2048, **opts)
    xb, sb = fvc.solve(f, code, grid_blocks=2048, **opts)
    xc, sc = fvc.solve(f, code, grid_blocks=1024, **opts)
    assert sa["grid_blocks"] == 2048 and sc["grid_blocks"] == 1024
    assert sa["status"] == sc["status"] == 0
    assert np.array_equal(xa, xb) and sa["iterations"] == sb["iterations"]
    assert sa["true_relres"] <= 1e-10 and sc["true_relres"] <= 1e-10
    assert _iters_close(sa["iterations"], sc["iterations"])
