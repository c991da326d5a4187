"""Mixed precision: FP32 matrix and inner PCG on the Jacobi-scaled system,
FP64 solution and FP64 residual recomputation.  It must reach the FP64
true relative residual 1e-10, including on the reference's 1e60-pinned
pressure systems (which overflow FP32 unless scaled first), and its error
is stated against the FP64 solution."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from fortranfvcell_amd import problems as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,build,lay", [
    ("poisson64_ell", lambda: P.config_poisson(64), "ell"),
    ("poisson64_csrd", lambda: P.config_poisson(64), "csrd"),
    ("lognormal_quad48_ell", lambda: P.config_pressure_lognormal(48), "ell"),
    ("lognormal_tri24_ell3", lambda: P.config_pressure_lognormal(24, triangles=True), "ell"),
])
def test_mixed_pcg_reaches_fp64_residual(fvc, name, build, lay):
    s = build()
    f = P.to_ell(s) if lay == "ell" else P.to_csrd(s)
    print(f"mixed {name} neq={s.neq}", flush=True)
    A = P.scipy_csr(s)
    xd = sla.spsolve(A.tocsc(), s.b)
    xm, sm = fvc.solve(f, fvc.PCG, tol=1e-10, precision=fvc.PREC_MIXED)
    x64, s64 = fvc.solve(f, fvc.PCG, tol=1e-10, stop_rule=fvc.STOP_RELRES)
    assert sm["status"] == 0, sm
    r = np.linalg.norm(s.b - A @ xm) / np.linalg.norm(s.b)
    assert r <= 1e-10, r
    assert abs(sm["true_relres"] - r) <= 0.05 * r
    e_m = np.linalg.norm(xm - xd) / np.linalg.norm(xd)
    e_64 = np.linalg.norm(x64 - xd) / np.linalg.norm(xd)
    print(f"  iterations mixed={sm['iterations']} (updates {sm['inner_iterations']}) fp64={s64['iterations']}"
          f" err mixed={e_m:.2e} fp64={e_64:.2e} vs_fp64={np.linalg.norm(xm - x64) / np.linalg.norm(x64):.2e}",
          flush=True)
    # the parity bar of the FP64 path, ||x - x_ref|| / ||x_ref|| <= 1e-8 at
    # residual 1e-10, against the direct solution
    assert e_m <= 1e-8
    # reliable updates keep the FP32 recursion within a small factor of FP64 CG
    assert sm["iterations"] <= 2 * s64["iterations"] + 20


def test_mixed_engines_bit_identical(fvc):
    s = P.config_pressure_lognormal(40)
    f = P.to_ell(s)
    xp, sp = fvc.solve(f, fvc.PCG, precision=fvc.PREC_MIXED, engine=fvc.ENGINE_PERSISTENT)
    xk, sk = fvc.solve(f, fvc.PCG, precision=fvc.PREC_MIXED, engine=fvc.ENGINE_KERNELS)
    assert sp["status"] == 0 and sk["status"] == 0
    assert np.array_equal(xp, xk) and sp["iterations"] == sk["iterations"]
    assert sp["inner_iterations"] == sk["inner_iterations"]


@pytest.mark.parametrize("triangles", [False, True], ids=["ell4", "ell3"])
def test_mixed_column_offsets_int16_and_int32(fvc, triangles):
    """The FP32 inner iterations read ELLPACK columns as int16 offsets j - i
    when every nonzero slot's offset fits, else the int32 columns.  The
    mesh's own numbering (offsets up to a grid row) takes the first path; a
    random renumbering of the same system (offsets up to its size, above
    32,767) takes the second.  Both meet the bar of the FP64 path."""
    s = P.config_pressure_lognormal(200 if not triangles else 140, triangles=triangles)
    n, w = s.neq, s.w
    rng = np.random.default_rng(11)
    p = rng.permutation(n)
    q = np.empty(n, np.int64)
    q[p] = np.arange(n)
    nb = s.nbr[p]
    inner = nb >= 0
    ja = np.where(inner, q[np.where(inner, nb, 0)] + 1, 1).astype(np.int32)
    al = np.where(inner, s.off[p], 0.0)
    renumbered = dict(matriz=4, neq=n, nad=int(inner.sum()), iax=np.array([w, n + 1], np.int32),
                      ja=np.ascontiguousarray(ja.ravel()), ad=s.ad[p].copy(),
                      al=np.ascontiguousarray(al.ravel()), au=None, b=s.b[p].copy(), w=w)
    offsets = np.abs(ja.astype(np.int64) - 1 - np.arange(n)[:, None])[inner]
    assert offsets.max() > 32767
    own = np.abs(s.nbr - np.arange(n)[:, None])[s.nbr >= 0]
    assert own.max() <= 32767
    A = P.scipy_csr(s)
    xd = sla.spsolve(A.tocsc(), s.b)
    print(f"column offsets neq={n} w={w} renumbered max offset {offsets.max()}", flush=True)
    for f, xref, col16 in ((P.to_ell(s), xd, True), (renumbered, xd[p], False)):
        x, st = fvc.solve(f, fvc.PCG, tol=1e-10, precision=fvc.PREC_MIXED)
        assert st["status"] == 0, st
        assert bool(st["flags"] & fvc.FLAG_COL16) == col16, st
        assert np.linalg.norm(x - xref) <= 1e-8 * np.linalg.norm(xref)


def test_mixed_maxit_and_warm_start(fvc):
    # maxit stops the FP32 iteration count; a warm start from the converged
    # solution needs no iteration at all
    s = P.config_pressure_lognormal(32)
    f = P.to_ell(s)
    for engine in (fvc.ENGINE_PERSISTENT, fvc.ENGINE_KERNELS):
        _, st = fvc.solve(f, fvc.PCG, precision=fvc.PREC_MIXED, engine=engine, maxit=5)
        assert st["status"] == 1 and st["iterations"] == 5, st
        x, st = fvc.solve(f, fvc.PCG, precision=fvc.PREC_MIXED, engine=engine, stop_rule=fvc.STOP_RELRES)
        assert st["status"] == 0
        x2, st2 = fvc.solve(f, fvc.PCG, x0=x, precision=fvc.PREC_MIXED, engine=engine, x0_rule=fvc.X0_GIVEN)
        assert st2["status"] == 0 and st2["iterations"] == 0 and np.array_equal(x2, x)


def test_mixed_unsupported_combinations(fvc):
    sp = P.config_poisson(16)
    with pytest.raises(fvc.FvcError):
        fvc.solve(P.to_csrd(sp, sym=True), fvc.PCG, precision=fvc.PREC_MIXED, raise_on_failure=True)
