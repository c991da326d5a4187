"""Mixed-precision BiCGSTAB and GMRES(k): FP32 matrix and inner iterations
on the Jacobi-scaled system, FP64 solution, FP64 residual recomputation, and
for GMRES the Hessenberg matrix, Givens rotations and back substitution in
FP64.  Both stop on the FP64 true residual.  Held to the FP64 path's bar:
true relative residual <= 1e-10 and ||x - x_ref|| / ||x_ref|| <= 1e-8
against the reference's solutions (tests/golden: the C2 configurations at
64^2 and 128^2, the 1024^2 statistics at the ladder tolerance)."""
import os

import numpy as np
import pytest

from fortranfvcell_amd import problems as P
from test_golden import CONFIG_KEYS, REF_FIXTURES, config_system

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _relres(s, x):
    return np.linalg.norm(s.b - P.scipy_csr(s) @ x) / np.linalg.norm(s.b)


@pytest.mark.parametrize("key", [k for k in CONFIG_KEYS if k.startswith("c2_")])
@pytest.mark.parametrize("layout", ["ell", "csrd"])
def test_mixed_krylov_vs_reference(fvc, key, layout):
    s, f = config_system(key)
    if layout == "csrd":
        f = P.to_csrd(s)
    xref = REF_FIXTURES[key + "_x"]
    code = fvc.GMRES if key.endswith("gmres30") else fvc.BICGSTAB
    x, st = fvc.solve(f, code, tol=1e-10, precision=fvc.PREC_MIXED, gmres_k=30, x0_rule=fvc.X0_ZERO)
    r = _relres(s, x)
    err = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    print(f"mixed {key} {layout}: {st['iterations']} it, {st['inner_iterations']} (refreshes/steps), "
          f"relres {r:.2e}, vs reference {err:.1e}, flags {st['flags']}", flush=True)
    assert st["status"] == 0 and r <= 1e-10 and err <= 1e-8
    if layout == "ell":
        assert st["flags"] & fvc.FLAG_COL16


def test_mixed_bicgstab_engines_bit_identical(fvc):
    s = P.config_advdiff(96)
    f = P.to_ell(s)
    out = []
    for engine in (fvc.ENGINE_PERSISTENT, fvc.ENGINE_KERNELS):
        x, st = fvc.solve(f, fvc.BICGSTAB, tol=1e-10, precision=fvc.PREC_MIXED, engine=engine,
                          block_size=256, grid_blocks=64, x0_rule=fvc.X0_ZERO)
        assert st["status"] == 0 and _relres(s, x) <= 1e-10
        out.append((x, st))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1]["iterations"] == out[1][1]["iterations"]


def test_mixed_krylov_triangles_and_reliable_updates(fvc):
    """Width-3 ELLPACK (triangles) and a small refresh interval: many
    reliable updates, still the FP64 bar."""
    mesh = P.tri_mesh(80, jitter=0.1, seed=P.SEED_BASE + 9)
    s = P.advdiff_system(mesh, peclet_max=50.0, seed=P.SEED_BASE + 9, dt=0.005, under_relax=0.8)
    import scipy.sparse.linalg as sla
    xd = sla.spsolve(P.scipy_csr(s).tocsc(), s.b)
    for code, kw in ((fvc.BICGSTAB, dict(refresh_every=5)), (fvc.GMRES, dict(gmres_k=8))):
        x, st = fvc.solve(P.to_ell(s), code, tol=1e-10, precision=fvc.PREC_MIXED, x0_rule=fvc.X0_ZERO, **kw)
        err = np.linalg.norm(x - xd) / np.linalg.norm(xd)
        print(f"triangles code {code}: {st['iterations']} it ({st['inner_iterations']}), relres "
              f"{_relres(s, x):.2e}, vs direct {err:.1e}", flush=True)
        assert st["status"] == 0 and _relres(s, x) <= 1e-10 and err <= 1e-8
        if code == fvc.BICGSTAB:
            assert st["inner_iterations"] >= st["iterations"] // 5


def test_mixed_krylov_1024(fvc):
    lad = np.load(os.path.join(HERE, "golden", "ref_ladder_1024.npz"), allow_pickle=False)
    s = P.config_advdiff(1024)
    ds = fvc.DeviceSystem(P.to_ell(s))
    for name, code in (("advdiff1024_ell_bicgstab", fvc.BICGSTAB), ("advdiff1024_ell_gmres30", fvc.GMRES)):
        st64 = ds.solve(code, tol=1e-10, stop_rule=fvc.STOP_RELRES, x0_rule=fvc.X0_ZERO, gmres_k=30)
        st = ds.solve(code, tol=1e-10, precision=fvc.PREC_MIXED, x0_rule=fvc.X0_ZERO, gmres_k=30)
        x = ds.x.get()
        r = _relres(s, x)
        print(f"{name} mixed: {st['iterations']} it ({st['inner_iterations']}), {st['seconds'] * 1e3:.2f} ms "
              f"(FP64 {st64['seconds'] * 1e3:.2f} ms), relres {r:.2e}", flush=True)
        assert st["status"] == 0 and r <= 1e-10
        xsq = float(lad[name + "_xsq"])
        assert abs((x * x).sum() - xsq) <= 1e-8 * xsq
        idx = lad[name + "_idx"]
        assert np.max(np.abs(x[idx] - lad[name + "_xsample"])) <= 1e-7 * np.sqrt(xsq / s.neq)


def test_mixed_krylov_refusals(fvc):
    s = P.config_advdiff(16)
    with pytest.raises(fvc.FvcError):
        fvc.solve(P.to_ell(s), fvc.GMRES, precision=fvc.PREC_MIXED, gmres_k=40, raise_on_failure=True)
    with pytest.raises(fvc.FvcError):
        fvc.solve(P.to_ell(s), fvc.GMRES, precision=fvc.PREC_MIXED, engine=fvc.ENGINE_KERNELS,
                  raise_on_failure=True)
    with pytest.raises(fvc.FvcError):
        fvc.solve(P.to_csrd(P.config_poisson(16), sym=True), fvc.BICGSTAB, precision=fvc.PREC_MIXED,
                  raise_on_failure=True)


# Mixed-precision GMRES(k) with lagged CGS2 (FVC_ORTH_CGS2_LAG): FP32 basis,
# candidate and SpMV on the scaled FP32 matrix, every dot product and the
# Hessenberg / Givens / back substitution in FP64, one kernel per phase.

@pytest.mark.parametrize("key", [k for k in CONFIG_KEYS if k.startswith("c2_") and k.endswith("gmres30")])
@pytest.mark.parametrize("layout", ["ell", "csrd"])
def test_mixed_lagged_gmres_vs_reference(fvc, key, layout):
    s, f = config_system(key)
    if layout == "csrd":
        f = P.to_csrd(s)
    xref = REF_FIXTURES[key + "_x"]
    x, st = fvc.solve(f, fvc.GMRES, tol=1e-10, precision=fvc.PREC_MIXED, gmres_k=30,
                      gmres_orth=fvc.ORTH_CGS2_LAG)
    r = _relres(s, x)
    err = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    print(f"mixed lagged gmres {key} {layout}: {st['iterations']} cycles, {st['inner_iterations']} steps, "
          f"relres {r:.2e}, vs reference {err:.1e}", flush=True)
    assert st["status"] == 0 and r <= 1e-10 and err <= 1e-8
    assert st["engine"] == fvc.ENGINE_KERNELS
    x2, _ = fvc.solve(f, fvc.GMRES, tol=1e-10, precision=fvc.PREC_MIXED, gmres_k=30,
                      gmres_orth=fvc.ORTH_CGS2_LAG)
    assert np.array_equal(x, x2)   # deterministic


def test_mixed_lagged_gmres_triangles_and_1024(fvc):
    mesh = P.tri_mesh(80, jitter=0.1, seed=P.SEED_BASE + 9)
    s = P.advdiff_system(mesh, peclet_max=50.0, seed=P.SEED_BASE + 9, dt=0.005, under_relax=0.8)
    import scipy.sparse.linalg as sla
    xd = sla.spsolve(P.scipy_csr(s).tocsc(), s.b)
    for k in (8, 64):
        x, st = fvc.solve(P.to_ell(s), fvc.GMRES, tol=1e-10, precision=fvc.PREC_MIXED, gmres_k=k,
                          gmres_orth=fvc.ORTH_CGS2_LAG)
        err = np.linalg.norm(x - xd) / np.linalg.norm(xd)
        print(f"triangles k={k}: {st['iterations']} cycles ({st['inner_iterations']}), relres "
              f"{_relres(s, x):.2e}, vs direct {err:.1e}", flush=True)
        assert st["status"] == 0 and _relres(s, x) <= 1e-10 and err <= 1e-8
    lad = np.load(os.path.join(HERE, "golden", "ref_ladder_1024.npz"), allow_pickle=False)
    s = P.config_advdiff(1024)
    ds = fvc.DeviceSystem(P.to_ell(s))
    name = "advdiff1024_ell_gmres30"
    st64 = ds.solve(fvc.GMRES, tol=1e-10, stop_rule=fvc.STOP_RELRES, gmres_k=30, gmres_orth=fvc.ORTH_CGS2_LAG)
    st = ds.solve(fvc.GMRES, tol=1e-10, precision=fvc.PREC_MIXED, gmres_k=30, gmres_orth=fvc.ORTH_CGS2_LAG)
    x = ds.x.get()
    r = _relres(s, x)
    print(f"{name} mixed lagged: {st['iterations']} cycles ({st['inner_iterations']}), "
          f"{st['seconds'] * 1e3:.2f} ms (FP64 lagged {st64['seconds'] * 1e3:.2f} ms), relres {r:.2e}", flush=True)
    assert st["status"] == 0 and r <= 1e-10
    xsq = float(lad[name + "_xsq"])
    assert abs((x * x).sum() - xsq) <= 1e-8 * xsq
    idx = lad[name + "_idx"]
    assert np.max(np.abs(x[idx] - lad[name + "_xsample"])) <= 1e-7 * np.sqrt(xsq / s.neq)


def test_mixed_lagged_gmres_refusals(fvc):
    s = P.config_advdiff(16)
    f = P.to_ell(s)
    with pytest.raises(fvc.FvcError):
        fvc.solve(f, fvc.GMRES, precision=fvc.PREC_MIXED, gmres_k=65, gmres_orth=fvc.ORTH_CGS2_LAG,
                  raise_on_failure=True)
    with pytest.raises(fvc.FvcError):
        fvc.solve(f, fvc.GMRES, precision=fvc.PREC_MIXED, gmres_orth=fvc.ORTH_CGS2_LAG,
                  engine=fvc.ENGINE_PERSISTENT, raise_on_failure=True)
