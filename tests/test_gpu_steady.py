"""The advection-diffusion operator at the reference's defaults: no
under-relaxation (underU = 1, src/Main.f:219) and no transient term, so
a_P = sum a_nb + sum F_f (largest cell Peclet number 100), at 256x256 cells:
the largest size on which Jacobi-BiCGSTAB converges on it (at 512^2 and
1024^2 it stalls above residual 1e-3 or breaks down), and on which the
sequential reference's GMRES(30) (~4,100 restarts) finishes in minutes.
Against the reference's solutions at the ladder tolerance
(tests/golden/ref_steady.npz): iterations, energy and solution statistics.
This system is far less well conditioned than the momentum form: the
reference's own BiCGSTAB and GMRES solutions, both at residual <= 1e-10,
differ by 1.2e-8 in ||x||, so solutions are compared at 1e-7."""
import os

import numpy as np
import pytest

from fortranfvcell_amd import problems as P

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_steady.npz")


@pytest.fixture(scope="module")
def steady():
    import fortranfvcell_amd as F
    s = P.config_advdiff(256, dt=None, under_relax=1.0)
    return s, P.scipy_csr(s), F.DeviceSystem(P.to_ell(s)), np.load(FIX, allow_pickle=False)


def _check(name, ref, s, A, x, st, it_band):
    r = np.linalg.norm(s.b - A @ x) / np.linalg.norm(s.b)
    it_ref = int(ref[name + "_iterations"])
    print(f"{name}: gpu {st['iterations']} it ({st['inner_iterations']}), reference {it_ref}, relres {r:.2e}, "
          f"{st['seconds'] * 1e3:.0f} ms", flush=True)
    assert st["status"] == 0 and r <= 1e-10
    if it_band is not None:
        assert abs(st["iterations"] - it_ref) <= it_band * it_ref, (st["iterations"], it_ref)
    xsq = float(ref[name + "_xsq"])
    assert abs((x * x).sum() - xsq) <= 1e-7 * xsq
    idx = ref[name + "_idx"]
    assert np.max(np.abs(x[idx] - ref[name + "_xsample"])) <= 1e-6 * np.sqrt(xsq / s.neq)


def test_steady_bicgstab(fvc, steady):
    s, A, ds, ref = steady
    name = "steady256_ell_bicgstab"
    st = ds.solve(fvc.BICGSTAB, tol=float(ref[name + "_tol"]), maxit=100000, x0_rule=fvc.X0_ZERO)
    _check(name, ref, s, A, ds.x.get(), st, 0.03)
    assert abs(st["energy"] - float(ref[name + "_energy"])) <= 1e-7 * abs(float(ref[name + "_energy"]))
    st = ds.solve(fvc.BICGSTAB, tol=1e-10, maxit=100000, x0_rule=fvc.X0_ZERO, stop_rule=fvc.STOP_RELRES)
    _check(name, ref, s, A, ds.x.get(), st, None)


def test_steady_gmres30(fvc, steady):
    s, A, ds, ref = steady
    name = "steady256_ell_gmres30"
    st = ds.solve(fvc.GMRES, tol=float(ref[name + "_tol"]), gmres_k=30, maxit=20000)
    _check(name, ref, s, A, ds.x.get(), st, 0.03)
    st = ds.solve(fvc.GMRES, tol=1e-10, gmres_k=30, maxit=20000, stop_rule=fvc.STOP_RELRES,
                  gmres_orth=fvc.ORTH_CGS2)
    _check(name, ref, s, A, ds.x.get(), st, None)
