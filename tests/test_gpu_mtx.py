"""Systems exported by the reference's own writer (exportCoo,
src/ExportCoo.f:23-62, compiled into oracle/_ref/libmvf_dropin.so), read
back with fortranfvcell_amd.mtx and solved on the GPU: SpMV against the
oracle on the original arrays (the file keeps 16 significant digits, so to
1e-15), solutions against a direct solve of the original system at the
parity bar."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle
from fortranfvcell_amd import mtx
from fortranfvcell_amd import problems as P
from oracle import export

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not export.have_dropin(),
                                 reason="oracle/_ref/libmvf_dropin.so is built only where the reference sources are")]


@pytest.mark.parametrize("name", ["advdiff48_bicgstab", "lognormal40_pcg", "advdiff40_gmres"])
def test_gpu_on_reference_exported_systems(fvc, tmp_path, name):
    s, code, kw = {"advdiff48_bicgstab": (P.config_advdiff(48), fvc.BICGSTAB, {}),
                   "lognormal40_pcg": (P.config_pressure_lognormal(40), fvc.PCG, {}),
                   "advdiff40_gmres": (P.config_advdiff(40, seed=P.SEED_BASE + 12), fvc.GMRES,
                                       dict(gmres_k=30))}[name]
    f = P.to_csrd(s)
    path, _ = export.ref_export_coo(f, 1, True, str(tmp_path))
    g = mtx.read_system(path[:-len(".mtx")])
    assert g["neq"] == s.neq and g["nad"] == f["nad"]
    x = np.linspace(-1.0, 1.0, s.neq)
    y = fvc.matvec(g, x)
    yo = oracle.matvec(f, x)
    assert np.max(np.abs(y - yo)) <= 1e-15 * np.abs(yo).max() * 8
    xd = sla.spsolve(P.scipy_csr(s).tocsc(), s.b)
    xg, st = fvc.solve(g, code, tol=1e-10, stop_rule=fvc.STOP_RELRES, x0_rule=fvc.X0_ZERO, **kw)
    err = np.linalg.norm(xg - xd) / np.linalg.norm(xd)
    print(f"{name}: {st['iterations']} it, relres {st['true_relres']:.2e}, |x - x_direct|/|x_direct| {err:.1e}",
          flush=True)
    assert st["status"] == 0 and err <= 1e-8
