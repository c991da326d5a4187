"""fvc_solve_dev_noalloc, the solve whose scratch lives in the library's
fixed per-device workspace instead of device memory it allocates: called
through the C ABI, two calls in a row with the second on a smaller system,
for each method (and mixed precision), against fvc_solve_dev on the same
systems to the last bit (same kernels, same workspace layout; only its base
address differs).  A solve too large for the workspace, or a storage format
other than ELLPACK, is refused with FVC_UNSUPPORTED and a message, and the
workspace serves the next call as before."""
import numpy as np
import pytest

from fortranfvcell_amd import problems as P

pytestmark = pytest.mark.gpu


def _cases(fvc):
    return [
        ("pcg", fvc.PCG, P.config_pressure_lognormal(384), P.config_pressure_lognormal(200), {}),
        ("bicgstab", fvc.BICGSTAB, P.config_advdiff(320), P.config_advdiff(150), {}),
        ("gmres10", fvc.GMRES, P.config_advdiff(160), P.config_advdiff(100), dict(gmres_k=10)),
        ("gmres10_lagged_cgs2", fvc.GMRES, P.config_advdiff(160), P.config_advdiff(100),
         dict(gmres_k=10, gmres_orth=fvc.ORTH_CGS2_LAG)),
        ("gmres10_lagged_cgs2_mixed", fvc.GMRES, P.config_advdiff(160), P.config_advdiff(100),
         dict(gmres_k=10, gmres_orth=fvc.ORTH_CGS2_LAG, precision=fvc.PREC_MIXED)),
        ("pcg_mixed", fvc.PCG, P.config_pressure_lognormal(300), P.config_pressure_lognormal(120),
         dict(precision=fvc.PREC_MIXED)),
    ]


def test_noalloc_matches_solve_dev(fvc):
    for name, code, big, small, kw in _cases(fvc):
        systems = [fvc.DeviceSystem(P.to_ell(s)) for s in (big, small)]
        opts = dict(tol=1e-10, stop_rule=fvc.STOP_RELRES, x0_rule=fvc.X0_ZERO, **kw)
        print(f"{name}: neq {big.neq} then {small.neq}", flush=True)
        # two calls in a row on the fixed workspace, the second smaller
        got = []
        for ds in systems:
            st = ds.solve(code, noalloc=True, **opts)
            got.append((ds.x.get(), st))
        for ds, (x, st) in zip(systems, got):
            st_ref = ds.solve(code, **opts)
            x_ref = ds.x.get()
            print(f"  neq {ds.neq}: {st['iterations']} it, relres {st['true_relres']:.2e}", flush=True)
            assert st["status"] == 0 and st_ref["status"] == 0, (st, st_ref)
            assert np.array_equal(x, x_ref)
            for key in ("iterations", "inner_iterations", "energy", "true_relres", "grid_blocks", "flags"):
                assert st[key] == st_ref[key], (key, st[key], st_ref[key])


def test_noalloc_refusals(fvc):
    s = P.config_advdiff(1024)   # ~9 vectors of 8 MiB: beyond the 32 MiB workspace
    ds = fvc.DeviceSystem(P.to_ell(s))
    with pytest.raises(fvc.FvcError) as e:
        ds.solve(fvc.BICGSTAB, noalloc=True)
    assert e.value.code == 4 and "workspace" in str(e.value)
    sc = P.config_advdiff(64)
    dc = fvc.DeviceSystem(P.to_csrd(sc))
    with pytest.raises(fvc.FvcError) as e:
        dc.solve(fvc.BICGSTAB, noalloc=True)
    assert e.value.code == 4 and "ELLPACK" in str(e.value)
    # the workspace still serves the next call
    de = fvc.DeviceSystem(P.to_ell(sc))
    opts = dict(tol=1e-10, stop_rule=fvc.STOP_RELRES, x0_rule=fvc.X0_ZERO)
    st = de.solve(fvc.BICGSTAB, noalloc=True, **opts)
    x = de.x.get()
    assert st["status"] == 0
    st_ref = de.solve(fvc.BICGSTAB, **opts)
    assert np.array_equal(x, de.x.get()) and st["iterations"] == st_ref["iterations"]
