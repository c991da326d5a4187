"""Project autotuner: it times real capped solves, picks a layout and launch
shape, caches the choice per size bucket, and the tuned options give the
same solution as the defaults."""
import numpy as np
import pytest

from fortranfvcell_amd import problems as P

pytestmark = pytest.mark.gpu


def test_tune_single_system(fvc, tmp_path, monkeypatch):
    from fortranfvcell_amd import autotune
    monkeypatch.setattr(autotune, "CACHE_PATH", str(tmp_path / "tune.json"))
    s = P.config_advdiff(192)
    fmts = {"ell": P.to_ell(s), "csrd": P.to_csrd(s)}
    print(f"tune advdiff192 neq={s.neq}", flush=True)
    best = autotune.tune(fmts, fvc.BICGSTAB, iters=30, verbose=True)
    assert best["layout"] in fmts
    assert np.isfinite(best["s_per_iter"]) and best["s_per_iter"] > 0
    o = best["options"]
    assert set(o) == {"block_size", "grid_blocks", "barrier"}
    # cached: the second call returns the stored choice without timing
    assert autotune.tune(fmts, fvc.BICGSTAB, iters=30) == best
    ds_t = fvc.DeviceSystem(fmts[best["layout"]])
    st_t = ds_t.solve(fvc.BICGSTAB, tol=1e-12, x0_rule=fvc.X0_ZERO, stop_rule=fvc.STOP_RELRES, **o)
    ds_d = fvc.DeviceSystem(fmts["ell"])
    st_d = ds_d.solve(fvc.BICGSTAB, tol=1e-12, x0_rule=fvc.X0_ZERO, stop_rule=fvc.STOP_RELRES)
    assert st_t["status"] == 0 and st_d["status"] == 0
    xt, xd = ds_t.x.get(), ds_d.x.get()
    assert np.linalg.norm(xt - xd) <= 1e-8 * np.linalg.norm(xd)


def test_tune_batch(fvc):
    # layout (ELLPACK or CSRD), threads per block, allocation exponent and
    # rows per wavefront, each probe a whole batched solve; the winner
    # solves every system
    from fortranfvcell_amd import autotune
    sides = P.config_batch_sides(24, 8, 64, 3)
    systems = [P.pressure_system(P.quad_mesh(int(n)), seed=100 + i) for i, n in enumerate(sides)]
    layouts = {"ell": fvc.concat_ell([P.to_ell(s) for s in systems]),
               "csrd": fvc.concat_csrd([P.to_csrd(s) for s in systems])}
    best = autotune.tune_batch(layouts, verbose=True)
    assert best["layout"] in layouts and best["options"]["block_size"] in autotune.BLOCK_SIZES
    tried = {(q["layout"], q["options"].get("batch_rows_per_wave", 0)) for q in best["trials"]}
    assert ("csrd", best["options"].get("batch_rows_per_wave", 0)) in tried
    assert {r for l, r in tried if l == "ell"} >= set(autotune.ROWS_PER_WAVE)
    launches = {q["options"].get("batch_launch_systems", 0) for q in best["trials"] if q["layout"] == "ell"}
    assert launches >= {-(-24 // d) for d in autotune.LAUNCH_DIVISORS}
    assert best["seconds"] == min(q["seconds"] for q in best["trials"])
    x, sts = fvc.solve_batch(layouts[best["layout"]], tol=1e-10, stop_rule=fvc.STOP_RELRES, **best["options"])
    assert all(st["status"] == 0 for st in sts)
