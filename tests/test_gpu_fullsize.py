"""Full-size parity: the GPU solves the 1024x1024 systems and matches the
statistics the reference's own solver produced on them (iteration count,
energy norm, sampled entries of x; tests/golden/ref_solver_1024_stats.npz)."""
import os

import numpy as np
import pytest

from fortranfvcell_amd import problems as P

pytestmark = pytest.mark.gpu

STATS = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "ref_solver_1024_stats.npz"), allow_pickle=False)


@pytest.mark.parametrize("name,build,code", [
    ("poisson1024_ell_pcg", lambda: P.config_poisson(1024), 1),
    ("advdiff1024_ell_bicgstab", lambda: P.config_advdiff(1024), 2),
])
def test_fullsize_vs_reference_statistics(fvc, name, build, code):
    s = build()
    f = P.to_ell(s)
    print(f"full size {name}", flush=True)
    x = np.zeros(s.neq)
    st = fvc.solver(f["ad"], None, f["al"], x, np.ascontiguousarray(f["b"]), f["iax"], f["ja"], s.neq,
                    f["nad"], True, 1e-10, 50000, code, 4, 2, omp_solver=False)
    it_ref = int(STATS[name + "_iterations"])
    print(f"  gpu {st['iterations']} it, reference {it_ref} it, {st['seconds'] * 1e3:.1f} ms", flush=True)
    assert abs(st["iterations"] - it_ref) <= max(2, int(0.01 * it_ref))
    e_ref = float(STATS[name + "_energy"])
    assert abs(st["energy"] - e_ref) <= 1e-8 * abs(e_ref)
    idx = STATS[name + "_idx"]
    xs = STATS[name + "_xsample"]
    scale = np.sqrt(STATS[name + "_xsq"] / s.neq)
    assert np.max(np.abs(x[idx] - xs)) <= 1e-8 * scale * np.sqrt(s.neq) ** 0 * 10
    assert abs(np.sqrt((x * x).sum()) - np.sqrt(STATS[name + "_xsq"])) <= 1e-8 * np.sqrt(STATS[name + "_xsq"])


def test_large_bicgstab_grid_and_row_order(fvc):
    """Above 24 rows per thread at four blocks of 256 per CU (6.76M cells
    here), BiCGSTAB runs on two blocks per CU, with XCD row tiles.  The solve
    reaches the true residual bar and agrees with the four-per-CU grid and
    with one contiguous chunk of rows per block."""
    s = P.config_advdiff(2600)
    f = P.to_ell(s)
    A = P.scipy_csr(s)
    ds = fvc.DeviceSystem(f)
    opts = dict(tol=1e-10, stop_rule=fvc.STOP_RELRES, x0_rule=fvc.X0_ZERO)
    runs = []
    for extra in ({}, {"grid_blocks": 1024}, {"row_order": 1}):
        st = ds.solve(fvc.BICGSTAB, **opts, **extra)
        x = ds.x.get()
        r = np.linalg.norm(s.b - A @ x) / np.linalg.norm(s.b)
        print(f"  neq={s.neq} {extra or 'default'}: G={st['grid_blocks']} {st['iterations']} it "
              f"relres={r:.2e} {st['seconds'] * 1e3:.1f} ms", flush=True)
        assert st["status"] == 0 and r <= 1e-10, (extra, r)
        runs.append((x, st))
    cus = fvc.device_info()["compute_units"]
    per_cu = 2 if s.neq >= 24 * 4 * cus * 256 else 4
    assert runs[0][1]["grid_blocks"] == per_cu * cus and runs[0][1]["block_size"] == 256
    for x, _ in runs[1:]:
        assert np.linalg.norm(x - runs[0][0]) <= 1e-8 * np.linalg.norm(runs[0][0])


def test_large_mixed_pcg_grid(fvc):
    """From 16 rows per thread at four blocks of 256 per CU (the 2048x2048
    lognormal pressure system), mixed PCG runs on three blocks per CU, reads
    int16 column offsets and reaches the FP64 true residual bar; its
    solution agrees with FP64 PCG's to what that residual leaves free."""
    s = P.config_pressure_lognormal(2048)
    f = P.to_ell(s)
    A = P.scipy_csr(s)
    ds = fvc.DeviceSystem(f)
    opts = dict(tol=1e-10, stop_rule=fvc.STOP_RELRES, x0_rule=fvc.X0_ZERO)
    st64 = ds.solve(fvc.PCG, **opts)
    x64 = ds.x.get()
    stm = ds.solve(fvc.PCG, precision=fvc.PREC_MIXED, **opts)
    xm = ds.x.get()
    r = np.linalg.norm(s.b - A @ xm) / np.linalg.norm(s.b)
    d = np.linalg.norm(xm - x64) / np.linalg.norm(x64)
    print(f"  neq={s.neq}: fp64 {st64['iterations']} it {st64['seconds']:.2f} s, mixed {stm['iterations']} it "
          f"{stm['seconds']:.2f} s G={stm['grid_blocks']}, relres {r:.2e}, vs fp64 {d:.2e}", flush=True)
    assert st64["status"] == 0 and stm["status"] == 0
    assert stm["flags"] & fvc.FLAG_COL16 and stm["flags"] & fvc.FLAG_TILES, stm
    assert r <= 1e-10
    cus = fvc.device_info()["compute_units"]
    per_cu = 3 if s.neq >= 16 * 4 * cus * 256 else 4
    assert stm["grid_blocks"] == per_cu * cus
    assert d <= 1e-7
