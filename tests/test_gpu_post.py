"""The post-solve step on the GPU (fvc_gform_dev, fvc_simple_update_dev)
against the oracle (oracle/post.py, itself bit-identical to the reference's
gform and simpleUpdate): bit for bit, on quads and triangles, every
wall-condition branch, up to the 1M-cell mesh; in place; and an index out of
range gives NaN for its cell instead of a read out of bounds."""
import numpy as np
import pytest

from fortranfvcell_amd import problems as P
from oracle import post as O

pytestmark = pytest.mark.gpu


def _mesh(kind, n):
    return P.quad_mesh(n, jitter=0.15, seed=21) if kind == "quad" else P.tri_mesh(n, jitter=0.15, seed=22)


@pytest.mark.parametrize("kind,n", [("quad", 17), ("tri", 13), ("quad", 1024)])
@pytest.mark.parametrize("wall_bc", [(3, 6, 1, 3), (1, 1, 6, 3)])
def test_gform_bit_identical_to_oracle(fvc, kind, n, wall_bc):
    from fortranfvcell_amd.post import DeviceMesh
    m = _mesh(kind, n)
    a = P.ref_mesh_arrays(m, wall_bc)
    rng = np.random.default_rng(n)
    u = rng.normal(size=m.neq)
    g = rng.normal(size=(m.neq, 2))
    print(f"gform {kind} n={n} cells={m.neq} walls={wall_bc}", flush=True)
    want, wantf = O.gform_p(a["x"], a["ix"], a["nelcon"], a["pedge"], u, g)
    dm = DeviceMesh(a["x"], a["ix"], a["nelcon"], a["pedge"])
    du, dg, df = fvc.DeviceArray.from_host(u), fvc.DeviceArray.from_host(g), fvc.DeviceArray.empty(m.neq)
    dm.gform(du, dg, df)
    fvc.load().fvc_synchronize()
    assert np.array_equal(dg.get(), want)
    assert np.array_equal(df.get(), wantf)


def test_simple_update_bit_identical_to_oracle(fvc):
    from fortranfvcell_amd.post import simple_update
    rng = np.random.default_rng(3)
    n = 100_003
    w, d, gp = rng.normal(size=(n, 2)), rng.uniform(0.1, 2.0, size=(n, 2)), rng.normal(size=(n, 2))
    pre, u1, u2, pc = (rng.normal(size=n) for _ in range(4))
    for simple_r in (False, True):
        want_w, want_p = O.simple_update(w, pre, u1, u2, pc, gp, d, 0.3, simple_r)
        dw, dpre = fvc.DeviceArray.from_host(w), fvc.DeviceArray.from_host(pre)
        args = [fvc.DeviceArray.from_host(v) for v in (u1, u2, pc, gp, d)]
        simple_update(n, dw, dpre, *args, 0.7, 0.3, simple_r)
        fvc.load().fvc_synchronize()
        assert np.array_equal(dw.get(), want_w)
        assert np.array_equal(dpre.get(), want_p)


def test_gform_index_out_of_range_gives_nan(fvc):
    # DeviceMesh refuses such arrays on the host; the C ABI checks on the
    # device: every index is clamped before its load and the cell reads NaN
    m = _mesh("quad", 8)
    a = P.ref_mesh_arrays(m)
    nelcon = a["nelcon"].copy()
    ix = a["ix"].copy()
    nelcon[5, 1] = m.neq + 7
    ix[9, 2] = len(a["x"]) + 3
    u, g = np.ones(m.neq), np.zeros((m.neq, 2))
    d = {k: fvc.DeviceArray.from_host(v) for k, v in
         dict(x=a["x"], ix=ix, nelcon=nelcon, pedge=a["pedge"], u=u, g=g).items()}
    f = fvc.DeviceArray.empty(m.neq)
    L = fvc.load()
    from fortranfvcell_amd import _ptr
    fvc.check(L.fvc_gform_dev(m.neq, len(a["x"]), 4, _ptr(d["x"]), _ptr(d["ix"]), _ptr(d["nelcon"]),
                              _ptr(d["pedge"]), _ptr(d["u"]), _ptr(d["g"]), _ptr(f), None))
    L.fvc_synchronize()
    gg, ff = d["g"].get(), f.get()
    bad = np.zeros(m.neq, bool)
    bad[5] = True
    bad |= np.any(nelcon == 10, axis=1) | (np.arange(m.neq) == 9)   # cells that gather cell 10's nodes
    assert np.all(np.isnan(gg[5])) and np.all(np.isnan(gg[9])) and np.isnan(ff[5])
    assert np.all(np.isfinite(gg[~bad])) and np.all(ff[~bad] == 1.0)
