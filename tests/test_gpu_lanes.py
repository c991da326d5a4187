"""The lane neighbour exchange of the fused PCG and BiCGSTAB phases: the SpMV
operand of columns i-1 and i+1 comes from the adjacent lanes of the
wavefront (DPP wave shifts) instead of a gather of every vector it is formed
from.  It is the same expression on the same values, so a solve must be
bit-identical with and without it: iterates, iteration count and energy.
Sizes leave partial wavefronts at the ends of the row ranges; both row
orders and both engines are covered, on quads (ELLPACK w=4, CSRD) and on
triangles (ELLPACK w=3, whose numbering puts only some neighbours at i+-1),
and against the oracle."""
import numpy as np
import pytest

import oracle
from fortranfvcell_amd import problems as P

pytestmark = pytest.mark.gpu


def _systems(method):
    if method == "pcg":
        yield "quad", P.config_pressure_lognormal(203)
        yield "tri", P.config_pressure_lognormal(141, triangles=True)
    else:
        yield "quad", P.config_advdiff(203)
        mesh = P.tri_mesh(141, jitter=0.1, seed=P.SEED_BASE + 9)
        yield "tri", P.advdiff_system(mesh, peclet_max=50.0, seed=P.SEED_BASE + 9, dt=0.005, under_relax=0.8)


@pytest.mark.parametrize("method", ["pcg", "bicgstab"])
def test_lane_exchange_bit_identical(fvc, method):
    code = fvc.PCG if method == "pcg" else fvc.BICGSTAB
    for mesh, s in _systems(method):
        layouts = [("ell", P.to_ell(s))] + ([("csrd", P.to_csrd(s))] if mesh == "quad" else [])
        for lay, f in layouts:
            for row_order in (1, 2):
                for engine in (fvc.ENGINE_KERNELS, fvc.ENGINE_PERSISTENT):
                    print(f"{method} {mesh} {lay} neq={s.neq} row_order={row_order} engine={engine}", flush=True)
                    res = []
                    for lx in (1, 2):
                        x, st = fvc.solve(f, code, tol=1e-10, x0_rule=fvc.X0_ZERO, row_order=row_order,
                                          engine=engine, lane_exchange=lx)
                        assert st["status"] == 0, st
                        res.append((x, st))
                    (x1, s1), (x2, s2) = res
                    assert np.array_equal(x1, x2)
                    assert s1["iterations"] == s2["iterations"] and s1["energy"] == s2["energy"]


@pytest.mark.parametrize("method", ["pcg", "bicgstab"])
def test_lane_exchange_vs_oracle(fvc, method):
    code = fvc.PCG if method == "pcg" else fvc.BICGSTAB
    for mesh, s in _systems(method):
        f = P.to_ell(s)
        xo, _, so = oracle.solver(f, code, omp=True, tol=1e-10)
        x, st = fvc.solve(f, code, tol=1e-10, lane_exchange=1)
        err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        print(f"{method} {mesh}: {st['iterations']} it (oracle {so['iterations']}), |x - x_o|/|x_o| {err:.1e}",
              flush=True)
        assert err <= 1e-8
        assert abs(st["iterations"] - so["iterations"]) <= max(2, int(0.03 * so["iterations"]))
