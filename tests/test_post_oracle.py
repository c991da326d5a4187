"""The oracle of the post-solve step (oracle/post.py), pinned: bit-identical
to the reference's own gform and simpleUpdate (compiled from its sources
into oracle/_ref/libmvf_dropin.so) on jittered quad and triangle meshes
with every boundary-condition branch, and exact, as Green-Gauss over the
neighbour centres must be, on linear fields."""
import numpy as np
import pytest

from fortranfvcell_amd import problems as P
from oracle import post

MESHES = {"quad": lambda: P.quad_mesh(9, jitter=0.15, seed=11),
          "tri": lambda: P.tri_mesh(8, jitter=0.15, seed=12)}

needs_ref = pytest.mark.skipif(not post.have_dropin(),
                               reason="oracle/_ref/libmvf_dropin.so is built only where the reference sources are")


def _fields(m, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=m.neq), rng.normal(size=(m.neq, 2))


@needs_ref
@pytest.mark.parametrize("kind", sorted(MESHES))
@pytest.mark.parametrize("wall_bc", [(3, 6, 1, 3), (6, 1, 3, 1)])
def test_gform_matches_the_reference_bit_for_bit(kind, wall_bc):
    m = MESHES[kind]()
    a = P.ref_mesh_arrays(m, wall_bc)
    u, g = _fields(m, 5)
    go, fo = post.gform_p(a["x"], a["ix"], a["nelcon"], a["pedge"], u, g)
    gr, fr = post.ref_gform_p(a["x"], a["ix"], a["nelcon"], a["pedge"], u, g, a["cell_type"])
    assert np.array_equal(go, gr)
    assert np.array_equal(fo, fr)
    assert np.all(np.isfinite(go)) and np.abs(go).max() > 0


@needs_ref
def test_simple_update_matches_the_reference_bit_for_bit():
    rng = np.random.default_rng(7)
    n = 333
    w, d, gp = rng.normal(size=(n, 2)), rng.uniform(0.1, 2.0, size=(n, 2)), rng.normal(size=(n, 2))
    pre, u1, u2, pc = (rng.normal(size=n) for _ in range(4))
    for simple_r in (False, True):
        wo, po = post.simple_update(w, pre, u1, u2, pc, gp, d, 0.3, simple_r)
        wr, pr = post.ref_simple_update(w, pre, u1, u2, pc, gp, d, 0.7, 0.3, simple_r)
        assert np.array_equal(wo, wr) and np.array_equal(po, pr)
        assert np.array_equal(po, pre) == simple_r


@pytest.mark.parametrize("kind", sorted(MESHES))
def test_gform_is_exact_on_linear_fields(kind):
    # the trapezoid rule on each side is exact for a linear field, so the
    # Green-Gauss gradient over the neighbour centres is too, wherever the
    # neighbours are real cells
    m = MESHES[kind]()
    a = P.ref_mesh_arrays(m)
    c = post._centroid(a["x"][a["ix"][:, :m.w] - 1], m.w)
    u = 0.7 * c[:, 0] - 1.3 * c[:, 1] + 2.0
    g, _ = post.gform_p(a["x"], a["ix"], a["nelcon"], a["pedge"], u, np.zeros((m.neq, 2)))
    inner = np.all(a["nelcon"] > 0, axis=1)
    assert inner.sum() > m.neq // 3
    assert np.allclose(g[inner], [0.7, -1.3], rtol=0, atol=1e-12)


def test_gform_boundary_branches():
    # a wall cell's ghost value: 0 (prescribed pressure correction), its own
    # value (zero gradient) or the old gradient extrapolated; changing the
    # old gradient moves only cells on extrapolating walls
    m = P.quad_mesh(6, jitter=0.1, seed=2)
    a = P.ref_mesh_arrays(m, (3, 6, 1, 3))
    u, g = _fields(m, 9)
    g1, _ = post.gform_p(a["x"], a["ix"], a["nelcon"], a["pedge"], u, g)
    g2, _ = post.gform_p(a["x"], a["ix"], a["nelcon"], a["pedge"], u, g + 1.0)
    moved = np.any(g1 != g2, axis=1)
    on_extrap = np.any(a["pedge"][:, :4] == 1, axis=1)
    assert np.array_equal(moved, on_extrap) and on_extrap.any()
