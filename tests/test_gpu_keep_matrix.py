"""fvc_keep_matrix: the host-array drop-in keeps A on the device between the
two pressure-correction solves of a SIMPLE step (src/Simple.f:590,621; pform
rebuilds only the right-hand side in between, src/Simple.f:607-615).

The kept solve must be bit-identical to a fresh one; the flag must be
one-shot; a different matrix must be uploaded even under the flag; and the
flag must really skip the upload (shown by changing A's contents in place,
which the contract forbids: the kept solve then still sees the old matrix)."""
import numpy as np
import pytest

from fortranfvcell_amd import problems as P

pytestmark = pytest.mark.gpu


def _arrays(fmt):
    f64 = lambda a: np.ascontiguousarray(np.array(a, dtype=np.float64))
    i32 = lambda a: np.ascontiguousarray(np.array(a, dtype=np.int32))
    al = f64(fmt["al"])
    return dict(ad=f64(fmt["ad"]), al=al, au=al if fmt.get("au") is None else f64(fmt["au"]),
                iax=i32(fmt["iax"]), ja=i32(fmt["ja"]), neq=fmt["neq"], nad=fmt["nad"],
                matriz=fmt["matriz"])


def _solve(fvc, a, b, code):
    x = np.zeros(a["neq"])
    st = fvc.solver(a["ad"], a["au"], a["al"], x, b, a["iax"], a["ja"], a["neq"], a["nad"],
                    code == fvc.BICGSTAB, 1e-10, 20000, code, a["matriz"])
    assert st["status"] == 0
    return x


@pytest.mark.parametrize("layout", ["ell", "csrd"])
def test_kept_solve_is_bit_identical_to_a_fresh_one(fvc, layout):
    s = P.config_pressure_lognormal(48)
    a = _arrays(P.to_ell(s) if layout == "ell" else P.to_csrd(s))
    b1 = np.array(s.b)
    b2 = np.random.default_rng(5).normal(size=s.neq)
    fresh2 = _solve(fvc, a, b2, fvc.PCG)
    _solve(fvc, a, b1, fvc.PCG)
    fvc.keep_matrix()
    kept2 = _solve(fvc, a, b2, fvc.PCG)
    assert np.array_equal(kept2, fresh2)


def test_keep_skips_the_upload_and_is_one_shot(fvc):
    s = P.config_pressure_lognormal(40)
    a = _arrays(P.to_ell(s))
    b = np.array(s.b)
    x_old = _solve(fvc, a, b, fvc.PCG)
    # change A in place (same addresses): a kept solve still sees the old A
    a["ad"] *= 1.5
    fvc.keep_matrix()
    assert np.array_equal(_solve(fvc, a, b, fvc.PCG), x_old)
    # one-shot: the next call uploads the changed A
    x_new = _solve(fvc, a, b, fvc.PCG)
    assert not np.array_equal(x_new, x_old)
    a2 = _arrays(P.to_ell(s))
    a2["ad"] *= 1.5
    assert np.array_equal(x_new, _solve(fvc, a2, b, fvc.PCG))


def test_keep_uploads_a_different_matrix(fvc):
    s1 = P.config_pressure_lognormal(40)
    s2 = P.config_advdiff(44, peclet_max=30.0)
    a1, a2 = _arrays(P.to_ell(s1)), _arrays(P.to_ell(s2))
    fresh = _solve(fvc, a2, np.array(s2.b), fvc.BICGSTAB)
    _solve(fvc, a1, np.array(s1.b), fvc.PCG)
    fvc.keep_matrix()
    assert np.array_equal(_solve(fvc, a2, np.array(s2.b), fvc.BICGSTAB), fresh)


def test_host_buffers_grow_and_shrink_across_systems(fvc):
    # the host path keeps grow-only device buffers: a large, a small and a
    # large system in turn each match their device-resident solve
    for n in (96, 24, 112, 24):
        s = P.config_pressure_lognormal(n)
        f = P.to_ell(s)
        a = _arrays(f)
        x = _solve(fvc, a, np.array(s.b), fvc.PCG)
        ds = fvc.DeviceSystem(f)
        ds.solve(fvc.PCG, tol=1e-10, maxit=20000, x0_rule=fvc.X0_ZERO)
        assert np.array_equal(x, ds.x.get()), n
