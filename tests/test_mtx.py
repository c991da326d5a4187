"""MatrixMarket interchange in the reference's export layout
(src/ExportCoo.f, src/CsrToCoo.f:135-179): Fortran E24.16 values, (3i9) and
(2i9) headers, the diagonal first in each row; exact round trips."""
import numpy as np
import scipy.sparse as sp
import pytest

import oracle
from fortranfvcell_amd import mtx
from fortranfvcell_amd import problems as P


@pytest.mark.parametrize("x,text", [
    (4.0, "  0.4000000000000000E+01"),
    (-1.5e-3, " -0.1500000000000000E-02"),
    (1e60, "  0.9999999999999999E+60"),       # the pinning penalty: 1e60 is not a double
    (0.0, "  0.0000000000000000E+00"),
    (1.0 / 3.0, "  0.3333333333333333E+00"),
    (2.5e-120, "  0.2500000000000000-119"),    # Ew.d drops the E past two digits
])
def test_fortran_e24_16(x, text):
    assert mtx.fortran_e(x) == text
    assert len(text) == 24


def test_file_layout_matches_the_reference(tmp_path):
    s = P.config_poisson(2)
    f = P.to_csrd(s)
    a, b = mtx.write_system(f, str(tmp_path / "p"))
    lines = open(a).read().splitlines()
    assert lines[0] == "%%MatrixMarket matrix coordinate real general"
    assert lines[1] == f"{4:9d}{4:9d}{4 + f['nad']:9d}"
    # row 1: the diagonal first, then its off-diagonals in storage order
    assert lines[2] == f"{1:9d}{1:9d}" + mtx.fortran_e(f["ad"][0])
    assert lines[3].startswith(f"{1:9d}{int(f['ja'][0]):9d}")
    blines = open(b).read().splitlines()
    assert blines[0] == "%%MatrixMarket matrix array real general" and blines[1] == f"{4:9d}{1:9d}"
    assert len(blines) == 2 + 4


@pytest.mark.parametrize("lay", ["ell", "csrd", "csrc"])
def test_exact_round_trip(tmp_path, lay):
    s = P.config_advdiff(12, peclet_max=40.0)
    f = {"ell": P.to_ell, "csrd": P.to_csrd, "csrc": lambda q: P.to_csrc(q, unsym=True)}[lay](s)
    mtx.write_system(f, str(tmp_path / "s"), exact=True)
    g = mtx.read_system(str(tmp_path / "s"))
    want = P.to_csrd(s)
    for key in ("ad", "iax", "ja", "al", "b"):
        assert np.array_equal(np.asarray(g[key]), np.asarray(want[key])), key


def test_symmetric_header_mirrors_the_lower_triangle(tmp_path):
    s = P.config_pressure_lognormal(8)
    f = P.to_csrd(s, sym=True)
    a, _ = mtx.write_system(f, str(tmp_path / "sym"), exact=True)
    assert open(a).readline().strip().endswith("symmetric")
    g = mtx.read_system(str(tmp_path / "sym"))
    want = P.to_csrd(s)
    for key in ("ad", "iax", "ja"):
        assert np.array_equal(np.asarray(g[key]), np.asarray(want[key])), key
    # the file holds the lower triangle; the reader mirrors it, so the read
    # matrix is exactly symmetric and its lower entries are the stored ones
    G = sp.csr_matrix((g["al"], g["ja"] - 1, g["iax"] - 1), shape=(g["neq"],) * 2)
    W = sp.csr_matrix((want["al"], want["ja"] - 1, want["iax"] - 1), shape=(g["neq"],) * 2)
    assert (G != G.T).nnz == 0
    assert (sp.tril(G) != sp.tril(W)).nnz == 0
    assert abs(G - W).max() <= 1e-15 * abs(W).max()        # the original is symmetric to rounding


def test_reference_format_solves_like_the_original(tmp_path, orc):
    # 16 significant digits: the read system differs in the last bit of some
    # values; the reference's solver (the oracle) gives the same solution to 1e-12
    s = P.config_advdiff(16, peclet_max=60.0)
    f = P.to_csrd(s)
    mtx.write_system(f, str(tmp_path / "r"))
    g = mtx.read_system(str(tmp_path / "r"))
    assert np.allclose(g["al"], f["al"], rtol=1e-15, atol=0)
    x0, _, _ = orc.krylov("bicgstab", f, tol=1e-14)
    x1, _, _ = orc.krylov("bicgstab", g, tol=1e-14)
    assert np.linalg.norm(x1 - x0) <= 1e-12 * np.linalg.norm(x0)
