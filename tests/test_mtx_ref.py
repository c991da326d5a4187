"""The MatrixMarket writer (fortranfvcell_amd.mtx) pinned to the reference's
own export (exportCoo -> csrToCoo -> writeCoo, src/ExportCoo.f:23-62,
src/CsrToCoo.f:29-179, compiled from its sources into
oracle/_ref/libmvf_dropin.so): byte for byte, matrix and right-hand side,
for the storages the reference exports (CSRD general and symmetric, CSRC
symmetric), on systems with negative, tiny and 1e60 (pinned cell) values.
For a nonsymmetric CSRC matrix the reference writes only the diagonal and the
stored lower pattern (its col > row branch never fires there); the
project's writer adds the upper entries, and its other lines are the
reference's."""
import os

import numpy as np
import pytest

from fortranfvcell_amd import mtx
from fortranfvcell_amd import problems as P
from oracle import export

pytestmark = pytest.mark.skipif(not export.have_dropin(),
                                reason="oracle/_ref/libmvf_dropin.so is built only where the reference sources are")

SYSTEMS = {
    "advdiff": lambda: P.config_advdiff(12),
    "pressure_pinned": lambda: P.config_poisson(10),
    "lognormal_tri": lambda: P.config_pressure_lognormal(7, triangles=True),
}


def _read(p):
    with open(p) as fh:
        return fh.read()


@pytest.mark.parametrize("name", sorted(SYSTEMS))
@pytest.mark.parametrize("layout", ["csrd", "csrd_sym", "csrc_sym"])
def test_writer_matches_the_reference_byte_for_byte(tmp_path, name, layout):
    s = SYSTEMS[name]()
    if layout == "csrd":
        f, code, unsym = P.to_csrd(s), 1, True
    elif layout == "csrd_sym":
        f, code, unsym = P.to_csrd(s, sym=True), 1, False
    else:
        f, code, unsym = P.to_csrc(s, unsym=False), 3, False
    ra, rb = export.ref_export_coo(f, code, unsym, str(tmp_path))
    oa, ob = mtx.write_system(f, str(tmp_path / "ours"), unsym=unsym)
    assert _read(oa) == _read(ra)
    assert _read(ob) == _read(rb)


def test_csrc_unsymmetric_adds_the_upper_triangle(tmp_path):
    s = P.config_advdiff(9)
    f = P.to_csrc(s, unsym=True)
    ra, _ = export.ref_export_coo(f, 3, True, str(tmp_path))
    oa, _ = mtx.write_system(f, str(tmp_path / "ours"), unsym=True)
    ref = _read(ra).splitlines()
    ours = _read(oa).splitlines()
    assert ref[0] == ours[0]
    assert int(ref[1].split()[2]) == s.neq + f["nad"]
    assert int(ours[1].split()[2]) == s.neq + 2 * f["nad"]
    lower = [ln for ln in ours[2:] if int(ln.split()[1]) <= int(ln.split()[0])]
    assert lower == ref[2:]


def test_reference_export_reads_back_to_the_system(tmp_path):
    """The reference's file, read by mtx.read_system, is the same operator to
    E24.16's 16 significant digits, and b likewise."""
    s = P.config_advdiff(11)
    f = P.to_csrd(s)
    ra, _ = export.ref_export_coo(f, 1, True, str(tmp_path))
    g = mtx.read_system(ra[:-len(".mtx")])
    A = P.scipy_csr(s)
    B = P.scipy_csr(P.System(g["neq"], 0, g["ad"], np.zeros((g["neq"], 0)), np.zeros((g["neq"], 0), int),
                             g["b"]))
    import scipy.sparse as sp
    ia, ja = g["iax"].astype(np.int64) - 1, g["ja"].astype(np.int64) - 1
    rows = np.repeat(np.arange(g["neq"]), np.diff(ia))
    B = B + sp.csr_matrix((g["al"], (rows, ja)), shape=A.shape)
    assert abs(A - B).max() <= 1e-15 * abs(A).max()
    assert np.max(np.abs(g["b"] - s.b)) <= 1e-15 * np.abs(s.b).max()
