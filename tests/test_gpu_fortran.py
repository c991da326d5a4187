"""The Fortran drop-in: a Fortran program calling solver(...) and getRes(...)
with the reference's interface (src/Solver.f:28,368), linked against
libfvcell_fortran.so instead of Solver.o, solves on the GPU and matches
the oracle; its unit-13 log line has the reference's format."""
import os
import re
import subprocess

import numpy as np
import pytest

from fortranfvcell_amd import problems as P

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "drive_solver.exe")


def _run(tmp_path, f, code, unsym, omp, tol=1e-10, maxit=20000, x0=None):
    neq = f["neq"]
    iax, ja = np.asarray(f["iax"], np.int32), np.asarray(f["ja"], np.int32)
    ad = np.asarray(f["ad"], np.float64)
    al = np.zeros(0) if f.get("al") is None else np.asarray(f["al"], np.float64)
    au = np.zeros(0) if f.get("au") is None else np.asarray(f["au"], np.float64)
    x0 = np.zeros(neq) if x0 is None else x0
    hdr = np.array([code, f["matriz"], int(unsym), int(omp), neq, f["nad"], maxit, 2,
                    len(iax), len(ja), len(ad), len(al), len(au)], np.int32)
    fin, fout, flog = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "solver.log"
    with open(fin, "wb") as fh:
        for a in (hdr, np.array([tol]), iax, ja, ad, al, au, np.asarray(f["b"], np.float64), x0):
            fh.write(np.ascontiguousarray(a).tobytes())
    r = subprocess.run([EXE, str(fin), str(fout), str(flog)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(fout, "rb").read()
    x = np.frombuffer(raw[: 8 * neq], np.float64)
    status, iters = np.frombuffer(raw[8 * neq: 8 * neq + 8], np.int32)
    energy = np.frombuffer(raw[8 * neq + 8: 8 * neq + 16], np.float64)[0]
    u = np.frombuffer(raw[8 * neq + 16: 16 * neq + 16], np.float64)
    timers = np.frombuffer(raw[16 * neq + 16: 16 * neq + 48], np.float64)
    _run.timers = dict(zip(("time", "matvectime", "dottime", "precontime"), timers))
    # the Fortran caller's second solve: solverKeepMatrix, then the same A with 2 b
    _run.x_kept = np.frombuffer(raw[16 * neq + 48: 24 * neq + 48], np.float64)
    return x, int(status), int(iters), energy, u, open(flog).read(), r.stdout


@pytest.mark.parametrize("code,lay,omp", [(1, "ell", True), (1, "csrd", False), (2, "ell", True),
                                          (2, "csrd", True), (3, "csrd", False)])
def test_fortran_dropin_matches_oracle(fvc, orc, tmp_path, code, lay, omp):
    s = P.config_poisson(24) if code == 1 else P.config_advdiff(24, peclet_max=60.0)
    f = P.to_ell(s) if lay == "ell" else P.to_csrd(s)
    print(f"fortran drop-in code={code} {lay} omp={omp}", flush=True)
    xo, _, so = orc.solver(f, code, omp=omp)
    x, status, iters, energy, u, logtxt, out = _run(tmp_path, f, code, True, omp)
    assert status == 0
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    assert abs(iters - so["iterations"]) <= max(2, int(0.03 * so["iterations"]))
    assert abs(energy - so["energy"]) <= 1e-8 * abs(so["energy"])
    # getRes with num(i) = neq+1-i: u(i) = x(num(i))
    assert np.array_equal(u, x[::-1])
    # solverKeepMatrix + solve with 2 b on the kept A: exactly 2 x
    assert np.array_equal(_run.x_kept, 2.0 * x)
    # COMMON /timectrl/: time = this solve's time, added to matvectime
    tm = _run.timers
    assert tm["time"] > 0 and tm["matvectime"] == tm["time"]
    assert tm["dottime"] == 0 and tm["precontime"] == 0
    label = {(1, True): "PCG OMP      : ", (1, False): "PCG:       ", (2, True): "PBICGSTAB OMP: ",
             (2, False): "PBICGSTAB: ", (3, False): "GMRES: "}[(code, omp)]
    # write(13,'(a,a,i9,a,d20.10,a,d20.10)') label,"it",j," energy norm ",energy," tol ",tol
    # (d20.10 right-justifies the value in 20 columns)
    pat = re.escape(label) + r"it[ 0-9]{9} energy norm +-?0\.\d{10}D[+-]\d\d tol +0\.1000000000D-09"
    assert re.fullmatch(pat, logtxt.rstrip("\n")), logtxt


def test_fortran_dropin_stops_like_the_reference(fvc, tmp_path):
    s = P.config_poisson(16)
    f = P.to_ell(s)
    neq = f["neq"]
    iax, ja, ad, al = f["iax"], f["ja"], f["ad"], f["al"]
    hdr = np.array([1, 4, 1, 1, neq, f["nad"], 3, 2, len(iax), len(ja), len(ad), len(al), 0], np.int32)
    fin, fout, flog = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "log"
    with open(fin, "wb") as fh:
        for a in (hdr, np.array([1e-10]), iax, ja, ad, al, s.b, np.zeros(neq)):
            fh.write(np.ascontiguousarray(a).tobytes())
    r = subprocess.run([EXE, str(fin), str(fout), str(flog)], capture_output=True, text=True, timeout=120)
    # maxit = 3: the reference prints its warning and STOPs (src/Iterativos.f:76-77)
    assert "no convergence reached after" in r.stdout
    assert not os.path.exists(fout)
