"""The reference program's own objects linked against the drop-in: every
source of the reference Makefile except src/Solver.f (replaced) and
src/Main.f (not compilable with amdflang: it calls the GNU intrinsic iargc)
links with no undefined symbol, and the callers of the solver bind to the
drop-in's solver_/getres_."""
import os
import shutil
import subprocess

import pytest

import oracle

REF = "/root/reference/src"
pytestmark = pytest.mark.skipif(not (os.path.isdir(REF) and shutil.which("amdflang") and shutil.which("nm")),
                                reason="needs the reference sources and amdflang (build container only)")


def _nm(path, *flags):
    return subprocess.run(["nm", *flags, path], capture_output=True, text=True, check=True).stdout


def test_reference_objects_link_against_dropin():
    pkg = os.path.join(os.path.dirname(oracle.HERE), "fortranfvcell_amd")
    if not os.path.exists(os.path.join(pkg, "libfvcell_fortran.a")):
        pytest.skip("drop-in not built")
    subprocess.run(["make", "-s", "-C", oracle.HERE, "dropin"], check=True)
    so = oracle.DROPIN_PATH
    assert os.path.exists(so)
    dyn = _nm(so, "-D")
    defined = {l.split()[-1] for l in dyn.splitlines() if " T " in l}
    undefined = {l.split()[-1] for l in dyn.splitlines() if " U " in l}
    assert {"solver_", "getres_"} <= defined
    assert "fvc_solver" in undefined and "fvc_getres" in undefined     # bound to libfvcell.so
    prog = os.path.join(oracle.HERE, "_ref", "prog")
    # the call sites (src/Simple.f:488-694, src/Transporte.f:111) reference solver_
    for obj in ("Simple.o", "Transporte.o"):
        syms = _nm(os.path.join(prog, obj))
        assert " U solver_" in syms, obj
    assert " U getres_" in _nm(os.path.join(prog, "Simple.o"))
    # and no object of the reference defines it: the drop-in's archive does
    for obj in os.listdir(prog):
        if obj.endswith(".o"):
            assert " T solver_" not in _nm(os.path.join(prog, obj)), obj
