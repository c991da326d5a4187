"""The C ABI library loads and exports every symbol include/fvcell.h
declares; the Fortran drop-in exports the reference's solver/getRes.
CPU only: nothing here calls into the GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fvcell.h")
LIB = os.path.join(ROOT, "fortranfvcell_amd", "libfvcell.so")
FLIB = os.path.join(ROOT, "fortranfvcell_amd", "libfvcell_fortran.a")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fvc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_drop_in():
    names = declared()
    for must in ("fvc_solver", "fvc_solve", "fvc_solve_dev", "fvc_matvec", "fvc_dot", "fvc_getres"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from fortranfvcell_amd import _lib
    assert sorted(_lib.EXPORTS) == declared()


def test_fortran_dropin_exports_solver_and_getres():
    out = subprocess.run(["nm", "--defined-only", FLIB], capture_output=True, text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if " T " in line}
    # gfortran/flang external names of src/Solver.f:28 and :368
    assert "solver_" in syms and "getres_" in syms


def _header_layout(tmp_path):
    """sizeof and field offsets of fvc_options and fvc_stats as a C compiler
    lays out include/fvcell.h."""
    from fortranfvcell_amd import _lib
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {',
             'printf("options %zu\\n", sizeof(fvc_options));', 'printf("stats %zu\\n", sizeof(fvc_stats));']
    for struct, cls in (("fvc_options", _lib.Options), ("fvc_stats", _lib.Stats)):
        for name, _ in cls._fields_:
            lines.append(f'printf("{struct}.{name} %zu\\n", offsetof({struct}, {name}));')
    lines += ['return 0;', '}']
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return dict(line.rsplit(" ", 1) for line in out.splitlines())


def test_options_and_stats_layout_match_header(tmp_path):
    from fortranfvcell_amd import _lib
    lay = _header_layout(tmp_path)
    assert C.sizeof(_lib.Options) == int(lay["options"])
    assert C.sizeof(_lib.Stats) == int(lay["stats"])
    for struct, cls in (("fvc_options", _lib.Options), ("fvc_stats", _lib.Stats)):
        for name, _ in cls._fields_:
            assert getattr(cls, name).offset == int(lay[f"{struct}.{name}"]), (struct, name)
    o = _lib.default_options()
    assert o.x0_rule == _lib.X0_AS_OPENMP and o.gmres_k == 15 and o.engine == _lib.ENGINE_AUTO
    assert o.row_order == 0
    assert _lib.default_options(row_order=2).row_order == 2
    # a misspelt option is an error, not a silently ignored attribute
    with pytest.raises(TypeError, match="row_ordr"):
        _lib.default_options(row_ordr=2)
