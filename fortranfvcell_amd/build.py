"""Builds the native parts in-tree:

  libfvcell.so          hipcc, gfx950: kernels + C ABI (include/fvcell.h)
  libfvcell_fortran.a   amdflang: the ISO_C_BINDING drop-in for the
                        reference's Solver.o (subroutines solver, getRes),
                        linked into the calling program together with
                        -lfvcell

Run ``python -m fortranfvcell_amd.build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ARCH = os.environ.get("FVC_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # reference-order rounding: no FMA contraction of products into sums
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return r


# translation units: the host code and small kernels, then one per family of
# solver kernels (explicitly instantiated there, see csrc/fvc_instances.h)
HIP_UNITS = ("fvcell", "k_cg", "k_bicg", "k_gmres", "k_mixed", "k_batch")


def build_hip() -> str:
    """Compiles the units in parallel (hipcc -c each), then links libfvcell.so."""
    out = os.path.join(HERE, "libfvcell.so")
    objdir = os.path.join(HERE, "csrc", "_obj")
    os.makedirs(objdir, exist_ok=True)
    flags = [f for f in HIP_FLAGS if f != "-shared"]
    procs = []
    for u in HIP_UNITS:
        obj = os.path.join(objdir, u + ".o")
        cmd = ["hipcc", *flags, "-c", "-o", obj, os.path.join(HERE, "csrc", u + ".hip")]
        procs.append((cmd, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                                  text=True)))
    objs = []
    for cmd, obj, pr in procs:
        log = pr.communicate()[0]
        if pr.returncode != 0:
            sys.stderr.write(log)
            raise RuntimeError(f"build step failed: {' '.join(cmd)}")
        objs.append(obj)
    tmp = out + ".tmp"
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_fortran() -> str | None:
    fc = shutil.which("amdflang") or (os.path.exists("/opt/rocm/bin/amdflang") and "/opt/rocm/bin/amdflang")
    if not fc:
        return None
    out = os.path.join(HERE, "libfvcell_fortran.a")
    objdir = os.path.join(HERE, "fortran", "_obj")
    os.makedirs(objdir, exist_ok=True)
    src = os.path.join(HERE, "fortran", "fvc_solver.f90")
    obj = os.path.join(objdir, "fvc_solver.o")
    _run([fc, "-O2", "-fPIC", "-module-dir", objdir, "-c", src, "-o", obj])
    # a static archive: linked into the calling program it shares that
    # program's Fortran runtime (unit 13, STOP), exactly like Solver.o did
    tmp = out + ".tmp"
    if os.path.exists(tmp):
        os.remove(tmp)
    _run(["ar", "rcs", tmp, obj])
    os.replace(tmp, out)
    stale = os.path.join(HERE, "libfvcell_fortran.so")
    if os.path.exists(stale):
        os.remove(stale)
    return out


def build_fortran_driver() -> str | None:
    """tests/fortran/drive_solver.exe: a Fortran caller of solver(...) linked
    against the drop-in, used by the GPU tests."""
    fc = shutil.which("amdflang") or (os.path.exists("/opt/rocm/bin/amdflang") and "/opt/rocm/bin/amdflang")
    if not fc:
        return None
    tdir = os.path.join(ROOT, "tests", "fortran")
    out = os.path.join(tdir, "drive_solver.exe")
    objdir = os.path.join(HERE, "fortran", "_obj")
    os.makedirs(objdir, exist_ok=True)
    _run([fc, "-O2", "-module-dir", objdir, os.path.join(tdir, "drive_solver.f90"), "-o", out,
          os.path.join(HERE, "libfvcell_fortran.a"), f"-L{HERE}", "-lfvcell", f"-Wl,-rpath,{HERE}",
          "-Wl,-rpath,$ORIGIN/../../fortranfvcell_amd"])
    return out


def build() -> None:
    build_hip()
    build_fortran()
    build_fortran_driver()


if __name__ == "__main__":
    build()
    print("built", os.listdir(HERE))
