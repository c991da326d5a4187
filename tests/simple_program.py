"""The reference program's own SIMPLE time step, simple() (src/Simple.f:
103-900), driven through ctypes on a lid-driven cavity, from the reference's
objects compiled from its sources (oracle/Makefile):

* oracle/_ref/libmvf_ref.so    -- linked with the reference's src/Solver.f;
* oracle/_ref/libmvf_dropin.so -- linked with the drop-in instead
  (libfvcell_fortran.a: solver_ and getres_ on the GPU).

Everything else -- the cell routines, pform, gform, updateVel, simpleUpdate,
the residual norms and the SIMPLE loop -- is the reference's own code in
both.  Test infrastructure: the arrays are set up as the reference's
Main.f and datastruc (src/Datastruct.f:150-176) set them up for ELLPACK
(au and al the same array), on a mesh from fortranfvcell_amd.problems.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from fortranfvcell_amd import problems as P
from fortranfvcell_amd.post import ell_pattern

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libmvf_ref.so")
DROPIN_LIB = os.path.join(ROOT, "oracle", "_ref", "libmvf_dropin.so")
POOL_WORDS = 1 << 24


class _Simple(C.Structure):
    # COMMON /simple_/ (include/simple.fi)
    _fields_ = [("underU", C.c_double), ("underP", C.c_double), ("underPc", C.c_double),
                ("underRo", C.c_double), ("g", C.c_double * 3), ("solvSimpleU1", C.c_double),
                ("solvSimpleU2", C.c_double), ("solvSimplePc", C.c_double), ("solvEnergy", C.c_double),
                ("cPc0", C.c_double), ("cPc", C.c_double), ("massa0", C.c_double), ("massa", C.c_double),
                ("fluxoM", C.c_double), ("intVel", C.c_int32), ("maxItSimple", C.c_int32),
                ("simpleC", C.c_int32), ("simpleR", C.c_int32), ("singularPressure", C.c_int32),
                ("vMass", C.c_int32), ("tDinamico", C.c_int32), ("sPressure", C.c_int32),
                ("bBodyPressure", C.c_int32), ("closed", C.c_int32), ("skewnessCorrection", C.c_int32)]


class _Omp(C.Structure):
    # COMMON /omp_control/ (include/openmp.fi:16-18)
    _fields_ = [("num_threads", C.c_int32), ("openmpCell", C.c_int32), ("openmpSolver", C.c_int32),
                ("nThreadsCell", C.c_int32), ("nThreadsSolver", C.c_int32), ("nThreadsDot", C.c_int32),
                ("nThreadsMatVec", C.c_int32)]


def _quiet(fn):
    """Runs fn with the process's stdout sent to /dev/null (the reference
    prints every SIMPLE iteration)."""
    saved, null = os.dup(1), os.open(os.devnull, os.O_WRONLY)
    try:
        os.dup2(null, 1)
        return fn()
    finally:
        os.dup2(saved, 1)
        os.close(saved)
        os.close(null)


_loaded: dict = {}


def load(path: str):
    """The library with its memory pool set up (src/Malloc.f:init_malloc),
    once per process: a second init_malloc STOPs the program."""
    if path not in _loaded:
        L = C.CDLL(path, mode=C.RTLD_LOCAL)
        C.c_int64.in_dll(L, "_QMmallocEmaxmem").value = POOL_WORDS
        _quiet(L.init_malloc_)
        _loaded[path] = L
    return _loaded[path]


class Cavity:
    """A lid-driven cavity on an n x n jittered quad mesh in the reference's
    arrays: no-slip walls (pedge 0), the lid moving at speed 1 (pedge 1 with
    sedge(1) = 1), the pressure-correction level pinned on cell 1 (pedge -1
    of the cell), viscosity nu, density 1, the TVD limiter `limiter`."""

    def __init__(self, n: int = 24, jitter: float = 0.1, nu: float = 0.01, limiter: int = 5, seed: int = 5):
        m = P.quad_mesh(n, jitter=jitter, seed=seed)
        a = P.ref_mesh_arrays(m, (0, 0, 1, 0))
        self.numel, self.nen = a["nelcon"].shape
        self.nnode = len(a["x"])
        self.x, self.ix, self.nelcon = a["x"], a["ix"], a["nelcon"]
        self.pedge = a["pedge"].copy()
        self.pedge[0, self.nen] = -1
        self.sedge = np.zeros((self.numel, self.nen + 1, 3))
        self.sedge[:, :self.nen, 0] = np.where(self.pedge[:, :self.nen] == 1, 1.0, 0.0)
        self.e = np.zeros((1, 10))
        self.e[0, 0], self.e[0, 1], self.e[0, 5], self.e[0, 6] = nu, 1.0, 1.0, limiter
        self.num = np.arange(1, self.numel + 1, dtype=np.int32)
        self.ja = ell_pattern(self.num, self.nelcon)
        self.nad = int((self.nelcon > 0).sum())


def run(L, cav: Cavity, steps: int, dt: float, its: int, tol: float = 1e-12, under_u: float = 0.7,
        under_pc: float = 0.3, workdir: str = ".") -> dict:
    """steps time steps of simple() with its SIMPLE iterations each, BiCGSTAB
    for the momentum and PCG for the pressure correction (the reference's
    defaults, src/Main.f:192-203), ELLPACK everywhere.  Returns the fields."""
    sim = _Simple.in_dll(L, "simple__")
    sim.underU, sim.underP, sim.underPc, sim.underRo = under_u, 1.0, under_pc, 1.0
    sim.solvSimpleU1 = sim.solvSimpleU2 = sim.solvSimplePc = sim.solvEnergy = 0.0   # no early exit
    sim.cPc0 = sim.cPc = sim.massa0 = sim.massa = 1.0
    sim.fluxoM = 0.0
    sim.maxItSimple = its
    sim.simpleC = sim.simpleR = sim.singularPressure = sim.vMass = sim.tDinamico = 0
    sim.bBodyPressure = sim.closed = sim.skewnessCorrection = 0
    omp = _Omp.in_dll(L, "omp_control_")
    omp.openmpCell = omp.openmpSolver = 0
    n, nen, nnode = cav.numel, cav.nen, cav.nnode
    f = lambda *s: np.zeros(s)
    i32 = lambda *s: np.zeros(s, dtype=np.int32)
    iax = np.array([nen, n + 1], dtype=np.int32)
    arrays = dict(
        x=np.ascontiguousarray(cav.x), ix=np.ascontiguousarray(cav.ix), e=cav.e.copy(),
        ie=np.array([65], np.int32), nelcon=np.ascontiguousarray(cav.nelcon), pedge=cav.pedge.copy(),
        sedge=cav.sedge.copy(), pedgeE=cav.pedge.copy(), sedgeE=f(n, nen + 1, 1), w=f(n, 2), w0=f(n, 2),
        wP=f(n, 2), num=cav.num.copy(), ls=f(n, nen, 2), gradU1=f(n, 2), gradU2=f(n, 2), gradP=f(n, 2),
        gradPc=f(n, 2), gradE=f(n, 2), div=f(n), mP=f(n, 20), iM=f(n, 4), fluxlU1=f(n), fluxlU2=f(n),
        fluxlPc=f(n), fluxlE=f(n), rCu1=f(n), rCu2=f(n), rCpC=f(n), rCe=f(n))
    for s in ("U1", "U2", "Pc"):
        al = f(n, nen)
        arrays.update({f"ad{s}": f(n), f"au{s}": al, f"al{s}": al, f"b{s}": f(n), f"b{s}0": f(n),
                       f"ia{s}": iax.copy(), f"ja{s}": cav.ja.copy()})
    arrays.update(adE=f(n), auE=f(4 * n), alE=f(4 * n), bE=f(n), bE0=f(n), iaE=i32(n + 1), jaE=i32(4 * n),
                  u1=f(n), u2=f(n), p=f(n), pC=f(n), pC1=f(n), en=f(n), en0=f(n), rO=np.ones((n, 3)),
                  un=f(nnode, 3), mdf=f(nnode), md=i32(nnode), ddU=f(n, 2), temp=f(n), sx=f(n),
                  eddyVisc=f(n), cs=f(n), yPlus=f(n))
    order = ["x", "ix", "e", "ie", "nelcon", "pedge", "sedge", "pedgeE", "sedgeE", "w", "w0", "wP", "num",
             "ls", "gradU1", "gradU2", "gradP", "gradPc", "gradE", "div", "mP", "iM", "fluxlU1", "fluxlU2",
             "fluxlPc", "fluxlE", "rCu1", "rCu2", "rCpC", "rCe",
             "adU1", "auU1", "alU1", "bU1", "bU10", "iaU1", "jaU1",
             "adU2", "auU2", "alU2", "bU2", "bU20", "iaU2", "jaU2",
             "adPc", "auPc", "alPc", "bPc", "iaPc", "jaPc",
             "adE", "auE", "alE", "bE", "bE0", "iaE", "jaE",
             "u1", "u2", "p", "pC", "pC1", "en", "en0", "rO", "un", "mdf", "md", "ddU", "temp", "sx",
             "eddyVisc", "cs", "yPlus"]
    ptrs = [arrays[k].ctypes.data_as(C.c_void_p) for k in order]
    ri = lambda v: C.byref(C.c_int32(int(v)))
    dt_ = C.c_double(dt)
    t_ = C.c_double(0.0)
    istep = C.c_int32(0)
    outs = [C.c_double(0.0) for _ in range(10)]      # cfl, re, kEnergy, pr, disfiltro, transSubGrid,
    prename = C.create_string_buffer(b"cavity".ljust(80), 80)   # kEnergySub, grashof, vol, (spare)
    nad = cav.nad
    L.simple_.restype = None
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        for _ in range(steps):
            istep.value += 1
            t_.value += dt_.value
            args = ptrs + [ri(n), ri(nnode), ri(2), ri(nen), ri(nen), ri(3), ri(1), C.byref(dt_), C.byref(t_),
                           ri(4), ri(4), ri(4), ri(1), ri(n), ri(n), ri(n), ri(n), ri(nad), ri(nad), ri(nad),
                           ri(nad), ri(2), ri(2), ri(1), ri(2), C.byref(C.c_double(tol)),
                           C.byref(C.c_double(tol)), ri(50000), ri(21), ri(0), ri(0), C.byref(istep),
                           *[C.byref(o) for o in outs[:9]], ri(1), ri(1), prename, ri(22), ri(0), ri(0),
                           C.c_size_t(80)]
            _quiet(lambda: L.simple_(*args))
    finally:
        os.chdir(cwd)
    return dict(w=arrays["w"].copy(), p=arrays["p"].copy(), pC=arrays["pC"].copy(), div=arrays["div"].copy(),
                rCpC=arrays["rCpC"].copy(), rCu1=arrays["rCu1"].copy(), gradP=arrays["gradP"].copy())


def run_isolated(lib: str, n: int, steps: int, its: int, dt: float, workdir: str, timeout: float = 600) -> dict:
    """run() in a child process: the reference STOPs the program (exit status
    0) on a solver that does not converge and on other errors, which would
    otherwise end the calling process as if it had succeeded.  The child
    writes the fields to workdir/fields.npz; their absence is the failure."""
    import subprocess
    import sys
    out = os.path.join(workdir, "fields.npz")
    if os.path.exists(out):
        os.remove(out)
    env = dict(os.environ, PYTHONPATH=os.pathsep.join([ROOT, os.environ.get("PYTHONPATH", "")]).rstrip(os.pathsep))
    r = subprocess.run([sys.executable, os.path.abspath(__file__), lib, str(n), str(steps), str(its), repr(dt), out],
                       cwd=workdir, capture_output=True, text=True, timeout=timeout, env=env)
    if r.returncode != 0 or not os.path.exists(out):
        tail = (r.stdout + r.stderr)[-3000:]
        raise RuntimeError(f"simple() in {os.path.basename(lib)} ended without its fields "
                           f"(exit status {r.returncode}):\n{tail}")
    with np.load(out, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


if __name__ == "__main__":
    import sys
    lib, n, steps, its, dt, out = sys.argv[1:7]
    fields = run(load(lib), Cavity(int(n)), steps=int(steps), dt=float(dt), its=int(its),
                 workdir=os.path.dirname(os.path.abspath(out)))
    with open("/proc/self/maps") as fh:
        fields["fvcell_loaded"] = np.array("libfvcell.so" in fh.read())   # the HIP library ran the solves
    np.savez(out, **fields)
