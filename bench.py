"""Benchmark: Krylov solves/s at 1M cells to relative residual 1e-10, and
SpMV GB/s against the HBM peak, on one MI355X.

Workload (one "step" = one solve, inputs resident in HBM): the nonsymmetric
upwind advection-diffusion system in the momentum form the reference's
SIMPLE step hands to its solver (src/CellibSimple.f:347-355: transient term
rho area/dt with dt = 0.005, under-relaxation underU = 0.8) on a 1024x1024
jittered quad mesh (1,048,576 cells, largest cell Peclet number 100, ELLPACK
width 4, FP64), Jacobi-BiCGSTAB to ||b - A x||_2 / ||b||_2 <= 1e-10.  The
same line reports: GMRES(30) on that system to the same true residual (MGS,
CGS2 and lagged CGS2); the same system at the reference's defaults (underU =
1, no transient term: the steady a_P = sum a_nb + sum F_f), which Jacobi
Krylov solvers converge on up to 256x256 cells; Jacobi-PCG on the pinned
pressure system; the plain ELLPACK SpMV (warm and with the 256 MB last-level
cache defeated by rotating buffers); FP64 against mixed-precision PCG on the
lognormal pressure system, and FP64 against mixed-precision BiCGSTAB and
GMRES(30) on the momentum system; the 256-system batch (FP64, mixed, one at
a time) with its roofline; the roofline of the dominant kernel; and the
reference's own OpenMP Fortran solver (compiled from its sources,
oracle/_ref) timed on the host cores as the CPU baseline, at the tolerance
where its own stopping rule reaches the same true residual.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --spmv-only R      (rotated SpMV launches only, for counters)
       python bench.py --c3b N [--maxit M] (FP64 vs mixed PCG on 2 N^2 triangles)
       python bench.py --gmres-only [--steps K] (GMRES(30) per orthogonalisation, for a profile)
       python bench.py --batch-only [--steps K] (the 256-system batch, for a profile)
       python bench.py --calibrate-copy MiB  (a plain copy, to scale the byte counters)
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(neq: int, nnz_off: int, vbytes: int = 8) -> int:
    """Algorithmic bytes of one SpMV: ad, x, y per row; value + int32 column
    per stored off-diagonal (padding excluded)."""
    return neq * 3 * vbytes + nnz_off * (vbytes + 4)


def bicgstab_iter_bytes(neq: int, nnz_off: int) -> int:
    """Algorithmic bytes of one iteration of the two-phase BiCGSTAB (the
    default recurrence).  P: the matrix (ad + value/column per off-diagonal),
    reads s, t, p, v, x, rhat and writes x, b, p, v; Q: the matrix, reads b,
    v, rhat and writes s, t.  Each vector element counted once (neighbour
    gathers are cache hits); the Jacobi diagonal is the matrix's ad."""
    per_matrix = 8 * neq + 12 * nnz_off
    return 2 * per_matrix + (10 + 5) * 8 * neq


def bicgstab_solve_bytes(neq: int, nnz_off: int, its: int) -> int:
    """its iterations + the P phase that completes the last one + setup
    (b, ad read; x, r, p, b written) + final (matrix, x, b)."""
    p_phase = 8 * neq + 12 * nnz_off + 10 * 8 * neq
    return its * bicgstab_iter_bytes(neq, nnz_off) + p_phase + 6 * 8 * neq + (8 * neq + 12 * nnz_off + 2 * 8 * neq)


def pcg_iter_bytes(neq: int, nnz_off: int) -> int:
    """One iteration of the single-reduction PCG (the default recurrence):
    the matrix, reads r, w, s, p, x and writes p, s, x, r, w."""
    return 8 * neq + 12 * nnz_off + 10 * 8 * neq


def gmres_cycle_bytes(neq: int, nnz_off: int, steps: int, cgs2: bool = False) -> int:
    """Arnoldi step j: one SpMV plus (5j + 6) vector streams with modified
    Gram-Schmidt (per projection j: read w, g_j, g_j+1, write w); with CGS2
    (3j + 5): read the j basis columns in each of the three passes, w read
    and written in each, column j normalised once."""
    per = (lambda j: 3 * j + 5) if cgs2 else (lambda j: 5 * j + 6)
    return sum(spmv_bytes(neq, nnz_off) + per(j) * 8 * neq for j in range(1, steps + 1))


def gmres_lag_cycle_bytes(neq: int, nnz_off: int, steps: int, k: int) -> int:
    """A cycle of `steps` Arnoldi steps with lagged CGS2: D_j for j = 1 ..
    steps + 1 (the SpMV u -> w, the j - 1 basis columns; step k + 1 reads u and
    the k columns only) and B_j for j = 1 .. steps (reads the j - 1 columns,
    u and w; writes column j and u)."""
    b = 0
    for j in range(1, steps + 2):
        nv = j - 1
        b += (spmv_bytes(neq, nnz_off) if j <= k else 8 * neq) + 8 * nv * neq
        if j <= steps:
            b += 8 * neq * (nv + 2) + 16 * neq
    return b


def gmres_solve_bytes(neq: int, nnz_off: int, cycles: int, steps: int, k: int, cgs2: bool,
                      lag: bool = False) -> int:
    """All Arnoldi steps (cycles of k steps, the last one shorter), one
    residual SpMV per cycle start (and one to confirm the last), and each
    cycle's update of x (its columns and x read, x written)."""
    full, last = divmod(steps, k)
    if lag:
        b = full * gmres_lag_cycle_bytes(neq, nnz_off, k, k) + (gmres_lag_cycle_bytes(neq, nnz_off, last, k)
                                                                 if last else 0)
    else:
        b = full * gmres_cycle_bytes(neq, nnz_off, k, cgs2) + gmres_cycle_bytes(neq, nnz_off, last, cgs2)
    return b + (cycles + 1) * (spmv_bytes(neq, nnz_off) + 16 * neq) + (steps + 2 * cycles) * 8 * neq


def pcg_ref_iter_bytes(neq: int, nnz_off: int) -> int:
    """One iteration of the reference-order PCG phases (batched FP64):
    phase A's SpMV reads the matrix (ad and the off-diagonals), phase B reads
    the Jacobi diagonal m = ad again (z = r/m), and 88 bytes of vector
    traffic per row (A: z, p read, p, q written; B: x, p, r, q read, x, r, z
    written)."""
    return 8 * neq + 12 * nnz_off + 8 * neq + 88 * neq


def mixed_inner_bytes(neq: int, nnz_off: int, col16: bool) -> int:
    """One FP32 inner iteration of mixed PCG: FP32 values and int16 column
    offsets (or int32 columns), dsq, the row's 16-byte record read and the
    next one written."""
    return (6 if col16 else 8) * nnz_off + 36 * neq


def read_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 counters, if recorded
    under profiles/ (line 'hbm_bytes_per_launch: N')."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", open(p).read())
    return float(m.group(1)) if m else None


def time_spmv(F, ds, neq, nnz, reps=50, rotate=1):
    """Average SpMV kernel time (HIP events on the launch stream) over reps."""
    import ctypes as C
    L = F.load()
    systems = [ds] + [F.DeviceSystem(ds._fmt) for _ in range(rotate - 1)]
    xs = [F.DeviceArray.from_host(np.linspace(-1, 1, neq)) for _ in range(rotate)]
    ys = [F.DeviceArray.empty(neq) for _ in range(rotate)]
    for i in range(rotate):
        systems[i].matvec(xs[i], ys[i])
    e0, e1 = C.c_void_p(), C.c_void_p()
    L.fvc_event_create(C.byref(e0))
    L.fvc_event_create(C.byref(e1))
    L.fvc_synchronize()
    L.fvc_event_record(e0, None)
    for r in range(reps):
        i = r % rotate
        systems[i].matvec(xs[i], ys[i])
    L.fvc_event_record(e1, None)
    ms = C.c_float()
    L.fvc_event_elapsed_ms(e0, e1, C.byref(ms))
    L.fvc_event_destroy(e0)
    L.fvc_event_destroy(e1)
    for s in systems[1:]:
        del s
    t = ms.value * 1e-3 / reps
    return t, spmv_bytes(neq, nnz) / t / 1e9


def ladder_tol(n: int, method: str) -> float:
    """The tolerance at which the reference's own stopping rule leaves a true
    relative residual <= 1e-10 on the n x n momentum system (its BiCGSTAB and
    GMRES rules are not the residual), as recorded by the reference run in
    tests/golden/ref_ladder_1024.npz; 1e-10 for other sizes."""
    p = os.path.join(ROOT, "tests", "golden", "ref_ladder_1024.npz")
    if n != 1024 or not os.path.exists(p):
        return 1e-10
    return float(np.load(p, allow_pickle=False)[f"advdiff1024_ell_{method}_tol"])


def cpu_baseline(fmt, A, b, method, tol, budget_s=10.0, threads=None):
    """The reference Fortran (oracle/_ref) with OpenMP on the host cores: the
    solver's own OpenMP routines (pbicgstab_omp + matvec_ellpack_omp +
    dot_ompL4), repeated for about budget_s seconds, at `tol` under its own
    stopping rule.  Reports its iteration count (from its unit-13 log line)
    and the true relative residual it reached.  Falls back to the sequential
    C restatement."""
    import oracle
    if threads is None:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    if oracle.have_ref():
        oracle.ref()
        oracle.ref_set_omp(1, threads)
        x0 = np.zeros(fmt["neq"])
        t0 = time.perf_counter()
        (x, _, _), _, its, _ = oracle.ref_logged(lambda: oracle.krylov(
            method, fmt, tol=tol, use_ref=True, omp=True, nthreads=threads, x0=x0, init=True))
        first = time.perf_counter() - t0
        relres = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
        sample_solves = max(1, min(1000, int(budget_s / max(first, 1e-6))))
        t0 = time.perf_counter()
        for _ in range(sample_solves):
            oracle.krylov(method, fmt, tol=tol, use_ref=True, omp=True, nthreads=threads, x0=x0, init=True)
        dt = (time.perf_counter() - t0) / sample_solves
        return {"value": 1.0 / dt, "unit": "solves/s", "cores": threads, "kind": "reference",
                "sample": f"{sample_solves} full {method} solve(s) of the same system, reference "
                          f"Fortran (pbicgstab_omp/matvec_ellpack_omp/dot_ompL4) built from its "
                          f"sources, {threads} OpenMP threads, its own stopping rule at tol={tol:g} "
                          f"(the tolerance where that rule reaches true relative residual 1e-10): "
                          f"{its} iterations, true relative residual {relres:.2e}",
                "seconds_per_solve": dt, "iterations": its, "true_relres": relres, "tol": tol}
    t0 = time.perf_counter()
    x, _, st = oracle.krylov(method, fmt, tol=tol)
    dt = time.perf_counter() - t0
    relres = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
    return {"value": 1.0 / dt, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"1 full {method} solve, sequential C restatement, {st['iterations']} iterations, "
                      f"true relative residual {relres:.2e}", "seconds_per_solve": dt,
            "iterations": st["iterations"], "true_relres": relres, "tol": tol}


def gform_bytes(numel: int, nnode: int, nen: int = 4) -> int:
    """Algorithmic bytes of the pressure-correction gradient: per cell ix
    (nen+1 int32), nelcon (nen), pedge (nen+1), u, the old and new gradient
    and fluxl; each node's coordinates once (neighbour gathers are cache
    hits)."""
    return numel * (4 * (3 * nen + 2) + 8 + 32 + 8) + nnode * 16


def assemble_pc_bytes(numel: int, nnode: int, nen: int = 4) -> int:
    """Algorithmic bytes of the pressure-correction assembly: per cell ix,
    nelcon, pedge (as gform), num, w, d, gradp and p in; ad, al (nen), b and
    the mass residual out; each node's coordinates once (the wall values in
    sedge are read for wall edges only and not counted)."""
    return numel * (4 * (3 * nen + 2) + 4 + 56 + 8 * (nen + 3)) + nnode * 16


def bench_post(F, n: int, reps: int = 50) -> dict:
    """fvc_assemble_pc_dev, fvc_gform_dev and fvc_simple_update_dev on the
    n x n jittered quad mesh (HIP events on the launch stream)."""
    import ctypes as C
    from fortranfvcell_amd import problems as P
    from fortranfvcell_amd.post import DeviceMesh, simple_update
    m = P.quad_mesh(n, jitter=0.1, seed=P.SEED_BASE + 6)
    a = P.ref_mesh_arrays(m)
    dm = DeviceMesh(a["x"], a["ix"], a["nelcon"], a["pedge"])
    rng = np.random.default_rng(1)
    u = F.DeviceArray.from_host(rng.normal(size=m.neq))
    g = F.DeviceArray.from_host(np.zeros((m.neq, 2)))
    fl = F.DeviceArray.empty(m.neq)
    L = F.load()
    e0, e1 = C.c_void_p(), C.c_void_p()
    L.fvc_event_create(C.byref(e0))
    L.fvc_event_create(C.byref(e1))
    dm.gform(u, g, fl)
    L.fvc_synchronize()
    L.fvc_event_record(e0, None)
    for _ in range(reps):
        dm.gform(u, g, fl)
    L.fvc_event_record(e1, None)
    ms = C.c_float()
    L.fvc_event_elapsed_ms(e0, e1, C.byref(ms))
    t_g = ms.value * 1e-3 / reps
    vecs = [F.DeviceArray.from_host(rng.normal(size=s)) for s in ((m.neq, 2), m.neq, m.neq, m.neq, m.neq,
                                                                   (m.neq, 2), (m.neq, 2))]
    w, pre, u1, u2, pc, gp, d = vecs
    simple_update(m.neq, w, pre, u1, u2, pc, gp, d, 0.7, 0.3, False)
    L.fvc_synchronize()
    L.fvc_event_record(e0, None)
    for _ in range(reps):
        simple_update(m.neq, w, pre, u1, u2, pc, gp, d, 0.7, 0.3, False)
    L.fvc_event_record(e1, None)
    L.fvc_event_elapsed_ms(e0, e1, C.byref(ms))
    t_s = ms.value * 1e-3 / reps
    # the pressure-correction assembly (intVel 5, walls: inflow, outflow,
    # parabolic, zero gradient; the level pinned on cell 1)
    a2 = P.ref_mesh_arrays(m, (2, 3, 4, 6))
    pedge = a2["pedge"].copy()
    pedge[0, m.w] = -1
    dm2 = DeviceMesh(a2["x"], a2["ix"], a2["nelcon"], pedge)
    f = {"num": np.arange(1, m.neq + 1, dtype=np.int32), "e": np.array([[0.0, 1.0] + [0.0] * 8]),
         "sedge": rng.normal(size=(m.neq, m.w + 1, 3)), "w": rng.normal(size=(m.neq, 2)),
         "d": rng.uniform(0.5, 1.5, size=(m.neq, 2)), "gradp": rng.normal(size=(m.neq, 2)),
         "p": rng.normal(size=m.neq)}
    df = {k: F.DeviceArray.from_host(v) for k, v in f.items()}
    outs = (F.DeviceArray.empty(m.neq), F.DeviceArray.empty((m.neq, m.w)), F.DeviceArray.empty(m.neq),
            F.DeviceArray.empty(m.neq))
    pc_args = (df["num"], df["e"], df["sedge"], df["w"], df["d"], df["gradp"], df["p"], *outs)
    dm2.assemble_pc(*pc_args)
    L.fvc_synchronize()
    L.fvc_event_record(e0, None)
    for _ in range(reps):
        dm2.assemble_pc(*pc_args)
    L.fvc_event_record(e1, None)
    L.fvc_event_elapsed_ms(e0, e1, C.byref(ms))
    t_a = ms.value * 1e-3 / reps
    L.fvc_event_destroy(e0)
    L.fvc_event_destroy(e1)
    ab = assemble_pc_bytes(m.neq, len(a["x"]))
    gb = gform_bytes(m.neq, len(a["x"]))
    sb = m.neq * (16 + 8 + 8 + 8 + 8 + 16 + 16 + 8)     # w out, pre r/w, u1, u2, preC, gradPc, d
    return {"cells": m.neq, "assemble_pc_us": t_a * 1e6, "assemble_pc_GBs": ab / t_a / 1e9, "assemble_pc_bytes": ab,
            "gform_us": t_g * 1e6, "gform_GBs": gb / t_g / 1e9, "gform_bytes": gb,
            "simple_update_us": t_s * 1e6, "simple_update_GBs": sb / t_s / 1e9}


def calibrate_copy(mib: int, reps: int) -> None:
    """A plain device copy of a known size: rocprofv3's FETCH_SIZE and
    WRITE_SIZE for it give the scale of those counters (bytes each reads per
    byte moved), applied to the solver kernels' counts."""
    import ctypes as C
    import fortranfvcell_amd as F
    F.init(0)
    L = F.load()
    nbytes = mib << 20
    src = F.DeviceArray.empty(nbytes // 8)
    dst = F.DeviceArray.empty(nbytes // 8)
    L.fvc_memset_dev(C.c_void_p(src.ptr), 1, nbytes)
    e0, e1 = C.c_void_p(), C.c_void_p()
    L.fvc_event_create(C.byref(e0))
    L.fvc_event_create(C.byref(e1))
    L.fvc_event_record(e0, None)
    for _ in range(reps):
        F.check(L.fvc_memcpy_dtod(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes, None))
    L.fvc_event_record(e1, None)
    ms = C.c_float()
    L.fvc_event_elapsed_ms(e0, e1, C.byref(ms))
    t = ms.value * 1e-3 / reps
    print(json.dumps({"copy_bytes_each_way": nbytes, "reps": reps, "us_per_copy": t * 1e6,
                      "GBs_read_plus_write": 2 * nbytes / t / 1e9}), flush=True)


def spmv_only(n: int, reps: int) -> None:
    """Only the ELLPACK SpMV of the n x n momentum system, reps launches over
    4 rotating copies of A, x, y (> 256 MB, so each launch streams from HBM):
    the kernel whose compulsory bytes are known exactly, to check the scale of
    the hardware byte counters on a gather kernel."""
    import fortranfvcell_amd as F
    from fortranfvcell_amd import problems as P
    F.init(0)
    s = P.config_advdiff(n)
    ell = P.to_ell(s)
    ds = F.DeviceSystem(ell)
    ds._fmt = ell
    t, gbs = time_spmv(F, ds, s.neq, ell["nad"], reps=reps, rotate=4)
    print(json.dumps({"kernel": "spmv_kernel<0>", "neq": s.neq, "launches": reps + 4, "rotate": 4,
                      "compulsory_bytes": spmv_bytes(s.neq, ell["nad"]), "us": t * 1e6, "GBs": gbs}),
          flush=True)


def gmres_only(n: int, steps: int) -> None:
    """Only GMRES(30) on the n x n momentum system to true relative residual
    1e-10, `steps` solves per orthogonalisation (MGS, CGS2, lagged CGS2, and
    lagged CGS2 in mixed precision): the kernels of each, for a profile."""
    import fortranfvcell_amd as F
    from fortranfvcell_amd import problems as P
    F.init(0)
    s = P.config_advdiff(n)
    ell = P.to_ell(s)
    ds = F.DeviceSystem(ell)
    out = {}
    for name, kw in (("mgs", dict(gmres_orth=F.ORTH_MGS)), ("cgs2", dict(gmres_orth=F.ORTH_CGS2)),
                     ("cgs2_lag", dict(gmres_orth=F.ORTH_CGS2_LAG)),
                     ("cgs2_lag_mixed", dict(gmres_orth=F.ORTH_CGS2_LAG, precision=F.PREC_MIXED))):
        ts = []
        for _ in range(steps):
            st = ds.solve(F.GMRES, tol=1e-10, gmres_k=30, stop_rule=F.STOP_RELRES, **kw)
            ts.append(st["seconds"])
        gb = gmres_solve_bytes(s.neq, ell["nad"], st["iterations"], st["inner_iterations"], 30,
                               kw["gmres_orth"] == F.ORTH_CGS2, lag=kw["gmres_orth"] == F.ORTH_CGS2_LAG)
        out[name] = {"cycles": st["iterations"], "arnoldi_steps": st["inner_iterations"],
                     "ms_per_solve": float(np.mean(ts)) * 1e3, "true_relres": st["true_relres"],
                     "status": st["status"],
                     "GBs_algorithmic": (gb / float(np.mean(ts)) / 1e9) if "precision" not in kw else None}
    print(json.dumps({"workload": f"GMRES(30), {n}x{n} momentum system, true relative residual 1e-10",
                      "neq": s.neq, "steps": steps, **out}), flush=True)


def batch_only(steps: int) -> None:
    """Only the 256-system batch (pressure systems, sides log-uniform in
    [32, 512], each to true relative residual 1e-10), FP64 and mixed, at the
    options the autotuner picks on MI355X (1,024 threads, work exponent 1.5;
    FP64 64 systems per launch, mixed 86), `steps` batched solves each: the
    batch kernels, for a profile."""
    import fortranfvcell_amd as F
    from fortranfvcell_amd import problems as P
    F.init(0)
    fmts = [P.to_ell(q) for q in P.config_batch()]
    sizes = [(int(f["neq"]), int(f["nad"])) for f in fmts]
    db = F.DeviceBatch(F.concat_ell(fmts))
    out = {}
    for name, prec, per in (("fp64", F.PREC_FP64, 64), ("mixed", F.PREC_MIXED, 86)):
        kw = dict(tol=1e-10, precision=prec, stop_rule=F.STOP_RELRES, block_size=1024, batch_work_exp10=15,
                  batch_launch_systems=per)
        ts = []
        for _ in range(steps):
            sts = db.solve(**kw)
            ts.append(sts[0]["seconds"])
        if prec == F.PREC_FP64:
            nb = sum(q["iterations"] * pcg_ref_iter_bytes(ne, nz) for q, (ne, nz) in zip(sts, sizes))
        else:
            nb = sum(q["iterations"] * mixed_inner_bytes(ne, nz, bool(q["flags"] & F.FLAG_COL16))
                     + q["inner_iterations"] * (spmv_bytes(ne, nz) + 40 * ne) for q, (ne, nz) in zip(sts, sizes))
        t = float(np.mean(ts))
        out[name] = {"ms": t * 1e3, "solves_per_s": len(sts) / t, "algorithmic_bytes": nb,
                     "GBs_algorithmic": nb / t / 1e9, "frac": nb / t / 1e9 / HBM_PEAK_GBS,
                     "max_iterations": max(q["iterations"] for q in sts),
                     "all_converged": all(q["status"] == 0 for q in sts)}
    print(json.dumps({"workload": "256 pressure systems, sides log-uniform in [32, 512], true relative residual 1e-10",
                      "rows": int(sum(ne for ne, _ in sizes)), "steps": steps, **out}), flush=True)


def bench_c3b(n: int, steps: int, maxit: int = 200000) -> None:
    """FP64 and mixed-precision Jacobi-PCG (FP32 matrix and inner iterations,
    FP64 residual recomputation) to true relative residual 1e-10 on the
    lognormal(0.5) pressure system of an n x n quad mesh split into 2 n^2
    triangles (ELLPACK w = 3); the error of the mixed solution is stated
    against the FP64 GPU solution (itself checked against the oracle at the
    sizes the oracle finishes)."""
    import fortranfvcell_amd as F
    from fortranfvcell_amd import problems as P
    F.init(0)
    t0 = time.perf_counter()
    s = P.config_pressure_lognormal(n, triangles=True)
    ell = P.to_ell(s)
    log(f"generated {s.neq} triangle cells in {time.perf_counter() - t0:.1f}s")
    ds = F.DeviceSystem(ell)
    del ell
    opts = dict(stop_rule=F.STOP_RELRES, x0_rule=F.X0_ZERO)
    out = {}
    for name, prec in (("fp64", F.PREC_FP64), ("mixed", F.PREC_MIXED)):
        ts = []
        for _ in range(steps):
            st = ds.solve(F.PCG, tol=1e-10, maxit=maxit, precision=prec, **opts)
            ts.append(st["seconds"])
            log(name, st["iterations"], st["inner_iterations"], f"{st['seconds']:.2f}s", st["status"])
        x = ds.x.get()
        out[name] = {"iterations": st["iterations"], "refinements": st["inner_iterations"],
                     "s_per_solve": float(np.mean(ts)), "true_relres": st["true_relres"],
                     "status": st["status"], "_x": x}
    x64 = out["fp64"].pop("_x")
    xm = out["mixed"].pop("_x")
    out["mixed"]["rel_err_vs_fp64"] = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))
    nnz = int(s.nbr.__ge__(0).sum())
    print(json.dumps({"workload": f"lognormal pressure, {n}x{n} quads split in 2 triangles, ELLPACK w=3",
                      "neq": s.neq, "nnz_offdiag": nnz, "steps": steps, **out,
                      "speedup_mixed_over_fp64": out["fp64"]["s_per_solve"] / out["mixed"]["s_per_solve"]}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024, help="mesh side (n*n cells)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel on MIB MiB (hardware byte-counter reference)")
    ap.add_argument("--c3b", type=int, default=0, metavar="N",
                    help="only run FP64 against mixed-precision PCG on the lognormal pressure system "
                         "of an N x N quad mesh split into 2 N^2 triangles (N = 4096: 33.5M cells)")
    ap.add_argument("--maxit", type=int, default=200000, help="--c3b: iteration cap of each solve")
    ap.add_argument("--spmv-only", type=int, default=0, metavar="R",
                    help="only run R rotated SpMV launches of the momentum system (byte-counter check)")
    ap.add_argument("--gmres-only", action="store_true",
                    help="only run GMRES(30) with each orthogonalisation, --steps solves each (for a profile)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only run the 256-system batch, FP64 and mixed, --steps batched solves each (for a profile)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy, max(args.steps, 1))
    if args.c3b:
        return bench_c3b(args.c3b, max(args.steps, 1), args.maxit)
    if args.spmv_only:
        return spmv_only(args.n, args.spmv_only)
    if args.gmres_only:
        return gmres_only(args.n, max(args.steps, 1))
    if args.batch_only:
        return batch_only(max(args.steps, 1))
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU only (asked for --gpus {args.gpus}); "
              "there is no multi-GPU path to measure", file=sys.stderr)
        sys.exit(2)

    import fortranfvcell_amd as F
    from fortranfvcell_amd import problems as P

    F.init(0)
    info = F.device_info()
    log("device", info)
    n = args.n
    t0 = time.perf_counter()
    s = P.config_advdiff(n)
    ell = P.to_ell(s)
    log(f"generated advection-diffusion {n}x{n} in {time.perf_counter() - t0:.1f}s")
    neq, nnz = s.neq, ell["nad"]
    ds = F.DeviceSystem(ell)
    ds._fmt = ell
    tol = 1e-10
    opts = dict(stop_rule=F.STOP_RELRES, x0_rule=F.X0_ZERO)

    for _ in range(args.warmup):
        st = ds.solve(F.BICGSTAB, tol=tol, **opts)
    F.load().fvc_synchronize()
    kern = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = ds.solve(F.BICGSTAB, tol=tol, **opts)
        kern.append(st["seconds"])
    F.load().fvc_synchronize()
    wall = time.perf_counter() - t0
    assert st["status"] == 0, st
    x = ds.x.get()
    relres = float(np.linalg.norm(s.b - P.scipy_csr(s) @ x) / np.linalg.norm(s.b))
    its = st["iterations"]
    solve_bytes = bicgstab_solve_bytes(neq, nnz, its)
    launches = 2 * its + 4   # SETUP, P/Q per iteration, the closing P and Q, FINAL (+ a no-row flush)
    kt = float(np.mean(kern))
    achieved = solve_bytes / kt / 1e9
    log(f"BiCGSTAB: {its} it, {kt * 1e3:.3f} ms/solve kernel, wall {wall / args.steps * 1e3:.3f} ms, relres {relres:.2e}")

    result = {
        "metric": "Krylov solves/s at 1M cells to 1e-10 resid; SpMV GB/s vs HBM peak",
        "value": args.steps / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"advection-diffusion in the momentum form of a SIMPLE step (transient "
                               f"term area/dt, dt={s.meta['dt']:g}; under-relaxation underU="
                               f"{s.meta['under_relax']:g}): {n}x{n} jittered quads, upwind, largest cell "
                               f"Peclet {s.meta['peclet_max']:g}, ELLPACK w=4 FP64, Jacobi-BiCGSTAB to true "
                               "relative residual 1e-10",
                   "dt": s.meta["dt"], "under_relax": s.meta["under_relax"], "peclet_max": s.meta["peclet_max"],
                   "neq": neq, "nnz_offdiag": nnz, "method": "bicgstab", "format": "ellpack",
                   "iterations": its, "true_relres": relres, "engine": st["engine"],
                   "grid_blocks": st["grid_blocks"], "block_size": st["block_size"],
                   "device": info["name"]},
        # the dominant kernel is bicg2_kernel<0> (ELLPACK width 4): one launch
        # per phase.  achieved = algorithmic bytes per launch / event time per
        # launch on the solve's stream, i.e. bytes per solve / time per solve
        # (launch gaps and the host's done-checks included, so a lower bound
        # of the kernel-only rate rocprofv3 reports)
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic("bicg2_kernel_advdiff1024_traffic.txt"),
                     "kernel": "fvc::bicg2_kernel<0> (two-phase BiCGSTAB, one launch per phase)",
                     "launches_per_solve": launches,
                     "algorithmic_bytes_per_launch": solve_bytes / launches,
                     "us_per_launch": kt / launches * 1e6,
                     "solve_ms": kt * 1e3},
    }

    if not args.no_extra:
        extra = {}
        # plain SpMV, warm and cache-defeating rotation of 4 copies (> 256 MB)
        t_w, gbs_w = time_spmv(F, ds, neq, nnz, reps=200, rotate=1)
        t_f, gbs_f = time_spmv(F, ds, neq, nnz, reps=200, rotate=4)
        extra["spmv_ell_fp64"] = {"us_warm": t_w * 1e6, "GBs_warm": gbs_w, "us_rotated": t_f * 1e6,
                                  "GBs_rotated": gbs_f, "frac_rotated": gbs_f / HBM_PEAK_GBS,
                                  "bytes": spmv_bytes(neq, nnz)}
        log("spmv", extra["spmv_ell_fp64"])
        # the reference's own BiCGSTAB stopping rule at the tolerance where it
        # reaches true relative residual 1e-10: the CPU baseline's exact work
        ltol = ladder_tol(n, "bicgstab")
        ds.solve(F.BICGSTAB, tol=ltol, x0_rule=F.X0_ZERO)
        st_l = ds.solve(F.BICGSTAB, tol=ltol, x0_rule=F.X0_ZERO)
        extra["bicgstab_reference_rule"] = {"tol": ltol, "iterations": st_l["iterations"],
                                            "ms_per_solve": st_l["seconds"] * 1e3,
                                            "solves_per_s": 1.0 / st_l["seconds"],
                                            "true_relres": st_l["true_relres"]}
        log("bicgstab reference rule", extra["bicgstab_reference_rule"])
        # GMRES(30) on the same system, to the same true residual: modified
        # Gram-Schmidt in the reference's order, CGS2, and CGS2 lagged by one
        # step (one reduction per Arnoldi step); bytes by each one's schedule
        gm = {}
        for oname, orth in (("mgs", F.ORTH_MGS), ("cgs2", F.ORTH_CGS2), ("cgs2_lag", F.ORTH_CGS2_LAG)):
            ds.solve(F.GMRES, tol=tol, gmres_k=30, stop_rule=F.STOP_RELRES, gmres_orth=orth)
            st_g = min((ds.solve(F.GMRES, tol=tol, gmres_k=30, stop_rule=F.STOP_RELRES, gmres_orth=orth)
                        for _ in range(3)), key=lambda q: q["seconds"])
            gb = gmres_solve_bytes(neq, nnz, st_g["iterations"], st_g["inner_iterations"], 30, orth == F.ORTH_CGS2,
                                   lag=orth == F.ORTH_CGS2_LAG)
            gm[oname] = {"cycles": st_g["iterations"], "arnoldi_steps": st_g["inner_iterations"],
                         "ms_per_solve": st_g["seconds"] * 1e3, "solves_per_s": 1.0 / st_g["seconds"],
                         "true_relres": st_g["true_relres"], "status": st_g["status"],
                         "GBs_algorithmic": gb / st_g["seconds"] / 1e9,
                         "frac": gb / st_g["seconds"] / 1e9 / HBM_PEAK_GBS, "block_size": st_g["block_size"],
                         "grid_blocks": st_g["grid_blocks"]}
        extra["gmres30"] = gm
        log("gmres30", gm)
        # the reference's defaults: underU = 1 (src/Main.f:219), no transient
        # term.  Jacobi-BiCGSTAB stops converging on this system from 512^2
        # cells on, so it is measured at 256^2, where both methods converge
        # and the reference's solutions pin them (tests/golden/ref_steady.npz)
        ns = min(n, 256)
        ss = P.config_advdiff(ns, dt=None, under_relax=1.0)
        ells = P.to_ell(ss)
        dss = F.DeviceSystem(ells)
        st_sb = dss.solve(F.BICGSTAB, tol=tol, maxit=100000, **opts)
        sbb = bicgstab_solve_bytes(ss.neq, ells["nad"], st_sb["iterations"])
        st_sg = dss.solve(F.GMRES, tol=tol, gmres_k=30, maxit=20000, stop_rule=F.STOP_RELRES,
                          gmres_orth=F.ORTH_CGS2_LAG)
        extra["steady_reference_defaults"] = {
            "workload": f"the same advection-diffusion operator with underU=1 and no transient term, "
                        f"{ns}x{ns} cells",
            "bicgstab": {"iterations": st_sb["iterations"], "ms_per_solve": st_sb["seconds"] * 1e3,
                         "true_relres": st_sb["true_relres"], "status": st_sb["status"],
                         "GBs_algorithmic": sbb / st_sb["seconds"] / 1e9},
            "gmres30_cgs2_lag": {"cycles": st_sg["iterations"], "arnoldi_steps": st_sg["inner_iterations"],
                             "ms_per_solve": st_sg["seconds"] * 1e3, "true_relres": st_sg["true_relres"],
                             "status": st_sg["status"]},
            "note": "Jacobi-BiCGSTAB does not converge on this system at 512^2 and 1024^2 cells "
                    "(no true residual below 1e-3 within 6e4-1.2e5 iterations)"}
        log("steady", extra["steady_reference_defaults"])
        # Jacobi-PCG on the 1M-cell pinned pressure system
        sp = P.config_poisson(n)
        ellp = P.to_ell(sp)
        dsp = F.DeviceSystem(ellp)
        st_p = dsp.solve(F.PCG, tol=tol, **opts)
        st_p = dsp.solve(F.PCG, tol=tol, **opts)
        pb = st_p["iterations"] * pcg_iter_bytes(sp.neq, ellp["nad"]) + 2 * spmv_bytes(sp.neq, ellp["nad"])
        extra["pcg_pressure"] = {"iterations": st_p["iterations"], "ms_per_solve": st_p["seconds"] * 1e3,
                                 "solves_per_s": 1.0 / st_p["seconds"], "true_relres": st_p["true_relres"],
                                 "GBs_algorithmic": pb / st_p["seconds"] / 1e9,
                                 "us_per_iteration": st_p["seconds"] / st_p["iterations"] * 1e6}
        log("pcg", extra["pcg_pressure"])
        # mixed precision (FP32 inner PCG, FP64 residuals and refinement) against
        # FP64 on the 1M-cell lognormal-conductivity pressure system
        sl = P.config_pressure_lognormal(n)
        dsl = F.DeviceSystem(P.to_ell(sl))
        Al = P.scipy_csr(sl)
        mixed = {}
        for name, prec in (("fp64", F.PREC_FP64), ("mixed", F.PREC_MIXED)):
            dsl.solve(F.PCG, tol=tol, precision=prec, **opts)
            stl = dsl.solve(F.PCG, tol=tol, precision=prec, **opts)
            xl = dsl.x.get()
            mixed[name] = {"iterations": stl["iterations"], "refinements": stl["inner_iterations"],
                           "ms_per_solve": stl["seconds"] * 1e3, "solves_per_s": 1.0 / stl["seconds"],
                           "true_relres": float(np.linalg.norm(sl.b - Al @ xl) / np.linalg.norm(sl.b)),
                           "status": stl["status"], "_x": xl}
        x64 = mixed["fp64"].pop("_x")
        xm = mixed["mixed"].pop("_x")
        mixed["mixed"]["rel_err_vs_fp64"] = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))
        mixed["speedup_mixed_over_fp64"] = mixed["fp64"]["ms_per_solve"] / mixed["mixed"]["ms_per_solve"]
        extra["pcg_lognormal_fp64_vs_mixed"] = mixed
        log("mixed", mixed)
        # mixed-precision BiCGSTAB and GMRES(30) against FP64 on the headline
        # system and on the same operator at 2n x 2n cells (4M: beyond the
        # 256 MB Infinity Cache), all to true relative residual 1e-10
        mk = {}
        for nn in (n, 2 * n):
            if nn == n:
                dm, sm = ds, s
            else:
                sm = P.config_advdiff(nn)
                dm = F.DeviceSystem(P.to_ell(sm))
            Am = P.scipy_csr(sm)
            row = {"neq": sm.neq}
            # GMRES: mixed precision (FP32 basis) against FP64, with CGS2 and
            # with lagged CGS2
            for mname, code, fp64kw in (("bicgstab", F.BICGSTAB, {}),
                                        ("gmres30", F.GMRES, dict(gmres_orth=F.ORTH_CGS2))):
                res = {}
                runs = [("fp64", F.PREC_FP64, dict(stop_rule=F.STOP_RELRES, **fp64kw)),
                        ("mixed", F.PREC_MIXED, dict(fp64kw))]
                if code == F.GMRES:
                    runs += [("fp64_cgs2_lag", F.PREC_FP64, dict(stop_rule=F.STOP_RELRES, gmres_orth=F.ORTH_CGS2_LAG)),
                             ("mixed_cgs2_lag", F.PREC_MIXED, dict(gmres_orth=F.ORTH_CGS2_LAG))]
                for pname, prec, kw in runs:
                    skw = dict(tol=tol, x0_rule=F.X0_ZERO, gmres_k=30, precision=prec, **kw)
                    dm.solve(code, **skw)
                    stm = dm.solve(code, **skw)
                    xm_ = dm.x.get()
                    res[pname] = {"iterations": stm["iterations"], "inner_iterations": stm["inner_iterations"],
                                  "ms_per_solve": stm["seconds"] * 1e3, "solves_per_s": 1.0 / stm["seconds"],
                                  "true_relres": float(np.linalg.norm(sm.b - Am @ xm_) / np.linalg.norm(sm.b)),
                                  "status": stm["status"], "_x": xm_}
                x64_ = res["fp64"].pop("_x")
                xmx_ = res["mixed"].pop("_x")
                res["mixed"]["rel_err_vs_fp64"] = float(np.linalg.norm(xmx_ - x64_) / np.linalg.norm(x64_))
                res["speedup_mixed_over_fp64"] = res["fp64"]["ms_per_solve"] / res["mixed"]["ms_per_solve"]
                if code == F.GMRES:
                    xl_ = res["fp64_cgs2_lag"].pop("_x")
                    xml_ = res["mixed_cgs2_lag"].pop("_x")
                    res["mixed_cgs2_lag"]["rel_err_vs_fp64"] = float(np.linalg.norm(xml_ - xl_) / np.linalg.norm(xl_))
                    res["speedup_mixed_over_fp64_cgs2_lag"] = (res["fp64_cgs2_lag"]["ms_per_solve"]
                                                               / res["mixed_cgs2_lag"]["ms_per_solve"])
                row[mname] = res
            mk[f"{int(round(sm.neq ** 0.5))}x{int(round(sm.neq ** 0.5))}"] = row
        extra["advdiff_fp64_vs_mixed"] = mk
        log("mixed krylov", mk)
        # batching: 256 uneven pressure systems (sides log-uniform in [32, 512])
        # in one persistent launch, FP64 and mixed, against the same systems
        # solved one at a time in FP64 (all resident in HBM)
        systems = P.config_batch()
        fmts = [P.to_ell(q) for q in systems]
        layouts = {"ell": F.concat_ell(fmts), "csrd": F.concat_csrd([P.to_csrd(q) for q in systems])}
        batch = layouts["ell"]
        bres = {"systems": len(fmts), "rows": int(batch["row_offsets"][-1])}
        from fortranfvcell_amd import autotune
        sizes = [(int(f["neq"]), int(f["nad"])) for f in fmts]
        for name, prec in (("fp64", F.PREC_FP64), ("mixed", F.PREC_MIXED)):
            # layout (ELLPACK or CSRD), threads per block, block allocation
            # exponent and rows per wavefront by the project autotuner, timing
            # whole batched solves of this batch
            tb = autotune.tune_batch(layouts, precision=prec)
            tuned = tb["options"]
            db = F.DeviceBatch(layouts[tb["layout"]])
            db.solve(tol=tol, precision=prec, stop_rule=F.STOP_RELRES, **tuned)
            sts = db.solve(tol=tol, precision=prec, stop_rule=F.STOP_RELRES, **tuned)
            t = sts[0]["seconds"]
            # algorithmic bytes: every system's iterations at its own size (the
            # mixed refreshes each add an FP64 residual: SpMV + 5 vectors)
            if prec == F.PREC_FP64:
                nb = sum(q["iterations"] * pcg_ref_iter_bytes(ne, nz) for q, (ne, nz) in zip(sts, sizes))
            else:
                nb = sum(q["iterations"] * mixed_inner_bytes(ne, nz, bool(q["flags"] & F.FLAG_COL16))
                         + q["inner_iterations"] * (spmv_bytes(ne, nz) + 40 * ne) for q, (ne, nz) in zip(sts, sizes))
            bres[name] = {"ms": t * 1e3, "solves_per_s": len(sts) / t,
                          "max_iterations": max(q["iterations"] for q in sts),
                          "all_converged": all(q["status"] == 0 for q in sts), "tuned": tuned,
                          "layout": tb["layout"],
                          "trials_ms": {f"{q['layout']} {q['options']}": round(q["seconds"] * 1e3, 2)
                                        for q in tb["trials"]},
                          "algorithmic_bytes": nb, "GBs_algorithmic": nb / t / 1e9,
                          "frac": nb / t / 1e9 / HBM_PEAK_GBS}
        single = 0.0
        for f in fmts:
            d1 = F.DeviceSystem(f)
            single += d1.solve(F.PCG, tol=tol, **opts)["seconds"]
        bres["one_at_a_time_fp64"] = {"ms": single * 1e3, "solves_per_s": len(fmts) / single}
        extra["batch256"] = bres
        log("batch", bres)
        # the post-solve step on the 1M-cell quad mesh: the pressure-correction
        # gradient (gform) and simpleUpdate, on resident arrays
        extra["post_solve"] = bench_post(F, n)
        log("post", extra["post_solve"])
        # the drop-in as the Fortran program calls it: host arrays in, x out
        # (H2D of the matrix and vectors and D2H of x inside every solve)
        F.solve(ell, F.BICGSTAB, tol=tol, **opts)
        t0 = time.perf_counter()
        nd = 5
        for _ in range(nd):
            _, sth = F.solve(ell, F.BICGSTAB, tol=tol, **opts)
        dt = (time.perf_counter() - t0) / nd
        # the same with the matrix kept on the device (fvc_keep_matrix): the
        # second pressure-correction solve of a SIMPLE step, where only b changes
        t0 = time.perf_counter()
        for _ in range(nd):
            F.keep_matrix()
            F.solve(ell, F.BICGSTAB, tol=tol, **opts)
        dtk = (time.perf_counter() - t0) / nd
        extra["dropin_host_arrays"] = {"solves_per_s": 1.0 / dt, "ms_per_solve": dt * 1e3,
                                       "device_ms_per_solve": sth["seconds"] * 1e3,
                                       "kept_matrix_solves_per_s": 1.0 / dtk,
                                       "kept_matrix_ms_per_solve": dtk * 1e3,
                                       "note": "fvc_solve with host arrays: upload of A, b, x0 and download "
                                               "of x included (PCIe), or of b, x0 and x only with the matrix "
                                               "kept; value above is device-resident"}
        log("dropin", extra["dropin_host_arrays"])
        result["extra"] = extra

    if not args.no_cpu:
        A = P.scipy_csr(s)
        ltol = ladder_tol(n, "bicgstab")
        result["cpu_baseline"] = cpu_baseline(ell, A, s.b, "bicgstab", ltol)
        result["cpu_baseline"]["host_cpus"] = os.cpu_count()
        log("cpu", result["cpu_baseline"])
        if not args.no_extra:
            one = cpu_baseline(ell, A, s.b, "bicgstab", ltol, budget_s=5.0, threads=1)
            result.setdefault("extra", {})["cpu_reference_1thread"] = one
            log("cpu 1 thread", one)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
