"""Row (f)1: the reference program's own SIMPLE time step, simple()
(src/Simple.f:103-900, calling solver at :488,499,590), run on a lid-driven
cavity twice -- linked with the reference's src/Solver.f
(oracle/_ref/libmvf_ref.so) and linked with the drop-in, whose solver_ runs
on the GPU (oracle/_ref/libmvf_dropin.so) -- with everything else the
reference's own compiled code.  After several time steps of several SIMPLE
iterations the velocity and pressure fields agree to the solver tolerance's
level: the two solvers stop on the same rule at tol = 1e-12 but sum in
different orders and, by default, the drop-in uses its fused recurrences."""
import os

import numpy as np
import pytest

import simple_program as S

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(S.REF_LIB) and os.path.exists(S.DROPIN_LIB)),
                                 reason="oracle/_ref/libmvf_{ref,dropin}.so are built only where the "
                                        "reference sources are (make -C oracle dropin refprog)")]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n,steps,its", [(24, 3, 4), (48, 2, 5)])
def test_simple_with_the_drop_in_matches_the_reference_solver(fvc, tmp_path, n, steps, its):
    cav = S.Cavity(n)
    print(f"simple() cavity {n}x{n}: {steps} steps x {its} SIMPLE iterations", flush=True)
    (tmp_path / "ref").mkdir()
    (tmp_path / "gpu").mkdir()
    # each program in a child process: a reference STOP must not end pytest
    ref = S.run_isolated(S.REF_LIB, n, steps, its, 0.05, str(tmp_path / "ref"))
    gpu = S.run_isolated(S.DROPIN_LIB, n, steps, its, 0.05, str(tmp_path / "gpu"))
    assert bool(gpu["fvcell_loaded"]) and not bool(ref["fvcell_loaded"])   # the drop-in's solves ran on the GPU
    errs = {k: _rel(gpu[k], ref[k]) for k in ("w", "p", "pC", "gradP")}
    print(f"relative differences {errs}; |w| max {np.abs(ref['w']).max():.3f}", flush=True)
    assert np.abs(ref["w"]).max() > 0.05           # the lid has set the cavity moving
    for k, e in errs.items():
        assert e <= 1e-8, (k, e)
