"""The driver of the reference program's simple() (tests/simple_program.py)
on the CPU, with the reference's own solver: it runs, and the cavity behaves
as a lid-driven cavity must -- the flow follows the lid under it, turns back
along the bottom, and each time step drives the mass residual of the
pressure correction down across its SIMPLE iterations."""
import os

import numpy as np
import pytest

import simple_program as S

pytestmark = pytest.mark.skipif(not os.path.exists(S.REF_LIB),
                                reason="oracle/_ref/libmvf_ref.so is built only where the reference sources are")


def test_reference_simple_runs_a_cavity(tmp_path):
    cav = S.Cavity(16)
    r = S.run_isolated(S.REF_LIB, 16, 3, 4, 0.05, str(tmp_path))
    assert all(np.all(np.isfinite(v)) for v in r.values())
    xc = cav.x[cav.ix[:, :4] - 1].mean(axis=1)
    top = xc[:, 1] > xc[:, 1].max() - 1e-9 - (xc[:, 1].max() - xc[:, 1].min()) / 15
    bottom = xc[:, 1] < xc[:, 1].min() + (xc[:, 1].max() - xc[:, 1].min()) / 15
    assert r["w"][top, 0].mean() > 0.1                 # dragged along by the lid
    assert r["w"][bottom, 0].mean() < 0.0              # the return flow
    assert np.abs(r["rCpC"]).max() < 1e-2              # mass residual of the last p' assembly
