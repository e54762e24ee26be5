"""Failure handling of the step, forced from the host: a non-finite value
planted in the velocity (a host write, no kernel is made to fault) must end
the step with BLZ_E_NUMERIC instead of stepping on silently. The reference
prints and exits on device errors (src/utils/helper_cuda.h:576-591) and has
no residual; here the residual check and the CFL max reduction map NaN to
+inf and the step reports it."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from helpers import random_state, settings  # noqa: E402

pytestmark = pytest.mark.gpu

B = pytest.importorskip("blaze_amd")


def nan_state(s, which=B.FIELD_U, value=np.nan):
    st = random_state(s, seed=3)
    a = st[which].copy()
    a[a.shape[0] // 2, a.shape[1] // 2, a.shape[2] // 2] = value
    st[which] = a
    return st


@pytest.mark.parametrize("value", [np.nan, np.inf], ids=["nan", "inf"])
def test_nonfinite_velocity_ends_residual_driven_solve(value):
    print(f"[failure] residual path {value}", flush=True)
    s = settings(B, (32, 24, 20), max_iter_solve=200, residual_tol=1e-4, residual_check_every=10)
    g = B.Solver(s)
    for w, a in nan_state(s, value=value).items():
        g.set_field(w, a)
    with pytest.raises(B.BlazeError) as ei:
        g.step(0.0)
    assert ei.value.code == B.E_NUMERIC
    assert "not finite" in str(ei.value)
    # the solve stopped at its first check instead of sweeping to the cap
    assert g.stats().rbgs_iters == 0


def test_nonfinite_velocity_caught_by_cfl_reduction():
    print("[failure] CFL path", flush=True)
    s = settings(B, (32, 24, 20), max_iter_solve=20, cfl_number=0.5, max_substeps=4)
    g = B.Solver(s)
    for w, a in nan_state(s, which=B.FIELD_W).items():
        g.set_field(w, a)
    with pytest.raises(B.BlazeError) as ei:
        g.step(0.0)
    assert ei.value.code == B.E_NUMERIC
    assert "velocity" in str(ei.value)


def test_finite_state_after_numeric_error_steps_again():
    """The handle stays usable: finite fields written back step normally."""
    print("[failure] recovery", flush=True)
    s = settings(B, (32, 24, 20), max_iter_solve=40, residual_tol=1e-4, residual_check_every=10)
    g = B.Solver(s)
    for w, a in nan_state(s).items():
        g.set_field(w, a)
    with pytest.raises(B.BlazeError):
        g.step(0.0)
    for w, a in random_state(s, seed=4).items():
        g.set_field(w, a)
    g.step(0.0)
    st = g.stats()
    assert np.isfinite(st.residual) and st.rbgs_iters > 0
