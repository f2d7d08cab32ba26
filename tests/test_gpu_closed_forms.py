"""The HIP solver against the hand-derivable solver steps of
tests/golden/solver_closed_forms.json: the same known answers the FP64
oracle reproduces to 1e-12 (tests/test_oracle_closed_forms.py), here held
to 1e-5 relative (normwise per field), the FP32-state bar; the time steps
are formed in FP64 on the device and held to 1e-12."""
import pytest

from closed_form_cases import CASES, check, gpu_runner

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("name", CASES)
def test_hip_matches_closed_form(ctx, name):
    errs = check(name, gpu_runner(ctx), TOL, tol_dt=1e-12)
    print(name, {k: f"{v:.1e}" for k, v in errs.items()})
