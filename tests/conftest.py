import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs through libsphg's HIP kernels)")


def have_gpu() -> bool:
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def ctx():
    """One libsphg context on device 0.  On a machine with a GPU a failure
    here is an error, not a skip: the solver has no CPU path."""
    if not have_gpu():
        pytest.skip("no GPU in this container")
    from sph_amd import sphg
    c = sphg.Context(0)
    yield c
    c.close()


def rel_norm(a, b) -> float:
    """max |a - b| / max |b| over a field (normwise relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max() if b.size else 0.0
    num = np.abs(a - b).max() if b.size else 0.0
    return float(num / den) if den > 0 else float(num)


def rel_elem(a, b) -> float:
    """max |a - b| / |b| elementwise (for fields bounded away from zero)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.abs(b))) if b.size else 0.0


def dev_err(g_drho, o_rho, rho0, floor: float = 0.0) -> float:
    """Normwise relative error of a density deviation: the solver's stored
    rho - rho0 (FIELD_DRHO / FIELD_DRHOSTAR) against the oracle's FP64
    density minus rho0, max|a - b| / max(max|b|, floor).  The deviation, not
    the density near rho0, is what the pressure solvers run on.  `floor`
    (the density solve's own tolerance eta rho0 where one applies) keeps a
    block at rest, whose deviations are all near zero, from dividing by
    nothing: 1e-5 of eta rho0 = 1e-4 kg/m^3 at eta = 1e-4 is 1e-9 of rho0."""
    a = np.asarray(g_drho, dtype=np.float64)
    b = np.asarray(o_rho, dtype=np.float64) - float(rho0)
    if not b.size:
        return 0.0
    den = max(float(np.abs(b).max()), float(floor))
    num = float(np.abs(a - b).max())
    return num / den if den > 0 else num
