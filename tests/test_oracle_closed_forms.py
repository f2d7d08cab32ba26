"""The FP64 oracle against hand-derivable solver steps.

tests/golden/make_solver_closed_forms.py writes one step of each case in
40-digit decimal arithmetic, term by term from the reference's formulas
(DFSPH alpha, kappa, kappa_v, the velocity sweeps and the CFL rule; the
WCSPH pressure force and XSPH; the PCISPH pressure update and Spiky
pressure acceleration; Akinci cohesion, curvature and adhesion).  The
oracle must reproduce each to 1e-12, and each planted restatement mistake
(oracle_plant) must make the cases it concerns fail.
"""
import pytest

import oracle.pyoracle as po
from closed_form_cases import CASES, PLANT_CASES, check, oracle_runner

TOL = 1e-12


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_closed_form(name):
    errs = check(name, oracle_runner, TOL)
    assert errs


def test_every_plant_has_cases():
    assert sorted(PLANT_CASES) == sorted(p for p in po.PLANTS if p != "NONE")


@pytest.mark.parametrize("plant", sorted(PLANT_CASES))
def test_planted_mistake_fails_its_cases(plant):
    po.plant(plant)
    try:
        for name in PLANT_CASES[plant]:
            with pytest.raises(AssertionError):
                check(name, oracle_runner, TOL)
    finally:
        po.plant("NONE")
    # and the faithful oracle passes them again
    for name in PLANT_CASES[plant]:
        check(name, oracle_runner, TOL)
