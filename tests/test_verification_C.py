"""Known answer of PISM's regression test 15 (verification test C, isothermal
SIA with time-dependent SMB, 5000 years; test/regression/test_15.sh:19-28):
the oracle reproduces every printed digit of the error table for 31x31 and
41x41 grids. This pins the whole SIA + mass-continuity + time-step path of
the oracle end to end against the reference's own expected output."""
import json
import os

import numpy as np
import pytest

import verification as V
import workloads as W

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "regression_outputs.json")))["test_C"]


def run_oracle_test_c(oracle, Mx):
    grid, P, r, E0 = V.test_c_setup(Mx, Mx)
    n2 = Mx * Mx
    H = np.zeros(n2)
    b = np.zeros(n2)
    mask, h = W.geometry_mask_surface(P, b, H)
    st = dict(H=H, b=b, h=h, mask=mask)
    E = np.full(n2 * grid.Mz, E0)

    def step(t, tte):
        _, M = V.exactC(t, r)
        dt, _, _ = oracle.step(grid, P, st, E, full=True, smb=M * P.ice_density, time_to_end=tte)
        return dt

    t, steps = V.run_time_loop(step)
    return V.geometry_errors(grid, st["H"], r, t), steps


@pytest.mark.parametrize("Mx,key", [(31, "Mx31"), (41, "Mx41")])
def test_oracle_reproduces_test_15_table(oracle, Mx, key):
    errs, steps = run_oracle_test_c(oracle, Mx)
    assert ["%.6f" % e for e in errs] == GOLD[key]
    assert steps == 84  # 83 steps of 60 a and the 20 a to the end of the run


def test_exact_C_closed_form():
    """exactC (exactTestsABCDE.c:80-107): margin R0 (t/t0)^2, dome height H0 t/t0."""
    t0 = 15208.0 * V.SPERA_EXACT
    H, M = V.exactC(t0, np.array([0.0, 749e3, 751e3]))
    assert H[0] == pytest.approx(3600.0, rel=1e-14)
    assert H[1] > 0 and H[2] == 0.0
    assert M[0] == pytest.approx(5.0 / t0 * 3600.0, rel=1e-14)
