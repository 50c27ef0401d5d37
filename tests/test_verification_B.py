"""PISM's regression test 12 (test/regression/test_12.sh): mass conservation
within rounding error on verification test B (Halfar's similarity solution,
isothermal SIA, zero SMB, moving margin). 31x31x31 grid, years 1000-6000,
max_dt 25 a, ice volume every 25 a; the reference's check is
diff(ivol).max() < 10^(floor(log10(max ivol)) - 14), i.e. no increase beyond
the 14th digit. Here on the oracle; tests/test_gpu_mass.py runs it on the HIP
path. The closed form itself is pinned by its own invariants."""
import math

import numpy as np
import pytest

import verification as V
import workloads as W


def run_oracle_test_b(oracle, Mx=31, t_start=1000.0, years=5000.0):
    grid, P, r, E0 = V.test_b_setup(Mx, Mx)
    n2 = Mx * Mx
    H = V.exactB(t_start * V.SECPERA, r)
    b = np.zeros(n2)
    mask, h = W.geometry_mask_surface(P, b, H)
    st = dict(H=H, b=b, h=h, mask=mask)
    E = np.full(n2 * grid.Mz, E0)
    zero = np.zeros(n2)
    ivol = []

    def step(t, tte):
        dt, _, _ = oracle.step(grid, P, st, E, full=True, smb=zero, time_to_end=tte)
        return dt

    steps = V.run_reporting_loop(step, lambda t: ivol.append(V.ice_volume(grid, st["H"])), t_start, years, 25.0)
    return np.array(ivol), st["H"], steps


def test_oracle_test_12_mass_conservation(oracle):
    ivol, H, steps = run_oracle_test_b(oracle)
    assert len(ivol) == 201 and steps >= 200
    threshold = 10.0 ** (math.floor(math.log10(ivol.max())) - 14)
    assert np.diff(ivol).max() < threshold
    assert np.all(H >= 0.0) and H.max() > 1000.0


def test_exact_B_closed_form():
    """exactB (exactTestsABCDE.c:56-77): H0 at the dome and margin R0 at t0;
    the volume of the similarity solution does not change with time
    (H ~ t^-1/9, R ~ t^1/18, so H R^2 ~ t^0)."""
    t0 = 422.45 * V.SPERA_EXACT
    H = V.exactB(t0, np.array([0.0, 749e3, 751e3]))
    assert H[0] == pytest.approx(3600.0, rel=1e-14)
    assert H[1] > 0 and H[2] == 0.0
    x = np.linspace(-1.2e6, 1.2e6, 1201)
    X, Y = np.meshgrid(x, x, indexing="ij")
    r = np.sqrt(X * X + Y * Y).ravel()
    v1 = V.exactB(1000 * V.SECPERA, r).sum()
    v2 = V.exactB(6000 * V.SECPERA, r).sum()
    assert abs(v2 - v1) < 2e-3 * v1
    assert V.exactB(6000 * V.SECPERA, np.array([0.0]))[0] < V.exactB(1000 * V.SECPERA, np.array([0.0]))[0]
