"""Known answer of PISM's regression test 14 (verification test E, isothermal
SIA with SIA-type sliding, 100 years; test/regression/test_14.sh:17-35): the
oracle reproduces every printed digit of the geometry and basal-velocity
error tables for 21x21 and 41x41 grids. This pins SIA_Sliding (its own
surface gradient and the test E sliding law), the sliding velocity in SIAFD's
u, v and in mass continuity (upwind flux, 2-D CFL), end to end against the
reference's own expected output.

The loop restates pismv's (iceModel.cc:601-794, iceCompModel.cc:799-812):
SIA_Sliding then SIAFD (StressBalance::update, PISMStressBalance.cc:89-96),
dt from the diffusivity and the 2-D CFL, mass continuity, surface and mask,
then H = 0 beyond 750 km (reset_thickness_tests_AE, iceCompModel.cc:478-497)
without a new surface/mask update.
"""
import json
import os

import numpy as np
import pytest

import verification as V
import workloads as W

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "regression_outputs.json")))["test_E"]


def run_oracle_test_e(oracle, Mx, years=100.0):
    grid, P, x, y, r, E0 = V.test_e_setup(Mx, Mx)
    X, Y = np.meshgrid(x, y, indexing="ij")
    H0, M, _, _, _ = V.exactE(X.ravel(), Y.ravel())
    b = np.zeros(Mx * Mx)
    mask, h = W.geometry_mask_surface(P, b, H0)
    st = dict(H=H0.copy(), b=b, h=h, mask=mask)
    E = np.full(Mx * Mx * grid.Mz, E0)
    smb = M * P.ice_density  # kg m-2 s-1 (PSVerification.cc:161-162, 169-220)
    last = {}

    def step(t, tte):
        ub, vb, _, _, _ = oracle.sia_sliding(grid, P, "test_E", x, y, st["H"], st["h"], st["b"], st["mask"], E)
        out = oracle.sia_update(grid, P, st["H"], st["b"], st["h"], st["mask"], E, full=True, ub=ub, vb=vb)
        dt = oracle.max_timestep(grid, P, out["D_max"], tte, ub=ub, vb=vb, mask=st["mask"])
        Hn, _ = oracle.mass_continuity(grid, P, dt, st["H"], out["flux"], st["mask"], smb=smb, ub=ub, vb=vb)
        mask, h = oracle.update_mask_surface(grid, P, st["b"], Hn)
        Hn[r > 750e3] = 0.0
        st.update(H=Hn, mask=mask, h=h)
        last["ub"], last["vb"] = ub, vb
        return dt

    t, steps = V.run_time_loop(step, years=years)
    geo = V.geometry_errors_E(grid, st["H"], x, y)
    vel = V.basal_velocity_errors(grid, st["H"], last["ub"], last["vb"], x, y)
    return geo, vel, steps


def fmt_vels(v):
    return ["%.4f" % v[0], "%.5f" % v[1], "%.5f" % v[2], "%.4f" % v[3], "%.4f" % v[4]]


@pytest.mark.parametrize("Mx,key", [(21, "Mx21"), (41, "Mx41")])
def test_oracle_reproduces_test_14_table(oracle, Mx, key):
    geo, vel, steps = run_oracle_test_e(oracle, Mx)
    assert ["%.6f" % e for e in geo] == GOLD[key]["geometry"]
    assert fmt_vels(vel) == GOLD[key]["base_vels"]


def test_exact_E_closed_form():
    """exactE (exactTestsABCDE.c:183-251): test A's profile outside the sliding
    region, no sliding there, ub = -mu rho g H h_x inside, reflection symmetry."""
    x = np.array([0.0, 300e3, -300e3, 300e3, 760e3])
    y = np.array([0.0, 100e3, 100e3, -100e3, 0.0])
    H, M, mu, ub, vb = V.exactE(x, y)
    assert H[0] > 3000 and H[4] == 0.0 and mu[0] == 0.0 and ub[0] == 0.0
    assert mu[1] > 0 and ub[1] > 0 and vb[1] > 0          # flow away from the dome
    assert ub[2] == -ub[1] and vb[2] == vb[1] and vb[3] == -vb[1]
    assert H[1] == H[2] == H[3]
