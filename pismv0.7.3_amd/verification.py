"""Verification test C of pismv (isothermal SIA with time-dependent SMB) on the
HIP path: the known-answer end-to-end check of SIA + mass continuity.

Restated from the reference (paths relative to its root):
* exactC, the Bueler et al. (2005) similarity solution with lambda = 5
  (src/verif/tests/exactTestsABCDE.c:80-107);
* the test C setup of pismv / IceCompModel: Lx = 1000 km, Lz = 4000 m, EQUAL
  levels, non-periodic grid (src/pismv.cc:64-136); zero bed, arr flow law with
  T chosen so that A = 1e-16 Pa-3 a-1, enhancement 1, no bed smoothing, dry,
  cold enthalpy converter, no basal melt in continuity
  (src/verif/iceCompModel.cc:55-81, 125-154, 302-377);
* SMB = exactC's M at the start of each step (src/verif/PSVerification.cc:169-220);
* the run loop: dt = min(60 a, end of run, diffusivity limit), Time::step
  snapping to the end (src/base/util/PISMTime.cc:212-221);
* the error norms printed by pismv (src/verif/iceCompModel.cc:595-744, 965-972).
"""
import math

import numpy as np

import sia_types as T

SECPERA = 3.15569259747e7  # IceCompModel::secpera (iceCompModel.cc:53), the model clock
SPERA_EXACT = 31556926.0   # SperA of the closed forms (exactTestsABCDE.c:26)


def exactC(t, r):
    """exactTestsABCDE.c:80-107 (vectorised over r). Returns (H, M), M in m/s.
    Note the closed form's own year (SperA = 31556926 s) differs from the
    model's (3.15569259747e7 s); t is model seconds, as in the reference."""
    n, H0, R0 = 3.0, 3600.0, 750000.0
    lam, alpha, beta = 5.0, -1.0, 2.0
    t0 = 15208.0 * SPERA_EXACT
    r = np.asarray(r, dtype=np.float64)
    Rmargin = R0 * math.pow(t / t0, beta)
    H = np.zeros_like(r)
    inside = r < Rmargin
    if inside.any():
        H[inside] = H0 * math.pow(t / t0, -alpha) * np.power(
            1.0 - np.power(math.pow(t / t0, -beta) * (r[inside] / R0), (n + 1) / n), n / (2 * n + 1))
    if t > 0.1 * SPERA_EXACT:
        M = (lam / t) * H
    else:
        Rm = R0 * math.pow(0.1 * SPERA_EXACT / t0, beta)
        M = np.where(r < Rm, 5 * H0 / t0, 0.0)
    return H, M


def test_c_setup(Mx, My, Mz=31, Lx=1000e3, Lz=4000.0):
    """pismv -test C grid: the test defaults are replaced by
    vertical_grid_from_options (pismv.cc:152-157, IceGrid.cc:1232-1240), so the
    levels use the configured grid_ice_vertical_spacing = quadratic with
    grid_lambda = 4 (pism_config.cdl:1289, 1334)."""
    dx = T.horizontal_spacing(Lx, Mx)
    dy = T.horizontal_spacing(Lx, My)
    x = T.horizontal_coordinates(Lx, Mx)
    y = T.horizontal_coordinates(Lx, My)
    z = T.vertical_levels(Lz, Mz, "quadratic", 4.0)
    grid = T.make_grid(Mx, My, dx, dy, z)
    P = T.make_params(sia_flow_law="arr", sia_enhancement_factor=1.0, bed_smoother_range=0.0,
                      is_dry_simulation=1, include_bmr_in_continuity=0, enthalpy_converter="cold")
    # T0 from PatersonBuddCold::tempFromSoftness (flowlaws.hh:279-281), A0 = 1e-16 / secpera
    A0 = 1.0e-16 / SECPERA
    T0 = -P.Paterson_Budd_Q_cold / (P.ideal_gas_constant * (math.log(A0) - math.log(P.Paterson_Budd_A_cold)))
    # compute_enthalpy_cold with the cold converter: E = c_i (T - T_0) (enthalpyConverter.cc:331-335)
    E0 = P.ice_specific_heat_capacity * (T0 - P.enthalpy_converter_reference_temperature)
    X, Y = np.meshgrid(x, y, indexing="ij")
    r = np.sqrt(X * X + Y * Y)  # radius(), IceGrid.cc:966-968
    return grid, P, r.ravel(), E0


def geometry_errors(grid, H, r, t):
    """IceCompModel::computeGeometryErrors for test C (iceCompModel.cc:595-744)
    and the printed ratios (965-972): prcntVOL, maxH, avH, relmaxETA."""
    Mx, My = grid.Mx, grid.My
    Hexact, _ = exactC(t, r)
    a = grid.dx * grid.dy * 1e-3 * 1e-3
    m = (2.0 * 3.0 + 2.0) / 3.0
    vol = float(np.sum(a * H[H > 0] * 1e-3))
    volexact = float(np.sum(a * Hexact[Hexact > 0] * 1e-3))
    Herr = float(np.max(np.abs(H - Hexact)))
    etaerr = float(np.max(np.abs(np.power(H, m) - np.power(Hexact, m))))
    avHerr = float(np.sum(np.abs(H - Hexact))) / (Mx * My)
    c = ((Mx - 1) // 2) * My + (My - 1) // 2
    domeHexact = float(Hexact[c])
    volerr = abs(vol - volexact)
    return (100 * volerr / volexact, Herr, avHerr, etaerr / math.pow(domeHexact, m))


def run_time_loop(step_fn, years=5000.0, max_dt_years=60.0):
    """IceModel::run's loop (iceModel.cc:853) with Time::step (PISMTime.cc:212-221).
    step_fn(t, time_to_end) performs one model step and returns the dt taken."""
    t, t_end = 0.0, years * SECPERA
    steps = 0
    while t < t_end:
        dt = step_fn(t, t_end - t)
        t += dt
        if t_end > t and t_end - t < 1e-3:
            t = t_end
        steps += 1
    return t, steps


def run_test_c_gpu(Mx, My, Mz=31, years=5000.0):
    """Test C through the HIP library (C ABI). Returns (errors, steps, H)."""
    import pism_sia as S
    import workloads as W
    grid, P, r, E0 = test_c_setup(Mx, My, Mz)
    n2 = Mx * My
    H = np.zeros(n2)
    b = np.zeros(n2)
    mask, h = W.geometry_mask_surface(P, b, H)
    dev = dict(H=S.to_device(H), b=S.to_device(b), h=S.to_device(h), mask=S.to_device(mask),
               E=S.to_device(np.full(n2 * Mz, E0)))
    smb_dev = S.to_device(np.zeros(n2))
    sia = S.SIAFD(grid, P)
    geom = S.geometry(dev["H"], dev["h"], dev["mask"], dev["b"], smb=smb_dev)

    def step(t, tte):
        sia.update(dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], fast=False)
        dt = sia.max_timestep(tte)
        _, M = exactC(t, r)
        smb_dev.copy_(S.to_device(M * P.ice_density))  # kg m-2 s-1 (PSVerification.cc:161-162)
        sia.mass_continuity_step(geom, dt, accounting=False)
        return dt

    t, steps = run_time_loop(step, years)
    Hf = dev["H"].cpu().numpy()
    return geometry_errors(grid, Hf, r, t), steps, Hf
