"""Verification tests B, C and E of pismv on the HIP path: the known-answer
end-to-end checks of SIA + mass continuity (C: isothermal SIA with
time-dependent SMB), of the SIA-type sliding path (E: test_14.sh) and the
mass-conservation invariant of test_12.sh (B: Halfar's solution, zero SMB).

Restated from the reference (paths relative to its root):
* exactC, the Bueler et al. (2005) similarity solution with lambda = 5
  (src/verif/tests/exactTestsABCDE.c:80-107), and exactB, the lambda = 0
  member (Halfar 1983; exactTestsABCDE.c:56-77);
* the test B domain: Lx = 1200 km, Lz = 4000 m (src/pismv.cc:91-97), the rest
  as test C (initTestABCDEH, src/verif/iceCompModel.cc:274-280, 302-377);
* test_12.sh: -Mx 31 -My 31 -Mz 31 -ys 1000 -y 5000 -max_dt 25, ice volume
  every 25 a from year 1000 (test/regression/test_12.sh:14-20), and its check
  diff(ivol).max() < 10^(floor(log10 max ivol) - 14) (test_12.sh:27-44);
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


def exactB(t, r):
    """exactTestsABCDE.c:56-77 (vectorised over r): H of Halfar's solution,
    alpha = 1/9, beta = 1/18, t0 = 422.45 a of the closed form's year; M = 0."""
    n, H0, R0 = 3.0, 3600.0, 750000.0
    alpha, beta = 1.0 / 9.0, 1.0 / 18.0
    t0 = 422.45 * SPERA_EXACT
    r = np.asarray(r, dtype=np.float64)
    Rmargin = R0 * math.pow(t / t0, beta)
    H = np.zeros_like(r)
    inside = r < Rmargin
    if inside.any():
        H[inside] = H0 * math.pow(t / t0, -alpha) * np.power(
            1.0 - np.power(math.pow(t / t0, -beta) * (r[inside] / R0), (n + 1) / n), n / (2 * n + 1))
    return H


def test_b_setup(Mx, My, Mz=31, max_dt_years=25.0):
    """pismv -test B: test C's setup on the 2400 km domain (pismv.cc:91-97),
    with -max_dt (the maximum_time_step parameter, in seconds)."""
    grid, P, r, E0 = test_c_setup(Mx, My, Mz, Lx=1200e3)
    P.maximum_time_step = max_dt_years * SECPERA
    return grid, P, r, E0


def ice_volume(grid, H):
    """IceModel::ice_volume: the sum of H dx dy over the cells (m^3)."""
    return float(np.sum(H)) * grid.dx * grid.dy


def run_reporting_loop(step_fn, report_fn, t_start_years, years, every_years):
    """The run loop from t_start for `years`, every step ending no later than
    the next reporting time (t_start + k every_years; reporting times cap the
    step as -ts_times does, IceModel::max_timestep). step_fn(t, time_to_end)
    takes one step and returns its dt; report_fn(t) runs at the start and at
    each reporting time. Returns the number of steps."""
    t = t_start_years * SECPERA
    t_end = (t_start_years + years) * SECPERA
    report_fn(t)
    k, steps = 1, 0
    while t < t_end:
        t_mark = min(t_end, (t_start_years + k * every_years) * SECPERA)
        t += step_fn(t, t_mark - t)
        steps += 1
        if t_mark - t < 1e-3:
            t = t_mark
            report_fn(t)
            k += 1
    return steps


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


def exactE(x, y):
    """exactTestsABCDE.c:183-251 (vectorised). Returns (H, M, mu, ub, vb) in
    SI units; the closed form's year is SperA = 31556926 s."""
    L, M0 = 750000.0, 0.3 / SPERA_EXACT
    g, rho, n = 9.81, 910.0, 3.0
    A = 1.0e-16 / SPERA_EXACT
    Gamma = 2 * (rho * g) ** n * A / (n + 2)
    mu_max, r1, r2 = 2.5e-11, 200e3, 700e3
    theta1, theta2 = 10 * (math.pi / 180), 40 * (math.pi / 180)
    rbot, thetabot = (r2 - r1) * (r2 - r1), (theta2 - theta1) * (theta2 - theta1)
    xa, ya = np.abs(x), np.abs(y)
    sgnx = np.where(np.asarray(x) < 0, -1.0, 1.0)
    sgny = np.where(np.asarray(y) < 0, -1.0, 1.0)
    r = np.sqrt(xa * xa + ya * ya)
    m, q = 2.0 * n + 2.0, 1.0 + 1.0 / n
    C = ((2.0 ** (n - 1)) * M0 / Gamma) ** (1.0 / m)
    with np.errstate(divide="ignore", invalid="ignore"):
        chi = L ** q - r ** q
        H = np.where(r < L, C * np.power(np.maximum(chi, 0.0), n / m), 0.0)
        theta = np.where(xa < 1.0, math.pi / 2.0, np.arctan(ya / np.where(xa < 1.0, 1.0, xa)))
        inside = (r < L) & ~((r <= r1) | (r >= r2) | (theta <= theta1) | (theta >= theta2))
        mufactor = mu_max * (4.0 * (theta - theta1) * (theta2 - theta) / thetabot)
        mu = np.where(inside, mufactor * (4.0 * (r - r1) * (r2 - r) / rbot), 0.0)
        dchidr = -q * np.power(r, 1.0 / n)
        dhdr = ((n * C) / m) * np.power(np.where(inside, chi, 1.0), (n - m) / m) * dchidr
        P = rho * g * H
        hx, hy = dhdr * np.cos(theta), dhdr * np.sin(theta)
        ub = np.where(inside, sgnx * (-mu * P * hx), 0.0)
        vb = np.where(inside, sgny * (-mu * P * hy), 0.0)
        d2hdr2 = dhdr * (-((n + 2.0) / m) * dchidr / np.where(inside, chi, 1.0) + 1.0 / (n * r))
        dmudr = mufactor * 4.0 * (r1 + r2 - 2.0 * r) / rbot
        Mb = -P * ((mu / r) * H * dhdr + dmudr * H * dhdr + 2.0 * mu * dhdr * dhdr + mu * H * d2hdr2)
        M = np.where(inside, M0 + Mb, M0)
    return H, M, mu, ub, vb


def test_e_setup(Mx, My, Mz=3, Lx=800e3, Lz=4000.0):
    """pismv -test E (pismv.cc:84-88; iceCompModel.cc:196-213, 302-377): the
    test C settings on an 800 km domain, sliding from SIA_Sliding in
    verification mode; returns (grid, params, x, y, r, E0)."""
    x = T.horizontal_coordinates(Lx, Mx)
    y = T.horizontal_coordinates(Lx, My)
    z = T.vertical_levels(Lz, Mz, "quadratic", 4.0)
    grid = T.make_grid(Mx, My, T.horizontal_spacing(Lx, Mx), T.horizontal_spacing(Lx, My), z)
    P = T.make_params(sia_flow_law="arr", sia_enhancement_factor=1.0, bed_smoother_range=0.0,
                      is_dry_simulation=1, include_bmr_in_continuity=0, enthalpy_converter="cold")
    A0 = 1.0e-16 / SECPERA
    T0 = -P.Paterson_Budd_Q_cold / (P.ideal_gas_constant * (math.log(A0) - math.log(P.Paterson_Budd_A_cold)))
    E0 = P.ice_specific_heat_capacity * (T0 - P.enthalpy_converter_reference_temperature)
    X, Y = np.meshgrid(x, y, indexing="ij")
    return grid, P, x, y, np.sqrt(X * X + Y * Y).ravel(), E0


def geometry_errors_E(grid, H, x, y):
    """computeGeometryErrors with exactE (iceCompModel.cc:595-744)."""
    X, Y = np.meshgrid(x, y, indexing="ij")
    Hexact = exactE(X.ravel(), Y.ravel())[0]
    return _geometry_errors(grid, H, Hexact)


def basal_velocity_errors(grid, H, ub, vb, x, y):
    """IceCompModel::computeBasalVelocityErrors (iceCompModel.cc:744-796) and the
    printed line (iceCompModel.cc:1115-1127): maxvector, avvector (m/a),
    prcntavvec, maxub, maxvb (m/a); the year of m/a is the model's."""
    X, Y = np.meshgrid(x, y, indexing="ij")
    _, _, _, ube, vbe = exactE(X.ravel(), Y.ravel())
    icy = H > 0.0
    uberr = np.abs(ub - ube)[icy]
    vberr = np.abs(vb - vbe)[icy]
    vecerr = np.sqrt(uberr * uberr + vberr * vberr)
    avvecerr = float(np.sum(vecerr)) / (grid.Mx * grid.My)
    xpeak, ypeak = 450e3 * math.cos(25.0 * (math.pi / 180.0)), 450e3 * math.sin(25.0 * (math.pi / 180.0))
    _, _, _, ubp, vbp = exactE(np.array([xpeak]), np.array([ypeak]))
    exactmaxspeed = math.sqrt(ubp[0] ** 2 + vbp[0] ** 2)
    return (float(vecerr.max()) * SECPERA, avvecerr * SECPERA, (avvecerr / exactmaxspeed) * 100.0,
            float(uberr.max()) * SECPERA, float(vberr.max()) * SECPERA)


def _geometry_errors(grid, H, Hexact):
    Mx, My = grid.Mx, grid.My
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


def run_test_b_gpu(Mx=31, t_start_years=1000.0, years=5000.0, every_years=25.0):
    """test_12.sh's run of test B through the HIP library (C ABI): full SIA
    updates, dt (at most 25 a, capped at the reporting times), mass continuity
    with zero SMB. Returns (ice volumes at the reporting times, final H, steps)."""
    import pism_sia as S
    import workloads as W
    grid, P, r, E0 = test_b_setup(Mx, Mx)
    n2 = Mx * Mx
    H0 = exactB(t_start_years * SECPERA, r)
    b = np.zeros(n2)
    mask, h = W.geometry_mask_surface(P, b, H0)
    dev = dict(H=S.to_device(H0), b=S.to_device(b), h=S.to_device(h), mask=S.to_device(mask),
               E=S.to_device(np.full(n2 * grid.Mz, E0)))
    sia = S.SIAFD(grid, P)
    geom = S.geometry(dev["H"], dev["h"], dev["mask"], dev["b"], smb=S.to_device(np.zeros(n2)))
    ivol = []

    def step(t, tte):
        sia.update(dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], fast=False)
        dt = sia.max_timestep(tte)
        sia.mass_continuity_step(geom, dt, accounting=False)
        return dt

    steps = run_reporting_loop(step, lambda t: ivol.append(ice_volume(grid, dev["H"].cpu().numpy())),
                               t_start_years, years, every_years)
    Hf = dev["H"].cpu().numpy()
    sia.close()
    return np.array(ivol), Hf, steps


def run_test_e_gpu(Mx, years=100.0):
    """Test E through the HIP library (C ABI): SIA_Sliding (sia_sliding_update),
    SIAFD with the sliding velocity, dt, mass continuity, then H = 0 beyond
    750 km (iceCompModel.cc:478-497). Returns (geometry errors, base-velocity
    errors, steps)."""
    import pism_sia as S
    import torch
    import workloads as W
    grid, P, x, y, r, E0 = test_e_setup(Mx, Mx)
    X, Y = np.meshgrid(x, y, indexing="ij")
    H0, M, _, _, _ = exactE(X.ravel(), Y.ravel())
    b = np.zeros(Mx * Mx)
    mask, h = W.geometry_mask_surface(P, b, H0)
    dev = dict(H=S.to_device(H0), b=S.to_device(b), h=S.to_device(h), mask=S.to_device(mask),
               E=S.to_device(np.full(Mx * Mx * grid.Mz, E0)))
    smb = S.to_device(M * P.ice_density)
    outside = torch.from_numpy(r > 750e3).to("cuda")
    sia = S.SIAFD(grid, P)
    geom = S.geometry(dev["H"], dev["h"], dev["mask"], dev["b"], smb=smb)
    last = {}

    def step(t, tte):
        ub, vb, _ = sia.sliding(dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], mode="test_E", x=x, y=y)
        sia.update(dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], fast=False, ub=ub, vb=vb)
        dt = sia.max_timestep(tte)
        sia.mass_continuity_step(geom, dt, accounting=False)
        dev["H"][outside] = 0.0
        last["ub"], last["vb"] = ub, vb
        return dt

    t, steps = run_time_loop(step, years)
    Hf = dev["H"].cpu().numpy()
    geo = geometry_errors_E(grid, Hf, x, y)
    vel = basal_velocity_errors(grid, Hf, last["ub"].cpu().numpy(), last["vb"].cpu().numpy(), x, y)
    sia.close()
    return geo, vel, steps
