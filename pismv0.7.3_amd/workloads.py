"""Synthetic SIA workloads (numpy), in PISM's field layout.

* ``dome``: the 61x61x31 radial dome on a flat bed (Vialov profile
  H = 3600 (1 - (r/750 km)^(4/3))^(3/8)), isothermal Glen n = 3,
  A = 1e-16 Pa-3 a-1, dx = 25 km, 31 equal levels to 4000 m.
* ``greenland_like``: the 1501x2801x401 polythermal sheet at dx = 1 km: a bed of
  200 seeded Gaussian bumps, a smooth seeded dome (max 3200 m) with ice-free
  margins, cold interior columns above a temperate basal layer, GPBLD, e = 3.

Every random number comes from a counter-based hash of (seed, stream, index),
so any window of the large grid can be generated on its own and the HIP
generator (csrc/workload.hip) produces the same values on the device.
"""
import math

import numpy as np

import sia_types as T

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G1 = np.uint64(0x9E3779B97F4A7C15)
_G2 = np.uint64(0xD1B54A32D192ED03)


def _mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform01(seed, stream, idx):
    """U[0,1) from the splitmix64 finaliser of seed*G1 + stream*G2 + idx."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = np.uint64(seed) * _G1 + np.uint64(stream) * _G2 + idx
    return (_mix64(key) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# streams of the counter-based generator
S_BUMP_A, S_BUMP_W, S_BUMP_CX, S_BUMP_CY, S_DOME, S_TEMPERATE, S_OMEGA, S_TSURF = range(8)


class Workload:
    """Grid + parameters + fields (flat float64 arrays in PISM layout)."""

    def __init__(self, name, grid, params, fields, full_shape, window):
        self.name = name
        self.grid = grid
        self.params = params
        self.fields = fields          # H, b, h, mask, E
        self.full_shape = full_shape  # (Mx, My, Mz) of the whole workload
        self.window = window          # (i0, i1, j0, j1)


def geometry_mask_surface(params, b, H):
    """GeometryCalculator::compute (Mask.hh:82-118), vectorised."""
    alpha = 1 - params.ice_density / params.sea_water_density
    hgrounded = b + H
    hfloating = params.sea_level + alpha * H
    is_floating = hfloating > hgrounded + params.mask_is_floating_thickness_standard
    ice_free = H < params.mask_icefree_thickness_standard
    floating = is_floating & (not params.is_dry_simulation)
    h = np.where(floating, hfloating, hgrounded)
    mask = np.where(floating, np.where(ice_free, T.MASK_ICE_FREE_OCEAN, T.MASK_FLOATING),
                    np.where(ice_free, T.MASK_ICE_FREE_BEDROCK, T.MASK_GROUNDED)).astype(np.float64)
    return mask, h


# ------------------------------------------------------------------ dome
DOME_A = 1.0e-16 / T.SECONDS_PER_YEAR  # EISMINT / test A softness, exactTestsABCDE.c:35


def dome(Mx=61, My=61, L=750e3, Mz=31, Lz=4000.0, H0=3600.0, R=750e3, **param_overrides):
    dx = T.horizontal_spacing(L, Mx)
    dy = T.horizontal_spacing(L, My)
    x = T.horizontal_coordinates(L, Mx)
    y = T.horizontal_coordinates(L, My)
    z = T.vertical_levels(Lz, Mz)
    X, Y = np.meshgrid(x, y, indexing="ij")
    r = np.sqrt(X * X + Y * Y)
    H = np.where(r < R, H0 * np.power(np.maximum(1.0 - np.power(r / R, 4.0 / 3.0), 0.0), 3.0 / 8.0), 0.0)
    opts = dict(sia_flow_law="isothermal_glen", ice_softness=DOME_A, sia_enhancement_factor=1.0,
                is_dry_simulation=1)
    opts.update(param_overrides)
    P = T.make_params(**opts)
    b = np.zeros_like(H)
    mask, h = geometry_mask_surface(P, b, H)
    # isothermal Glen ignores enthalpy; use a cold constant (-20 C)
    E = np.full((Mx, My, Mz), P.ice_specific_heat_capacity * (253.15 - P.enthalpy_converter_reference_temperature))
    grid = T.make_grid(Mx, My, dx, dy, z)
    fields = dict(H=H.ravel(), b=b.ravel(), h=h.ravel(), mask=mask.ravel(), E=E.ravel())
    return Workload("dome", grid, P, fields, (Mx, My, Mz), (0, Mx, 0, My))


def vialov_smb(params, H0=3600.0, L=750e3, n=3.0):
    """Steady accumulation of the Vialov profile (exactTestsABCDE.c:28-53):
    M0 = Gamma (H0/sqrt(L))^(2n+2) / 2^(n-1), Gamma = 2 (rho g)^n A / (n+2)."""
    rho, g = params.ice_density, params.standard_gravity
    Gamma = 2 * (rho * g) ** n * params.ice_softness / (n + 2)
    return Gamma * (H0 / math.sqrt(L)) ** (2 * n + 2) / 2 ** (n - 1)


# ------------------------------------------------------------------ greenland-like
# z_spacing: "equal" (the benchmark sheet) or "quadratic" (PISM's default,
# grid_ice_vertical_spacing, cdl:1289, with grid_lambda = 4, cdl:1334);
# dy: None for dy = dx (the benchmark sheet)
GL_DEFAULT = dict(Mx=1501, My=2801, dx=1000.0, Mz=401, Lz=4000.0, seed=20261203,
                  n_bumps=200, Hmax=3200.0, z_spacing="equal", dy=None)


def gl_params(**overrides):
    opts = dict(sia_flow_law="gpbld", sia_enhancement_factor=3.0, sia_Glen_exponent=3.0,
                is_dry_simulation=1)
    opts.update(overrides)
    return T.make_params(**opts)


def gl_bumps(seed, n_bumps, Lx, Ly):
    m = np.arange(n_bumps, dtype=np.uint64)
    a = -500.0 + 2000.0 * uniform01(seed, S_BUMP_A, m)
    w = 5e3 + 75e3 * uniform01(seed, S_BUMP_W, m)
    cx = -Lx + 2 * Lx * uniform01(seed, S_BUMP_CX, m)
    cy = -Ly + 2 * Ly * uniform01(seed, S_BUMP_CY, m)
    return a, w, cx, cy


def gl_dome_shape(seed):
    u = uniform01(seed, S_DOME, np.arange(4, dtype=np.uint64))
    return dict(x0f=-0.05 + 0.1 * u[0], y0f=-0.05 + 0.1 * u[1],
                ph1=2 * math.pi * u[2], ph2=2 * math.pi * u[3])


def greenland_like(i0=0, i1=None, j0=0, j1=None, **kw):
    """Window [i0,i1) x [j0,j1) of the 1501x2801x401 polythermal workload."""
    cfg = dict(GL_DEFAULT)
    cfg.update({k: v for k, v in kw.items() if k in cfg})
    poverrides = {k: v for k, v in kw.items() if k not in cfg}
    Mx, My, dx, Mz, Lz, seed = cfg["Mx"], cfg["My"], cfg["dx"], cfg["Mz"], cfg["Lz"], cfg["seed"]
    dy = dx if cfg["dy"] is None else cfg["dy"]
    i1 = Mx if i1 is None else i1
    j1 = My if j1 is None else j1
    P = gl_params(**poverrides)
    Lx, Ly = 0.5 * (Mx - 1) * dx, 0.5 * (My - 1) * dy
    ii = np.arange(i0, i1)
    jj = np.arange(j0, j1)
    x = -Lx + ii * dx
    y = -Ly + jj * dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    # bed
    a, w, cx, cy = gl_bumps(seed, cfg["n_bumps"], Lx, Ly)
    b = np.zeros_like(X)
    for m in range(cfg["n_bumps"]):
        d2 = (X - cx[m]) ** 2 + (Y - cy[m]) ** 2
        b += a[m] * np.exp(-d2 / (2.0 * w[m] * w[m]))
    # thickness: perturbed elliptical Vialov-type dome
    s = gl_dome_shape(seed)
    Rx, Ry = 0.78 * Lx, 0.78 * Ly
    xn = (X - s["x0f"] * Lx) / Rx
    yn = (Y - s["y0f"] * Ly) / Ry
    rho = np.sqrt(xn * xn + yn * yn)
    phi = np.arctan2(yn, xn)
    sc = 1.0 + 0.06 * np.sin(3.0 * phi + s["ph1"]) + 0.04 * np.sin(5.0 * phi + s["ph2"])
    r = rho / sc
    H = np.where(r < 1.0, cfg["Hmax"] * np.power(np.maximum(1.0 - np.power(r, 4.0 / 3.0), 0.0), 0.375), 0.0)
    mask, h = geometry_mask_surface(P, b, H)
    # enthalpy: temperate basal layer (fraction U[0,0.1] of H, omega U[0,0.01]),
    # cold above with T from T_s at the surface to T_m(p_cts) - 1 K at the CTS
    col = (ii[:, None].astype(np.uint64) * np.uint64(My) + jj[None, :].astype(np.uint64))
    f_t = 0.1 * uniform01(seed, S_TEMPERATE, col)
    omega = 0.01 * uniform01(seed, S_OMEGA, col)
    T_s = 238.15 + 10.0 * uniform01(seed, S_TSURF, col)
    z = T.vertical_levels(Lz, Mz, cfg["z_spacing"])
    E = enthalpy_columns(P, H, f_t, omega, T_s, z)
    grid = T.make_grid(i1 - i0, j1 - j0, dx, dy, z)
    fields = dict(H=H.ravel(), b=b.ravel(), h=h.ravel(), mask=mask.ravel(), E=E.ravel())
    return Workload("greenland_like", grid, P, fields, (Mx, My, Mz), (i0, i1, j0, j1))


def enthalpy_columns(P, H, f_t, omega, T_s, z):
    """Polythermal enthalpy E[i,j,k] with the default EnthalpyConverter formulas
    (enthalpyConverter.cc:92-143, 256-281)."""
    ci, T0, L = P.ice_specific_heat_capacity, P.enthalpy_converter_reference_temperature, P.water_latent_heat_fusion
    rho_g = P.ice_density * P.standard_gravity
    beta, Tmelt, pair = P.beta_CC, P.water_melting_point_temperature, P.surface_pressure
    z_t = f_t * H
    p_t = np.where(H - z_t > 0.0, pair + rho_g * (H - z_t), pair)
    T_base = Tmelt - beta * p_t - 1.0
    E_air = ci * (T_s - T0)
    E = np.empty(H.shape + (len(z),))
    denom = np.where(H - z_t > 0.0, H - z_t, 1.0)
    for k, zk in enumerate(z):
        depth = H - zk
        p = np.where(depth > 0.0, pair + rho_g * depth, pair)
        E_cts = ci * ((Tmelt - beta * p) - T0)
        E_temp = E_cts + omega * L
        frac = np.clip(depth / denom, 0.0, 1.0)
        T_cold = T_s + (T_base - T_s) * frac * frac
        E_cold = ci * (T_cold - T0)
        Ek = np.where(zk <= z_t, E_temp, E_cold)
        Ek = np.where(zk > H, E_air, Ek)
        E[..., k] = np.where(H > 0.0, Ek, E_air)
    return E


# ------------------------------------------------------------------ marine window
def marine_window(i0=140, i1=204, j0=1350, j1=1422, float_quantile=0.15, Mz=401, **kw):
    """A window of the polythermal sheet with is_dry_simulation = 0 and a sea
    level that floats part of the ice: the sea level is the float_quantile
    quantile, over the icy cells, of the level at which a cell starts to float
    (b + (rho_i / rho_w) H + 1 m, Mask.hh:82-118). The default window holds
    grounded and floating ice, ice-free ocean and ice-free land."""
    base = greenland_like(i0, i1, j0, j1, Mz=Mz, **kw)
    H, b = base.fields["H"], base.fields["b"]
    P = base.params
    icy = H > 0.0
    fl = b + (P.ice_density / P.sea_water_density) * H + P.mask_is_floating_thickness_standard
    sea_level = float(np.quantile(fl[icy], float_quantile))
    return greenland_like(i0, i1, j0, j1, Mz=Mz, is_dry_simulation=0, sea_level=sea_level, **kw)


def sliding_field(w, speed=100.0 / T.SECONDS_PER_YEAR):
    """A smooth synthetic sliding (SSA) velocity (u_b, v_b) of the given
    speed scale on every cell of workload w, flat arrays in m s-1: it drives
    the upwind flux and the 2-D CFL limit of mass continuity."""
    Mx, My = w.grid.Mx, w.grid.My
    ii, jj = np.meshgrid(np.arange(Mx), np.arange(My), indexing="ij")
    u = speed * np.sin(2.0 * math.pi * ii / Mx + 0.3) * np.cos(2.0 * math.pi * jj / My)
    v = speed * np.cos(2.0 * math.pi * ii / Mx) * np.sin(2.0 * math.pi * jj / My + 0.7)
    return u.ravel(), v.ravel()


# ------------------------------------------------------------------ config-1 grid, polythermal
def dome_polythermal(seed, Mx=61, My=61, L=750e3, Mz=31, Lz=4000.0, H0=3600.0, R=750e3, **param_overrides):
    """The 61x61x31 dome geometry (flat bed, Vialov profile, dx = 25 km, 31
    equal levels to 4000 m) with GPBLD and a seeded polythermal enthalpy of
    the greenland_like kind (temperate basal layer of U[0, 0.1] H with water
    fraction U[0, 0.01], cold above from a surface temperature U[238, 248] K):
    the surrogate's form of the dome, whose isothermal Glen law it does not
    model."""
    dx = T.horizontal_spacing(L, Mx)
    dy = T.horizontal_spacing(L, My)
    x = T.horizontal_coordinates(L, Mx)
    y = T.horizontal_coordinates(L, My)
    z = T.vertical_levels(Lz, Mz)
    X, Y = np.meshgrid(x, y, indexing="ij")
    r = np.sqrt(X * X + Y * Y)
    H = np.where(r < R, H0 * np.power(np.maximum(1.0 - np.power(r / R, 4.0 / 3.0), 0.0), 3.0 / 8.0), 0.0)
    opts = dict(sia_flow_law="gpbld", sia_enhancement_factor=1.0, is_dry_simulation=1)
    opts.update(param_overrides)
    P = T.make_params(**opts)
    b = np.zeros_like(H)
    mask, h = geometry_mask_surface(P, b, H)
    col = np.arange(Mx * My, dtype=np.uint64).reshape(Mx, My)
    f_t = 0.1 * uniform01(seed, S_TEMPERATE, col)
    omega = 0.01 * uniform01(seed, S_OMEGA, col)
    T_s = 238.15 + 10.0 * uniform01(seed, S_TSURF, col)
    E = enthalpy_columns(P, H, f_t, omega, T_s, z)
    grid = T.make_grid(Mx, My, dx, dy, z)
    fields = dict(H=H.ravel(), b=b.ravel(), h=h.ravel(), mask=mask.ravel(), E=E.ravel())
    return Workload("dome_polythermal", grid, P, fields, (Mx, My, Mz), (0, Mx, 0, My))


# ------------------------------------------------------------------ ensemble workload
ENS_SIZES = (201, 401, 601, 801)


def ensemble_members(n=64, seed=2026, base=1000):
    """The ensemble workload: n members with seeded sizes Mx, My in ENS_SIZES,
    dx ~ U[2, 5] km, Mz ~ U{101..201} (Lz 4000 m), GPBLD with e ~ U[1, 5] and
    n in {3, 4}, each a polythermal sheet of its own seed (base + member)."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n):
        Mx, My = (int(v) for v in rng.choice(ENS_SIZES, 2))
        out.append(dict(Mx=Mx, My=My, dx=float(rng.uniform(2e3, 5e3)), Mz=int(rng.integers(101, 202)),
                        e=float(rng.uniform(1.0, 5.0)), n=float(rng.choice([3.0, 4.0])), seed=base + m))
    return out
