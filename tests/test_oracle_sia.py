"""Pin the oracle's SIA, vertical-velocity and mass-continuity restatement with
what the mathematics fixes (no reference vectors exist for SIAFD outputs):

* exact discrete trapezoid sums (Euler-Maclaurin closed forms) for D and I on
  a planar slab with isothermal Glen ice, n = 1 and n = 3 (SIAFD.cc:660-684, 863-868);
* u = -I grad h on the slab (SIAFD.cc:948-957) and w from linear u, v (exact
  trapezoid, PISMStressBalance.cc:262-285);
* planar surfaces: mahaffy/haseloff exact, eta exact for linear eta (SIAFD.cc:210-317);
* haseloff's zero-gradient rules at cliffs and marine margins (SIAFD.cc:404-438);
* the Vialov steady-state flux |Q| = M0 r / 2 on the 61x61 dome (exactTestsABCDE.c:28-53);
* mass conservation to 14 digits (test/regression/test_12.sh:33-44);
* the RuntimeError cases as status codes.
"""
import math

import numpy as np
import pytest

import sia_types as T
import workloads as W


def slab(n, Mx=12, My=10, H=1000.0, slope=2.0 ** -10, Mz=401, Lz=4000.0, **kw):
    """Uniform-thickness slab on a bed b = slope * x (exactly representable)."""
    dx = 1000.0
    z = T.vertical_levels(Lz, Mz)
    P = T.make_params(sia_flow_law="isothermal_glen", sia_Glen_exponent=float(n),
                      bed_smoother_range=0.0, surface_gradient_method="mahaffy", **kw)
    x = np.arange(Mx) * dx
    b = np.repeat((slope * x)[:, None], My, axis=1)
    Hf = np.full((Mx, My), H)
    h = b + Hf
    assert np.array_equal(h - b, Hf)
    mask = np.full((Mx, My), float(T.MASK_GROUNDED))
    E = np.zeros((Mx, My, Mz))
    grid = T.make_grid(Mx, My, dx, dx, z)
    return grid, P, dict(H=Hf.ravel(), b=b.ravel(), h=h.ravel(), mask=mask.ravel(), E=E.ravel())


def interior(Mx, My, m=2):
    return [(i, j) for i in range(m, Mx - m) for j in range(m, My - m)]


@pytest.mark.parametrize("n", [1, 3])
def test_slab_D_I_u_exact_trapezoid(oracle, n):
    H, s, dz = 1000.0, 2.0 ** -10, 10.0
    grid, P, f = slab(n, H=H, slope=s)
    out = oracle.sia_update(grid, P, f["H"], f["b"], f["h"], f["mask"], f["E"], full=True)
    Mx, My, Mz = grid.Mx, grid.My, grid.Mz
    rho_g = 910.0 * 9.81
    A = P.ice_softness
    c = 2.0 * A * s ** (n - 1) * rho_g ** n  # delta(z) = c (H - z)^n
    if n == 1:
        D_exact = c * (H ** 3 / 3 + dz * dz * H / 6)
    else:
        D_exact = c * (H ** 5 / 5 + dz * dz * H ** 3 / 3 - dz ** 4 * H / 30)
    flux = out["flux"].reshape(Mx, My, 2)
    Dst = out["D_stag"].reshape(Mx, My, 2)
    u = out["u"].reshape(Mx, My, Mz)
    v = out["v"].reshape(Mx, My, Mz)
    w = out["w"].reshape(Mx, My, Mz)
    I0 = out["I0"].reshape(Mx, My, Mz)
    z = T.vertical_levels(4000.0, Mz)
    zm = z[:101]
    if n == 1:
        I_exact = c * (H * zm - zm * zm / 2)  # linear delta: trapezoid exact
    else:
        I_exact = c * ((H ** 4 - (H - zm) ** 4) / 4 + dz * dz / 4 * (H ** 2 - (H - zm) ** 2))
    for i, j in interior(Mx, My):
        assert Dst[i, j, 0] == pytest.approx(D_exact, rel=1e-13)
        assert Dst[i, j, 1] == pytest.approx(D_exact, rel=1e-13)
        assert flux[i, j, 0] == pytest.approx(-D_exact * s, rel=1e-13)
        assert flux[i, j, 1] == 0.0
        np.testing.assert_allclose(I0[i, j, :101], I_exact, rtol=1e-13, atol=1e-30)
        assert np.all(I0[i, j, 101:] == I0[i, j, 100])  # constant above the ice
        np.testing.assert_allclose(u[i, j], -I0[i, j] * s, rtol=1e-14, atol=0)
        assert np.all(v[i, j] == 0.0)
    for i, j in interior(Mx, My, 3):
        assert np.all(np.abs(w[i, j]) <= 1e-25)  # uniform slab: no divergence
    # D_max excludes faces on the last row/column (SIAFD.cc:694-697)
    assert out["D_max"] >= D_exact * (1 - 1e-13)


def test_flux_converges_to_continuous_D_off_level(oracle):
    """H off the vertical levels: D -> 2A alpha^2 (rho g)^3 H^5/5 with O(dz^2) error."""
    H, s = 1234.5, 2.0 ** -9
    errs = []
    for Mz in (101, 201, 401):
        grid, P, f = slab(3, H=H, slope=s, Mz=Mz)
        out = oracle.sia_update(grid, P, f["H"], f["b"], f["h"], f["mask"], f["E"], full=False)
        D = out["D_stag"].reshape(grid.Mx, grid.My, 2)[4, 4, 0]
        D_cont = 2 * P.ice_softness * s * s * (910.0 * 9.81) ** 3 * H ** 5 / 5
        errs.append(abs(D - D_cont) / D_cont)
    assert errs[0] < 1e-2
    assert 3.0 < errs[0] / errs[1] < 5.0 and 3.0 < errs[1] / errs[2] < 5.0


def test_vertical_velocity_linear_fields_exact(oracle):
    """u = a x f(z), v = b y g(z), f, g linear in z: centred (and one-sided at a
    margin) differences and the trapezoid rule are exact: w = -(a F + b G)."""
    Mx, My, Mz = 9, 8, 21
    dx, dy = 1000.0, 2000.0
    z = T.vertical_levels(400.0, Mz)
    grid = T.make_grid(Mx, My, dx, dy, z)
    a, bb = 1e-9, -3e-9
    x = np.arange(Mx) * dx
    y = np.arange(My) * dy
    fz = 1.0 + 0.5 * z / 400.0
    gz = 2.0 - z / 400.0
    u = a * x[:, None, None] * fz[None, None, :] * np.ones((Mx, My, Mz))
    v = bb * y[None, :, None] * gz[None, None, :] * np.ones((Mx, My, Mz))
    mask = np.full((Mx, My), 2.0)
    mask[6, :] = 0.0  # ice-free row: one-sided differences next to it
    w = np.zeros(Mx * My * Mz)
    bmr = np.full(Mx * My, 1e-10)
    rc = oracle.lib().or_vertical_velocity(oracle.C.byref(grid),
                                          oracle._ptr(mask.ravel()), oracle._ptr(u.ravel()),
                                          oracle._ptr(v.ravel()), oracle._ptr(bmr), oracle._ptr(w))
    assert rc == 0
    w = w.reshape(Mx, My, Mz)
    F = z + 0.25 * z * z / 400.0
    G = 2.0 * z - 0.5 * z * z / 400.0
    expect = -1e-10 - (a * F + bb * G)
    for i in range(1, Mx - 1):
        for j in range(1, My - 1):
            if i == 6:
                # ice-free cell between icy neighbours: both x-sides are margins,
                # D_x = 0 (PISMStressBalance.cc:232-243)
                np.testing.assert_allclose(w[i, j], -1e-10 - bb * G, rtol=1e-12, atol=1e-22)
            else:
                np.testing.assert_allclose(w[i, j], expect, rtol=1e-12, atol=1e-22)


def plane(Mx=10, My=9, a=0.003, b=-0.002, method="mahaffy"):
    dx, dy = 1000.0, 1500.0
    z = T.vertical_levels(4000.0, 11)
    grid = T.make_grid(Mx, My, dx, dy, z)
    P = T.make_params(surface_gradient_method=method)
    X, Y = np.meshgrid(np.arange(Mx) * dx, np.arange(My) * dy, indexing="ij")
    return grid, P, X, Y


@pytest.mark.parametrize("method", ["mahaffy", "haseloff"])
def test_gradient_planar_surface_exact(oracle, method):
    grid, P, X, Y = plane(method=method)
    a, b = 0.003, -0.002
    h = 2000.0 + a * X + b * Y
    H = np.full_like(h, 1500.0)
    bed = h - H
    mask = np.full_like(h, 2.0)
    hx, hy = oracle.surface_gradient(grid, P, H.ravel(), bed.ravel(), h.ravel(), mask.ravel())
    hx = hx.reshape(grid.Mx, grid.My, 2)
    hy = hy.reshape(grid.Mx, grid.My, 2)
    for i, j in interior(grid.Mx, grid.My):
        np.testing.assert_allclose(hx[i, j], [a, a], rtol=1e-10)
        np.testing.assert_allclose(hy[i, j], [b, b], rtol=1e-10)


def test_gradient_eta_exact_for_linear_eta(oracle):
    """eta = H^((2n+2)/n) linear in x: grad h = n/(2n+2) eta^((-n-2)/(2n+2)) eta_x + grad b."""
    grid, P, X, Y = plane(method="eta")
    n = 3.0
    eta = 1e8 + 2e4 * X
    H = eta ** (n / (2 * n + 2))
    bed = 1e-3 * X - 2e-3 * Y
    h = H + bed
    mask = np.full_like(h, 2.0)
    hx, hy = oracle.surface_gradient(grid, P, H.ravel(), bed.ravel(), h.ravel(), mask.ravel())
    hx = hx.reshape(grid.Mx, grid.My, 2)
    hy = hy.reshape(grid.Mx, grid.My, 2)
    dx = 1000.0
    for i, j in interior(grid.Mx, grid.My):
        xs = X[i, j] + 0.5 * dx  # east face
        eta_s = 1e8 + 2e4 * xs
        exp_e = n / (2 * n + 2) * eta_s ** ((-n - 2) / (2 * n + 2)) * 2e4 + 1e-3
        assert hx[i, j, 0] == pytest.approx(exp_e, rel=1e-12)
        assert hy[i, j, 0] == pytest.approx(-2e-3, rel=1e-12)
        eta_n = 1e8 + 2e4 * X[i, j]  # north face has the same x
        exp_n = n / (2 * n + 2) * eta_n ** ((-n - 2) / (2 * n + 2)) * 2e4 + 1e-3
        assert hx[i, j, 1] == pytest.approx(exp_n, rel=1e-12)


def test_haseloff_cliff_and_marine_margin_rules(oracle):
    """Zero gradient across an ice cliff (ice-free bed above the ice surface) and across
    a floating/ice-free-ocean face; elsewhere haseloff equals mahaffy when no
    neighbouring weight is zero (SIAFD.cc:399-499)."""
    Mx, My = 10, 10
    grid = T.make_grid(Mx, My, 1000.0, 1000.0, T.vertical_levels(4000.0, 11))
    rng = np.random.default_rng(11)
    H = rng.uniform(100, 300, (Mx, My))
    bed = rng.uniform(0, 50, (Mx, My))
    mask = np.full((Mx, My), 2.0)
    # cliff: (5,4) ice-free with bed above h(4,4)
    H[5, 4] = 0.0
    mask[5, 4] = 0.0
    bed[5, 4] = 1000.0
    # marine margin: (2,7) floating, (2,8) ice-free ocean
    mask[2, 7] = 3.0
    mask[2, 8] = 4.0
    h = bed + H
    args = (H.ravel(), bed.ravel(), h.ravel(), mask.ravel())
    hxH, hyH = oracle.surface_gradient(grid, T.make_params(surface_gradient_method="haseloff"), *args)
    hxM, hyM = oracle.surface_gradient(grid, T.make_params(surface_gradient_method="mahaffy"), *args)
    hxH, hyH = hxH.reshape(Mx, My, 2), hyH.reshape(Mx, My, 2)
    hxM, hyM = hxM.reshape(Mx, My, 2), hyM.reshape(Mx, My, 2)
    assert hxH[4, 4, 0] == 0.0 and hxM[4, 4, 0] != 0.0          # cliff face (i+1/2)
    assert hxH[5, 4, 0] == 0.0  # the ice-free high cell is a cliff for its east neighbour too
    assert hyH[2, 7, 1] == 0.0 and hyM[2, 7, 1] != 0.0          # marine margin (j+1/2)
    # h_x on the north face of (4,3) averages 3 of the 4 east faces (one weight is 0)
    exp = (hxH[4, 3, 0] + hxH[3, 3, 0] + hxH[3, 4, 0] + hxH[4, 4, 0]) / 3.0
    assert hxH[4, 3, 1] == pytest.approx(exp, rel=1e-14)
    def pdist(a, b, M):
        d = abs(a - b) % M
        return min(d, M - d)
    special = [(5, 4), (2, 7), (2, 8)]
    far = [(i, j) for i in range(Mx) for j in range(My)
           if all(max(pdist(i, a, Mx), pdist(j, b, My)) > 2 for a, b in special)]
    assert far
    for i, j in far:
        # same differences, averaged in another order: equal up to rounding
        np.testing.assert_allclose(hxH[i, j], hxM[i, j], rtol=1e-13, atol=1e-17)
        np.testing.assert_allclose(hyH[i, j], hyM[i, j], rtol=1e-13, atol=1e-17)


def test_vialov_dome_flux_and_dt(oracle):
    """61x61 dome, isothermal Glen: steady flux |Q(r)| = M0 r / 2 and the
    diffusivity time step 0.12 * 2 / (D_max (2/dx^2)) (iMadaptive.cc:157-171)."""
    w = W.dome()
    f = w.fields
    out = oracle.sia_update(w.grid, w.params, f["H"], f["b"], f["h"], f["mask"], f["E"], full=True)
    M0 = W.vialov_smb(w.params)
    assert M0 * T.SECONDS_PER_YEAR == pytest.approx(0.634317, rel=1e-6)
    x = T.horizontal_coordinates(750e3, 61)
    fl = out["flux"].reshape(61, 61, 2)
    for i in range(34, 56):
        xs = x[i] + 12.5e3
        assert fl[i, 30, 0] == pytest.approx(M0 * xs / 2, rel=1e-2)
        assert fl[30, i, 1] == pytest.approx(M0 * xs / 2, rel=1e-2)
        assert fl[60 - i - 1, 30, 0] == pytest.approx(-M0 * xs / 2, rel=1e-2)
    assert 1.40 < out["D_max"] < 1.50
    dt = oracle.max_timestep(w.grid, w.params, out["D_max"])
    assert dt == pytest.approx(0.12 * 2.0 / (out["D_max"] * 2.0 / 25e3 ** 2), rel=1e-15)
    assert dt / T.SECONDS_PER_YEAR == pytest.approx(1.63, abs=0.02)
    # end of run limits dt; so does the 60-year cap
    assert oracle.max_timestep(w.grid, w.params, out["D_max"], 1000.0) == 1000.0
    assert oracle.max_timestep(w.grid, w.params, 0.0) == 60 * T.SECONDS_PER_YEAR


def test_mass_conservation_14_digits(oracle):
    """test_12.sh:33-44: with zero SMB the ice volume changes only by rounding."""
    w = W.dome(Mx=31, My=31, Mz=31)
    st = dict(w.fields)
    E = st.pop("E")
    area = w.grid.dx * w.grid.dy
    ivol = [st["H"].sum() * area]
    nonneg = 0.0
    for _ in range(20):
        dt, out, acc = oracle.step(w.grid, w.params, st, E, full=False)
        nonneg += acc[1]
        ivol.append(st["H"].sum() * area)
        # accounting: flux divergence sums to zero over the closed domain
        assert abs(acc[4]) < 1e-12 * ivol[0] * 910.0
    ivol = np.array(ivol)
    threshold = 10 ** (math.floor(math.log10(ivol.max())) - 14)
    d = np.diff(ivol) - 0.0
    assert np.abs(d).max() < threshold + nonneg / 910.0


def test_error_cases_become_status_codes(oracle):
    w = W.dome(Mx=11, My=11, Mz=11)
    f = dict(w.fields)
    H = f["H"].copy()
    H[5] = -1.0
    with pytest.raises(oracle.OracleError) as e:
        oracle.sia_update(w.grid, w.params, H, f["b"], f["h"], f["mask"], f["E"])
    assert e.value.code == 3
    w2 = W.dome(Mx=11, My=11, Mz=11, Lz=3000.0)  # H0 = 3600 > Lz
    f2 = w2.fields
    with pytest.raises(oracle.OracleError) as e:
        oracle.sia_update(w2.grid, w2.params, f2["H"], f2["b"], f2["h"], f2["mask"], f2["E"])
    assert e.value.code == 4
    P = T.make_params(surface_gradient_method=7)
    with pytest.raises(oracle.OracleError) as e:
        oracle.sia_update(w.grid, P, f["H"], f["b"], f["h"], f["mask"], f["E"])
    assert e.value.code == 6
    P = T.make_params(sia_flow_law="gk", compute_grain_size_using_age=1, do_age=0)
    with pytest.raises(oracle.OracleError) as e:
        oracle.sia_update(w.grid, P, f["H"], f["b"], f["h"], f["mask"], f["E"])
    assert e.value.code == 6
