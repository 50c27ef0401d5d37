"""Pin the CPU oracle against the known answers PISM's own tests hold.

* flow-law tables of test/regression/test_20 (gpbld), 21 (pb), 24 (hooke),
  25 (gk), 26 (arr), generated by src/software_tests/flowlaw_test.cc:62-90;
* the bed-smoother output of test_22 (src/software_tests/bedrough_test.cc);
* the enthalpy-converter assertions of test/enthalpy_converter.py:21-128;
* kBelowHeight's GSL binary-search semantics, by brute force.
"""
import json
import math
import os

import numpy as np
import pytest

import sia_types as T

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "regression_outputs.json")))


@pytest.mark.parametrize("law", ["gpbld", "pb", "hooke", "gk", "arr"])
def test_flow_table_matches_reference_golden(oracle, law):
    """flowlaw_test.cc:62-90 with the default EnthalpyConverter, printed with %10.6e."""
    tab = GOLD["flow_tables"][law]
    P = T.make_params(sia_flow_law=law)
    ec = oracle.EC(P)
    p = ec.pressure(2000.0)
    assert "pressure = %9.3e Pa = (hydrostatic at depth %7.2f m)" % (p, 2000.0) == tab["pressure_line"]
    Tm = ec.melting_temperature(p)
    row = 0
    for sigma in (1e4, 5e4, 1e5, 1.5e5):
        for jj, TpaC in enumerate((-30.0, -5.0, 0.0, 0.0)):
            Tk = Tm + TpaC
            omega = 0.005 if jj == 3 else 0.0
            E = ec.enthalpy(Tk, omega, p)
            f = oracle.flow(P, sigma, E, p, 1e-3)
            exp = tab["rows"][row]
            row += 1
            assert "%10.2e" % sigma == "%10.2e" % exp[0]
            assert "%10.3f" % Tk == "%10.3f" % exp[1]
            assert "%10.6e" % f == exp[4], (law, sigma, Tk, omega)
    assert row == 16


def test_isothermal_glen_and_arrwarm_closed_forms(oracle):
    """IsothermalGlen::flow = A sigma^(n-1) (flowlaws.hh:227-229); PatersonBuddWarm
    uses A_warm exp(-Q_warm/(R T)) with the absolute temperature (flowlaws.hh:296-326)."""
    P = T.make_params(sia_flow_law="isothermal_glen")
    for s in (1e3, 5e4, 2e5):
        assert oracle.flow(P, s, 1e5, 1e7, 1e-3) == pytest.approx(3.1689e-24 * s * s, rel=1e-15)
    P4 = T.make_params(sia_flow_law="isothermal_glen", sia_Glen_exponent=4.0)
    assert oracle.flow(P4, 5e4, 1e5, 1e7, 1e-3) == pytest.approx(3.1689e-24 * 5e4 ** 3, rel=1e-15)
    Pw = T.make_params(sia_flow_law="arrwarm")
    ec = oracle.EC(Pw)
    p = ec.pressure(1000.0)
    E = ec.enthalpy(260.0, 0.0, p)
    Tabs = ec.temperature(E, p)
    assert Tabs == pytest.approx(260.0, rel=1e-14)
    expect = 1.73e3 * math.exp(-13.9e4 / (8.31441 * Tabs)) * 1e5 ** 2
    assert oracle.flow(Pw, 1e5, E, p, 1e-3) == pytest.approx(expect, rel=1e-14)


def test_gpbld_branch_tie_goes_temperate(oracle):
    """GPBLD uses `enthalpy < E_s` for cold ice (flowlaws.cc:211), so E == E_s is
    temperate with omega = 0: A(273.15)."""
    P = T.make_params(sia_flow_law="gpbld")
    ec = oracle.EC(P)
    p = ec.pressure(1500.0)
    Es = ec.enthalpy_cts(p)
    A_temperate = 1.73e3 * math.exp(-13.9e4 / (8.31441 * 273.15))
    assert oracle.flow(P, 1.0, Es, p) == pytest.approx(A_temperate, rel=1e-15)
    # 1000 J/kg below: cold branch, T_pa = 273.15 - 1000/c_i, warm constants (T_pa >= 263.15)
    T_pa = (Es - 1000.0) / 2009.0 + 223.15 - ec.melting_temperature(p) + 273.15
    assert T_pa == pytest.approx(273.15 - 1000.0 / 2009.0, abs=1e-9)
    below = oracle.flow(P, 1.0, Es - 1000.0, p)
    assert below == pytest.approx(1.73e3 * math.exp(-13.9e4 / (8.31441 * T_pa)), rel=1e-14)
    # above: temperate, A(273.15) (1 + 181.25 omega), omega capped at 0.01
    om = 0.004
    assert oracle.flow(P, 1.0, Es + om * 3.34e5, p) == pytest.approx(A_temperate * (1 + 181.25 * om), rel=1e-12)
    assert oracle.flow(P, 1.0, Es + 0.05 * 3.34e5, p) == pytest.approx(A_temperate * (1 + 181.25 * 0.01), rel=1e-14)
    # below 263.15 K the cold constants take over
    E_c = ec.enthalpy(250.0, 0.0, p)
    T_pa = ec.pressure_adjusted_temperature(E_c, p)
    assert oracle.flow(P, 1.0, E_c, p) == pytest.approx(3.61e-13 * math.exp(-6.0e4 / (8.31441 * T_pa)), rel=1e-14)


def test_bed_smoother_matches_test_22(oracle):
    """bedrough_test.cc:53-139: 81x81 non-periodic grid on +-1200 km, sinusoidal
    bed, usurf = 1000 m, bed_smoother_range = 50 km."""
    g = GOLD["bedrough"]
    Mx = My = 81
    L = 1200e3
    x = T.horizontal_coordinates(L, Mx)
    y = T.horizontal_coordinates(L, My)
    dx = T.horizontal_spacing(L, Mx)
    X, Y = np.meshgrid(x, y, indexing="ij")
    topg = np.array([[400.0 * math.sin(2.0 * math.pi * X[i, j] / 600.0e3) +
                      100.0 * math.sin(2.0 * math.pi * (X[i, j] + 1.5 * Y[i, j]) / 40.0e3)
                      for j in range(My)] for i in range(Mx)])
    P = T.make_params(bed_smoother_range=50.0e3, sia_Glen_exponent=3.0)
    grid = T.make_grid(Mx, My, dx, dx, T.vertical_levels(4000.0, 31))
    sm = oracle.bed_smoother(grid, P, topg.ravel())
    th = oracle.theta(grid, P, sm, np.full(Mx * My, 1000.0))
    assert (sm["Nx"], sm["Ny"]) == (g["Nx"], g["Ny"])
    assert ["%.6f" % topg.min(), "%.6f" % topg.max()] == g["bed"]
    assert ["%.6f" % sm["topgsmooth"].min(), "%.6f" % sm["topgsmooth"].max()] == g["smoothed"]
    assert ["%.9f" % th.min(), "%.9f" % th.max()] == g["theta"]


def test_bed_smoother_inactive(oracle):
    """bed_smoother_range <= 0: topgsmooth = topg, theta = 1 (PISMBedSmoother.cc:98-106, 362-365)."""
    rng = np.random.default_rng(3)
    Mx, My = 9, 7
    grid = T.make_grid(Mx, My, 1e3, 1e3, T.vertical_levels(4000.0, 11))
    P = T.make_params(bed_smoother_range=0.0)
    topg = rng.uniform(-100, 100, Mx * My)
    sm = oracle.bed_smoother(grid, P, topg)
    assert np.array_equal(sm["topgsmooth"], topg) and sm["Nx"] == -1
    assert np.all(oracle.theta(grid, P, sm, topg + 500.0) == 1.0)


def test_bed_smoother_box_mean_brute_force(oracle):
    """Smoothed bed = mean over the in-grid (2N+1)^2 box; maxtl = max(0, max(b - bs));
    C2..C4 = scaled central moments (PISMBedSmoother.cc:162-269)."""
    rng = np.random.default_rng(7)
    Mx, My = 11, 8
    dx = 2000.0
    grid = T.make_grid(Mx, My, dx, dx, T.vertical_levels(4000.0, 11))
    P = T.make_params(bed_smoother_range=3000.0)  # N = ceil(1.5) = 2
    b = rng.uniform(-300, 300, (Mx, My))
    sm = oracle.bed_smoother(grid, P, b.ravel())
    assert sm["Nx"] == 2 and sm["Ny"] == 2
    n = 3.0
    k = (n + 2) / n
    s2 = k * (2 * n + 2) / (2 * n)
    s3 = s2 * (3 * n + 2) / (3 * n)
    s4 = s3 * (4 * n + 2) / (4 * n)
    for i in range(Mx):
        for j in range(My):
            box = b[max(0, i - 2):i + 3, max(0, j - 2):j + 3]
            bs = box.mean()
            tl = box - bs
            q = i * My + j
            assert sm["topgsmooth"][q] == pytest.approx(bs, rel=1e-13, abs=1e-10)
            assert sm["maxtl"][q] == pytest.approx(max(0.0, tl.max()), abs=1e-10)
            assert sm["C2"][q] == pytest.approx(s2 * (tl ** 2).mean(), rel=1e-12)
            assert sm["C3"][q] == pytest.approx(s3 * (tl ** 3).mean(), rel=1e-9, abs=1e-6)
            assert sm["C4"][q] == pytest.approx(s4 * (tl ** 4).mean(), rel=1e-12)


CONVERTERS = ["default", "cold", "kirchhoff"]


@pytest.mark.parametrize("conv", CONVERTERS)
def test_enthalpy_converter_reversibility(oracle, conv):
    """test/enthalpy_converter.py:21-56."""
    ec = oracle.EC(T.make_params(enthalpy_converter=conv))
    P = ec.pressure(1000.0)
    E = ec.enthalpy(250.0, 0.0, P)
    Tt = ec.temperature(E, P)
    om = ec.water_fraction(E, P)
    assert E == ec.enthalpy(Tt, om, P)
    assert om == 0.0
    T_m = ec.melting_temperature(P)
    E = ec.enthalpy(T_m, 0.1, P)
    Tt = ec.temperature(E, P)
    om = ec.water_fraction(E, P)
    assert E == ec.enthalpy(Tt, om, P)
    if ec.is_temperate(E, P):
        assert om == 0.1


@pytest.mark.parametrize("conv", CONVERTERS)
def test_enthalpy_converter_temperate_cts_water_fraction(oracle, conv):
    """test/enthalpy_converter.py:59-103."""
    ec = oracle.EC(T.make_params(enthalpy_converter=conv))
    P = ec.pressure(1000.0)
    T_m = ec.melting_temperature(P)
    E = ec.enthalpy(T_m, 0.005, P)
    if ec.is_temperate(E, P):
        for dE in (0, 100, 1000):
            assert ec.temperature(E + dE, P) == T_m
    assert ec.enthalpy(T_m, 0.0, P) == ec.enthalpy_cts(P)
    assert ec.water_fraction(ec.enthalpy_cts(P), P) == 0


def test_kirchhoff_enthalpy_of_water(oracle):
    """test/enthalpy_converter.py:106-128: E_w(T_m(p1)) - E_w(T_m(p0)) = c_w (T1 - T0)."""
    ec = oracle.EC(T.make_params(enthalpy_converter="kirchhoff"))
    p0, p1 = ec.pressure(0.0), ec.pressure(1000.0)
    T0, T1 = ec.melting_temperature(p0), ec.melting_temperature(p1)
    E0, E1 = ec.enthalpy(T0, 1.0, p0), ec.enthalpy(T1, 1.0, p1)
    assert abs((E1 - E0) - 4170.0 * (T1 - T0)) < 1e-9


def test_pressure_and_melting_point_values(oracle):
    """p = p_air + rho g d (d > 0) else p_air; T_m = 273.15 - beta p (enthalpyConverter.cc:92-130)."""
    ec = oracle.EC(T.make_params())
    assert ec.pressure(2000.0) == 910.0 * 9.81 * 2000.0
    assert ec.pressure(-5.0) == 0.0 and ec.pressure(0.0) == 0.0
    assert ec.melting_temperature(1.785e7) == 273.15 - 7.9e-8 * 1.785e7
    cold = oracle.EC(T.make_params(enthalpy_converter="cold"))
    assert cold.melting_temperature(1e7) == 273.15 and cold.water_fraction(1e9, 1e7) == 0.0


def test_kbelow_gsl_bsearch_semantics(oracle):
    """kBelowHeight = largest k with z_k <= height, clamped to Mz-2, errors beyond
    [-1e-6, Lz+1e-6] (IceGrid.cc:379-391). Brute force over every level and midpoints."""
    z = T.vertical_levels(4000.0, 31)
    Mz = len(z)
    for k in range(Mz):
        expect = min(k, Mz - 2)
        assert oracle.kbelow(z, z[k]) == expect
        if k < Mz - 1:
            assert oracle.kbelow(z, 0.5 * (z[k] + z[k + 1])) == k
            assert oracle.kbelow(z, np.nextafter(z[k + 1], 0)) == k
    assert oracle.kbelow(z, -0.5e-6) == 0
    assert oracle.kbelow(z, 4000.0 + 0.5e-6) == Mz - 2
    assert oracle.kbelow(z, -2e-6) < 0
    assert oracle.kbelow(z, 4000.0 + 2e-6) < 0
    zq = np.array([0.0, 1.0, 3.0, 7.0, 20.0])  # non-uniform levels
    for hgt, k in ((0.0, 0), (0.999, 0), (1.0, 1), (6.5, 2), (7.0, 3), (19.9, 3), (20.0, 3)):
        assert oracle.kbelow(zq, hgt) == k


def test_grain_size_vostok_interpolation(oracle):
    """SIAFD::grainSizeVostok (SIAFD.cc:980-1018): linear between the samples, constant outside."""
    P = T.make_params()
    ka = 1000.0 * T.SECONDS_PER_YEAR
    assert oracle.lib().or_grain_size_vostok(__import__("ctypes").byref(P), -1.0) == 1.8e-3
    g = lambda a: oracle.lib().or_grain_size_vostok(__import__("ctypes").byref(P), a * ka)
    assert g(0.0) == pytest.approx(1.8e-3)
    assert g(25.0) == pytest.approx(0.5 * (1.8e-3 + 2.2e-3), rel=1e-12)
    assert g(162.0) == pytest.approx(3.0e-3, rel=1e-12)
    assert g(2e4) == 1.0e-2
