"""CPU checks that pin the oracle's strain heating and hardness
(oracle/sia_oracle.c: or_strain_heating, or_hardness), which restate
StressBalance::compute_volumetric_strain_heating (PISMStressBalance.cc:370-500)
and FlowLaw::hardness_parameter (flowlaws.cc:103-105, 336-351).

Known answers used:
* Sigma = 2 e^(-1/n) B D^((1+1/n)) with 2 D^2 = D_ij D_ij of the shallow
  strain rate; for fields linear in x, y or z every difference the reference
  takes is exact, so D^2 has a closed form: vertical shear u = a z, v = b z
  gives D^2 = (a^2 + b^2) / 4; horizontal stretching u = al x, v = be y gives
  D^2 = ((al + be)^2 + al^2 + be^2) / 2 (w_z = -(u_x + v_y)); horizontal
  shear u = ga y, v = de x gives D^2 = (ga + de)^2 / 4;
* at an ice margin the icy cell takes one-sided differences, so junk in the
  ice-free neighbour must not change its Sigma;
* Sigma = 0 above kBelowHeight(H);
* hardness B = A^(-1/n): B^(-n) sigma^(n-1) equals the flow law (pinned by
  the reference's golden tables) for every law that shares softness_parameter.
"""
import numpy as np
import pytest

import sia_types as T

N_LAWS = ("isothermal_glen", "gpbld", "pb", "gk", "hooke", "arr", "arrwarm")


def slab(Mx=9, My=8, Mz=21, Lz=1000.0, dx=1000.0, dy=700.0, H=600.0):
    z = T.vertical_levels(Lz, Mz)
    grid = T.make_grid(Mx, My, dx, dy, z)
    Hf = np.full((Mx, My), H)
    mask = np.full((Mx, My), 2.0)
    return grid, z, Hf, mask


def closed_sigma(P, B, D2):
    n = P.sia_Glen_exponent
    return 2.0 * P.sia_enhancement_factor ** (-1.0 / n) * B * D2 ** (0.5 * (1.0 / n + 1.0))


@pytest.mark.parametrize("n,e", [(3.0, 1.0), (3.0, 4.5), (4.0, 2.0)])
def test_vertical_shear_closed_form(oracle, n, e):
    P = T.make_params(sia_flow_law="isothermal_glen", sia_Glen_exponent=n, sia_enhancement_factor=e)
    grid, z, H, mask = slab()
    Mx, My, Mz = grid.Mx, grid.My, grid.Mz
    a, b = 3e-10, -2e-10
    u = np.broadcast_to(a * z, (Mx, My, Mz)).copy()
    v = np.broadcast_to(b * z, (Mx, My, Mz)).copy()
    E = np.full((Mx, My, Mz), 1e5)
    S = oracle.strain_heating(grid, P, H, mask, E, u, v).reshape(Mx, My, Mz)
    ks = oracle.kbelow(z, 600.0)
    B = P.ice_softness ** (-1.0 / n)
    ref = closed_sigma(P, B, (a * a + b * b) / 4.0)
    np.testing.assert_allclose(S[:, :, :ks + 1], ref, rtol=1e-12)
    assert np.all(S[:, :, ks + 1:] == 0.0)


def test_horizontal_stretching_and_shear(oracle):
    P = T.make_params(sia_flow_law="isothermal_glen", sia_Glen_exponent=3.0, sia_enhancement_factor=2.0)
    grid, z, H, mask = slab()
    Mx, My, Mz = grid.Mx, grid.My, grid.Mz
    x = np.arange(Mx) * grid.dx
    y = np.arange(My) * grid.dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    E = np.full((Mx, My, Mz), 1e5)
    B = P.ice_softness ** (-1.0 / 3.0)
    inner = (slice(1, Mx - 1), slice(1, My - 1))  # the periodic wrap breaks linearity at the edges
    al, be = 2e-11, -5e-11
    u = np.repeat((al * X)[:, :, None], Mz, axis=2)
    v = np.repeat((be * Y)[:, :, None], Mz, axis=2)
    S = oracle.strain_heating(grid, P, H, mask, E, u, v).reshape(Mx, My, Mz)
    ks = oracle.kbelow(z, 600.0)
    ref = closed_sigma(P, B, 0.5 * ((al + be) ** 2 + al * al + be * be))
    np.testing.assert_allclose(S[inner][:, :, :ks + 1], ref, rtol=1e-11)
    ga, de = 4e-11, 1e-11
    u = np.repeat((ga * Y)[:, :, None], Mz, axis=2)
    v = np.repeat((de * X)[:, :, None], Mz, axis=2)
    S = oracle.strain_heating(grid, P, H, mask, E, u, v).reshape(Mx, My, Mz)
    ref = closed_sigma(P, B, (ga + de) ** 2 / 4.0)
    np.testing.assert_allclose(S[inner][:, :, :ks + 1], ref, rtol=1e-11)


def test_margin_takes_one_sided_differences(oracle):
    """Column i = 5 is ice-free: icy cells i = 4 and 6 must ignore its junk."""
    P = T.make_params(sia_flow_law="isothermal_glen", sia_Glen_exponent=3.0)
    grid, z, H, mask = slab(Mx=11)
    Mx, My, Mz = grid.Mx, grid.My, grid.Mz
    mask[5, :] = 0.0
    H[5, :] = 0.0
    al = 3e-11
    x = np.arange(Mx) * grid.dx
    u = np.repeat(np.repeat((al * x)[:, None], My, axis=1)[:, :, None], Mz, axis=2)
    v = np.zeros((Mx, My, Mz))
    u[5] = 1.0  # junk where there is no ice
    E = np.full((Mx, My, Mz), 1e5)
    S = oracle.strain_heating(grid, P, H, mask, E, u, v).reshape(Mx, My, Mz)
    ks = oracle.kbelow(z, 600.0)
    B = P.ice_softness ** (-1.0 / 3.0)
    ref = closed_sigma(P, B, 0.5 * (al * al + al * al))  # u_x = al, v_y = 0
    for i in (2, 3, 4, 6, 7, 8):
        np.testing.assert_allclose(S[i, :, :ks + 1], ref, rtol=1e-11, err_msg="column %d" % i)
    # H = 0 at i = 5: level 0 only, where both x-differences are switched off
    # and u is constant in z, so Sigma = 0
    assert np.all(S[5] == 0.0)


def test_zero_above_the_ice_and_on_a_level(oracle):
    P = T.make_params(sia_flow_law="isothermal_glen")
    z_level = 500.0
    grid, z, H, mask = slab(H=z_level)
    Mx, My, Mz = grid.Mx, grid.My, grid.Mz
    u = np.broadcast_to(1e-10 * z, (Mx, My, Mz)).copy()
    S = oracle.strain_heating(grid, P, H, mask, np.full((Mx, My, Mz), 1e5), u, 0 * u).reshape(Mx, My, Mz)
    ks = oracle.kbelow(z, z_level)
    assert z[ks] == z_level  # H exactly on a level: ks is that level
    assert np.all(S[:, :, ks + 1:] == 0.0) and np.all(S[:, :, :ks + 1] > 0.0)


@pytest.mark.parametrize("law", N_LAWS)
def test_hardness_is_softness_to_minus_one_over_n(oracle, law):
    P = T.make_params(sia_flow_law=law, sia_Glen_exponent=3.0)
    ec = oracle.EC(P)
    rng = np.random.default_rng(5)
    sigma = 7e4
    for _ in range(50):
        depth = rng.uniform(0, 3000)
        p = ec.pressure(depth)
        E = ec.enthalpy_cts(p) + rng.uniform(-9e4, 2e3)
        B = oracle.hardness(P, E, p)
        assert B > 0
        A = B ** -3.0
        if law in ("isothermal_glen", "gpbld"):
            # the same softness_parameter as the flow law
            assert A * sigma ** 2 == pytest.approx(oracle.flow(P, sigma, E, p), rel=1e-13)
        elif law in ("pb", "hooke"):
            # flow uses temp + beta p, hardness the converter's T_pa: equal to rounding
            assert A * sigma ** 2 == pytest.approx(oracle.flow(P, sigma, E, p), rel=1e-12)
        elif law == "gk":
            pb = T.make_params(sia_flow_law="pb", sia_Glen_exponent=3.0)
            assert A == pytest.approx(oracle.hardness(pb, E, p) ** -3.0, rel=1e-13)
        else:  # arr, arrwarm: pressure-adjusted T_pa in the hardness, absolute T in flow
            T_pa = ec.pressure_adjusted_temperature(E, p)
            Ac, Qc = (P.Paterson_Budd_A_cold, P.Paterson_Budd_Q_cold) if law == "arr" else \
                (P.Paterson_Budd_A_warm, P.Paterson_Budd_Q_warm)
            assert A == pytest.approx(Ac * np.exp(-Qc / (P.ideal_gas_constant * T_pa)), rel=1e-13)


def test_full_update_returns_strain_heating(oracle):
    """or_sia_update computes Sigma from its own u, v on full updates."""
    import workloads as W
    w = W.greenland_like(700, 716, 1180, 1196, Mz=101)
    f = w.fields
    o = oracle.sia_update(w.grid, w.params, f["H"], f["b"], f["h"], f["mask"], f["E"], full=True)
    S = oracle.strain_heating(w.grid, w.params, f["H"], f["mask"], f["E"], o["u"], o["v"])
    assert np.array_equal(o["Sigma"], S)
    assert np.all(np.isfinite(S)) and S.max() > 0
