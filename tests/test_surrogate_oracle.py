"""CPU checks of the surrogate's oracle (oracle/surrogate_oracle.py):

* the per-level regrouping D = C sum_k W_k p_k^n A_k reproduces the C oracle's
  D of SIAFD::compute_diffusive_flux to 1e-12 on polythermal sheet windows
  (this pins the level weights W_k, including the top partial cell);
* the vectorised GPBLD softness and enthalpy converter equal the C oracle's;
* the MLP backward equals central finite differences of the forward (float64);
* BF16 rounding is round-to-nearest-even;
* the fallback score's closed forms: zero without a correction, the one-cell
  value, invariance under a common scale of C Qb and C S_t, and the weighting
  by predicted shares (a cell with no cold levels adds nothing).
"""
import numpy as np
import pytest

import surrogate_oracle as SO
import workloads as W


@pytest.fixture(scope="module")
def window(oracle):
    w = W.greenland_like(700, 740, 1180, 1230)
    f = w.fields
    o = oracle.sia_update(w.grid, w.params, f["H"], f["b"], f["h"], f["mask"], f["E"], full=False)
    return w, o


def test_level_regrouping_reproduces_reference_D(window):
    w, o = window
    z = np.array([w.grid.z[k] for k in range(w.grid.Mz)])
    st = SO.staggered_inputs(w.grid, w.fields, o["thksmooth"], o["theta"], o["h_x"], o["h_y"])
    sel = (st["thk"] > 0) & (st["alpha"] > 0)
    assert sel.sum() > 2000
    c = SO.columns(w.params, z, st["thk"][sel], st["alpha"][sel], st["theta"][sel], st["E"][sel])
    D = o["D_stag"][sel]
    rel = np.abs(c["Dref"] - D) / D
    assert rel.max() < 1e-12, rel.max()
    # both branches of GPBLD occur, and most of D is in the cold cells or the temperate sum
    assert c["has"].any() and (c["St"] > 0).any()


def test_margin_window_thin_columns(oracle):
    w = W.greenland_like(150, 190, 1380, 1430, Mz=201, Lz=4000.0)
    f = w.fields
    o = oracle.sia_update(w.grid, w.params, f["H"], f["b"], f["h"], f["mask"], f["E"], full=False)
    z = np.array([w.grid.z[k] for k in range(w.grid.Mz)])
    st = SO.staggered_inputs(w.grid, f, o["thksmooth"], o["theta"], o["h_x"], o["h_y"])
    sel = (st["thk"] > 0) & (st["alpha"] > 0)
    c = SO.columns(w.params, z, st["thk"][sel], st["alpha"][sel], st["theta"][sel], st["E"][sel])
    D = o["D_stag"][sel]
    assert np.all(np.abs(c["Dref"] - D) <= 1e-12 * np.abs(D) + 1e-300)
    assert (c["ks"] < 5).any()  # thin columns with a partial top cell are covered


def test_vectorised_flow_law_matches_c_oracle(oracle):
    P = W.gl_params()
    ec = oracle.EC(P)
    rng = np.random.default_rng(3)
    depth = rng.uniform(0, 3500, 200)
    p = SO.ec_pressure(P, depth)
    Es = SO.ec_enthalpy_cts(P, p)
    for i in range(200):
        assert SO.ec_enthalpy_cts(P, p[i]) == pytest.approx(ec.enthalpy_cts(p[i]), rel=1e-15)
        E = Es[i] + rng.uniform(-80000, 3000)
        Tpa = ec.pressure_adjusted_temperature(E, p[i])
        if E < Es[i]:
            # GPBLD cold branch: flow = A(T_pa) sigma^(n-1)
            sigma = 1e5
            ref = oracle.flow(P, sigma, E, p[i], 1.0)
            mine = float(SO.softness_pb(P, np.array(Tpa))) * sigma ** 2
            assert mine == pytest.approx(ref, rel=1e-13)


def test_mlp_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.15, SO.NPARAM)
    B = 6
    X = np.zeros((B, SO.KIN))
    X[:, :SO.NIN] = rng.normal(0, 1, (B, SO.NIN))
    Tt = rng.normal(0, 0.3, (B, 32))
    S = rng.uniform(0, 0.1, (B, 32))
    Hm = (rng.uniform(size=(B, 32)) < 0.7).astype(float)
    loss, g = SO.loss_grad(w, X, Tt, S, Hm, rscale=0.05, lam=1e-3, rounding=False)
    idx = rng.choice(SO.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [SO.NPARAM - 1, SO.LAYOUT[1][1], SO.LAYOUT[8][1] + 5]])  # b5, b1, W5
    eps = 1e-6
    for k in idx:
        wp, wm = w.copy(), w.copy()
        wp[k] += eps
        wm[k] -= eps
        lp, _ = SO.loss_grad(wp, X, Tt, S, Hm, 0.05, 1e-3, rounding=False)
        lm, _ = SO.loss_grad(wm, X, Tt, S, Hm, 0.05, 1e-3, rounding=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-8, abs(fd)) + 1e-12, (k, fd, g[k])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -9, 3.0e-39], dtype=np.float32)
    r = SO.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                  # tie -> even
    assert r[2] == 1.0 + 2 ** -6        # tie -> even (up)
    assert r[3] == -2.5
    assert r[4] == 1.0
    assert r[5] == pytest.approx(3.0e-39, rel=1e-2)


def test_fallback_score_closed_forms():
    rng = np.random.default_rng(5)
    CQb = rng.uniform(0.0, 2.0, (6, 32))
    CSt = rng.uniform(0.0, 1.0, 6)
    # no correction: score 0
    assert np.all(SO.fallback_score(CQb, CSt, np.zeros((6, 32)), 1.0) == 0.0)
    # one cell with a correction x: q e^x |x| / (s + q e^x + the other cells)
    out = np.zeros((6, 32))
    out[:, 7] = 0.3
    x = 0.3 * 2.5
    expect = CQb[:, 7] * np.exp(x) * x / (CSt + CQb.sum(axis=1) - CQb[:, 7] + CQb[:, 7] * np.exp(x))
    np.testing.assert_allclose(SO.fallback_score(CQb, CSt, out, 2.5), expect, rtol=1e-14)
    # homogeneous of degree 0 in (C Qb, C S_t); |r| makes the sign of r irrelevant to the weight
    out = rng.normal(0.0, 0.2, (6, 32))
    s1 = SO.fallback_score(CQb, CSt, out, 1.5)
    np.testing.assert_allclose(SO.fallback_score(7.0 * CQb, 7.0 * CSt, out, 1.5), s1, rtol=1e-13)
    assert np.all(s1 > 0)
    # a cell without cold levels (C Qb = 0) contributes nothing, whatever its output
    CQb0 = CQb.copy()
    CQb0[:, 3] = 0.0
    out2 = out.copy()
    out2[:, 3] = 50.0
    np.testing.assert_allclose(SO.fallback_score(CQb0, CSt, out2, 1.5), SO.fallback_score(CQb0, CSt, out, 1.5),
                               rtol=1e-14)
