"""GPU checks of the benchmark configurations against the oracle, through the
C ABI:

* config 1: the 61x61x31 dome (isothermal Glen, A = 1e-16 Pa-3 a-1): one full
  update (D, flux, u, v, w, strain heating to the bars of test_gpu_parity.py;
  the device flux also against the Vialov closed form M0 r / 2 to 1 %),
  then 100 steps with the diffusivity/CFL-limited dt computed on the device;
  every step's dt and H against the oracle;
* config 2: a 96 x 96 window of the 1501x2801x401 polythermal sheet across
  the ice margin (GPBLD, e = 3): a full update, then 10 steps of full updates;
* config 3: the 64-member ensemble at its real shapes ({201..801}^2, 101-201
  levels, e U[1,5], n in {3,4}): 3 steps from a HIP graph, every member
  bitwise equal to the same member stepped alone, and the smallest member of
  each Glen exponent against the oracle's steps;
* config 4: the learned surrogate on the config-1 grid. The surrogate models
  GPBLD (its cold sum is a Paterson-Budd integral); on the dome's isothermal
  Glen law D is a closed form with nothing to learn, so the dome geometry and
  grid are run with GPBLD and a seeded polythermal enthalpy
  (workloads.dome_polythermal). A model trained on other seeds of that grid
  is judged on a held-out seed against the FP64 path and against the
  oracle's forward pass.

Tolerances: the parity bars of test_gpu_parity.py for updates; H after step k
to 1e-12 of max H times (k + 1) (the oracle and the device each run their own
trajectory); dt and model time to 1e-12 relative; ensemble members bitwise;
surrogate D to 1e-3 relative against FP64 and 1e-4 against the oracle's
forward pass.
"""
import numpy as np
import pytest

import sia_types as T
import training
import workloads as W
from parity_util import check_against_oracle, run_gpu, run_oracle, say
from surrogate_util import oracle_inference, plain_fast, surrogate_fast

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import pism_sia
    import torch
    assert torch.cuda.is_available()
    return pism_sia


def step_both(S, oracle, w, nsteps, full):
    """nsteps SIA-only steps on the device (sia_step: dt on the device) and in
    the oracle from the same initial state; asserts each step's dt, H, h and
    mask. Returns the device and oracle model times."""
    import torch
    st = {k: w.fields[k].copy() for k in ("H", "b", "h", "mask")}
    E = w.fields["E"]
    dev = {k: S.to_device(v) for k, v in st.items()}
    dE = S.to_device(E)
    sia = S.SIAFD(w.grid, w.params)
    sia.set_sync_check(False)
    geom = S.geometry(dev["H"], dev["h"], dev["mask"], dev["b"])
    t = torch.zeros(1, dtype=torch.float64, device="cuda")
    dt = torch.zeros(1, dtype=torch.float64, device="cuda")
    Hmax = st["H"].max()
    t_o = 0.0
    worst = 0.0
    for k in range(nsteps):
        sia.step(geom, dE, fast=not full, t_end=0.0, t_dev=t, dt_dev=dt)
        dt_o, _, _ = oracle.step(w.grid, w.params, st, E, full=full)
        t_o += dt_o
        torch.cuda.synchronize()
        dt_g = dt.item()
        assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (k, dt_g, dt_o)
        Hg = dev["H"].cpu().numpy()
        err = np.max(np.abs(Hg - st["H"]))
        worst = max(worst, err / Hmax / (k + 1))
        assert err <= 1e-12 * Hmax * (k + 1), (k, err)
        assert np.array_equal(dev["mask"].cpu().numpy(), st["mask"]), k
        np.testing.assert_allclose(dev["h"].cpu().numpy(), st["h"], rtol=0, atol=1e-12 * Hmax * (k + 1))
    sia.check_errors()
    sia.close()
    say("  %d steps: model time %.6e s (oracle %.6e s), max |dH| / (max H (k + 1)) %.1e" % (
        nsteps, t.item(), t_o, worst))
    assert abs(t.item() - t_o) <= 1e-12 * t_o
    return t.item(), t_o


def test_config1_dome_full_update_and_100_steps(S, oracle):
    say("[gpu] config 1: dome 61x61x31, full update + 100 steps against the oracle")
    w = W.dome()
    g = run_gpu(w)
    errs = check_against_oracle(g, run_oracle(oracle, w), w.grid, label="config 1")
    say("  full update, max rel errors:", {k: "%.1e" % v for k, v in errs.items()})
    # the closed form on the device's own flux: the Vialov profile's steady
    # flux |Q(r)| = M0 r / 2 (exactTestsABCDE.c:28-53) along both axes to 1 %,
    # as the discrete SIA reproduces it away from the dome and the margin
    M0 = W.vialov_smb(w.params)
    x = T.horizontal_coordinates(750e3, 61)
    fl = g["flux"]
    fl = (fl.cpu().numpy() if hasattr(fl, "cpu") else np.asarray(fl)).reshape(61, 61, 2)
    for i in range(34, 56):
        xs = x[i] + 12.5e3
        assert fl[i, 30, 0] == pytest.approx(M0 * xs / 2, rel=1e-2)
        assert fl[30, i, 1] == pytest.approx(M0 * xs / 2, rel=1e-2)
        assert fl[60 - i - 1, 30, 0] == pytest.approx(-M0 * xs / 2, rel=1e-2)
    step_both(S, oracle, w, 100, full=True)


def test_config2_sheet_window_full_update_and_10_steps(S, oracle):
    say("[gpu] config 2: 96x96x401 window of the sheet, full update + 10 steps against the oracle")
    w = W.greenland_like(140, 236, 1350, 1446)
    m = w.fields["mask"]
    assert (m == 0).any() and (m == 2).any() and w.fields["H"].max() > 1000.0
    errs = check_against_oracle(run_gpu(w), run_oracle(oracle, w), w.grid, label="config 2")
    say("  full update, max rel errors:", {k: "%.1e" % v for k, v in errs.items()})
    step_both(S, oracle, w, 10, full=True)


def test_config3_64_member_ensemble(S, oracle):
    """The 64 members at their real shapes; the graph of 3 steps is launched
    once. Members alone run sia_step (fast) with their own launches."""
    import torch
    spec = W.ensemble_members()
    say("[gpu] config 3: 64-member ensemble, %d columns" % sum(m["Mx"] * m["My"] for m in spec))
    legs = {"ens": [], "alone": []}
    init = []
    for m in spec:
        grid, params = training.member_grid_params(m)
        f = S.generate_greenland_like(grid, params, Mx_full=m["Mx"], My_full=m["My"], seed=m["seed"], dx=m["dx"])
        init.append({k: f[k].clone() for k in ("H", "b", "h", "mask")})
        for leg in legs:
            g = {k: f[k].clone() for k in ("H", "h", "mask")}
            g["b"], g["E"] = f["b"], f["E"]
            sia = S.SIAFD(grid, params)
            sia.set_sync_check(False)
            sia.update(g["H"], g["b"], g["h"], g["mask"], g["E"], fast=True)  # bed smoother, buffers
            legs[leg].append(dict(sia=sia, g=g, geom=S.geometry(g["H"], g["h"], g["mask"], g["b"]),
                                  t=torch.zeros(1, dtype=torch.float64, device="cuda"),
                                  dt=torch.zeros(1, dtype=torch.float64, device="cuda"), grid=grid, params=params))
    ens_leg = legs["ens"]
    ens = S.Ensemble([x["sia"] for x in ens_leg], [x["geom"] for x in ens_leg], [x["g"]["E"] for x in ens_leg])
    ens.step(3, graph=True)
    for x in legs["alone"]:
        for _ in range(3):
            x["sia"].step(x["geom"], x["g"]["E"], fast=True, t_end=0.0, t_dev=x["t"], dt_dev=x["dt"])
    torch.cuda.synchronize()
    ens.check_errors()
    for i, (a, b) in enumerate(zip(ens_leg, legs["alone"])):
        for k in ("H", "h", "mask"):
            assert torch.equal(a["g"][k], b["g"][k]), (i, k)
        assert np.array_equal(a["sia"].diffusive_flux(), b["sia"].diffusive_flux()), i
        assert ens.time(i)[0] == b["t"].item(), i
    # the smallest member of each Glen exponent against the oracle
    for n in (3.0, 4.0):
        i = min((k for k, m in enumerate(spec) if m["n"] == n), key=lambda k: spec[k]["Mx"] * spec[k]["My"])
        m = spec[i]
        say("  member %d (%dx%dx%d, e %.2f, n %g) against the oracle" % (i, m["Mx"], m["My"], m["Mz"], m["e"], n))
        st = {k: init[i][k].cpu().numpy() for k in ("H", "b", "h", "mask")}
        E = ens_leg[i]["g"]["E"].cpu().numpy()
        grid, params = ens_leg[i]["grid"], ens_leg[i]["params"]
        Hmax = st["H"].max()
        t_o = 0.0
        for _ in range(3):
            dt_o, _, _ = oracle.step(grid, params, st, E, full=False)
            t_o += dt_o
        Hg = ens_leg[i]["g"]["H"].cpu().numpy()
        err = np.max(np.abs(Hg - st["H"]))
        say("    max |dH| %.2e m (max H %.0f m), t %.6e s vs %.6e s" % (err, Hmax, ens.time(i)[0], t_o))
        assert err <= 3e-12 * Hmax
        assert abs(ens.time(i)[0] - t_o) <= 1e-12 * t_o
        assert np.array_equal(ens_leg[i]["g"]["mask"].cpu().numpy(), st["mask"])
    ens.close()
    for leg in legs.values():
        for x in leg:
            x["sia"].close()


def test_config4_surrogate_on_dome_grid(S, oracle):
    import torch
    say("[gpu] config 4: surrogate on the 61x61x31 grid (GPBLD, polythermal)")
    sg, n = training.train_dome_surrogate(S, torch, seeds=list(range(1, 121)))
    say("  trained on %d samples of 120 seeds" % n)
    w = W.dome_polythermal(9999)  # held out
    dev = {k: S.to_device(v) for k, v in w.fields.items()}
    ref = plain_fast(S, w, dev)
    # the FP64 update itself against the oracle
    o = run_oracle(oracle, w, full=False)
    rel = np.abs(ref["D"] - o["D_stag"])[o["D_stag"] > 0] / o["D_stag"][o["D_stag"] > 0]
    assert rel.max() <= 1e-12
    sur = surrogate_fast(S, sg, w, dev)
    zero, nsur, nfall = sur["counts"]
    say("  columns: zero %d, surrogate %d, FP64 fallback %d" % (zero, nsur, nfall))
    assert nsur > 0.8 * (nsur + nfall)
    pos = ref["D"] > 0
    err = np.abs(sur["flux"][pos] - ref["flux"][pos]) / np.abs(ref["flux"][pos])
    dmax_err = abs(sur["D_max"] - ref["D_max"]) / ref["D_max"]
    say("  flux vs FP64: median %.2e p99 %.2e max %.2e; D_max %.2e" % (
        np.median(err), np.percentile(err, 99), err.max(), dmax_err))
    assert err.max() <= 1e-3 and dmax_err <= 1e-3
    sel, Dor, inside = oracle_inference(sg, w, ref)
    Ds = sur["D"][sel]
    ok = inside & (Dor > 0)
    assert ok.sum() > 0.8 * len(sel)
    r = np.abs(Ds[ok] - Dor[ok]) / Dor[ok]
    say("  against the oracle's forward pass: max rel diff %.2e over %d columns" % (r.max(), ok.sum()))
    assert r.max() < 1e-4
    assert np.array_equal(Ds[~inside], ref["D"][sel][~inside])
    sg.close()
