"""GPU parity of the marine paths (is_dry_simulation = 0, the reference's
default, pism_config.cdl:136) against the oracle, through the C ABI.

The window of the polythermal sheet holds grounded and floating ice, ice-free
ocean and ice-free land (workloads.marine_window), with a smooth sliding
velocity, so that these branches run on the device:
* haseloff's marine-margin and cliff rules (SIAFD.cc:404-438, 441-499);
* get_smoothed_thk's floating branch (PISMBedSmoother.cc:325-329);
* adjust_flow's floating and ocean cases and the upwind SSA flux
  (iMgeometry.cc:172-415, 637-650), the ice-free-ocean source rule (692-697);
* GeometryCalculator's flotation criterion (Mask.hh:82-118).

Tolerances as in test_gpu_parity.py: 1e-12 relative for D, flux and the 2-D
fields; u, v, w to 1e-12 of their summand scales; H after step k to 1e-12 of
max H times (k + 1); dt to 1e-12; the mask bitwise; the accounting sums to
1e-12 of the size of their terms. Mass conservation with
zero SMB as in test/regression/test_12.sh:33-44: the ice volume changes by
the non-negativity rule's volume only, to 10^(floor(log10 V) - 14) m3 a step.
"""
import math

import numpy as np
import pytest

import workloads as W
from parity_util import check_against_oracle, run_gpu, run_oracle, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import pism_sia
    import torch
    assert torch.cuda.is_available()
    return pism_sia


@pytest.fixture(scope="module")
def marine():
    w = W.marine_window()
    m = w.fields["mask"]
    counts = {k: int((m == k).sum()) for k in (0, 2, 3, 4)}
    assert all(v > 0 for v in counts.values()), counts
    return w, W.sliding_field(w)


def test_marine_full_update(S, oracle, marine):
    w, (ub, vb) = marine
    say("[gpu] marine window %dx%dx%d, sea level %.1f m, full update" % (w.grid.Mx, w.grid.My, w.grid.Mz,
                                                                         w.params.sea_level))
    g = run_gpu(w, ub=ub, vb=vb)
    o = run_oracle(oracle, w, ub=ub, vb=vb)
    errs = check_against_oracle(g, o, w.grid, ub=ub, vb=vb, label="marine")
    say("  max rel errors:", {k: "%.1e" % v for k, v in errs.items()})
    Mx, My = w.grid.Mx, w.grid.My
    mask = w.fields["mask"].reshape(Mx, My)
    hx = o["h_x"].reshape(Mx, My, 2)
    # some marine-margin faces (floating next to ice-free ocean) have a zero gradient
    marg = (mask == 3) & (np.roll(mask, -1, axis=0) == 4)
    assert marg.any() and np.all(hx[marg, 0] == 0.0)
    # floating cells keep their thickness as thksmooth
    fl = (mask == 3).ravel()
    assert np.array_equal(o["thksmooth"][fl], w.fields["H"][fl])


def test_marine_steps_and_mass_conservation(S, oracle, marine):
    """10 steps (full and fast updates alternating) with the sliding velocity,
    zero SMB: H, mask, surface, dt and the accounting against the oracle, and
    test_12's conservation criterion on both."""
    import torch
    w, (ub, vb) = marine
    say("[gpu] marine window: 10 steps against the oracle")
    st = {k: w.fields[k].copy() for k in ("H", "b", "h", "mask")}
    E = w.fields["E"]
    dev = {k: S.to_device(v) for k, v in st.items()}
    dE, dub, dvb = S.to_device(E), S.to_device(ub), S.to_device(vb)
    sia = S.SIAFD(w.grid, w.params)
    geom = S.geometry(dev["H"], dev["h"], dev["mask"], dev["b"])
    area = w.grid.dx * w.grid.dy
    rho = w.params.ice_density
    Hmax = st["H"].max()
    vol_g, vol_o = [st["H"].sum() * area], [st["H"].sum() * area]
    nonneg_g, nonneg_o = [], []
    for k in range(10):
        full = (k % 2 == 0)
        sia.update(dev["H"], dev["b"], dev["h"], dev["mask"], dE, fast=not full, ub=dub, vb=dvb)
        dt_g = sia.max_timestep()
        acc_g = sia.mass_continuity_step(geom, dt_g)
        Hmax_k = st["H"].max()
        dt_o, out_o, acc_o = oracle.step(w.grid, w.params, st, E, full=full, ub=ub, vb=vb)
        torch.cuda.synchronize()
        Hg, hg, mg = dev["H"].cpu().numpy(), dev["h"].cpu().numpy(), dev["mask"].cpu().numpy()
        assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (k, dt_g, dt_o)
        assert np.max(np.abs(Hg - st["H"])) <= 1e-12 * Hmax * (k + 1), k
        np.testing.assert_allclose(hg, st["h"], rtol=0, atol=1e-12 * Hmax * (k + 1))
        assert np.array_equal(mg, st["mask"]), k
        # the accounting sums are sums over cells; the flux divergences cancel
        # between neighbours (zero SMB), so they are compared to 1e-12 of the
        # size of their terms, bounded by the face fluxes
        kg = dt_o * area * rho
        fl = out_o["flux"].reshape(-1, 2)
        scale = np.abs(acc_o) + np.abs(acc_o).max()
        scale[4] = 2.0 * (np.abs(fl[:, 0]).sum() / w.grid.dx + np.abs(fl[:, 1]).sum() / w.grid.dy) * kg
        scale[5] = 2.0 * (np.abs(ub).sum() / w.grid.dx + np.abs(vb).sum() / w.grid.dy) * Hmax_k * kg
        for c, (a, b) in enumerate(zip(acc_g, acc_o)):
            assert abs(a - b) <= 1e-12 * scale[c], (k, c, acc_g, acc_o)
        vol_g.append(Hg.sum() * area)
        vol_o.append(st["H"].sum() * area)
        nonneg_g.append(acc_g[1] / rho)
        nonneg_o.append(acc_o[1] / rho)
    sia.close()
    for vol, nn, who in ((vol_g, nonneg_g, "gpu"), (vol_o, nonneg_o, "oracle")):
        vol = np.array(vol)
        thr = 10 ** (math.floor(math.log10(vol.max())) - 14)
        resid = np.diff(vol) - np.array(nn)
        say("  %s: volume %.6e m3, max |dV - nonneg| %.2e m3 (threshold %.0e)" % (who, vol[-1],
                                                                               np.abs(resid).max(), thr))
        assert np.abs(resid).max() < thr, who
    assert max(nonneg_o) >= 0.0
