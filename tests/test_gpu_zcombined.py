"""GPU check of the combined configuration (config 5), reduced from 5000 to
200 steps: the learned surrogate inside the 64-member batched ensemble, i.e.
the 1501x2801x401 polythermal sheet plus the first 63 members of the
ensemble workload, with the binding bench.py --combined reports (the
surrogate on members of at least 301 levels: the sheet; FP64 elsewhere).

* Every 50 steps (one captured HIP graph) the sheet's surrogate flux is
  compared with an FP64 fast update of the same state on all of its
  staggered columns: max relative error <= 1e-3 at every check.
* Every member, the sheet included, is bitwise equal to the same member
  stepped alone through sia_step (its own launches; the sheet bound to the
  same surrogate), and so is its model time.

The surrogate is trained first, on the GPU, on 12 other seeds of the sheet
(pismv0.7.3_amd/training.py, the recipe bench.py uses).
"""
import numpy as np
import pytest

import sia_types as T
import training
import workloads as W
from parity_util import say

pytestmark = pytest.mark.gpu

STEPS, CHUNK, MIN_MZ = 200, 50, 301


@pytest.fixture(scope="module")
def S():
    import pism_sia
    import torch
    assert torch.cuda.is_available()
    return pism_sia


def test_config5_reduced_combined_run(S):
    import torch
    z = T.vertical_levels(4000.0, 401)
    sheet_grid = T.make_grid(1501, 2801, 1000.0, 1000.0, z)
    sheet_params = W.gl_params()
    say("[gpu] config 5 (200 steps): training the surrogate on 12 sheets")
    sg, info = training.train_sheet_surrogate(S, torch, sheet_grid, sheet_params, log=say)
    specs = [(sheet_grid, sheet_params, {})]
    for m in W.ensemble_members()[:63]:
        grid, params = training.member_grid_params(m)
        specs.append((grid, params, dict(Mx_full=m["Mx"], My_full=m["My"], seed=m["seed"], dx=m["dx"])))
    legs = {"ens": [], "alone": []}
    for grid, params, kw in specs:
        f = S.generate_greenland_like(grid, params, **kw)
        for leg in legs:
            g = {k: f[k].clone() for k in ("H", "h", "mask")}
            g["b"], g["E"] = f["b"], f["E"]
            sia = S.SIAFD(grid, params)
            sia.set_sync_check(False)
            sia.update(g["H"], g["b"], g["h"], g["mask"], g["E"], fast=True)  # bed smoother, buffers
            if grid.Mz >= MIN_MZ:
                sia.use_surrogate(sg)
            legs[leg].append(dict(sia=sia, g=g, geom=S.geometry(g["H"], g["h"], g["mask"], g["b"]), grid=grid,
                                  params=params, t=torch.zeros(1, dtype=torch.float64, device="cuda"),
                                  dt=torch.zeros(1, dtype=torch.float64, device="cuda")))
    ens_leg, alone = legs["ens"], legs["alone"]
    assert sum(1 for x in ens_leg if x["grid"].Mz >= MIN_MZ) == 1
    say("  %d members, %d columns" % (len(ens_leg), sum(x["grid"].Mx * x["grid"].My for x in ens_leg)))
    ens = S.Ensemble([x["sia"] for x in ens_leg], [x["geom"] for x in ens_leg], [x["g"]["E"] for x in ens_leg])
    # checkers of the sheet's current state: the surrogate and the FP64 fast update
    sheet = ens_leg[0]
    chk_s = S.SIAFD(sheet_grid, sheet_params)
    chk_s.use_surrogate(sg)
    chk_p = S.SIAFD(sheet_grid, sheet_params)
    worst = []
    for c in range(STEPS // CHUNK):
        ens.step(CHUNK, graph=True)
        for x in alone:
            for _ in range(CHUNK):
                x["sia"].step(x["geom"], x["g"]["E"], fast=True, t_end=0.0, t_dev=x["t"], dt_dev=x["dt"])
        torch.cuda.synchronize()
        ens.check_errors()
        for i, (a, b) in enumerate(zip(ens_leg, alone)):
            for k in ("H", "h", "mask"):
                assert torch.equal(a["g"][k], b["g"][k]), (c, i, k)
            assert ens.time(i)[0] == b["t"].item(), (c, i)
        g = sheet["g"]
        chk_s.update(g["H"], g["b"], g["h"], g["mask"], g["E"], fast=True)
        chk_p.update(g["H"], g["b"], g["h"], g["mask"], g["E"], fast=True)
        qs, qp = chk_s.diffusive_flux(), chk_p.diffusive_flux()
        nz = qp != 0
        err = float((np.abs(qs[nz] - qp[nz]) / np.abs(qp[nz])).max())
        worst.append(err)
        say("  step %d: sheet flux max rel err %.2e (counts %s)" % ((c + 1) * CHUNK, err, chk_s.surrogate_counts()))
        assert err <= 1e-3, (c, err)
    say("  max over the checks %.2e" % max(worst))
    ens.close()
    for h in (chk_s, chk_p):
        h.close()
    for leg in legs.values():
        for x in leg:
            x["sia"].close()
    sg.close()
