"""Training recipes of the learned column surrogate (include/sia_surrogate.h),
shared by bench.py and the GPU tests.

Samples always come from this library's FP64 path on fields that are not the
ones the model is judged on; training, the fallback-score calibration and
inference run on the GPU through the C ABI (pism_sia.Surrogate).
"""
import time

import sia_types as T
import workloads as W

# seeds of the 1501x2801x401 training sheets (the benchmark sheet's seed,
# 20261203, is never among them)
TRAIN_SEEDS = [101, 202, 303, 404, 505, 606, 707, 808, 909, 1010, 1111, 1212]
# share of a training field's surrogate columns whose fallback score exceeds
# the calibrated threshold (they are computed in FP64 instead)
SG_FALLBACK_FRACTION = 0.005


def train_sheet_surrogate(S, torch, grid, params, seeds=TRAIN_SEEDS, steps=12000, batch=4096, stream=None,
                          capacity=60_000_000, log=None):
    """A surrogate trained on the sheets of the given seeds (every staggered
    column of each, from the FP64 fast update), its score fallback calibrated
    at SG_FALLBACK_FRACTION on the last of them. Returns (surrogate, info)."""
    t0 = time.perf_counter()
    sg = S.Surrogate(seed=5, capacity=capacity, stream=stream)
    collector = S.SIAFD(grid, params, stream=stream)
    collector.set_sync_check(False)
    n = 0
    for k, seed in enumerate(seeds):
        ft = S.generate_greenland_like(grid, params, seed=seed)
        # each sheet has its own bed: a new bed state makes the handle smooth it again
        n = sg.collect(collector, ft["H"], ft["b"], ft["h"], ft["mask"], ft["E"], stride=1,
                       bed_state_counter=k + 1)
        del ft
    rc = S.lib().sia_check_errors(collector.h)
    assert rc == 0, S.lib().sia_last_error(collector.h)
    collector.close()
    torch.cuda.synchronize()
    t_collect = time.perf_counter() - t0
    sg.fit_normalization()
    t1 = time.perf_counter()
    loss = sg.train(steps, batch=batch, lr_start=2e-3, lr_end=1e-5, seed=9)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    # score fallback threshold from the last training sheet (never the benchmark sheet)
    ft = S.generate_greenland_like(grid, params, seed=seeds[-1])
    cal = S.SIAFD(grid, params, stream=stream)
    tau = sg.calibrate_fallback(cal, ft["H"], ft["b"], ft["h"], ft["mask"], ft["E"], SG_FALLBACK_FRACTION)
    cal.close()
    del ft
    info = {"samples": int(n), "sheets": len(seeds), "steps": steps, "batch": batch,
            "collect_s": round(t_collect, 2), "train_s": round(t_train, 2),
            "loss_first100": float(loss[:100].mean()), "loss_last100": float(loss[-100:].mean()),
            "fallback_score_tau": tau, "fallback_fraction_calibrated": SG_FALLBACK_FRACTION,
            "calibration_sheet_seed": seeds[-1]}
    if log:
        log("  %d samples (%.1f s), %d steps of %d (%.1f s), loss %.3e -> %.3e, fallback score tau %.3e" % (
            n, t_collect, steps, batch, t_train, info["loss_first100"], info["loss_last100"], tau))
    return sg, info


def train_dome_surrogate(S, torch, seeds, steps=4000, batch=2048, capacity=4_000_000):
    """A surrogate for the 61x61x31 dome grid (workloads.dome_polythermal):
    samples from the FP64 fast updates of the dome with the given enthalpy
    seeds, score fallback calibrated on the last of them. Returns
    (surrogate, samples)."""
    sg = S.Surrogate(seed=17, capacity=capacity)
    n = 0
    for seed in seeds:
        w = W.dome_polythermal(seed)
        dev = {k: S.to_device(v) for k, v in w.fields.items()}
        c = S.SIAFD(w.grid, w.params)
        n = sg.collect(c, dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], stride=1)
        c.close()
    sg.fit_normalization()
    sg.train(steps, batch=batch, lr_start=2e-3, lr_end=1e-5, seed=3)
    w = W.dome_polythermal(seeds[-1])
    dev = {k: S.to_device(v) for k, v in w.fields.items()}
    cal = S.SIAFD(w.grid, w.params)
    sg.calibrate_fallback(cal, dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], SG_FALLBACK_FRACTION)
    cal.close()
    torch.cuda.synchronize()
    return sg, int(n)


def member_grid_params(m, Lz=4000.0):
    """Grid and parameters of ensemble member spec m (workloads.ensemble_members)."""
    z = T.vertical_levels(Lz, m["Mz"])
    grid = T.make_grid(m["Mx"], m["My"], m["dx"], m["dx"], z)
    params = W.gl_params(sia_enhancement_factor=m["e"], sia_Glen_exponent=m["n"])
    return grid, params
