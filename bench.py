"""Benchmark: SIA column-updates/s on the 1501 x 2801 x 401 polythermal sheet,
learned surrogate vs plain FP64, with the surrogate's error stated.

One step is one SIA-only PISM time step on one MI355X, entirely on the device:
an SIA update, the diffusivity-limited dt, the explicit mass continuity step
and the surface/mask update (iceModel.cc:601-794 restricted to the SIA path).
value = columns (1501 * 2801) per step / seconds per step.

Default (--path surrogate): the learned column surrogate on fast (flux-only)
updates. It is first trained on the GPU (this library's forward, backward and
Adam kernels) on about 48M columns that the plain FP64 path computes on 12
other seeds of the sheet; the benchmark sheet is never trained on. The same
run times the plain FP64 step in fast and full mode, and states the
surrogate's error: D and flux of every column against the FP64 path on the
benchmark sheet. --path plain times the FP64 path alone (--mode full|fast).

Inputs are generated in HBM before the timed region (synthetic data of that
shape, see pismv0.7.3_amd/workloads.py:greenland_like). A timed region is K
steps between two synchronisations; per-stage durations come from hipEvents
that the library records on the stream its kernels run on (the reference's
Profiling event names, plus the surrogate's two kernels). The roofline entry
is that of the kernel stage with the largest time on the measured path:
algorithmic bytes (see algorithmic_bytes) over its average launch time; its
"traffic" is the copy-calibrated FETCH_SIZE + WRITE_SIZE per launch from the
committed rocprofv3 summary, when present.

The CPU baseline times the oracle (the CPU restatement of the reference) as
PISM runs on MPI ranks: oracle/sia_mpi under mpiexec, one rank per host core
(at most 16), a 2-D block decomposition with width-2 periodic ghosts and a
global max of D, on the whole sheet (fast updates) in the same update mode as
value; it is a reported baseline, not the target.

The default run also measures the batched ensemble (64 members of uneven
sizes, one kernel launch per stage for all members, 50 steps captured in a
HIP graph) and reports it under "ensemble", beside the same members stepped
with per-member launches.

It also runs the combined workload (the sheet + 63 ensemble members in one
batched ensemble, the surrogate bound to the sheet) for 200 steps with the
surrogate trained above, beside an all-plain FP64 run of the same seeds, and
reports it under "combined" (--no-combined skips it).

--combined runs the combined workload alone at full length and prints its own
line: 5000 steps in HIP graphs of 100, one surrogate trained per Glen
exponent, the surrogate's flux error checked in FP64 every 100 steps, beside
an all-plain FP64 run of the same seeds.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--path surrogate|plain]
                       [--mode full|fast] [--no-cpu-baseline] [--no-ensemble] [--no-combined]
                       [--copy-calibration GIB] [--combined]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(REPO, "pismv0.7.3_amd"), os.path.join(REPO, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

from training import SG_FALLBACK_FRACTION, TRAIN_SEEDS, train_sheet_surrogate  # noqa: E402

MX, MY, MZ, LZ, DX = 1501, 2801, 401, 4000.0, 1000.0
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(H, z, Mz, full=True):
    """The bytes an update must move at least once (the algorithmic bytes of
    the column-update unit): per map column c, with ks_c = kBelowHeight(H_c)
    (IceGrid.cc:379-391, clamped to Mz-2),
        B_fast(c) = 8 (ks_c + 1) + 88,   B_full(c) = B_fast(c) + 24 Mz:
    E read up to ks_c; 88 B of 2-D fields (H, b, h, mask and the five bed
    smoother fields read, the two fluxes written); u, v, w written at every
    level. Intermediates (delta, I, the surrogate's feature rows) are not
    algorithmic. Returns the step's bytes ("step") and the share each kernel
    stage must move:
      SIA flux (FP64 column kernel): E to ks_c + 88 B;
      SIA 3D hor. vel. (u, v, w and the strain heating Sigma in one pass):
        32 Mz B written + E to ks_c (Sigma) + 56 B of 2-D inputs;
      surrogate features: E to ks_c + 48 B (thksmooth, theta, h_x, h_y);
      surrogate MLP: 32 B (D and the flux of both staggered columns)."""
    ks = np.clip(np.searchsorted(z, H, side="right") - 1, 0, len(z) - 2)
    n = float(H.size)
    e_bytes = 8.0 * float((ks + 1).sum())
    out = {"step": e_bytes + 88.0 * n + (24.0 * Mz * n if full else 0.0), "SIA flux": e_bytes + 88.0 * n,
           "surrogate features": e_bytes + 48.0 * n, "surrogate MLP": 32.0 * n}
    if full:
        out["SIA 3D hor. vel."] = 32.0 * Mz * n + e_bytes + 56.0 * n
    return out


# the surrogate network's algorithmic flops per column: 2 (34*128 + 3*128*128 + 128*32)
SG_FLOPS_PER_COLUMN = 2.0 * (34 * 128 + 3 * 128 * 128 + 128 * 32)
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


# kernels behind each profiling stage (for the rocprof traffic lookup)
STAGE_KERNEL = {"SIA flux": "k_column", "SIA 3D hor. vel.": "k_velocity_tile",
                "surrogate features": "k_sg_features<false", "surrogate MLP": "k_sg_mlp"}


def stage_traffic(stage):
    """HBM bytes per launch of the stage's kernel from the committed PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    tfile = os.path.join(REPO, "profiles", "traffic_sheet_1501x2801x401.json")
    if not os.path.exists(tfile):
        return None
    t = json.load(open(tfile))
    for name, v in t.get("kernels", {}).items():
        if STAGE_KERNEL.get(stage, "?") in name:
            return round(v["bytes_per_launch"])
    return None


MPIEXEC = "/opt/conda/bin/mpiexec"
MPI_DRIVER = os.path.join(REPO, "oracle", "sia_mpi")
# window of the sheet timed by a full-update baseline (the whole sheet's 3-D
# work fields would not fit a host): a quarter of the sheet around its centre
FULL_WINDOW = (375, 700, 750, 1400)  # i0, j0, wx, wy


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(full=False, steps=3, warmup=1, ranks=None):
    """The oracle (CPU restatement of the reference) as PISM runs on MPI ranks:
    oracle/sia_mpi under MPICH's mpiexec, one rank per host core (at most 16,
    the host share of one GPU), a 2-D block decomposition of the map plane with
    periodic ghosts of width 2, the gradient's width-1 ghost update and
    MPI_Allreduce(MAX) of D_max, each rank generating its own block. One step
    is the SIA update (fast: the whole sheet; full: a quarter window, whose 3-D
    work fields fit the host) + dt + mass continuity + mask, the median of
    `steps` after `warmup`. Returns (rate, seconds, ranks, sample) or None
    with the reason when MPICH or the driver is missing."""
    import subprocess
    if not os.path.exists(MPI_DRIVER):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], capture_output=True)
    if not (os.path.exists(MPIEXEC) and os.path.exists(MPI_DRIVER)):
        return None, "MPICH (mpiexec) or the oracle's MPI driver is not available on this host"
    n = ranks or min(16, os.cpu_count() or 1)
    cmd = [MPIEXEC, "-n", str(n), MPI_DRIVER, "--steps", str(steps), "--warmup", str(warmup)]
    if full:
        i0, j0, wx, wy = FULL_WINDOW
        cmd += ["--full", "1", "--i0", str(i0), "--j0", str(j0), "--wx", str(wx), "--wy", str(wy)]
    t0 = time.perf_counter()
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=900, cwd="/tmp")
    wall = time.perf_counter() - t0
    line = [ln for ln in res.stdout.splitlines() if ln.startswith("{")]
    if res.returncode != 0 or not line:
        return None, "the MPI driver failed: " + (res.stdout + res.stderr)[-300:]
    r = json.loads(line[-1])
    sample = ("%s %dx%dx%d, %s updates; one step = SIA update + dt + mass continuity + mask; %d MPI ranks "
              "(MPICH) in a %dx%d block decomposition with width-2 periodic ghosts, MPI_Allreduce(MAX) of D_max; "
              "median of %d steps after %d warm-up (%.2f s/step); %s; %.1f s wall including generation"
              % ("quarter window of the sheet" if full else "whole sheet", r["Mx"], r["My"], r["Mz"], r["mode"],
                 r["ranks"], r["px"], r["py"], steps, warmup, r["median_step_s"], _cpu_model(), wall))
    return (r["column_updates_per_s"], r["median_step_s"] * steps * r["ranks"], r["ranks"], sample), None


def clone_geometry(S, f):
    g = {k: f[k].clone() for k in ("H", "h", "mask", "b")}
    return g, S.geometry(g["H"], g["h"], g["mask"], g["b"])


def time_steps(S, torch, sia, f, steps, warmup, fast):
    """K SIA-only steps of handle sia from the initial state f; returns
    (ms per step, per-stage ms averages, geometry used)."""
    g, geom = clone_geometry(S, f)
    stream = torch.cuda.current_stream()
    t_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    dt_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    for _ in range(warmup):
        sia.step(geom, f["E"], fast=fast, t_end=0.0, t_dev=t_dev, dt_dev=dt_dev)
    torch.cuda.synchronize()
    lib = S.lib()
    assert lib.sia_check_errors(sia.h) == 0, lib.sia_last_error(sia.h)
    sia.profiling_read()
    sia.profiling(True)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(steps):
        sia.step(geom, f["E"], fast=fast, t_end=0.0, t_dev=t_dev, dt_dev=dt_dev)
    end.record(stream)
    torch.cuda.synchronize()
    sia.profiling(False)
    stages = sia.profiling_read()
    assert lib.sia_check_errors(sia.h) == 0, lib.sia_last_error(sia.h)
    return start.elapsed_time(end) / steps, stages, g


def run_ensemble(S, torch, steps=50, reps=2):
    """One persistent-launch-per-stage ensemble step of 64 members, 50 steps
    captured in a HIP graph; beside it the same members stepped one by one
    (sia_step per member, its own launches)."""
    import sia_types as T
    import workloads as W
    spec = W.ensemble_members()
    members = []
    for m in spec:
        z = T.vertical_levels(LZ, m["Mz"])
        grid = T.make_grid(m["Mx"], m["My"], m["dx"], m["dx"], z)
        params = W.gl_params(sia_enhancement_factor=m["e"], sia_Glen_exponent=m["n"])
        f = S.generate_greenland_like(grid, params, Mx_full=m["Mx"], My_full=m["My"], seed=m["seed"], dx=m["dx"])
        sia = S.SIAFD(grid, params)
        sia.set_sync_check(False)
        sia.update(f["H"], f["b"], f["h"], f["mask"], f["E"], fast=True)  # bed smoother, buffers
        members.append((sia, S.geometry(f["H"], f["h"], f["mask"], f["b"]), f, m["Mx"] * m["My"]))
    torch.cuda.synchronize()
    cols = sum(c for *_, c in members)
    ens = S.Ensemble([m[0] for m in members], [m[1] for m in members], [m[2]["E"] for m in members])
    chosen, tile_ms = ens.autotune(steps=3)
    ens.step(steps, graph=True)  # capture + first replay
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        ens.step(steps, graph=True)
    b.record()
    torch.cuda.synchronize()
    ms_graph = a.elapsed_time(b) / (reps * steps)
    ens.check_errors()
    # the same members, one sia_step per member per step
    ts = [(torch.zeros(1, dtype=torch.float64, device="cuda"), torch.zeros(1, dtype=torch.float64, device="cuda"))
          for _ in members]
    k = 5
    a.record()
    for _ in range(k):
        for (sia, geom, f, _), (t, dt) in zip(members, ts):
            sia.step(geom, f["E"], fast=True, t_end=0.0, t_dev=t, dt_dev=dt)
    b.record()
    torch.cuda.synchronize()
    ms_loop = a.elapsed_time(b) / k
    ens.close()
    for sia, *_ in members:
        sia.close()
    return {"metric": "SIA column-updates/s, batched ensemble", "value": round(cols / (ms_graph / 1000.0), 1),
            "unit": "column-updates/s", "ms_per_step": round(ms_graph, 4), "steps_per_graph": steps,
            "members": len(members), "columns": cols, "tile_cols_per_wave": chosen,
            "autotune_ms_per_step": {str(k_): round(v, 4) for k_, v in tile_ms.items()},
            "per_member_launches_ms_per_step": round(ms_loop, 4),
            "speedup_vs_per_member_launches": round(ms_loop / ms_graph, 3),
            "workload": "64 seeded members: Mx, My in {201,401,601,801}, dx U[2,5] km, Mz U{101..201}, "
                        "GPBLD e U[1,5], n in {3,4}, polythermal sheets; one step = fast SIA update + dt + "
                        "mass continuity of every member", "dtype": "f64"}


COMBINED_STEPS, COMBINED_CHECK = 5000, 100
# The surrogate's cost per column is the MLP's (fixed) plus a feature pass in
# proportion to the column's levels, the FP64 kernel's is in proportion to the
# levels alone: measured on MI355X, the surrogate is ahead for the 401-level
# sheet (ks ~ 320) and behind for 101..201-level members (ks ~ 100). The
# depth policy binds it to members of at least this many levels.
SURROGATE_MIN_MZ = 301


def combined_members(S, torch):
    """The combined workload: the 1501x2801x401 sheet plus the first 63
    members of the ensemble workload, each with its grid, parameters and
    device fields (regenerated from the seeds on every call)."""
    import sia_types as T
    import workloads as W
    z = T.vertical_levels(LZ, MZ)
    out = [dict(grid=T.make_grid(MX, MY, DX, DX, z), params=W.gl_params(), n=3.0, z=z, Mz=MZ)]
    out[0]["f"] = S.generate_greenland_like(out[0]["grid"], out[0]["params"])
    for m in W.ensemble_members()[:63]:
        zm = T.vertical_levels(LZ, m["Mz"])
        grid = T.make_grid(m["Mx"], m["My"], m["dx"], m["dx"], zm)
        params = W.gl_params(sia_enhancement_factor=m["e"], sia_Glen_exponent=m["n"])
        f = S.generate_greenland_like(grid, params, Mx_full=m["Mx"], My_full=m["My"], seed=m["seed"], dx=m["dx"])
        out.append(dict(grid=grid, params=params, n=m["n"], z=zm, Mz=m["Mz"], f=f))
    for m in out:
        m["sia"] = S.SIAFD(m["grid"], m["params"])
        m["sia"].set_sync_check(False)
        f = m["f"]
        m["sia"].update(f["H"], f["b"], f["h"], f["mask"], f["E"], fast=True)  # bed smoother, buffers
        m["geom"] = S.geometry(f["H"], f["h"], f["mask"], f["b"])
        m["cols"] = m["grid"].Mx * m["grid"].My
    torch.cuda.synchronize()
    return out


def train_combined_surrogate(S, torch, n, train_steps, batch):
    """A surrogate for Glen exponent n, trained on sheets and ensemble-type
    members of other seeds than the combined run's."""
    import sia_types as T
    import workloads as W
    t0 = time.perf_counter()
    sources = []
    if n == 3.0:
        z = T.vertical_levels(LZ, MZ)
        sources += [(T.make_grid(MX, MY, DX, DX, z), W.gl_params(), dict(seed=sd)) for sd in TRAIN_SEEDS[:6]]
    for m in W.ensemble_members(n=64, seed=4242, base=50000):
        if m["n"] == n:
            zm = T.vertical_levels(LZ, m["Mz"])
            sources.append((T.make_grid(m["Mx"], m["My"], m["dx"], m["dx"], zm),
                            W.gl_params(sia_enhancement_factor=m["e"], sia_Glen_exponent=m["n"]),
                            dict(Mx_full=m["Mx"], My_full=m["My"], seed=m["seed"], dx=m["dx"])))
    sg = S.Surrogate(seed=int(n) + 5, capacity=45_000_000)
    total = 0
    for grid, params, kw in sources:
        ft = S.generate_greenland_like(grid, params, **kw)
        c = S.SIAFD(grid, params)
        c.set_sync_check(False)
        total = sg.collect(c, ft["H"], ft["b"], ft["h"], ft["mask"], ft["E"], stride=1)
        c.close()
        del ft
    torch.cuda.synchronize()
    t_collect = time.perf_counter() - t0
    sg.fit_normalization()
    t1 = time.perf_counter()
    loss = sg.train(train_steps, batch=batch, lr_start=2e-3, lr_end=1e-5, seed=9)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    # score fallback threshold from one of the training sources
    grid, params, kw = sources[0]
    ft = S.generate_greenland_like(grid, params, **kw)
    cal = S.SIAFD(grid, params)
    tau = sg.calibrate_fallback(cal, ft["H"], ft["b"], ft["h"], ft["mask"], ft["E"], SG_FALLBACK_FRACTION)
    cal.close()
    del ft
    info = {"glen_n": n, "samples": int(total), "sources": len(sources), "collect_s": round(t_collect, 2),
            "train_s": round(t_train, 2), "loss_first100": float(loss[:100].mean()),
            "loss_last100": float(loss[-100:].mean()), "fallback_score_tau": tau,
            "fallback_fraction_calibrated": SG_FALLBACK_FRACTION}
    log("  surrogate n=%g: %s" % (n, info))
    return sg, info


def flux_check(members, rng, all_columns):
    """The surrogate's flux error on the current state of every member: the
    member's surrogate update against an FP64 update of the same state, on a
    seeded 1 % of its staggered columns (and on all of them if all_columns)."""
    worst_s, worst_a, p9999 = 0.0, 0.0, 0.0
    for m in members:
        if m.get("sg") is None:
            continue
        sia, f = m["sia"], m["f"]
        sia.update(f["H"], f["b"], f["h"], f["mask"], f["E"], fast=True)
        q_s = sia.diffusive_flux()
        sia.use_surrogate(None)
        sia.update(f["H"], f["b"], f["h"], f["mask"], f["E"], fast=True)
        q_p = sia.diffusive_flux()
        sia.use_surrogate(m["sg"])
        pick = rng.choice(q_p.size, max(1, q_p.size // 100), replace=False)
        a, b = q_s[pick], q_p[pick]
        nz = b != 0
        if nz.any():
            worst_s = max(worst_s, float((np.abs(a[nz] - b[nz]) / np.abs(b[nz])).max()))
        if all_columns:
            st = error_stats(q_s, q_p)
            worst_a = max(worst_a, st["max"])
            p9999 = max(p9999, st["p99_99"])
    return worst_s, (worst_a if all_columns else None), (p9999 if all_columns else None)


def run_combined_leg(S, torch, sgs, steps, check_every, rng=None, min_mz=0):
    """steps ensemble steps of the combined members, as HIP graphs of
    check_every steps; with sgs, members of at least min_mz levels are bound
    to the surrogate of their Glen exponent and its FP64 check runs after
    each graph, outside the timed launches. Returns the timing, the checks
    and the final geometry."""
    members = combined_members(S, torch)
    if sgs is not None:
        for m in members:
            if m["Mz"] >= min_mz:
                m["sg"] = sgs[m["n"]]
                m["sia"].use_surrogate(m["sg"])
    alg = 0.0
    for m in members:
        alg += algorithmic_bytes(m["f"]["H"].cpu().numpy(), m["z"], m["Mz"], full=False)["step"]
    ens = S.Ensemble([m["sia"] for m in members], [m["geom"] for m in members], [m["f"]["E"] for m in members])
    chosen, _ = ens.autotune(steps=3)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms_total, checks = 0.0, []
    nchunks = steps // check_every
    for c in range(nchunks):
        torch.cuda.synchronize()
        a.record()
        ens.step(check_every, graph=True)
        b.record()
        torch.cuda.synchronize()
        ms_total += a.elapsed_time(b)
        if sgs is not None:
            ws, wa, p = flux_check(members, rng, all_columns=((c + 1) % 10 == 0 or c + 1 == nchunks))
            checks.append({"step": (c + 1) * check_every, "max_sampled": ws, "max_all": wa, "p99_99_all": p})
        if (c + 1) % 5 == 0:
            log("  %s: %d steps, %.2f ms/step%s" % ("surrogate" if sgs else "all-plain", (c + 1) * check_every,
                                                   ms_total / ((c + 1) * check_every),
                                                   "" if not checks else ", max sampled flux err %.2e" % max(
                                                       x["max_sampled"] for x in checks)))
    ens.check_errors()
    final = {"H": [m["f"]["H"].cpu().numpy() for m in members], "t": [ens.time(i)[0] for i in range(len(members))]}
    counts = [m["sia"].surrogate_counts() for m in members if m.get("sg") is not None] if sgs else None
    bound = sum(1 for m in members if m.get("sg") is not None)
    cols = sum(m["cols"] for m in members)
    ens.close()
    for m in members:
        m["sia"].close()
    del members
    torch.cuda.synchronize()
    return {"ms_per_step": ms_total / steps, "cols": cols, "alg_bytes": alg, "tile": chosen, "checks": checks,
            "final": final, "counts": counts, "bound": bound}


def leg_error(leg):
    cnt = np.array(leg["counts"]).sum(axis=0)
    return {"checks": len(leg["checks"]), "every_steps": COMBINED_CHECK, "sampled_fraction": 0.01,
            "max_rel_flux_err_sampled": max(c["max_sampled"] for c in leg["checks"]),
            "max_rel_flux_err_all_columns": max(c["max_all"] for c in leg["checks"] if c["max_all"] is not None),
            "p99_99_rel_flux_err_all_columns": max(c["p99_99_all"] for c in leg["checks"]
                                                   if c["p99_99_all"] is not None),
            "against": "an FP64 update of the same state of each bound member, every %d steps (all columns every "
                       "%d steps and at the last check)" % (COMBINED_CHECK, 10 * COMBINED_CHECK),
            "bound_columns_last_check": {"zero": int(cnt[0]), "surrogate": int(cnt[1]), "fp64_fallback": int(cnt[2])}}


def final_vs(leg, ref):
    dH = [np.abs(a - b) for a, b in zip(leg["final"]["H"], ref["final"]["H"])]
    vol = [abs(a.sum() - b.sum()) / b.sum() for a, b in zip(leg["final"]["H"], ref["final"]["H"])]
    yr = 3.15569259747e7
    return {"max_abs_dH_m": float(max(d.max() for d in dH)), "sheet_max_abs_dH_m": float(dH[0].max()),
            "max_rel_volume_diff": float(max(vol)),
            "sheet_model_years": {"this_run": leg["final"]["t"][0] / yr, "all_plain": ref["final"]["t"][0] / yr}}


def run_combined(S, torch, args):
    """The combined run: the learned surrogate inside the 64-member batched
    ensemble (the 1501x2801x401 sheet + 63 ensemble members), the surrogate's
    flux error checked in FP64 every 100 steps, beside an all-plain FP64 run
    of the same seeds. Two bindings are run: the depth policy (members of at
    least SURROGATE_MIN_MZ levels; the reported value) and every member."""
    steps = args.combined_steps
    if steps < COMBINED_CHECK or steps % COMBINED_CHECK:
        raise SystemExit("--combined-steps must be a positive multiple of %d" % COMBINED_CHECK)
    log("combined run: training one surrogate per Glen exponent")
    sgs, training = {}, []
    for n in (3.0, 4.0):
        sgs[n], info = train_combined_surrogate(S, torch, n, args.train_steps, args.train_batch)
        training.append(info)
    log("combined run: surrogate on members with Mz >= %d, %d steps" % (SURROGATE_MIN_MZ, steps))
    pol = run_combined_leg(S, torch, sgs, steps, COMBINED_CHECK, np.random.default_rng(20261203),
                           min_mz=SURROGATE_MIN_MZ)
    log("combined run: surrogate on every member, %d steps" % steps)
    allb = run_combined_leg(S, torch, sgs, steps, COMBINED_CHECK, np.random.default_rng(20261203))
    for sg in sgs.values():
        sg.close()
    log("combined run: all-plain FP64 ensemble, %d steps" % steps)
    pln = run_combined_leg(S, torch, None, steps, COMBINED_CHECK)
    cols = pol["cols"]
    gbs = pol["alg_bytes"] / (pol["ms_per_step"] / 1000.0) / 1e9
    out = {
        "metric": "SIA column-updates/s, combined run: surrogate inside the 64-member ensemble",
        "value": round(cols / (pol["ms_per_step"] / 1000.0), 1), "unit": "column-updates/s", "n_gpus": 1,
        "steps": steps, "ms_per_step": round(pol["ms_per_step"], 4), "higher_is_better": True,
        "dtype": "f64+bf16", "data": "synthetic (seeded sheets generated in HBM)",
        "config": {"workload": "1501x2801x401 greenland_like sheet (GPBLD e=3 n=3) + the first 63 members of the "
                               "64-member ensemble workload (Mx, My in {201,401,601,801}, Mz 101..201, e U[1,5], "
                               "n in {3,4}); one step = fast SIA update + dt + mass continuity of every member; "
                               "%d-step HIP graphs" % COMBINED_CHECK,
                   "members": 64, "columns": cols, "tile_cols_per_wave": pol["tile"],
                   "binding": "surrogate on members with Mz >= %d (%d bound), FP64 on the others"
                              % (SURROGATE_MIN_MZ, pol["bound"])},
        "speedup_vs_all_plain": round(pln["ms_per_step"] / pol["ms_per_step"], 3),
        "hbm": {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                "algorithmic_bytes_per_step": round(pol["alg_bytes"]),
                "bytes": "E of every column up to its largest face ks (8 B a level) + 88 B of 2-D data per column"},
        "surrogate_error": leg_error(pol),
        "final_state_vs_all_plain": final_vs(pol, pln),
        "all_plain_fp64": {"value": round(cols / (pln["ms_per_step"] / 1000.0), 1),
                           "ms_per_step": round(pln["ms_per_step"], 4), "tile_cols_per_wave": pln["tile"]},
        "every_member_bound": {"value": round(cols / (allb["ms_per_step"] / 1000.0), 1),
                               "ms_per_step": round(allb["ms_per_step"], 4), "bound": allb["bound"],
                               "speedup_vs_all_plain": round(pln["ms_per_step"] / allb["ms_per_step"], 3),
                               "surrogate_error": leg_error(allb),
                               "final_state_vs_all_plain": final_vs(allb, pln)},
        "training": training,
    }
    print(json.dumps(out), flush=True)


def run_combined_short(S, torch, sg, steps):
    """The combined workload for a few graphs of COMBINED_CHECK steps with the
    benchmark's sheet surrogate bound to the sheet (the depth policy of the
    full combined run: no member of 101-201 levels is bound), beside an
    all-plain FP64 run of the same seeds."""
    log("combined workload, %d steps: the surrogate on the sheet, FP64 on the 63 members" % steps)
    pol = run_combined_leg(S, torch, {3.0: sg}, steps, COMBINED_CHECK, np.random.default_rng(20261203),
                           min_mz=SURROGATE_MIN_MZ)
    log("combined workload, %d steps: all-plain FP64" % steps)
    pln = run_combined_leg(S, torch, None, steps, COMBINED_CHECK)
    cols = pol["cols"]
    return {
        "metric": "SIA column-updates/s, combined workload: surrogate inside the 64-member ensemble",
        "value": round(cols / (pol["ms_per_step"] / 1000.0), 1), "unit": "column-updates/s",
        "steps": steps, "ms_per_step": round(pol["ms_per_step"], 4),
        "workload": "the 1501x2801x401 sheet + the first 63 members of the ensemble workload (64 members, "
                    "%d columns); one step = fast SIA update + dt + mass continuity of every member; HIP graphs "
                    "of %d steps; the surrogate of this run bound to the sheet (members with Mz >= %d), FP64 on "
                    "the others" % (cols, COMBINED_CHECK, SURROGATE_MIN_MZ),
        "tile_cols_per_wave": pol["tile"],
        "speedup_vs_all_plain": round(pln["ms_per_step"] / pol["ms_per_step"], 3),
        "all_plain_fp64": {"value": round(cols / (pln["ms_per_step"] / 1000.0), 1),
                           "ms_per_step": round(pln["ms_per_step"], 4), "tile_cols_per_wave": pln["tile"]},
        "surrogate_error": leg_error(pol),
        "final_state_vs_all_plain": final_vs(pol, pln),
    }


def error_stats(a, b):
    """Relative error of a against b over the entries where b != 0."""
    nz = b != 0
    r = np.abs(a[nz] - b[nz]) / np.abs(b[nz])
    return {"median": float(np.median(r)), "p99": float(np.percentile(r, 99)),
            "p99_99": float(np.percentile(r, 99.99)), "max": float(r.max()), "n": int(nz.sum())}


def roofline_of(stages, alg, prefer):
    """Roofline of the stage with the largest time among those priced."""
    dom = max((k for k in prefer if k in alg and stages.get(k, (0, 0))[1] > 0), key=lambda k: stages[k][0])
    dom_ms = stages[dom][0] / stages[dom][1]
    achieved = alg[dom] / (dom_ms / 1000.0) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": stage_traffic(dom), "kernel": dom,
            "algorithmic_bytes": round(alg[dom]), "avg_launch_ms": round(dom_ms, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--path", choices=("surrogate", "plain"), default="surrogate")
    ap.add_argument("--mode", choices=("full", "fast"), default="full", help="update mode of --path plain")
    ap.add_argument("--train-steps", type=int, default=12000)
    ap.add_argument("--train-batch", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64-member ensemble measurement")
    ap.add_argument("--no-combined", action="store_true",
                    help="skip the 200-step combined-workload measurement of the default run")
    ap.add_argument("--combined", action="store_true",
                    help="run the combined workload instead (surrogate inside the 64-member ensemble, 5000 steps)")
    ap.add_argument("--combined-steps", type=int, default=COMBINED_STEPS,
                    help="steps of each combined-run leg (a multiple of %d)" % COMBINED_CHECK)
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="after the timed steps, run one plain device copy of this many GiB "
                         "(the reference kernel for scaling the rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the SIA path here runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    import torch
    import pism_sia as S
    import sia_types as T
    import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.combined:
        run_combined(S, torch, args)
        return
    z = T.vertical_levels(LZ, MZ)
    grid = T.make_grid(MX, MY, DX, DX, z)
    params = W.gl_params()
    stream = torch.cuda.current_stream()
    log("generating the %dx%dx%d sheet on the device" % (MX, MY, MZ))
    f = S.generate_greenland_like(grid, params)

    plain = S.SIAFD(grid, params, stream=stream.cuda_stream)
    plain.set_sync_check(False)
    results = {}
    for mode in (("full", "fast") if args.path == "surrogate" else (args.mode,)):
        ms, stages, _ = time_steps(S, torch, plain, f, args.steps, args.warmup, fast=(mode == "fast"))
        results[mode] = (ms, stages)
        log("plain FP64 %s: %.3f ms/step %s" % (mode, ms, {k: round(v[0] / v[1], 3) for k, v in stages.items() if v[1]}))

    out_extra = {}
    if args.path == "surrogate":
        log("training the surrogate on %d other sheets" % len(TRAIN_SEEDS))
        sg, train_info = train_sheet_surrogate(S, torch, grid, params, steps=args.train_steps,
                                               batch=args.train_batch, stream=stream.cuda_stream, log=log)
        sur = S.SIAFD(grid, params, stream=stream.cuda_stream)
        sur.set_sync_check(False)
        sur.use_surrogate(sg)
        ms_s, stages_s, _ = time_steps(S, torch, sur, f, args.steps, args.warmup, fast=True)
        log("surrogate fast: %.3f ms/step %s" % (ms_s, {k: round(v[0] / v[1], 3) for k, v in stages_s.items() if v[1]}))
        # error on the initial state of the benchmark sheet (never trained on)
        plain.update(f["H"], f["b"], f["h"], f["mask"], f["E"], fast=True)
        D0, Q0 = plain.diffusivity_staggered(), plain.diffusive_flux()
        sur.update(f["H"], f["b"], f["h"], f["mask"], f["E"], fast=True)
        D1, Q1 = sur.diffusivity_staggered(), sur.diffusive_flux()
        counts = sur.surrogate_counts()
        err = {"D_stag": error_stats(D1, D0), "flux": error_stats(Q1, Q0),
               "columns": {"zero": counts[0], "surrogate": counts[1], "fp64_fallback": counts[2]},
               "against": "plain FP64 fast update of the same state (the benchmark sheet, seed 20261203)"}
        log("  error vs FP64: D %s" % err["D_stag"])
        ms, stages = ms_s, stages_s
        ms_plain_fast = results["fast"][0]
        out_extra = {
            "plain_fp64": {m: {"value": round(MX * MY / (results[m][0] / 1000.0), 1),
                               "ms_per_step": round(results[m][0], 4),
                               "stages_ms": {k: round(v[0] / v[1], 4) for k, v in results[m][1].items() if v[1]}}
                           for m in results},
            "speedup_vs_plain_fast": round(ms_plain_fast / ms_s, 3),
            "surrogate_error": err,
            "training": train_info,
        }
        fast_mode = True
    else:
        ms, stages = results[args.mode]
        fast_mode = args.mode == "fast"
    value = MX * MY / (ms / 1000.0)

    if args.copy_calibration > 0:
        nel = int(args.copy_calibration * 2 ** 30) // 8
        a = torch.ones(nel, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log("copy calibration: %d bytes read and written" % (8 * nel))

    # roofline of the dominant kernel stage and of the whole step; algorithmic
    # bytes from the thickness of the initial state
    H0 = f["H"].cpu().numpy()
    alg = algorithmic_bytes(H0, z, MZ, full=not fast_mode)
    prefer = ["surrogate features", "surrogate MLP"] if args.path == "surrogate" else \
        [k for k in alg if k != "step"]
    roofline = roofline_of(stages, alg, prefer)
    step_roofline = {}
    for name, (ms_k, full_k) in {"value": (ms, not fast_mode),
                                 **({"plain_fp64_fast": (results["fast"][0], False),
                                     "plain_fp64_full": (results["full"][0], True)} if "full" in results and
                                    "fast" in results else {})}.items():
        b = algorithmic_bytes(H0, z, MZ, full=full_k)["step"]
        gbs = b / (ms_k / 1000.0) / 1e9
        step_roofline[name] = {"bytes": round(b), "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    if args.path == "surrogate":
        mlp_ms = stages["surrogate MLP"][0] / stages["surrogate MLP"][1]
        ncol = out_extra["surrogate_error"]["columns"]["surrogate"]
        tf = SG_FLOPS_PER_COLUMN * ncol / (mlp_ms / 1000.0) / 1e12
        out_extra["mlp_roofline"] = {"bound": "mfma", "achieved": round(tf, 1), "peak": BF16_PEAK_TFLOPS,
                                     "unit": "TFLOP/s", "frac": round(tf / BF16_PEAK_TFLOPS, 4),
                                     "kernel": "surrogate MLP", "avg_launch_ms": round(mlp_ms, 4)}

    if args.path == "surrogate" and not args.no_combined:
        out_extra["combined"] = run_combined_short(S, torch, sg, 2 * COMBINED_CHECK)
        log("  %s" % out_extra["combined"])
    if args.path == "surrogate":
        sur.close()
        sg.close()  # frees the training samples
    if not args.no_ensemble:
        log("64-member ensemble (one launch per stage, HIP graph of 50 steps)")
        out_extra["ensemble"] = run_ensemble(S, torch)
        log("  %s" % out_extra["ensemble"])

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (the oracle under mpiexec, one rank per host core)")
        res, why = cpu_baseline(full=not fast_mode)
        if res is None:
            cpu = {"value": None, "unit": "column-updates/s", "cores": 0, "kind": "port", "sample": why}
        else:
            v, secs, cores, sample = res
            cpu = {"value": round(v, 1), "unit": "column-updates/s", "cores": cores, "kind": "port",
                   "sample": sample + "; %.1f s of CPU work timed" % secs}
        log("  %s" % cpu)
    path = ("surrogate (BF16 MFMA MLP over FP64 per-level moments) on fast updates"
            if args.path == "surrogate" else "plain FP64 %s updates" % args.mode)
    out = {
        "metric": "SIA column-updates/s at 1501x2801x401: surrogate vs plain FP64 at stated err",
        "value": round(value, 1), "unit": "column-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+bf16" if args.path == "surrogate" else "f64",
        "data": "synthetic (seeded polythermal sheet generated in HBM)",
        "config": {"workload": "greenland_like sheet 1501x2801x401, dx=1 km, 200 seeded Gaussian bed bumps, "
                               "dome up to 3200 m, polythermal enthalpy; GPBLD e=3 n=3, haseloff gradient, "
                               "bed_smoother_range=5 km; one step = SIA %s update + dt + explicit mass "
                               "continuity" % ("fast" if fast_mode else "full"),
                   "path": path, "Mx": MX, "My": MY, "Mz": MZ, "mode": "fast" if fast_mode else "full"},
        "stages_ms": {k: round(v[0] / v[1], 4) for k, v in stages.items() if v[1]},
        "roofline": roofline,
        "step_roofline": step_roofline,
        "cpu_baseline": cpu,
    }
    out.update(out_extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
