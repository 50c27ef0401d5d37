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

The CPU baseline times the oracle (the CPU restatement of the reference, one
core) on seeded random windows cut from the same sheet, in the same update
mode as value; it is a reported baseline, not the target.

The default run also measures the batched ensemble (64 members of uneven
sizes, one kernel launch per stage for all members, 50 steps captured in a
HIP graph) and reports it under "ensemble", beside the same members stepped
with per-member launches.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--path surrogate|plain]
                       [--mode full|fast] [--no-cpu-baseline] [--no-ensemble] [--copy-calibration GIB]
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

MX, MY, MZ, LZ, DX = 1501, 2801, 401, 4000.0, 1000.0
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def staggered_ks(thks, z):
    """kBelowHeight (IceGrid.cc:379-391, clamped to Mz-2) of every staggered
    column, from the smoothed thickness the library used (thk = mean of the
    two cells, SIAFD.cc:631-633); -1 where thk == 0 (delta = 0, only I[0])."""
    Mx, My = thks.shape
    out = []
    for q2 in (np.roll(thks, -1, axis=0), np.roll(thks, -1, axis=1)):
        thk = 0.5 * (thks + q2)
        ks = np.clip(np.searchsorted(z, thk, side="right") - 1, 0, len(z) - 2)
        out.append(np.where(thk > 0.0, ks, -1))
    return np.stack(out, axis=-1)  # [Mx, My, 2]


def algorithmic_bytes(thks, z, Mz, full=True):
    """Bytes each stage must move at least once, for one update of the sheet.

    SIA flux (column kernel): E of each map column up to the largest ks of the
    four staggered columns that read it (8 B a level); I of each staggered
    column up to its ks (8 B a level, full updates); 2-D: thksmooth, theta,
    h_x, h_y read, flux, D_stag, ks written (88 B a map column).
    SIA 3D hor. vel. (fused u, v, w kernel): u, v, w written at every level
    (24 Mz B a map column); I of every staggered column up to its ks; ks,
    h_x, h_y and mask read (48 B a map column)."""
    ks = staggered_ks(thks, z)  # [Mx, My, 2]
    lev_I = np.where(ks < 0, 1, ks + 1)  # thk == 0 stores I[0] only
    # E levels of map column (i,j): faces (i,j,0), (i,j,1), (i-1,j,0), (i,j-1,1)
    kE = np.maximum.reduce([ks[:, :, 0], ks[:, :, 1], np.roll(ks[:, :, 0], 1, axis=0),
                            np.roll(ks[:, :, 1], 1, axis=1)])
    e_bytes = 8.0 * float(np.where(kE < 0, 0, kE + 1).sum())
    i_bytes = 8.0 * float(lev_I.sum())
    n = thks.size
    flux = e_bytes + 88.0 * n + (i_bytes if full else 0.0)
    out = {"SIA flux": flux}
    if full:
        out["SIA 3D hor. vel."] = 24.0 * Mz * n + i_bytes + 48.0 * n
    else:
        # surrogate: the feature pass reads E of the listed columns and writes
        # 272 B per staggered column (34 inputs, 32 cell baselines, C S_t);
        # the network reads those back and writes D and the flux
        nstag = float((ks >= 0).sum())
        out["surrogate features"] = e_bytes + 40.0 * nstag + 272.0 * nstag
        out["surrogate MLP"] = (272.0 + 4.0 + 16.0) * nstag
    return out


# the surrogate network's algorithmic flops per column: 2 (34*128 + 3*128*128 + 128*32)
SG_FLOPS_PER_COLUMN = 2.0 * (34 * 128 + 3 * 128 * 128 + 128 * 32)
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


# kernels behind each profiling stage (for the rocprof traffic lookup)
STAGE_KERNEL = {"SIA flux": "k_column", "SIA 3D hor. vel.": "k_velocity_uvw",
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


def cpu_baseline(f, grid_z, params, full=True, n_windows=96, win=128, seed=7):
    """The oracle (CPU restatement of the reference, one core) on seeded
    random windows cut from the same sheet: one update (full or fast) + dt +
    mass continuity per window. Returns column-updates/s, CPU seconds, sample."""
    import binding as oracle
    import sia_types as T
    oracle.build()
    rng = np.random.default_rng(seed)
    cols, secs = 0, 0.0
    Mz = len(grid_z)
    for _ in range(n_windows):
        i0 = int(rng.integers(0, MX - win))
        j0 = int(rng.integers(0, MY - win))
        st = {}
        for k in ("H", "b", "h", "mask"):
            st[k] = f[k].view(MX, MY)[i0:i0 + win, j0:j0 + win].contiguous().cpu().numpy().ravel()
        E = f["E"].view(MX, MY, Mz)[i0:i0 + win, j0:j0 + win].contiguous().cpu().numpy().ravel()
        grid = T.make_grid(win, win, DX, DX, grid_z)
        t0 = time.perf_counter()
        oracle.step(grid, params, st, E, full=full)
        secs += time.perf_counter() - t0
        cols += win * win
    return cols / secs, secs, "%d windows of %dx%d columns x %d levels at seeded random positions of the sheet, %s updates" % (
        n_windows, win, win, Mz, "full" if full else "fast")


TRAIN_SEEDS = [101, 202, 303, 404, 505, 606, 707, 808, 909, 1010, 1111, 1212]


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


ENS_SIZES = (201, 401, 601, 801)


def ensemble_members(n=64, seed=2026):
    """The ensemble workload: n members with seeded sizes Mx, My in ENS_SIZES,
    dx ~ U[2, 5] km, Mz ~ U{101..201} (Lz 4000 m), GPBLD with e ~ U[1, 5] and
    n in {3, 4}, each a polythermal sheet of its own seed."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n):
        Mx, My = (int(v) for v in rng.choice(ENS_SIZES, 2))
        out.append(dict(Mx=Mx, My=My, dx=float(rng.uniform(2e3, 5e3)), Mz=int(rng.integers(101, 202)),
                        e=float(rng.uniform(1.0, 5.0)), n=float(rng.choice([3.0, 4.0])), seed=1000 + m))
    return out


def run_ensemble(S, torch, steps=50, reps=2):
    """One persistent-launch-per-stage ensemble step of 64 members, 50 steps
    captured in a HIP graph; beside it the same members stepped one by one
    (sia_step per member, its own launches)."""
    import sia_types as T
    import workloads as W
    spec = ensemble_members()
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
        t0 = time.perf_counter()
        sg = S.Surrogate(seed=5, capacity=60_000_000, stream=stream.cuda_stream)
        collector = S.SIAFD(grid, params, stream=stream.cuda_stream)
        collector.set_sync_check(False)
        n = 0
        for seed in TRAIN_SEEDS:
            ft = S.generate_greenland_like(grid, params, seed=seed)
            n = sg.collect(collector, ft["H"], ft["b"], ft["h"], ft["mask"], ft["E"], stride=1)
            del ft
        collector.close()
        torch.cuda.synchronize()
        t_collect = time.perf_counter() - t0
        sg.fit_normalization()
        t1 = time.perf_counter()
        loss = sg.train(args.train_steps, batch=args.train_batch, lr_start=2e-3, lr_end=1e-5, seed=9)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t1
        log("  %d samples (%.1f s), %d steps of %d (%.1f s), loss %.3e -> %.3e" % (
            n, t_collect, args.train_steps, args.train_batch, t_train, loss[:100].mean(), loss[-100:].mean()))
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
            "training": {"samples": int(n), "sheets": len(TRAIN_SEEDS), "steps": args.train_steps,
                         "batch": args.train_batch, "collect_s": round(t_collect, 2), "train_s": round(t_train, 2),
                         "loss_first100": float(loss[:100].mean()), "loss_last100": float(loss[-100:].mean())},
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

    # roofline of the dominant kernel stage; algorithmic bytes from the
    # smoothed thickness of the initial state
    plain.update(f["H"], f["b"], f["h"], f["mask"], f["E"], fast=True)
    _, thks = plain.theta_thksmooth()
    alg = algorithmic_bytes(thks.reshape(MX, MY), z, MZ, full=not fast_mode)
    prefer = ["surrogate features", "surrogate MLP"] if args.path == "surrogate" else list(alg)
    roofline = roofline_of(stages, alg, prefer)
    if args.path == "surrogate":
        mlp_ms = stages["surrogate MLP"][0] / stages["surrogate MLP"][1]
        ncol = out_extra["surrogate_error"]["columns"]["surrogate"]
        tf = SG_FLOPS_PER_COLUMN * ncol / (mlp_ms / 1000.0) / 1e12
        out_extra["mlp_roofline"] = {"bound": "mfma", "achieved": round(tf, 1), "peak": BF16_PEAK_TFLOPS,
                                     "unit": "TFLOP/s", "frac": round(tf / BF16_PEAK_TFLOPS, 4),
                                     "kernel": "surrogate MLP", "avg_launch_ms": round(mlp_ms, 4)}

    if not args.no_ensemble:
        log("64-member ensemble (one launch per stage, HIP graph of 50 steps)")
        if args.path == "surrogate":
            sur.close()
            sg.close()  # frees the training samples
        out_extra["ensemble"] = run_ensemble(S, torch)
        log("  %s" % out_extra["ensemble"])

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle, 1 core)")
        v, secs, sample = cpu_baseline(f, z, params, full=not fast_mode)
        cpu = {"value": round(v, 1), "unit": "column-updates/s", "cores": 1, "kind": "port",
               "sample": sample + "; %.1f s of CPU work" % secs}
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
        "cpu_baseline": cpu,
    }
    out.update(out_extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
