"""Benchmark: SIA column-updates/s on the 1501 x 2801 x 401 polythermal sheet.

One step is one SIA-only PISM time step on one MI355X, entirely on the device
and in FP64: a full SIA update (surface gradient, flux and D, the column
integral I, 3-D u, v and w), the diffusivity-limited dt, the explicit mass
continuity step and the surface/mask update (iceModel.cc:601-794 restricted to
the SIA path). value = columns (1501 * 2801) per step / seconds per step.

Inputs are generated in HBM before the timed region (synthetic data of that
shape, see pismv0.7.3_amd/workloads.py:greenland_like). The timed region is K
steps between two synchronisations; per-stage durations come from hipEvents
that the library records on the stream its kernels run on (the reference's
Profiling event names). The roofline entry is that of the stage with the
largest time: algorithmic bytes (see algorithmic_bytes) over its average
launch time; its "traffic" is the copy-calibrated FETCH_SIZE + WRITE_SIZE per
launch from the committed rocprofv3 summary, when present.

The CPU baseline times the oracle (the CPU restatement of the reference, one
core) on seeded random windows cut from the same sheet; it is a reported
baseline, not the target.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode full|fast]
                       [--no-cpu-baseline] [--copy-calibration GIB]
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
    return out


# kernels behind each profiling stage (for the rocprof traffic lookup)
STAGE_KERNEL = {"SIA flux": "k_column", "SIA 3D hor. vel.": "k_velocity_uvw"}


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


def cpu_baseline(f, grid_z, params, n_windows=96, win=128, seed=7):
    """The oracle (CPU restatement of the reference, one core) on seeded
    random windows cut from the same sheet: full update + dt + mass
    continuity per window. Returns column-updates/s, CPU seconds, sample."""
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
        oracle.step(grid, params, st, E, full=True)
        secs += time.perf_counter() - t0
        cols += win * win
    return cols / secs, secs, "%d windows of %dx%d columns x %d levels at seeded random positions of the sheet" % (
        n_windows, win, win, Mz)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=("full", "fast"), default="full")
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
    log("generating the %dx%dx%d sheet on the device" % (MX, MY, MZ))
    f = S.generate_greenland_like(grid, params)
    stream = torch.cuda.current_stream()
    sia = S.SIAFD(grid, params, stream=stream.cuda_stream)
    sia.set_sync_check(False)
    # the CPU baseline windows are cut from the initial sheet (E is never modified)
    f0 = None
    if not args.no_cpu_baseline:
        f0 = {k: v.clone() for k, v in f.items() if k != "E"}
        f0["E"] = f["E"]
    geom = S.geometry(f["H"], f["h"], f["mask"], f["b"])
    t_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    dt_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    fast = args.mode == "fast"
    log("warmup %d steps" % args.warmup)
    for _ in range(args.warmup):
        sia.step(geom, f["E"], fast=fast, t_end=0.0, t_dev=t_dev, dt_dev=dt_dev)
    torch.cuda.synchronize()
    lib = S.lib()
    rc = lib.sia_check_errors(sia.h)
    assert rc == 0, S.lib().sia_last_error(sia.h)
    sia.profiling_read()  # clear
    sia.profiling(True)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        sia.step(geom, f["E"], fast=fast, t_end=0.0, t_dev=t_dev, dt_dev=dt_dev)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    sia.profiling(False)
    stages = sia.profiling_read()
    assert lib.sia_check_errors(sia.h) == 0, S.lib().sia_last_error(sia.h)
    ms_total = start.elapsed_time(end)
    ms_per_step = ms_total / args.steps
    value = MX * MY / (ms_per_step / 1000.0)
    log("ms/step %.3f (wall %.3f s for %d steps)" % (ms_per_step, wall, args.steps))
    for k, (ms, n) in stages.items():
        if n:
            log("  %-18s %8.3f ms avg over %d" % (k, ms / n, n))

    if args.copy_calibration > 0:
        n = int(args.copy_calibration * 2 ** 30) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log("copy calibration: %d bytes read and written" % (8 * n))

    # roofline of the dominant stage (largest time among the kernel stages);
    # algorithmic bytes from the smoothed thickness of the last update
    _, thks = sia.theta_thksmooth()
    alg = algorithmic_bytes(thks.reshape(MX, MY), z, MZ, full=not fast)
    dom = max((k for k in alg if stages.get(k, (0, 0))[1] > 0), key=lambda k: stages[k][0])
    dom_ms = stages[dom][0] / stages[dom][1]
    achieved = alg[dom] / (dom_ms / 1000.0) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": stage_traffic(dom), "kernel": dom,
                "algorithmic_bytes": round(alg[dom]), "avg_launch_ms": round(dom_ms, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle, 1 core)")
        v, secs, sample = cpu_baseline(f0, z, params)
        cpu = {"value": round(v, 1), "unit": "column-updates/s", "cores": 1, "kind": "port",
               "sample": sample + "; %.1f s of CPU work" % secs}
    out = {
        "metric": "SIA column-updates/s at 1501x2801x401: surrogate vs plain FP64 at stated err",
        "value": round(value, 1), "unit": "column-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded polythermal sheet generated in HBM)",
        "config": {"workload": "greenland_like sheet 1501x2801x401, dx=1 km, 200 seeded Gaussian bed bumps, "
                               "dome up to 3200 m, polythermal enthalpy; GPBLD e=3 n=3, haseloff gradient, "
                               "bed_smoother_range=5 km; one step = SIA %s update + dt + explicit mass "
                               "continuity" % args.mode,
                   "path": "plain FP64", "Mx": MX, "My": MY, "Mz": MZ, "mode": args.mode},
        "stages_ms": {k: round(ms / n, 4) for k, (ms, n) in stages.items() if n},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
