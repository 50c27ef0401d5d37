"""Benchmark: SIA column-updates/s on the 1501 x 2801 x 401 polythermal sheet.

One step is one SIA-only PISM time step on one MI355X, entirely on the device
and in FP64: a full SIA update (surface gradient, flux and D, the column
integral I, 3-D u, v and w), the diffusivity-limited dt, the explicit mass
continuity step and the surface/mask update (iceModel.cc:601-794 restricted to
the SIA path). value = columns (1501 * 2801) per step / seconds per step.

Inputs are generated in HBM before the timed region (synthetic data of that
shape, see pismv0.7.3_amd/workloads.py:greenland_like). The timed region is K
steps between two synchronisations; per-stage durations come from hipEvents
recorded by the library on the same stream (the reference's Profiling event
names). The CPU baseline times the oracle (single core) on seeded windows of
the same sheet; it is a reported baseline, not the target.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode full|fast]
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


def kbelow_counts(H, z):
    """ks(H) + 1 per column: kBelowHeight (IceGrid.cc:379-391) clamped to Mz-2."""
    ks = np.searchsorted(z, H, side="right") - 1
    ks = np.clip(ks, 0, len(z) - 2)
    return ks + 1


def algorithmic_bytes(H, z, Mz):
    """Per-stage algorithmic bytes of one update over all columns:
    flux: E read to ks (8 (ks+1)) + 88 B of 2-D traffic per column;
    3-D hor. vel.: u, v written (16 Mz); vert. vel.: w written (8 Mz)."""
    n = H.size
    e_bytes = 8.0 * float(kbelow_counts(H, z).sum())
    return {"SIA flux": e_bytes + 88.0 * n, "SIA 3D hor. vel.": 16.0 * Mz * n,
            "SB vert. vel.": 8.0 * Mz * n}


def cpu_baseline(n_windows=32, win=96, seed=7):
    """Oracle (CPU restatement, one core) on seeded windows of the same sheet:
    full update + dt + mass continuity per window. Returns column-updates/s."""
    import binding as oracle
    import workloads as W
    oracle.build()
    rng = np.random.default_rng(seed)
    cols, secs = 0, 0.0
    for _ in range(n_windows):
        i0 = int(rng.integers(0, MX - win))
        j0 = int(rng.integers(0, MY - win))
        w = W.greenland_like(i0, i0 + win, j0, j0 + win)
        st = {k: v.copy() for k, v in w.fields.items()}
        E = st.pop("E")
        t0 = time.perf_counter()
        oracle.step(w.grid, w.params, st, E, full=True)
        secs += time.perf_counter() - t0
        cols += win * win
    return cols / secs, secs, "%d windows of %dx%d columns (x %d levels) at seeded random positions" % (
        n_windows, win, win, MZ)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=("full", "fast"), default="full")
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
    H0 = f["H"].cpu().numpy()
    stream = torch.cuda.current_stream()
    sia = S.SIAFD(grid, params, stream=stream.cuda_stream)
    sia.set_sync_check(False)
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
    assert lib.sia_check_errors(sia.h) == 0
    ms_total = start.elapsed_time(end)
    ms_per_step = ms_total / args.steps
    value = MX * MY / (ms_per_step / 1000.0)
    log("ms/step %.3f (wall %.3f s for %d steps)" % (ms_per_step, wall, args.steps))
    for k, (ms, n) in stages.items():
        if n:
            log("  %-18s %8.3f ms avg over %d" % (k, ms / n, n))

    # roofline of the dominant stage (largest total time among the kernel stages)
    alg = algorithmic_bytes(H0, z, MZ)
    if fast:
        alg = {"SIA flux": alg["SIA flux"]}
    dom = max((k for k in alg if stages.get(k, (0, 0))[1] > 0), key=lambda k: stages[k][0])
    dom_ms = stages[dom][0] / stages[dom][1]
    achieved = alg[dom] / (dom_ms / 1000.0) / 1e9
    traffic = None
    tfile = os.path.join(REPO, "profiles", "traffic.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile)).get(dom)
        if t:
            traffic = t.get("bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "algorithmic_bytes": alg[dom], "avg_launch_ms": round(dom_ms, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle, 1 core)")
        v, secs, sample = cpu_baseline()
        cpu = {"value": round(v, 1), "unit": "column-updates/s", "cores": 1, "kind": "port",
               "sample": sample + "; %.1f s of CPU work" % secs}
    out = {
        "metric": "SIA column-updates/s at 1501x2801x401 (plain FP64 %s update + dt + mass continuity)" % args.mode,
        "value": round(value, 1), "unit": "column-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded polythermal sheet generated in HBM)",
        "config": {"workload": "greenland_like 1501x2801x401, dx=1 km, GPBLD e=3 n=3, haseloff, "
                               "bed_smoother_range=5 km, SIA %s update + explicit mass continuity" % args.mode,
                   "Mx": MX, "My": MY, "Mz": MZ, "mode": args.mode},
        "stages_ms": {k: round(ms / n, 4) for k, (ms, n) in stages.items() if n},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
