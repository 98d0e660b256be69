#!/usr/bin/env python3
"""Throughput of the MI355X per-RK-stage hot path (one GPU).

Workload: supersonic isothermal-wall channel, 512 x 256 x 256 cells, bulk
Mach 1.5, Re_tau 550, gamma 1.4, Pr 0.72, power-law viscosity 0.76, 6th-order
KEEP central + WENO5 where the Ducros sensor exceeds 0.9, Wray RK3 at CFL 0.8,
constant-flow-rate forcing; seeded laminar Poiseuille + 10% noise initial
field (seed 0).  A step is one full RK3 time step (3 stages, CFL dt recomputed
every step); the field is resident in HBM before timing starts.

metric: Gcell-updates/s per RK stage = nx*ny*nz * 3 * steps / seconds / 1e9.

Prints one JSON line.  The roofline object covers the kernel family with the
largest total time, from HIP events recorded on the solver's own stream
around every launch; the CPU baseline times the oracle (a restatement of the
reference CPU kernels) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "streams-2_amd"))

HBM_PEAK_GBS = 8000.0

# Compulsory bytes of each kernel family per unit (FP64).  "cell" = interior
# cell, "point" = point of the ghosted box.  Each array the kernel touches is
# counted once: reads of w_aux(1:6)+sensor = 56 B, fl read-modify-write of
# 5 variables = 80 B, etc.
ALGO_BYTES = {
    "aux": (96.0, "point"),      # read w(5) 40 B + write w_aux(1:7) 56 B
    "euler_x": (96.0, "cell"),   # read w_aux(1:6,8) 56 B + write fl(5) 40 B
    "visc": (120.0, "cell"),     # read w_aux(2,3,4,6,7) 40 B + fl(2:5) rw 64 B + write w_aux(9,10) 16 B
    "euler_y": (136.0, "cell"),  # read w_aux(1:6,8) 56 B + fl(5) rw 80 B
    "euler_z": (136.0, "cell"),
    "sensor": (24.0, "cell"),    # read w_aux(9,10) + write w_aux(8)
    "visc_div": (144.0, "cell"), # w_aux(2,3,4,7,10) 40 + fl(2:5) rw 64 + fl_prev(1,2) 16 + w(1,2) 16 + T 8
    "update": (168.0, "cell"),   # fl_prev 40 + fl 40 + w rw 80 + u 8
    "dt": (40.0, "cell"),        # read w(5)
}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of the same workload (scaled by the copy
    probe's factors), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("workload") != workload:
        return None
    return d.get("bytes_per_launch", {}).get(kernel)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--retau", type=float, default=550.0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target duration of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain device copy of 2 GiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def copy_probe():
    """A plain 2 GiB device-to-device copy, launched a few times: the HBM
    counters of this dispatch against its known 2 GiB read + 2 GiB write give
    the scale factor applied to the solver kernels' FETCH_SIZE/WRITE_SIZE."""
    import torch

    n = 1 << 28
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n}))


def cpu_baseline(args):
    """Oracle (plain C restatement of the reference CPU kernels, OpenMP over
    k) on a slab of the same channel: full nx x ny, nz reduced to 16."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle import Oracle
    from streams_mi355x import make_channel

    nzs = 16
    c = make_channel(nx=args.nx, ny=args.ny, nz=nzs, retau=args.retau)
    o = Oracle(c.params, c.grid)
    o.w[:] = c.w0
    o.init_run()
    o.advance(1)  # first touch / warm caches
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.advance(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= args.cpu_seconds or steps >= 20:
            break
    rate = args.nx * args.ny * nzs * 3 * steps / el / 1e9
    return {"value": rate, "unit": "Gcell-updates/s per RK stage", "cores": threads, "kind": "port",
            "sample": f"oracle, channel {args.nx}x{args.ny}x{nzs} slab of the same workload, {steps} RK3 steps "
                      f"in {el:.1f} s, {threads} OpenMP threads"}


def main():
    args = parse()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the hot path runs on a single MI355X; multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    from streams_mi355x import make_channel
    from streams_mi355x.solver import ChannelSolver

    nx, ny, nz = args.nx, args.ny, args.nz
    c = make_channel(nx=nx, ny=ny, nz=nz, retau=args.retau)
    g = ChannelSolver(c.params, c.grid)
    g.copy_from_field(c.w0)
    g.init_run()
    g.advance(args.warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.advance(args.steps)
    g.synchronize()
    el = time.perf_counter() - t0
    cells = nx * ny * nz
    value = cells * 3 * args.steps / el / 1e9
    res = g.compute_residual()
    finite = res["residual_rhou"] == res["residual_rhou"]
    weno_frac = [float(x) for x in g.count_weno()]

    # per-kernel timing on the solver's stream (separate, short pass)
    kt = g.kernel_times(min(20, max(3, args.steps // 10)))
    nvol = (nx + 6) * (ny + 6) * (nz + 6)
    fam_total = {k: ms * n for k, (ms, n) in kt.items()}
    dom = max((k for k in kt if k in ALGO_BYTES), key=lambda k: fam_total[k])
    per, unit = ALGO_BYTES[dom]
    units = nvol if unit == "point" else cells
    ms = kt[dom][0]
    achieved = per * units / (ms * 1e-3) / 1e9
    per_kernel = {k: {"ms": round(v[0], 4), "launches": v[1],
                      "GB/s": (round(ALGO_BYTES[k][0] * (nvol if ALGO_BYTES[k][1] == "point" else cells)
                                     / (v[0] * 1e-3) / 1e9, 1) if k in ALGO_BYTES else None)}
                  for k, v in kt.items()}

    out = {
        "metric": "Gcell-updates/s per RK stage (FP64)",
        "value": value,
        "unit": "Gcell-updates/s per RK stage",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded laminar Poiseuille + 10% uniform velocity noise, seed 0)",
        "config": {"workload": f"channel {nx}x{ny}x{nz}, M_b 1.5, Re_tau {args.retau:g}, FP64 plain, "
                               "KEEP6 + WENO5 (Ducros > 0.9), Wray RK3, CFL 0.8",
                   "grid": [nx, ny, nz], "precision": "fp64", "finite": bool(finite),
                   "weno_face_fraction_xyz": weno_frac},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"channel_{nx}x{ny}x{nz}_fp64", dom),
                     "algorithmic_bytes_per_launch": per * units, "avg_launch_ms": ms},
        "kernels": per_kernel,
    }
    g.close()
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args)
        except Exception as e:  # the oracle is optional on a bare box
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
