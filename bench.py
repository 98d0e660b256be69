#!/usr/bin/env python3
"""Throughput of the MI355X per-RK-stage hot path (one GPU).

Workload: supersonic isothermal-wall channel, 512 x 256 x 256 cells, bulk
Mach 1.5, Re_tau 550, gamma 1.4, Pr 0.72, power-law viscosity 0.76, 6th-order
KEEP central + WENO5 where the Ducros sensor exceeds 0.9, Wray RK3 at CFL 0.8,
constant-flow-rate forcing; seeded laminar Poiseuille + 10% noise initial
field (seed 0).  A step is one full RK3 time step (3 stages, CFL dt recomputed
every step); the field is resident in HBM before timing starts.

metric: Gcell-updates/s per RK stage = nx*ny*nz * 3 * steps / seconds / 1e9.

Prints one JSON line.  The roofline object covers the kernel with the
largest total time among those with a byte count (one launch per stage: the
z flux march, its WENO fix-up timed apart as "weno_fix"), from HIP events
recorded on the solver's own stream around every launch (the CFL reduction
of each step rides on the first stage's aux pass, so it is timed within
"aux"); its traffic comes from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
summary of the same workload.  The CPU baseline times the oracle (a
restatement of the reference CPU kernels) on a bounded sample of the same
workload, on every core the process may use.

Two more blocks of the same line:
* "mixed": the mixed-precision form on the same workload, its speedup, its
  relative L2 error per conserved variable against the FP64 form after 100
  steps, and the FP64 form's own difference after 100 steps from an initial
  field perturbed by 1e-7, the scale against which that error is read.
* "serving": the serving workload of 64 seeded channel solves of uneven
  sizes (grids uniform in [64,256] x [64,128] x [64,128], bulk Mach 0.3-3.0,
  Re_tau 100-500, Poisson arrivals, 200 steps each) through the library's
  continuous-batching server, with p50/p99 time-to-solution, against a
  sample of the same requests run one after another through the plain path;
  the same stream with batched launches (solves up to 2^20 cells share one
  launch per kernel per lane), and a stream of 128 small solves (grids of
  32-64 cells a side) served unbatched and batched.  Batched results are
  checked bit for bit against the unbatched ones.

--workload sbli runs the shock / boundary-layer box instead (1536 x 320 x
192, recycling inflow, imposed oblique shock, conservative viscous form) in
FP64 and mixed and prints that workload's own JSON line: rates, WENO face
fractions, per-family times and the mixed form's error after 100 steps.
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
    "euler_xy": (96.0, "cell"),  # x and y in one pass: read w_aux(1:6,8) 56 B + write fl(5) 40 B
    "visc": (120.0, "cell"),     # read w_aux(2,3,4,6,7) 40 B + fl(2:5) rw 64 B + write w_aux(9,10) 16 B
    "euler_y": (136.0, "cell"),  # read w_aux(1:6,8) 56 B + fl(5) rw 80 B
    "euler_z": (136.0, "cell"),
    # z march + viscous divergence + force_rhs partials: the march's 136 B, mu 8, div/3 8, fl_prev(1,2) 16, w(1,2) 16
    "euler_zv": (184.0, "cell"),
    "sensor": (24.0, "cell"),    # read w_aux(9,10) + write w_aux(8)
    "visc_div": (144.0, "cell"), # w_aux(2,3,4,7,10) 40 + fl(2:5) rw 64 + fl_prev(1,2) 16 + w(1,2) 16 + T 8
    # fl_prev 40 + fl 40 + w rw 80; the first Wray stage (rho = 0) skips fl_prev:
    # (120 + 160 + 160) / 3 per launch over the three stages of a step
    "update": (440.0 / 3.0, "cell"),
    "dt": (40.0, "cell"),        # read w(5)
}


def progress(msg):
    """One line per finished phase on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def timed_run(g, c, warmup, steps):
    """init + warmup steps untimed, then ``steps`` RK3 steps timed with the
    solver's stream drained on both sides; returns seconds."""
    g.copy_from_field(c.w0)
    g.init_run()
    g.advance(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.advance(steps)
    g.synchronize()
    return time.perf_counter() - t0


def mixed_form(args, c, value_fp64):
    """The mixed form (FP32 w_aux, right-hand sides, central fluxes and
    viscous stresses; FP64 state, T/p recovery, shock sensor, WENO, RK and
    reductions) on the same workload: its rate, its relative L2 error per
    conserved variable against the FP64 form after 100 steps from the same
    field, and, for scale, the FP64 form's own difference after 100 steps
    when its initial field is perturbed by 1e-7."""
    import numpy as np
    from streams_mi355x import PREC_MIXED
    from streams_mi355x.params import params_copy
    from streams_mi355x.solver import ChannelSolver

    pm = params_copy(c.params)
    pm.precision = PREC_MIXED
    gm = ChannelSolver(pm, c.grid)
    el = timed_run(gm, c, args.warmup, args.steps)
    value = args.nx * args.ny * args.nz * 3 * args.steps / el / 1e9
    kt = gm.kernel_times(min(20, max(3, args.steps // 10)))
    nerr = 100
    fields = []
    # mixed, FP64, and FP64 from the same field perturbed by 1e-7 (relative,
    # seeded): the last pair measures how far the switched scheme itself
    # carries a perturbation of FP32 size over the same 100 steps
    pert = c.w0 * (1.0 + 1e-7 * (np.random.default_rng(1).random(c.w0.shape) - 0.5))
    for g, w0 in ((gm, c.w0), (None, c.w0), (None, pert)):
        if g is None:
            g = ChannelSolver(c.params, c.grid)
        g.copy_from_field(w0)
        g.init_run()
        g.advance(nerr)
        fields.append(g.copy_to_field()[:, 3:-3, 3:-3, 3:-3])
        g.close()
    def rel(a, b):
        return [float(np.linalg.norm(a[m] - b[m]) / np.linalg.norm(b[m])) for m in range(5)]
    return {"value": value, "unit": "Gcell-updates/s per RK stage", "ms_per_step": el / args.steps * 1e3,
            "speedup_vs_fp64": value / value_fp64, "rel_l2_vs_fp64_after_100_steps": rel(fields[0], fields[1]),
            "fp64_rel_l2_under_1e-7_perturbation_after_100_steps": rel(fields[2], fields[1]),
            "kernels_ms": {k: round(v[0], 4) for k, v in kt.items()}}


def sbli_workload(args):
    """Oblique shock / boundary-layer interaction at the scale of the SBLI
    configuration: 1536x320x192 nodes over 70 x 12 x 6.5 (delta = 1), Mach 2,
    Re 200, adiabatic-recovery wall, recycling-rescaling inflow with the
    random inflow plane, supersonic outflow with the characteristic
    treatment, the oblique shock of an 8 deg wedge (37.21 deg) imposed at the
    top and meeting the wall at x = 40, with the non-reflecting top,
    periodic span, conservative viscous form, WENO5 where the Ducros sensor
    exceeds 0.2.  The timed steps start from a field that the incident shock
    already crosses (seeded BL profile plus 5 % velocity noise, shock jump
    imposed), so the WENO region of the shock is present throughout.  FP64
    and mixed rates, WENO face fractions, per-family times and the mixed
    form's error against FP64 after 100 steps from the same field."""
    import numpy as np
    from streams_mi355x import PREC_MIXED
    from streams_mi355x.cases import bl_recycling_case, shock_angle_from_deflection
    from streams_mi355x.params import params_copy
    from streams_mi355x.solver import ChannelSolver

    nx, ny, nz = args.nx, args.ny, args.nz
    beta = shock_angle_from_deflection(2.0, 8.0)
    c, rc = bl_recycling_case(nx, ny, nz, shock_angle_deg=beta, xshock_imp=40.0, developed_shock=True)
    p = c.params
    progress(f"sbli case built: {nx}x{ny}x{nz}")
    cells = nx * ny * nz
    out = {"metric": "Gcell-updates/s per RK stage (SBLI box, FP64 and mixed)",
           "config": {"workload": f"SBLI box {nx}x{ny}x{nz}, {args.steps} steps, M 2, Re 200, recycling inflow "
                                  f"(rand_type 2), 8 deg wedge shock ({beta:.3f} deg) meeting the wall at x = 40, "
                                  "developed-shock start, sensor > 0.2, conservative viscous form"}}
    fields = []
    for name, prec in (("fp64", 0), ("mixed", PREC_MIXED)):
        pp = params_copy(p)
        pp.precision = prec
        g = ChannelSolver(pp, c.grid)
        g.set_recycling(rc)
        g.copy_from_field(c.w0)
        g.init_run()
        g.advance(args.warmup)
        g.synchronize()
        t0 = time.perf_counter()
        g.advance(args.steps)
        g.synchronize()
        el = time.perf_counter() - t0
        res = g.compute_residual()
        kt = g.kernel_times(3)
        # FP64 byte counts; without channel forcing the viscous divergence
        # terms need no force_rhs inputs (104 B/cell alone, 152 B/cell folded
        # into the z march)
        nbytes = dict(ALGO_BYTES, visc_div=(104.0, "cell"), euler_zv=(152.0, "cell"))
        nvol = (nx + 6) * (ny + 6) * (nz + 6)

        def gbs(k, ms):
            if prec != 0 or k not in nbytes:
                return None
            per, unit = nbytes[k]
            return round(per * (nvol if unit == "point" else cells) / (ms * 1e-3) / 1e9, 1)
        out[name] = {"value": cells * 3 * args.steps / el / 1e9, "ms_per_step": el / args.steps * 1e3,
                     "finite": res["residual_rhou"] == res["residual_rhou"],
                     "weno_face_fraction_xyz": [float(x) for x in g.count_weno()],
                     "kernels": {k: {"ms": round(v[0], 4), "launches": v[1], "GB/s": gbs(k, v[0])}
                                 for k, v in kt.items()}}
        progress(f"sbli {name}: {out[name]['value']:.3f} Gcell-updates/s per RK stage, "
                 f"WENO {out[name]['weno_face_fraction_xyz']}")
        g.close()
        # the same 100 steps from the initial field for the error
        g = ChannelSolver(pp, c.grid)
        g.set_recycling(rc)
        g.copy_from_field(c.w0)
        g.init_run()
        g.advance(100)
        fields.append(g.copy_to_field()[:, 3:-3, 3:-3, 3:-3])
        g.close()
    out["mixed"]["speedup_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    out["mixed"]["rel_l2_vs_fp64_after_100_steps"] = [
        float(np.linalg.norm(fields[1][m] - fields[0][m]) / np.linalg.norm(fields[0][m])) for m in range(5)]
    return out


def serving(args):
    """Serving workload: a seeded stream of independent channel solves
    (grids uniform in [64,256]x[64,128]x[64,128], bulk Mach 0.3-3.0, Re_tau
    100-500, Poisson arrivals), each advancing ``--serve-steps`` RK3 steps,
    FP64.  Continuous batching through the library's server against the same
    kind of requests run one after another through the plain path; then the
    same stream in the mixed form, with each request's error against its
    FP64 result."""
    from streams_mi355x.serve import cell_updates, run_sequential, run_serving, serving_workload

    reqs = serving_workload(args.serve_requests, args.serve_steps, rate=args.serve_rate, seed=0)
    progress(f"serving workload of {len(reqs)} requests built")
    seq_n = min(len(reqs), args.serve_seq_sample)
    seq = run_sequential(reqs[:seq_n])
    progress(f"sequential sample done: {seq['gcups']:.3f} Gcell-updates/s")
    hbm = int(args.serve_hbm_gib * (1 << 30))
    srv = run_serving(reqs, hbm_bytes=hbm, max_slots=args.serve_slots, steps_per_graph=10, autotune=True,
                      want_fields=True)
    progress(f"served: {srv['gcups']:.3f} Gcell-updates/s, p50 {srv['p50_s']:.2f} s, p99 {srv['p99_s']:.2f} s")
    batched = serving_batched(args, reqs, srv, hbm)
    # the same requests in the mixed form, each compared with its FP64 served
    # result (bitwise the FP64 plain solve of that request)
    import numpy as np
    from streams_mi355x import PREC_MIXED
    reqs_m = serving_workload(args.serve_requests, args.serve_steps, rate=args.serve_rate, seed=0,
                              precision=PREC_MIXED)
    srv_m = run_serving(reqs_m, hbm_bytes=hbm, max_slots=args.serve_slots, steps_per_graph=10, autotune=True,
                        want_fields=True)
    errs = []
    for a, b in zip(srv["results"], srv_m["results"]):
        if a["error"] is None and b["error"] is None:
            wa, wb = a["w"][:, 3:-3, 3:-3, 3:-3], b["w"][:, 3:-3, 3:-3, 3:-3]
            errs.append(max(float(np.linalg.norm(wb[m] - wa[m]) / np.linalg.norm(wa[m])) for m in range(5)))
    progress(f"served mixed: {srv_m['gcups']:.3f} Gcell-updates/s, p50 {srv_m['p50_s']:.2f} s, "
             f"p99 {srv_m['p99_s']:.2f} s, worst rel L2 vs FP64 {max(errs):.2e}")
    mixed = {"value": srv_m["gcups"], "p50_s": srv_m["p50_s"], "p99_s": srv_m["p99_s"],
             "failed_requests": srv_m["failed"], "speedup_vs_fp64_served": srv_m["gcups"] / srv["gcups"],
             "rel_l2_vs_fp64_per_request": {"requests": len(errs), "median": float(np.median(errs)),
                                            "p90": float(np.percentile(errs, 90)), "max": float(max(errs))},
             "note": "worst conserved variable per request after its steps, against the FP64 served result"}
    return {"requests": len(reqs), "steps_per_request": args.serve_steps, "arrival_rate_per_s": args.serve_rate,
            "value": srv["gcups"], "unit": "Gcell-updates/s per RK stage", "p50_s": srv["p50_s"],
            "p99_s": srv["p99_s"], "mean_queue_s": srv["mean_queue_s"], "makespan_s": srv["makespan_s"],
            "failed_requests": srv["failed"],
            "failed_note": "requests whose residual turned NaN (the oracle diverges on them too, within the same 10 steps); "
                           "value and latencies cover the completed requests",
            "sequential_value": seq["gcups"], "sequential_sample": f"first {seq_n} requests, plain path, one by one",
            "speedup_vs_sequential": srv["gcups"] / seq["gcups"], "slots": args.serve_slots,
            "hbm_capacity_gib": args.serve_hbm_gib, "tuned_shape_classes": srv["stats"]["tuned_classes"],
            "total_cell_updates": cell_updates([r for r, o in zip(reqs, srv["results"]) if o["error"] is None]),
            "mixed": mixed, "batched": batched, "small_grid_stream": serving_small(args, hbm)}


def _same_fields(a_results, b_results):
    import numpy as np
    return all((a["error"] is None) == (b["error"] is None) and
               (a["error"] is not None or np.array_equal(a["w"][:, 3:-3, 3:-3, 3:-3], b["w"][:, 3:-3, 3:-3, 3:-3]))
               for a, b in zip(a_results, b_results))


def _batch_summary(r):
    st = r["stats"]
    n = max(1, st["batched_chunks"])
    return {"value": r["gcups"], "p50_s": r["p50_s"], "p99_s": r["p99_s"], "failed_requests": r["failed"],
            "batched_chunks": st["batched_chunks"], "solves_per_batched_chunk": st["batched_solves"] / n,
            "host_ms_per_batched_chunk": (st["batch_layout_us"] + st["batch_launch_us"]) / n / 1e3}


def serving_batched(args, reqs, srv, hbm):
    """The same stream with batched launches: solves of at most
    --serve-batch-max-cells cells are bound to --serve-batch-lanes lanes and
    each lane issues one launch per kernel over a table of its solves.
    Every result must equal the unbatched server's bit for bit."""
    from streams_mi355x.serve import run_serving

    r = run_serving(reqs, hbm_bytes=hbm, max_slots=args.serve_slots, steps_per_graph=10, autotune=True,
                    want_fields=True, batch=args.serve_batch_lanes, batch_max_cells=args.serve_batch_max_cells)
    out = _batch_summary(r)
    out.update({"lanes": args.serve_batch_lanes, "max_cells": args.serve_batch_max_cells,
                "same_results_as_unbatched": _same_fields(srv["results"], r["results"]),
                "speedup_vs_unbatched": r["gcups"] / srv["gcups"]})
    progress(f"served batched: {r['gcups']:.3f} Gcell-updates/s, p50 {r['p50_s']:.2f} s, p99 {r['p99_s']:.2f} s, "
             f"same results {out['same_results_as_unbatched']}")
    return out


def serving_small(args, hbm):
    """A stream of small solves (128 requests, grids uniform in [32,64]^3,
    Poisson arrivals at 64/s, --serve-steps steps each, FP64), where one
    solve's launches cannot fill the device: served unbatched (32 slots on
    their own streams) and batched (the same slots in --serve-batch-lanes
    lanes, no cell limit).  Results must agree bit for bit."""
    from streams_mi355x.serve import run_serving, serving_workload

    small = serving_workload(128, args.serve_steps, rate=64.0, seed=1, nx_range=(32, 64), nyz_range=(32, 64))
    k = args.serve_steps_per_graph
    a = run_serving(small, hbm_bytes=hbm, max_slots=32, steps_per_graph=k, autotune=True, want_fields=True)
    b = run_serving(small, hbm_bytes=hbm, max_slots=32, steps_per_graph=k, autotune=True, want_fields=True,
                    batch=args.serve_batch_lanes)
    out = {"requests": len(small), "grids": "uniform in [32,64]^3", "arrival_rate_per_s": 64.0, "slots": 32,
           "steps_per_graph": k,
           "unbatched": {"value": a["gcups"], "p50_s": a["p50_s"], "p99_s": a["p99_s"], "failed_requests": a["failed"]},
           "batched": _batch_summary(b), "lanes": args.serve_batch_lanes,
           "speedup_batched_vs_unbatched": b["gcups"] / a["gcups"],
           "same_results": _same_fields(a["results"], b["results"]), "unit": "Gcell-updates/s per RK stage"}
    progress(f"small-grid stream: unbatched {a['gcups']:.3f}, batched {b['gcups']:.3f} Gcell-updates/s, "
             f"same results {out['same_results']}")
    return out


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed RK3 steps (the GPU-scale channel runs 1000)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--retau", type=float, default=550.0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target duration of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-steps", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=16.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-hbm-gib", type=float, default=64.0)
    ap.add_argument("--serve-seq-sample", type=int, default=64)
    ap.add_argument("--serve-batch-lanes", type=int, default=2, help="batch lanes of the batched serving runs")
    ap.add_argument("--serve-steps-per-graph", type=int, default=10,
                    help="steps per captured graph of the small-grid stream (0: chunks of 10 steps, no graphs)")
    ap.add_argument("--serve-batch-max-cells", type=int, default=1 << 20,
                    help="solves of more cells keep their own streams in the batched run of the main stream")
    ap.add_argument("--workload", choices=("channel", "sbli", "serving-small"), default="channel",
                    help="sbli: the shock / boundary-layer box (1536x320x192 unless --nx/--ny/--nz say otherwise), "
                         "FP64 and mixed, printed as its own JSON line; serving-small: only the small-grid "
                         "serving stream, unbatched and batched, as its own JSON line")
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


def host_cpu():
    """CPU model, logical CPUs of the host and the CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    out = {"model": model, "logical_cpus": os.cpu_count(), "affinity_cpus": allowed,
           "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    # lscpu's topology lines, when the tool is there
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keep = ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)", "NUMA node(s)",
                "CPU max MHz", "Hypervisor vendor")
        out["lscpu"] = {k.strip(): v.strip() for k, v in (ln.split(":", 1) for ln in txt.splitlines() if ":" in ln)
                        if k.strip() in keep}
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return out


def cpu_baseline(args):
    """Oracle (plain C restatement of the reference CPU kernels, OpenMP over
    k) on a slab of the same channel: full nx x ny, nz reduced to 16.  Threads:
    OMP_NUM_THREADS when the environment sets it (the share of the host's
    cores given to this process), else every CPU in the affinity mask."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    host = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or host["affinity_cpus"]
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
    return {"value": rate, "unit": "Gcell-updates/s per RK stage", "cores": threads, "kind": "port", "host": host,
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
    if args.workload == "sbli":
        if (args.nx, args.ny, args.nz) == (512, 256, 256):
            args.nx, args.ny, args.nz = 1536, 320, 192
        print(json.dumps(sbli_workload(args)))
        return
    if args.workload == "serving-small":
        out = serving_small(args, int(args.serve_hbm_gib * (1 << 30)))
        out["workload"] = "serving-small"
        print(json.dumps(out))
        return
    from streams_mi355x import make_channel
    from streams_mi355x.solver import ChannelSolver

    nx, ny, nz = args.nx, args.ny, args.nz
    c = make_channel(nx=nx, ny=ny, nz=nz, retau=args.retau)
    g = ChannelSolver(c.params, c.grid)
    el = timed_run(g, c, args.warmup, args.steps)
    cells = nx * ny * nz
    value = cells * 3 * args.steps / el / 1e9
    progress(f"fp64 timed: {value:.3f} Gcell-updates/s per RK stage")
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
        "metric": "Gcell-updates/s per RK stage, FP64 vs mixed; p50/p99 latency of batched solves",
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
        "config": {"workload": f"channel {nx}x{ny}x{nz}, {args.steps} steps, M_b 1.5, Re_tau {args.retau:g}, FP64 plain, "
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
    if not args.no_mixed:
        out["mixed"] = mixed_form(args, c, value)
        progress(f"mixed: {out['mixed']['value']:.3f} Gcell-updates/s per RK stage")
    if not args.no_serve:
        out["serving"] = serving(args)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args)
            progress(f"cpu baseline: {out['cpu_baseline']['value']:.5f} Gcell-updates/s per RK stage")
        except Exception as e:  # the oracle is optional on a bare box
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
