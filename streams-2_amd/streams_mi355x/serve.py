"""Serving loop of many independent flow solves on one GPU (host side).

``Server`` wraps the continuous-batching server of the C library
(st_server_*): requests are admitted as slots and HBM free up, run
concurrently on their own streams from captured graphs (or, with
``batch=True``, the chunks of all running solves of one launch sequence as
one sequence of batched launches over a table of the solves), and are
returned with their submit / admit / done times.  ``serving_workload`` draws the
seeded stream of uneven channel solves, ``run_serving`` replays it with
Poisson arrivals and reports throughput and latency percentiles.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from dataclasses import dataclass

import numpy as np

from . import capi
from .cases import ChannelCase, make_channel
from .params import StParams, StScalars


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Server:
    def __init__(self, hbm_bytes: int, max_slots: int = 8, steps_per_graph: int = 10, autotune: bool = True,
                 batch: int = 0, batch_max_cells: int = 0):
        """``batch``: number of batch lanes (0: every solve on its own
        stream); solves of more than ``batch_max_cells`` interior cells (0:
        no limit) keep their own streams."""
        self.L = capi.lib()
        h = C.c_void_p()
        capi.check_server(self.L.st_server_create(int(hbm_bytes), max_slots, steps_per_graph, 1 if autotune else 0,
                                                  C.byref(h)))
        self.h = h
        self._shapes = {}
        if batch:
            capi.check_server(self.L.st_server_set_batching(self.h, int(batch), int(batch_max_cells)))

    def submit(self, params: StParams, grid, w0: np.ndarray, nsteps: int, time0: float = 0.0, icyc0: int = 0,
               checkpoint_every: int = 0, checkpoint_dir: str | None = None) -> int:
        """Queue a solve of ``nsteps`` steps from ``w0`` (ghosted), starting at
        ``time0`` / step ``icyc0`` (a resumed request), with rst.bin and
        field_info0.dat written to ``checkpoint_dir`` every
        ``checkpoint_every`` steps and after the last one."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in grid.arrays()]
        w0 = np.ascontiguousarray(w0, dtype=np.float64)
        t = C.c_int64()
        cdir = None if checkpoint_dir is None else os.fspath(checkpoint_dir).encode()
        capi.check_server(self.L.st_server_submit2(self.h, C.byref(params), *[_dp(a) for a in arrs], _dp(w0), nsteps,
                                                   float(time0), int(icyc0), int(checkpoint_every), cdir, C.byref(t)))
        self._shapes[t.value] = w0.shape
        return t.value

    def wait(self, ticket: int, timeout: float = -1.0, want_field: bool = True):
        shape = self._shapes.pop(ticket)
        w = np.empty(shape) if want_field else None
        sc = StScalars()
        tm = (C.c_double * 3)()
        capi.check_server(self.L.st_server_wait(self.h, ticket, timeout, _dp(w) if want_field else None,
                                                C.byref(sc), tm))
        return w, sc.as_dict(), (tm[0], tm[1], tm[2])

    def stats(self):
        out = (C.c_int64 * 6)()
        capi.check_server(self.L.st_server_stats(self.h, out))
        st = dict(zip(("queued", "running", "completed", "failed", "hbm_bytes", "tuned_classes"), list(out)))
        b = (C.c_int64 * 7)()
        capi.check_server(self.L.st_server_batch_stats(self.h, b))
        st.update(zip(("batched_chunks", "batched_launches", "batched_solves", "batch_layouts",
                       "batch_layout_us", "batch_launch_us", "batched_graph_chunks"), list(b)))
        return st

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.st_server_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Request:
    case: ChannelCase
    nsteps: int
    arrival: float   # seconds after the start of the stream


def serving_workload(n: int = 64, nsteps: int = 200, rate: float = 8.0, seed: int = 0, precision: int = 0,
                     nx_range=(64, 256), nyz_range=(64, 128), mach_range=(0.3, 3.0), retau_range=(100.0, 500.0),
                     multiple: int = 8):
    """Seeded stream of independent channel solves: grid extents uniform in
    the given ranges (rounded to ``multiple``), bulk Mach and Re_tau uniform,
    Poisson arrivals at ``rate`` requests/s."""
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for _ in range(n):
        nx = int(round(rng.uniform(*nx_range) / multiple)) * multiple
        ny = int(round(rng.uniform(*nyz_range) / multiple)) * multiple
        nz = int(round(rng.uniform(*nyz_range) / multiple)) * multiple
        mach = float(rng.uniform(*mach_range))
        retau = float(rng.uniform(*retau_range))
        seed_i = int(rng.integers(0, 2 ** 31 - 1))
        c = make_channel(nx=nx, ny=ny, nz=nz, mach=mach, retau=retau, seed=seed_i, precision=precision,
                         noise_source="pcg")
        t += float(rng.exponential(1.0 / rate))
        reqs.append(Request(c, nsteps, t))
    return reqs


def cell_updates(reqs) -> float:
    return float(sum(r.case.params.nx * r.case.params.ny * r.case.params.nz * 3 * r.nsteps for r in reqs))


def run_serving(reqs, hbm_bytes: int, max_slots: int = 8, steps_per_graph: int = 10, autotune: bool = True,
                want_fields: bool = False, timeout: float = 600.0, batch: int = 0, batch_max_cells: int = 0):
    """Submit each request at its arrival time, wait for all; returns a dict
    with throughput (Gcell-updates/s per RK stage over the makespan from the
    first arrival), p50/p99 time-to-solution and the per-request results.
    A request that fails (a NaN residual, a device error) is returned with
    its ``error`` and counted in ``failed``; the others are unaffected, and
    throughput and latencies cover the completed requests.  A request not
    finished ``timeout`` seconds after the wait for it began raises."""
    out = []
    with Server(hbm_bytes, max_slots, steps_per_graph, autotune, batch, batch_max_cells) as sv:
        t0 = time.perf_counter()
        tickets = []
        for r in reqs:
            delay = r.arrival - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets.append(sv.submit(r.case.params, r.case.grid, r.case.w0, r.nsteps))
        for r, tk in zip(reqs, tickets):
            try:
                w, sc, tm = sv.wait(tk, timeout, want_fields)
                out.append({"w": w, "scalars": sc, "submit": tm[0], "admit": tm[1], "done": tm[2], "error": None})
            except capi.StreamsError as e:
                if e.code == capi.ST_ERR_STATE:   # time-out: the server is stuck, not the request
                    raise
                out.append({"w": None, "scalars": None, "error": str(e)})
        stats = sv.stats()
    ok = [o for o, r in zip(out, reqs) if o["error"] is None]
    done_reqs = [r for o, r in zip(out, reqs) if o["error"] is None]
    if not ok:
        return {"gcups": 0.0, "failed": len(out), "results": out, "stats": stats}
    first = min(o["submit"] for o in ok)
    last = max(o["done"] for o in ok)
    lat = np.array([o["done"] - o["submit"] for o in ok])
    return {"gcups": cell_updates(done_reqs) / (last - first) / 1e9, "makespan_s": last - first,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_queue_s": float(np.mean([o["admit"] - o["submit"] for o in ok])),
            "failed": len(out) - len(ok), "results": out, "stats": stats}


def run_sequential(reqs, want_fields: bool = False):
    """The same requests one after another through the plain path
    (ChannelSolver: one stream, no graphs, no overlap).  A request whose
    residual is NaN is returned with its ``error``; throughput counts the
    completed requests."""
    from .solver import ChannelSolver

    out = []
    t0 = time.perf_counter()
    for r in reqs:
        with ChannelSolver(r.case.params, r.case.grid) as g:
            g.copy_from_field(r.case.w0)
            g.init_run()
            g.advance(r.nsteps)
            try:
                sc, err = g.compute_residual(), None
            except capi.StreamsError as e:
                if e.code != capi.ST_ERR_NAN:
                    raise
                sc, err = None, str(e)
            w = g.copy_to_field() if want_fields else None
        out.append({"w": w, "scalars": sc, "error": err})
    el = time.perf_counter() - t0
    done = [r for r, o in zip(reqs, out) if o["error"] is None]
    return {"gcups": cell_updates(done) / el / 1e9, "seconds": el, "failed": len(reqs) - len(done), "results": out}
