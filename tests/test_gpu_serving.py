"""The serving loop: continuous batching of independent solves.

Each result must be the same alone (plain ChannelSolver path) and in the
mix (server: concurrent streams, captured graphs, autotuned launch shapes,
batched launches over a table of solves), and admission must respect the
HBM capacity given to the server.
"""
import numpy as np
import pytest

from helpers import interior
from streams_mi355x import PREC_MIXED, make_channel
from streams_mi355x.capi import lib
from streams_mi355x.serve import Request, Server, run_serving, run_sequential
from streams_mi355x.params import StParams

pytestmark = pytest.mark.gpu


def _requests(precision=0):
    shapes = [(32, 24, 16), (48, 32, 24), (40, 24, 32), (64, 32, 16), (32, 40, 24), (56, 24, 24)]
    reqs = []
    for n, (nx, ny, nz) in enumerate(shapes):
        c = make_channel(nx=nx, ny=ny, nz=nz, mach=0.8 + 0.3 * n, retau=150.0 + 40 * n, seed=n,
                         sensor_threshold=0.5 if n % 2 else 0.9, precision=precision)
        reqs.append(Request(c, 24 + 2 * n, 0.0))
    return reqs


@pytest.mark.parametrize("precision", [0, PREC_MIXED])
def test_same_alone_and_in_the_mix(precision):
    reqs = _requests(precision)
    seq = run_sequential(reqs, want_fields=True)
    srv = run_serving(reqs, hbm_bytes=8 << 30, max_slots=4, steps_per_graph=10, autotune=True, want_fields=True)
    for n, (a, b) in enumerate(zip(seq["results"], srv["results"])):
        assert np.array_equal(interior(a["w"]), interior(b["w"])), n
        assert a["scalars"]["icyc"] == b["scalars"]["icyc"] == reqs[n].nsteps
        assert a["scalars"]["residual_rhou"] == b["scalars"]["residual_rhou"]


def test_server_without_graphs_matches_the_plain_path():
    """steps_per_graph 0: plain st_advance chunks, no capture."""
    reqs = _requests()[:3]
    seq = run_sequential(reqs, want_fields=True)
    srv = run_serving(reqs, hbm_bytes=8 << 30, max_slots=2, steps_per_graph=0, autotune=False, want_fields=True,
                      timeout=120.0)
    for n, (a, b) in enumerate(zip(seq["results"], srv["results"])):
        assert np.array_equal(interior(a["w"]), interior(b["w"])), n
        assert b["scalars"]["icyc"] == reqs[n].nsteps


def test_a_diverging_request_fails_alone():
    """A NaN in one request's field fails that request with ST_ERR_NAN; the
    requests beside it finish bit-identical to their plain solves."""
    from streams_mi355x.capi import ST_ERR_NAN

    reqs = _requests()[:4]
    bad = 1
    c = reqs[bad].case
    w0 = c.w0.copy()
    w0[0, 3 + 2, 3 + 5, 3 + 7] = np.nan
    reqs[bad] = Request(type(c)(c.case, c.grid, c.params, w0), reqs[bad].nsteps, 0.0)
    seq = run_sequential(reqs, want_fields=True)
    srv = run_serving(reqs, hbm_bytes=8 << 30, max_slots=4, steps_per_graph=10, autotune=True, want_fields=True,
                      timeout=120.0)
    assert srv["failed"] == 1 and seq["failed"] == 1
    assert f"[st error {ST_ERR_NAN}]" in srv["results"][bad]["error"]
    for n, (a, b) in enumerate(zip(seq["results"], srv["results"])):
        if n != bad:
            assert b["error"] is None
            assert np.array_equal(interior(a["w"]), interior(b["w"])), n


def test_admission_respects_the_hbm_capacity():
    reqs = _requests()
    foot = [lib().st_solver_footprint(r.case.params) for r in reqs]
    cap = max(foot) + min(foot)  # room for two solves at most
    srv = run_serving(reqs, hbm_bytes=cap, max_slots=6, steps_per_graph=4, autotune=False)
    ev = []
    for f, o in zip(foot, srv["results"]):
        ev.append((o["admit"], f))
        ev.append((o["done"], -f))
    used = peak = 0
    for _, d in sorted(ev, key=lambda e: (e[0], e[1])):
        used += d
        peak = max(peak, used)
    assert peak <= cap
    assert srv["stats"]["completed"] == len(reqs) and srv["stats"]["failed"] == 0


def test_request_larger_than_capacity_is_rejected():
    c = make_channel()
    with Server(hbm_bytes=1 << 20, max_slots=2) as sv:
        with pytest.raises(Exception):
            sv.submit(c.params, c.grid, c.w0, 2)


def test_checkpoint_files_and_resume(tmp_path):
    """Per-request checkpoints: a served request writes rst.bin and
    field_info0.dat every 10 steps and after its last step.  The files hold
    its final interior field bit for bit and its step and time; the request
    resumed from them through the server ends bit-identical to the same
    stop/restart through the plain path, and matches the oracle continued
    from the same files."""
    from helpers import fill_ghosts_wrap, rel_l2
    from oracle import Oracle
    from streams_mi355x.io import read_field_info, read_restart
    from streams_mi355x.solver import ChannelSolver

    reqs = _requests()[:3]
    dirs = [tmp_path / f"req{n}" for n in range(len(reqs))]
    with Server(hbm_bytes=8 << 30, max_slots=3, steps_per_graph=10, autotune=True) as sv:
        tk = [sv.submit(r.case.params, r.case.grid, r.case.w0, 20, checkpoint_every=10, checkpoint_dir=str(d))
              for r, d in zip(reqs, dirs)]
        first = [sv.wait(t, 120.0) for t in tk]
    resumed_w = []
    for r, d, (w, sc, _) in zip(reqs, dirs, first):
        p = r.case.params
        wi = read_restart(d / "rst.bin", p.nx, p.ny, p.nz)
        info = read_field_info(d / "field_info0.dat")
        assert np.array_equal(wi, interior(w))
        assert info["icyc0"] == sc["icyc"] == 20 and info["time0"] == sc["time"]
        w0 = np.zeros_like(r.case.w0)
        w0[:, 3:-3, 3:-3, 3:-3] = wi
        resumed_w.append((fill_ghosts_wrap(w0), info))
    with Server(hbm_bytes=8 << 30, max_slots=3, steps_per_graph=10, autotune=True) as sv:
        tk = [sv.submit(r.case.params, r.case.grid, w0, 10, time0=info["time0"], icyc0=info["icyc0"])
              for r, (w0, info) in zip(reqs, resumed_w)]
        second = [sv.wait(t, 120.0) for t in tk]
    for n, (r, (w0, info), (w, sc, _)) in enumerate(zip(reqs, resumed_w, second)):
        assert sc["icyc"] == 30
        with ChannelSolver(r.case.params, r.case.grid) as g:
            g.copy_from_field(w0)
            g.set_time(info["time0"], 0.0, info["icyc0"])
            g.init_run()
            g.advance(10)
            assert np.array_equal(interior(g.copy_to_field()), interior(w)), n
            assert g.scalars()["time"] == sc["time"]
        if n == 0:
            o = Oracle(r.case.params, r.case.grid)
            o.w[:] = w0
            o.set_time(info["time0"], 0.0, info["icyc0"])
            o.init_run()
            o.advance(10)
            errs = [rel_l2(interior(w)[m], interior(o.w)[m]) for m in range(5)]
            assert max(errs) <= 1e-10, errs


def test_checkpoint_period_must_be_whole_graphs(tmp_path):
    c = make_channel()
    with Server(hbm_bytes=8 << 30, max_slots=1, steps_per_graph=10) as sv:
        with pytest.raises(Exception):
            sv.submit(c.params, c.grid, c.w0, 20, checkpoint_every=7, checkpoint_dir=str(tmp_path))


# ------------------------------------------------------------ batched launches

@pytest.mark.parametrize("precision,lanes,max_cells", [(0, 1, 0), (PREC_MIXED, 1, 0), (0, 2, 0), (0, 2, 33000)])
def test_batched_same_alone_and_in_the_mix(precision, lanes, max_cells):
    """Batched launches: the chunks of all running solves go out as one
    launch per kernel over a table of the solves.  Each result equals its
    plain solve bit for bit; the final chunks shorter than a graph run on
    the solves' own streams, so both paths are exercised (with a cell
    limit, the larger solves keep their own streams throughout)."""
    reqs = _requests(precision)
    seq = run_sequential(reqs, want_fields=True)
    srv = run_serving(reqs, hbm_bytes=8 << 30, max_slots=6, steps_per_graph=10, autotune=False, want_fields=True,
                      timeout=120.0, batch=lanes, batch_max_cells=max_cells)
    st = srv["stats"]
    print("batched:", st)
    assert st["batched_chunks"] > 0
    for n, (a, b) in enumerate(zip(seq["results"], srv["results"])):
        assert b["error"] is None, (n, b["error"])
        assert np.array_equal(interior(a["w"]), interior(b["w"])), n
        assert b["scalars"]["icyc"] == reqs[n].nsteps
        assert a["scalars"]["residual_rhou"] == b["scalars"]["residual_rhou"]


def test_batched_with_online_tuning_matches_the_plain_path():
    """Tuning chunks run on the solves' own streams; once a class is tuned
    its solves join the batch with the tuned launch shape."""
    reqs = _requests()
    reqs = [Request(r.case, 60, 0.0) for r in reqs]
    seq = run_sequential(reqs, want_fields=True)
    srv = run_serving(reqs, hbm_bytes=8 << 30, max_slots=6, steps_per_graph=4, autotune=True, want_fields=True,
                      timeout=120.0, batch=1)
    print("batched with tuning:", srv["stats"])
    assert srv["stats"]["tuned_classes"] >= 1 and srv["stats"]["batched_chunks"] > 0
    for n, (a, b) in enumerate(zip(seq["results"], srv["results"])):
        assert np.array_equal(interior(a["w"]), interior(b["w"])), n


def test_batched_chunks_replay_from_graphs():
    """A lane whose solves stay the same reuses its layout and replays the
    chunk from a captured HIP graph; results are unchanged."""
    reqs = [Request(r.case, 48, 0.0) for r in _requests()[:3]]
    seq = run_sequential(reqs, want_fields=True)
    srv = run_serving(reqs, hbm_bytes=8 << 30, max_slots=3, steps_per_graph=4, autotune=False, want_fields=True,
                      timeout=120.0, batch=1)
    st = srv["stats"]
    print("graph replays:", st)
    assert st["batched_graph_chunks"] > 0
    for n, (a, b) in enumerate(zip(seq["results"], srv["results"])):
        assert np.array_equal(interior(a["w"]), interior(b["w"])), n


def test_batched_without_graphs():
    """steps_per_graph 0: batched chunks of 10 steps enqueued launch by
    launch, nothing captured."""
    reqs = _requests()[:4]
    seq = run_sequential(reqs, want_fields=True)
    srv = run_serving(reqs, hbm_bytes=8 << 30, max_slots=4, steps_per_graph=0, autotune=False, want_fields=True,
                      timeout=120.0, batch=1)
    st = srv["stats"]
    assert st["batched_chunks"] > 0 and st["batched_graph_chunks"] == 0
    for n, (a, b) in enumerate(zip(seq["results"], srv["results"])):
        assert np.array_equal(interior(a["w"]), interior(b["w"])), n


def test_batched_lane_with_two_launch_sequences():
    """FP64 and mixed-form solves in one lane: two launch sequences, batched
    in turns; every result equals its plain solve."""
    reqs = _requests(0)[:3] + _requests(PREC_MIXED)[3:6]
    seq = run_sequential(reqs, want_fields=True)
    srv = run_serving(reqs, hbm_bytes=8 << 30, max_slots=6, steps_per_graph=10, autotune=False, want_fields=True,
                      timeout=120.0, batch=1)
    st = srv["stats"]
    assert st["batched_chunks"] > 0 and st["batch_layouts"] > 1
    for n, (a, b) in enumerate(zip(seq["results"], srv["results"])):
        assert b["error"] is None, (n, b["error"])
        assert np.array_equal(interior(a["w"]), interior(b["w"])), n


def test_batched_diverging_request_fails_alone():
    from streams_mi355x.capi import ST_ERR_NAN

    reqs = _requests()[:4]
    bad = 2
    c = reqs[bad].case
    w0 = c.w0.copy()
    w0[0, 3 + 2, 3 + 5, 3 + 7] = np.nan
    reqs[bad] = Request(type(c)(c.case, c.grid, c.params, w0), reqs[bad].nsteps, 0.0)
    seq = run_sequential(reqs, want_fields=True)
    srv = run_serving(reqs, hbm_bytes=8 << 30, max_slots=4, steps_per_graph=10, autotune=False, want_fields=True,
                      timeout=120.0, batch=1)
    assert srv["failed"] == 1 and srv["stats"]["batched_chunks"] > 0
    assert f"[st error {ST_ERR_NAN}]" in srv["results"][bad]["error"]
    for n, (a, b) in enumerate(zip(seq["results"], srv["results"])):
        if n != bad:
            assert b["error"] is None
            assert np.array_equal(interior(a["w"]), interior(b["w"])), n


def test_batched_checkpoints(tmp_path):
    """Checkpoints of batched solves: each request's rst.bin after its last
    step holds its final field, which equals its plain solve."""
    from streams_mi355x.io import read_restart

    reqs = _requests()[:3]
    dirs = [tmp_path / f"req{n}" for n in range(len(reqs))]
    with Server(hbm_bytes=8 << 30, max_slots=3, steps_per_graph=10, autotune=False, batch=1) as sv:
        tk = [sv.submit(r.case.params, r.case.grid, r.case.w0, 30, checkpoint_every=10, checkpoint_dir=str(d))
              for r, d in zip(reqs, dirs)]
        out = [sv.wait(t, 120.0) for t in tk]
        st = sv.stats()
    assert st["batched_chunks"] > 0
    seq = run_sequential([Request(r.case, 30, 0.0) for r in reqs], want_fields=True)
    for r, d, (w, sc, _), a in zip(reqs, dirs, out, seq["results"]):
        p = r.case.params
        assert np.array_equal(read_restart(d / "rst.bin", p.nx, p.ny, p.nz), interior(w))
        assert np.array_equal(interior(a["w"]), interior(w))
        assert sc["icyc"] == 30


def test_batching_is_set_before_the_first_submission():
    c = make_channel()
    with Server(hbm_bytes=8 << 30, max_slots=1, steps_per_graph=10) as sv:
        sv.submit(c.params, c.grid, c.w0, 2)
        from streams_mi355x.capi import check_server

        with pytest.raises(Exception):
            check_server(lib().st_server_set_batching(sv.h, 1, 0))
    with pytest.raises(Exception):
        Server(hbm_bytes=8 << 30, max_slots=1, steps_per_graph=3, batch=1)


def test_closing_with_batched_chunks_in_flight():
    """A server closed while batched chunks run drains its lanes before it
    releases the solves and the pool; a new server afterwards gives the plain
    results."""
    import time

    reqs = _requests()[:4]
    with Server(hbm_bytes=8 << 30, max_slots=4, steps_per_graph=10, autotune=False, batch=1) as sv:
        for r in reqs:
            sv.submit(r.case.params, r.case.grid, r.case.w0, 2000)
        t_end = time.time() + 60.0
        while sv.stats()["batched_chunks"] < 4 and time.time() < t_end:
            time.sleep(0.01)
        assert sv.stats()["batched_chunks"] >= 4
    seq = run_sequential(reqs[:2], want_fields=True)
    srv = run_serving(reqs[:2], hbm_bytes=8 << 30, max_slots=2, steps_per_graph=10, autotune=False,
                      want_fields=True, timeout=120.0, batch=1)
    for n, (a, b) in enumerate(zip(seq["results"], srv["results"])):
        assert np.array_equal(interior(a["w"]), interior(b["w"])), n
