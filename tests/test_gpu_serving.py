"""The serving loop: continuous batching of independent solves.

Each result must be the same alone (plain ChannelSolver path) and in the
mix (server: concurrent streams, captured graphs, autotuned launch shapes),
and admission must respect the HBM capacity given to the server.
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
