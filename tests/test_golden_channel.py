"""The reference's own known answer for the channel path.

tests/tests.ini:1-3 of the reference holds the last progress.out row of
examples/subsonic_channel (80 x 60 x 80, 10 steps, 2x1x2 MPI ranks):
conservative viscous form, Sutherland viscosity, central KEEP fluxes
(sensor threshold 1.01), isothermal walls, bulk forcing, CFL dt, residual
and min/max reductions.  The reference's runner (tests/run_all_tests.py)
compares the printed columns.  The fixture tests/golden/
subsonic_channel_progress.json was extracted from that file by
tests/golden/extract_tests_ini.py.
"""
import json
import math
import os

import pytest

from streams_mi355x.cases import make_subsonic_channel
from streams_mi355x.solver import progress_row, progress_text

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    with open(os.path.join(HERE, "golden", "subsonic_channel_progress.json")) as f:
        return json.load(f)


def _patch_steps(g):
    # the patch string 'controls,num_iter=10,controls,rand_type=0' sets
    # num_iter to 10; rand_type keeps the deck's value 2 (seeds rank+1)
    assert g["patch"].startswith("grid,nxmax=80;grid,nymax=60;grid,nzmax=80;controls,num_iter=10")
    return 10


def _last_digit(text):
    """One unit in the last printed digit of an es20.10 field."""
    mant, exp = text.split("E")
    return 10.0 ** (int(exp) - 10)


def test_oracle_reproduces_the_reference_progress_row():
    from oracle import Oracle

    g = _golden()
    c = make_subsonic_channel()
    o = Oracle(c.params, c.grid)
    o.w[:] = c.w0
    o.init_run()
    o.advance(_patch_steps(g))
    o.compute_residual()
    text = progress_text(progress_row(o.scalars(), c.params))
    assert text == g["row_text"], list(zip(g["columns"], text, g["row_text"]))


@pytest.mark.gpu
def test_gpu_path_reproduces_the_reference_progress_row():
    from streams_mi355x.solver import ChannelSolver

    g = _golden()
    c = make_subsonic_channel()
    with ChannelSolver(c.params, c.grid) as s:
        s.copy_from_field(c.w0)
        s.init_run()
        s.advance(_patch_steps(g))
        row = progress_row(s.compute_residual(), c.params)
    text = progress_text(row)
    assert text[0] == g["row_text"][0]
    for name, v, want in zip(g["columns"][1:], row[1:], g["row_text"][1:]):
        assert math.isclose(v, float(want), rel_tol=0.0, abs_tol=_last_digit(want)), (name, "%.10E" % v, want)
