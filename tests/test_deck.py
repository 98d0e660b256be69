"""The reference's input decks (singleideal.ini) as the library reads them.

The fixtures tests/golden/decks/*.ini are examples/<case>/singleideal.ini of
the reference (copied by tests/golden/copy_decks.py).  The subsonic deck with
the patch of tests/tests.ini must give exactly the case whose 10-step
progress row matches the reference's known answer
(test_golden_channel.py); the GPU run of the patched deck must print that
row.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from streams_mi355x.cases import make_subsonic_channel
from streams_mi355x.deck import apply_patch, case_from_deck, read_deck

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")


def _golden():
    with open(os.path.join(HERE, "golden", "subsonic_channel_progress.json")) as f:
        return json.load(f)


def _bytes(p):
    return C.string_at(C.addressof(p), C.sizeof(p))


def test_patched_subsonic_deck_is_the_known_answer_case():
    g = _golden()
    deck = apply_patch(read_deck(os.path.join(DECKS, "subsonic_channel.ini")), g["patch"])
    run = case_from_deck(deck)
    ref = make_subsonic_channel()
    assert run.num_iter == 10 and run.splits == (2, 1, 2) and run.print_control == 10
    assert _bytes(run.case.params) == _bytes(ref.params)
    for a, b in zip(run.case.grid.arrays(), ref.grid.arrays()):
        assert np.array_equal(a, b)
    assert np.array_equal(run.case.w0, ref.w0)


def test_patch_follows_the_reference_runner():
    """'controls,num_iter=10,controls,rand_type=0' splits at its first comma
    and '=': num_iter reads 10 (list-directed), and rand_type stays the
    deck's 2 under [numerics]."""
    g = _golden()
    deck = read_deck(os.path.join(DECKS, "subsonic_channel.ini"))
    p = apply_patch(deck, g["patch"])
    assert p["grid"]["nxmax"] == "80" and p["grid"]["nzmax"] == "80"
    assert p["controls"]["num_iter"] == "10,controls,rand_type=0"
    assert p["numerics"]["rand_type"] == "2"
    assert deck["grid"]["nxmax"] == "384"   # the original is untouched


def test_supersonic_deck():
    """Mach 2 on the bulk temperature with theta_wall 0.35 (bulk-T control
    on), Re_tau 490, sensor 0.2 (shock capturing active), CFL 0.7."""
    run = case_from_deck(os.path.join(DECKS, "supersonic_channel.ini"), nx=64, ny=48, nz=32)
    p = run.case.params
    assert (p.nx, p.ny, p.nz) == (64, 48, 32)
    assert run.num_iter == 100 and run.splits == (2, 1, 2)
    assert list(p.bctags) == [0, 0, 6, 6, 0, 0] and list(p.bctags_nr) == [0] * 6
    assert p.theta_wall == 0.35 and p.sensor_threshold == 0.2 and p.cfl == 0.7 and p.rand_type == 2
    assert p.conservative_viscous == 1 and p.visc_model == 2 and p.t_ref_dim == 220.0
    # Mach 2 is the bulk Mach number: u_b / c(T_b) = 2 with u_b = u0 (bulk velocity 1 in units of u0)
    gam = 1.4
    assert math.isclose(p.u0 / math.sqrt(gam * p.rgas0 * p.t_bulk_target), 2.0, rel_tol=1e-12)
    # (T_w - T_b) / (T_r - T_b) = theta_wall with T_r = T_b (1 + (gam-1)/2 r M_b^2)
    rfac = 0.72 ** (1.0 / 3.0)
    tr = p.t_bulk_target * (1.0 + (gam - 1.0) / 2.0 * rfac * 4.0)
    assert math.isclose((p.t_wall - p.t_bulk_target) / (tr - p.t_bulk_target), 0.35, rel_tol=1e-12)
    assert np.isfinite(run.case.w0).all()


@pytest.mark.parametrize("patch,msg", [("grid,grid_type=4", "grid_type"), ("flow,flow_init=1", "flow_init"),
                                       ("numerics,visc_order=4", "viscous order"),
                                       ("grid,grid_dim=2", "curvilinear"), ("grid,nxmax=81", "splits")])
def test_decks_outside_the_path_are_refused(patch, msg):
    deck = apply_patch(read_deck(os.path.join(DECKS, "subsonic_channel.ini")), patch)
    with pytest.raises(ValueError, match=msg):
        case_from_deck(deck, ny=16, nz=8) if "nxmax" in patch else case_from_deck(deck, nx=16, ny=16, nz=8)


@pytest.mark.gpu
def test_run_deck_prints_the_known_answer_row(tmp_path):
    from streams_mi355x.deck import run_deck
    from streams_mi355x.io import read_field_info, read_restart

    g = _golden()
    rows = run_deck(os.path.join(DECKS, "subsonic_channel.ini"), patch=g["patch"], out_dir=tmp_path)
    assert len(rows) == 1 and rows[0][0] == g["row_text"][0] == "10"
    for name, v, want in zip(g["columns"][1:], rows[0][1:], g["row_text"][1:]):
        mant, exp = want.split("E")
        assert math.isclose(float(v), float(want), rel_tol=0.0, abs_tol=10.0 ** (int(exp) - 10)), (name, v, want)
    info = read_field_info(tmp_path / "field_info0.dat")
    assert info["icyc0"] == 10
    w = read_restart(tmp_path / "rst.bin", 80, 60, 80)
    assert np.isfinite(w).all()
    assert (tmp_path / "progress.out").read_text().split() == rows[0]


@pytest.mark.gpu
def test_fortran_driver_runs_the_deck_as_the_python_runner_does(tmp_path):
    """The patched deck through the Fortran driver: rst.bin equals the one
    the ctypes runner writes, bit for bit."""
    from streams_mi355x.deck import run_deck, run_deck_fortran
    from streams_mi355x.fortran_case import DRIVER
    from streams_mi355x.io import read_restart

    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (make -C streams-2_amd/fortran)")
    g = _golden()
    deck = os.path.join(DECKS, "subsonic_channel.ini")
    run_deck(deck, patch=g["patch"], out_dir=tmp_path / "py")
    run_deck_fortran(deck, tmp_path / "f", patch=g["patch"])
    a = read_restart(tmp_path / "py" / "rst.bin", 80, 60, 80)
    b = read_restart(tmp_path / "f" / "rst.bin", 80, 60, 80)
    assert np.array_equal(a, b)
