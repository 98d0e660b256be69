"""The Fortran iso_c_binding path: record layout (CPU) and a full run of the
Fortran driver through the C ABI compared with the oracle (GPU)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from helpers import interior, rel_l2
from streams_mi355x.fortran_case import DRIVER, read_result, run_driver, write_case
from streams_mi355x.params import StParams, StScalars


def test_fortran_records_match_the_c_header():
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (make -C streams-2_amd/fortran)")
    out = subprocess.run([DRIVER, "--sizes"], capture_output=True, text=True, timeout=60).stdout.split()
    assert int(out[0]) == C.sizeof(StParams)
    assert int(out[1]) == C.sizeof(StScalars)


def test_fortran_reads_the_case_file(tmp_path):
    """flang transfers a bind(c) record component by component, so any
    implicit padding in st_params would shift every later read."""
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (make -C streams-2_amd/fortran)")
    from streams_mi355x import make_channel

    c = make_channel(nx=32, ny=24, nz=16)
    case = tmp_path / "case.bin"
    write_case(str(case), c.params, c.grid, c.w0, 7)
    out = subprocess.run([DRIVER, "--read", str(case)], capture_output=True, text=True, timeout=60).stdout.split()
    assert [int(v) for v in out[:5]] == [32, 24, 16, 3, 7]
    assert float(out[5]) == c.params.cfl
    assert abs(float(out[6]) - c.grid.yn.sum()) <= 1e-12
    assert abs(float(out[7]) - c.w0.sum()) <= 1e-12 * abs(c.w0.sum())


@pytest.mark.gpu
def test_fortran_driver_matches_oracle(tmp_path):
    from oracle import Oracle
    from streams_mi355x import make_channel

    from streams_mi355x.io import read_field_info, read_restart, read_stats

    c = make_channel(nx=32, ny=24, nz=16, sensor_threshold=0.0)
    p = c.params
    case = tmp_path / "case.bin"
    outp = tmp_path / "out.bin"
    write_case(str(case), p, c.grid, c.w0, 10)
    log = run_driver(str(case), str(outp), files_dir=str(tmp_path))
    w, s = read_result(str(outp), c.w0.shape)
    o = Oracle(p, c.grid)
    o.w[:] = c.w0
    o.init_run()
    o.advance(10)
    for m in range(5):
        assert rel_l2(interior(w)[m], interior(o.w)[m]) <= 1e-10, (m, log)
    assert s["icyc"] == 10
    # the reference's files, written by the Fortran side
    assert np.array_equal(read_restart(tmp_path / "rst.bin", p.nx, p.ny, p.nz), interior(w))
    info = read_field_info(tmp_path / "field_info0.dat")
    assert info["icyc0"] == 10 and info["time0"] == s["time"]
    ws = np.zeros((70, p.nx, p.ny), order="F")
    o.update_ghost(1)
    o.compute_stats(0, ws)
    got = read_stats(tmp_path / "stat.bin", p.nx, p.ny)
    for q in range(70):
        assert np.abs(got[q] - ws[q]).max() <= 1e-8 * max(np.abs(ws[q]).max(), 1e-3), q + 1


@pytest.mark.gpu
def test_fortran_backend_by_procedure_matches_oracle(tmp_path):
    """rk_sync_by_procedure of the Fortran backend object: the reference's
    stage procedure by procedure through the range-argument calls (euler_x
    split in two ranges), bulk forcing with bulk-temperature control, WENO
    faces present, against the oracle's rk_sync."""
    from oracle import Oracle
    from streams_mi355x import make_channel

    c = make_channel(nx=32, ny=24, nz=16, sensor_threshold=0.3, theta_wall=1.0)
    p = c.params
    case = tmp_path / "case.bin"
    outp = tmp_path / "out.bin"
    write_case(str(case), p, c.grid, c.w0, 6)
    log = run_driver(str(case), str(outp), mode="procedure")
    w, s = read_result(str(outp), c.w0.shape)
    o = Oracle(p, c.grid)
    o.w[:] = c.w0
    o.init_run()
    o.advance(6)
    errs = [rel_l2(interior(w)[m], interior(o.w)[m]) for m in range(5)]
    assert max(errs) <= 1e-10, (errs, log)
    assert s["icyc"] == 6
    assert abs(s["time"] - o.scalars()["time"]) <= 1e-12 * o.scalars()["time"]
