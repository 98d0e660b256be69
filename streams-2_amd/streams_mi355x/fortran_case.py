"""Binary case files of the Fortran driver (streams-2_amd/fortran/streams_mi355x_run).

Layout (native byte order, no record markers): st_params, int32 nsteps, the
13 grid/metric arrays of st_set_grid in argument order, then w (5 ghosted
variables, Fortran order).  The output holds w after the run and st_scalars.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from .params import StParams, StScalars

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fortran", "streams_mi355x_run")


def write_case(path: str, params: StParams, grid, w: np.ndarray, nsteps: int):
    with open(path, "wb") as f:
        f.write(bytes(params))
        f.write(np.int32(nsteps).tobytes())
        for a in grid.arrays():
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(w, dtype=np.float64).tobytes())


def read_result(path: str, shape):
    raw = open(path, "rb").read()
    n = int(np.prod(shape)) * 8
    w = np.frombuffer(raw[:n], dtype=np.float64).reshape(shape).copy()
    s = StScalars.from_buffer_copy(raw[n:n + C.sizeof(StScalars)])
    return w, s.as_dict()


def run_driver(case_path: str, out_path: str, timeout: float = 600.0, files_dir: str = None,
               mode: str = "fused") -> str:
    """Run the Fortran driver as a child process and return its output.
    With files_dir the driver also writes rst.bin, field_info0.dat and
    stat.bin there.  mode "procedure" steps through rk_sync_by_procedure."""
    cmd = [DRIVER, case_path, out_path] + ([files_dir or os.path.dirname(out_path), mode]
                                           if (files_dir or mode != "fused") else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran driver failed ({r.returncode}): {r.stdout}\n{r.stderr}")
    return r.stdout
