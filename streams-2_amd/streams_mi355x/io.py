"""Restart and statistics files in the reference's formats (host side).

* ``rst.bin`` (write_field / read_field, field.F90:1853-1966): raw native
  FP64, one nxmax x nymax x nzmax block per conserved variable (x fastest,
  interior cells only, no header), variables 1..5 in order.
* ``field_info0.dat`` (write_field_info / read_field_info,
  singleideal.F90:2546-2659): one INI section ``[field_info]`` with
  ``key = value`` lines (cfgio's print_cfg for a single section).  Values are
  written in the shortest round-trip form; the reference writes them with
  list-directed Fortran output, whose spelling depends on the compiler, so
  the file matches the reference's by content, not by bytes.
* ``stat.bin`` (write_stats / read_stats, singleideal.F90:1779-1822): raw
  native FP64, one nxmax x nymax block per statistic (x fastest), the 70
  statistics of compute_stats in order.

One process writes the whole grid, so the MPI-IO subarray views of the
reference reduce to these contiguous blocks.
"""
from __future__ import annotations

import os

import numpy as np

NV = 5
NV_STAT = 70
# keys of write_field_info in its order (singleideal.F90:2620-2659)
FIELD_INFO_KEYS = ("icyc0", "time0", "itav", "itslice_p3d", "time_from_last_rst", "time_from_last_write",
                   "time_from_last_stat", "time_from_last_slice", "time_from_last_probe",
                   "time_from_last_slice_p3d", "istore")
_INT_KEYS = {"icyc0", "itav", "itslice_p3d", "istore"}


def write_restart(path, w: np.ndarray, ng: int = 3):
    """rst.bin from a ghosted conserved field w[m, k, j, i] (the layout of
    copy_to_field) or an interior one (ng = 0)."""
    wi = w[:, ng:-ng, ng:-ng, ng:-ng] if ng else w
    # [m, k, j, i] C order == per variable (i fastest, then j, then k)
    np.ascontiguousarray(wi, dtype="<f8").tofile(os.fspath(path))


def read_restart(path, nx: int, ny: int, nz: int) -> np.ndarray:
    """Interior conserved field [m, k, j, i] from rst.bin."""
    raw = np.fromfile(os.fspath(path), dtype="<f8")
    if raw.size != NV * nx * ny * nz:
        raise ValueError(f"rst.bin holds {raw.size} values, expected {NV * nx * ny * nz}")
    return raw.reshape(NV, nz, ny, nx)


def write_field_info(path, info: dict):
    """field_info0.dat with the keys of write_field_info (missing keys 0)."""
    lines = ["[field_info]"]
    for key in FIELD_INFO_KEYS:
        v = info.get(key, 0)
        lines.append(f"{key} = {int(v)}" if key in _INT_KEYS else f"{key} = {float(v)!r}")
    with open(os.fspath(path), "w") as f:
        f.write("\n".join(lines) + "\n")


def read_field_info(path) -> dict:
    """The [field_info] keys of a field_info file (reference or ours)."""
    out, sect = {}, None
    with open(os.fspath(path)) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith(("#", ";")):
                continue
            if s.startswith("["):
                sect = s.strip("[]").strip()
                continue
            if sect == "field_info" and "=" in s:
                k, v = (t.strip() for t in s.split("=", 1))
                v = v.replace("d", "e").replace("D", "e")   # Fortran exponent letters
                out[k] = int(float(v)) if k in _INT_KEYS else float(v)
    if "time_from_last_slice_p3d" not in out and "time_from_last_slice" in out:
        out["time_from_last_slice_p3d"] = out["time_from_last_slice"]   # read_field_info's default
    return out


def write_stats(path, w_stat: np.ndarray):
    """stat.bin from w_stat(1:70, 1:nx, 1:ny)."""
    nv, nx, ny = w_stat.shape
    # block q = w_stat(q, 1:nx, 1:ny) with i fastest
    np.ascontiguousarray(np.transpose(w_stat, (0, 2, 1)), dtype="<f8").tofile(os.fspath(path))


def read_stats(path, nx: int, ny: int) -> np.ndarray:
    """w_stat(1:70, 1:nx, 1:ny) (Fortran order) from stat.bin."""
    raw = np.fromfile(os.fspath(path), dtype="<f8")
    if raw.size != NV_STAT * nx * ny:
        raise ValueError(f"stat.bin holds {raw.size} values, expected {NV_STAT * nx * ny}")
    return np.asfortranarray(np.transpose(raw.reshape(NV_STAT, ny, nx), (0, 2, 1)))
