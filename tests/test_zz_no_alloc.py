"""st_compute_dt_noalloc (the allocation-free CFL reduction) against
st_compute_dt of a solver holding the same state, on two calls in a row
(the second on a smaller grid) and on a third call back on the first grid.

The test is the caller: it allocates the device arrays itself through the
HIP runtime the library is linked against, and the entry point only reads
and writes them (its block maxima live in the library's static scratch)."""
import ctypes as C

import numpy as np
import pytest

from streams_mi355x import capi, make_channel
from streams_mi355x.solver import ChannelSolver


def _hip():
    capi.lib()  # loads libamdhip64 through the library's own dependency
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DeviceArrays:
    """Device copies of host float64 arrays, freed on close()."""

    def __init__(self, hip):
        self.hip = hip
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, n):
        out = np.empty(n, dtype=np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _reference_dt(c):
    s = ChannelSolver(c.params, c.grid)
    s.copy_from_field(c.w0)
    t = s.get_aux(6)  # the Newton start of T(e): t0 everywhere after creation
    dt = C.c_double(0.0)
    capi.check(capi.lib().st_compute_dt(s.h, C.byref(dt)))
    s.close()
    return dt.value, t


def _noalloc_dt(hip, c, t_start):
    g = c.grid
    d = DeviceArrays(hip)
    try:
        w = d.put(c.w0)
        nvol = t_start.size
        wa = np.zeros(10 * nvol)
        wa[5 * nvol:6 * nvol] = t_start.ravel()
        wa_d = d.put(wa)
        mets = [d.put(a) for a in (g.dcsidx, g.detady, g.dzitdz, g.dcsidxs, g.detadys, g.dzitdzs)]
        dt_d = d.put(np.full(1, -1.0))
        rc = capi.lib().st_compute_dt_noalloc(C.byref(c.params), None, w, wa_d, *mets, dt_d)
        capi.check(rc)
        assert hip.hipDeviceSynchronize() == 0
        return d.get(dt_d, 1)[0]
    finally:
        d.close()


@pytest.mark.gpu
def test_noalloc_dt_matches_solver_dt():
    hip = _hip()
    big = make_channel(nx=96, ny=64, nz=48, sensor_threshold=0.9)
    small = make_channel(nx=32, ny=24, nz=16, sensor_threshold=0.9, retau=180.0)
    ref_big, t_big = _reference_dt(big)
    ref_small, t_small = _reference_dt(small)
    assert ref_big > 0 and ref_small > 0 and ref_big != ref_small
    for case, ref, t in ((big, ref_big, t_big), (small, ref_small, t_small), (big, ref_big, t_big)):
        got = _noalloc_dt(hip, case, t)
        # a maximum is exact in any order: bitwise equal
        assert got == ref, (case.params.nx, got, ref)


@pytest.mark.gpu
def test_noalloc_dt_rejects_oversize_grid_without_launching():
    c = make_channel(nx=32, ny=24, nz=16)
    p = c.params
    p.nx, p.ny, p.nz = 4096, 2400, 3
    dummy = C.c_void_p(256)  # never dereferenced: the size check comes first
    rc = capi.lib().st_compute_dt_noalloc(C.byref(p), None, *([dummy] * 9))
    assert rc == capi.ST_ERR_ARG
    assert b"too large" in capi.lib().st_last_error()
