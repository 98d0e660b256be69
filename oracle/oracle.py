"""ctypes wrapper of the CPU oracle (test infrastructure only).

The oracle is the checker of the HIP path: only tests/, the smoke check and
the CPU-baseline leg of bench.py use it.  Build it with ``make -C oracle``.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.join(os.path.dirname(_HERE), "streams-2_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from streams_mi355x.params import StParams, StScalars  # noqa: E402

NV_STAT = 70  # statistics of compute_stats (singleideal.F90:2003-2179)

_LIB = None


def lib():
    """The oracle library.  STREAMS_ORACLE_LIB selects another build of the
    same sources (the mutation checks build copies with planted mistakes to
    show that the pins fail on them)."""
    global _LIB
    if _LIB is None:
        path = os.environ.get("STREAMS_ORACLE_LIB", os.path.join(_HERE, "liboracle.so"))
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = C.CDLL(path)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.or_create.restype = vp
        L.or_create.argtypes = [C.POINTER(StParams)]
        L.or_destroy.argtypes = [vp]
        L.or_set_grid.argtypes = [vp] + [dp] * 13
        for name in ("or_w", "or_w_aux", "or_fl", "or_fln", "or_fhat", "or_wallprop"):
            getattr(L, name).restype = dp
            getattr(L, name).argtypes = [vp]
        L.or_eoc.restype = ip
        L.or_eoc.argtypes = [vp]
        L.or_get_scalars.argtypes = [vp, C.POINTER(StScalars)]
        L.or_set_time.argtypes = [vp, C.c_double, C.c_double, C.c_int64]
        L.or_rk_coeffs.argtypes = [vp, dp, dp]
        L.or_get_deriv_coeffs.argtypes = [dp, dp, C.c_int, C.c_int]
        L.or_get_gamloc.restype = C.c_double
        L.or_get_gamloc.argtypes = [vp, C.c_double]
        L.or_get_temperature_from_e.restype = C.c_double
        L.or_get_temperature_from_e.argtypes = [vp, C.c_double, C.c_double]
        L.or_get_e_from_temperature.restype = C.c_double
        L.or_get_e_from_temperature.argtypes = [vp, C.c_double]
        L.or_compute_roe_average.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.or_eigenvectors.argtypes = [C.c_int, dp, dp, dp]
        L.or_wenorec_1d.argtypes = [dp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
        L.or_wenorec_1d_rusanov.argtypes = [dp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
        L.or_eval_aux.argtypes = [vp] + [C.c_int] * 6
        L.or_compute_aux.argtypes = [vp, C.c_int, C.c_int]
        for name in ("or_bcswap", "or_zero_flux", "or_euler_x", "or_euler_y", "or_euler_z",
                     "or_force_rhs", "or_update_field", "or_force_var", "or_compute_dt",
                     "or_rk_sync", "or_init_run", "or_compute_residual", "or_bc_nr"):
            getattr(L, name).argtypes = [vp]
        L.or_bcswap_var.argtypes = [vp, C.c_int]
        L.or_bcextr_var.argtypes = [vp, C.c_int]
        L.or_update_ghost.argtypes = [vp, C.c_int]
        L.or_init_flux.argtypes = [vp, C.c_double]
        L.or_update_flux.argtypes = [vp, C.c_double]
        L.or_euler_fluxes.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.or_euler_update.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.or_visflx.argtypes = [vp, C.c_int]
        L.or_advance.argtypes = [vp, C.c_int]
        L.or_count_weno.argtypes = [vp, dp]
        L.or_compute_stats.argtypes = [vp, C.c_int, dp]
        L.or_inflow_plane.restype = dp
        L.or_inflow_plane.argtypes = [vp]
        L.or_set_wmean.restype = C.c_int
        L.or_set_wmean.argtypes = [vp, dp]
        L.or_seed_random.argtypes = [vp]
        L.or_set_recycling.restype = C.c_int
        L.or_set_recycling.argtypes = [vp, C.c_int, C.c_double, C.c_double, dp, dp, ip, ip, dp, dp, dp, dp]
        _LIB = L
    return _LIB


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One flow solve on the CPU, arrays viewed as Fortran-order numpy arrays.

    ``w`` has shape (5, nz+2ng, ny+2ng, nx+2ng), i.e. index [m, k, j, i]."""

    def __init__(self, params: StParams, grid=None):
        self.L = lib()
        self.p = params
        self.h = self.L.or_create(C.byref(params))
        if not self.h:
            raise MemoryError("oracle allocation failed")
        nx, ny, nz, ng = params.nx, params.ny, params.nz, params.ng
        self.shape_g = (nz + 2 * ng, ny + 2 * ng, nx + 2 * ng)
        self.shape_i = (nz, ny, nx)
        nvol = int(np.prod(self.shape_g))
        nint = int(np.prod(self.shape_i))
        self.w = np.ctypeslib.as_array(self.L.or_w(self.h), shape=(5 * nvol,)).reshape((5,) + self.shape_g)
        self.w_aux = np.ctypeslib.as_array(self.L.or_w_aux(self.h), shape=(10 * nvol,)).reshape((10,) + self.shape_g)
        self.fl = np.ctypeslib.as_array(self.L.or_fl(self.h), shape=(5 * nint,)).reshape((5,) + self.shape_i)
        self.fln = np.ctypeslib.as_array(self.L.or_fln(self.h), shape=(5 * nint,)).reshape((5,) + self.shape_i)
        self.fhat = np.ctypeslib.as_array(self.L.or_fhat(self.h), shape=(5 * nvol,)).reshape((5,) + self.shape_g)
        neoc = 3 * (nx + 1) * (ny + 1) * (nz + 1)
        self.eoc = np.ctypeslib.as_array(self.L.or_eoc(self.h), shape=(neoc,)).reshape((3, nz + 1, ny + 1, nx + 1))
        if grid is not None:
            self.set_grid(grid)

    def set_grid(self, g):
        self._grid_keep = [np.ascontiguousarray(a, dtype=np.float64) for a in g.arrays()]
        self.L.or_set_grid(self.h, *[_dp(a) for a in self._grid_keep])

    def close(self):
        if self.h:
            self.L.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scalars(self) -> dict:
        s = StScalars()
        self.L.or_get_scalars(self.h, C.byref(s))
        return s.as_dict()

    def __getattr__(self, name):
        # or_<name>(handle, *args) forwarding for the simple entry points
        if name.startswith("_"):
            raise AttributeError(name)
        fn = getattr(lib(), "or_" + name)
        return lambda *a: fn(self.h, *a)

    def rk_coeffs(self):
        rho, gam = np.zeros(3), np.zeros(3)
        self.L.or_rk_coeffs(self.h, _dp(rho), _dp(gam))
        return rho, gam

    def set_recycling(self, r):
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (r.wmean, r.weta, r.yplus_inflow, r.eta_inflow, r.yplus_recyc, r.eta_recyc)]
        maps = [np.ascontiguousarray(a, dtype=np.int32) for a in (r.map_inn, r.map_out)]
        self._rc = (keep, maps)
        ip = C.POINTER(C.c_int)
        rc = self.L.or_set_recycling(self.h, r.i_recyc, r.betarecyc, r.glund1, _dp(keep[0]), _dp(keep[1]),
                                     maps[0].ctypes.data_as(ip), maps[1].ctypes.data_as(ip),
                                     *[_dp(a) for a in keep[2:]])
        if rc:
            raise ValueError("bad recycling data")

    def set_wmean(self, wmean: np.ndarray):
        self._wm = np.ascontiguousarray(wmean, dtype=np.float64)
        if self.L.or_set_wmean(self.h, _dp(self._wm)):
            raise MemoryError("wmean allocation failed")

    def inflow_plane(self) -> np.ndarray:
        """inflow_random_plane(1:ny, 1:nz, 1:3) as an array [c, k, j] (a view)."""
        ptr = self.L.or_inflow_plane(self.h)
        if not ptr:
            raise RuntimeError("no recycling inflow set")
        p = self.p
        return np.ctypeslib.as_array(ptr, shape=(3 * p.ny * p.nz,)).reshape((3, p.nz, p.ny))

    def compute_stats(self, itav: int, w_stat: np.ndarray):
        """compute_stats: fold this state's spanwise means into the running
        time average ``w_stat`` (70 x nx x ny, Fortran order, updated in
        place) as the itav-th sample."""
        assert w_stat.flags.f_contiguous and w_stat.dtype == np.float64
        assert w_stat.shape == (NV_STAT, self.p.nx, self.p.ny)
        self.L.or_compute_stats(self.h, int(itav), w_stat.ctypes.data_as(C.POINTER(C.c_double)))

    def count_weno(self):
        out = np.zeros(3)
        self.L.or_count_weno(self.h, _dp(out))
        return out
