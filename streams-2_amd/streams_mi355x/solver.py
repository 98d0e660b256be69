"""Host-side mirror of the reference backend object for one flow solve.

``ChannelSolver`` exposes the type-bound procedures of
``equation_singleideal_cpu_object`` (code/src/singleideal/singleideal_cpu.F90)
that the hot loop uses -- ``update_ghost``, ``compute_aux``, ``visflx``,
``euler_x/y/z``, ``compute_dt``, ``rk_sync``, ``compute_residual``,
``count_weno`` -- each a thin call through the C ABI of libstreams_mi355x.
Fields are numpy arrays in the reference's Fortran layout, indexed
[m, k, j, i] (ghosted extents nz+2ng, ny+2ng, nx+2ng).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .grid import Grid
from .params import StParams, StScalars

NV_STAT = 70  # statistics of compute_stats (singleideal.F90:2003-2179)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class ChannelSolver:
    def __init__(self, params: StParams, grid: Grid):
        self.L = capi.lib()
        self.p = params
        h = C.c_void_p()
        capi.check(self.L.st_solver_create(C.byref(params), C.byref(h)))
        self.h = h
        ng = params.ng
        self.shape_g = (params.nz + 2 * ng, params.ny + 2 * ng, params.nx + 2 * ng)
        self.shape_i = (params.nz, params.ny, params.nx)
        self._grid = [np.ascontiguousarray(a, dtype=np.float64) for a in grid.arrays()]
        capi.check(self.L.st_set_grid(self.h, *[_dp(a) for a in self._grid]))

    # -------------------------------------------------------------- lifecycle
    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.st_solver_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------------ data
    def copy_from_field(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.shape == (5,) + self.shape_g, w.shape
        capi.check(self.L.st_copy_from_field(self.h, _dp(w)))

    def copy_to_field(self) -> np.ndarray:
        w = np.empty((5,) + self.shape_g)
        capi.check(self.L.st_copy_to_field(self.h, _dp(w)))
        return w

    def get_aux(self, var: int) -> np.ndarray:
        a = np.empty(self.shape_g)
        capi.check(self.L.st_get_aux(self.h, var, _dp(a)))
        return a

    def set_aux(self, var: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        capi.check(self.L.st_set_aux(self.h, var, _dp(a)))

    def get_fl(self) -> np.ndarray:
        a = np.empty((5,) + self.shape_i)
        capi.check(self.L.st_get_fl(self.h, _dp(a)))
        return a

    def scalars(self) -> dict:
        s = StScalars()
        capi.check(self.L.st_get_scalars(self.h, C.byref(s)))
        return s.as_dict()

    def set_time(self, time: float, dt: float, icyc: int):
        capi.check(self.L.st_set_time(self.h, time, dt, icyc))

    # ------------------------------------------------------------ procedures
    def update_ghost(self, do_swap: int = 1):
        capi.check(self.L.st_update_ghost(self.h, do_swap))

    def bcswap(self):
        capi.check(self.L.st_bcswap(self.h))

    def compute_aux(self, central: int = 1, ghost: int = 1):
        capi.check(self.L.st_compute_aux(self.h, central, ghost))

    def visflx(self, mode: int):
        capi.check(self.L.st_visflx(self.h, mode))

    def eul_range(self):
        """eul_imin, eul_imax, eul_jmin, eul_jmax, eul_kmin, eul_kmax."""
        r = (C.c_int32 * 6)()
        capi.check(self.L.st_get_eul_range(self.h, r))
        return tuple(r)

    def euler_x(self, istart=None, iend=None, istart_face=None, iend_face=None, do_update=True):
        """euler_x(istart, iend, istart_face, iend_face, do_update); the
        defaults are the reference's: eul_imin..eul_imax, faces istart-1..iend."""
        rg = self.eul_range()
        istart = rg[0] if istart is None else istart
        iend = rg[1] if iend is None else iend
        istart_face = istart - 1 if istart_face is None else istart_face
        iend_face = iend if iend_face is None else iend_face
        capi.check(self.L.st_euler_x(self.h, istart, iend, istart_face, iend_face, int(bool(do_update))))

    def euler_y(self, jmin=None, jmax=None):
        rg = self.eul_range()
        capi.check(self.L.st_euler_y(self.h, rg[2] if jmin is None else jmin, rg[3] if jmax is None else jmax))

    def euler_z(self, kmin=None, kmax=None):
        rg = self.eul_range()
        capi.check(self.L.st_euler_z(self.h, rg[4] if kmin is None else kmin, rg[5] if kmax is None else kmax))

    def init_flux(self, rhodt: float):
        capi.check(self.L.st_init_flux(self.h, rhodt))

    def update_flux(self, gamdt: float):
        capi.check(self.L.st_update_flux(self.h, gamdt))

    def update_field(self):
        capi.check(self.L.st_update_field(self.h))

    def force_rhs(self):
        capi.check(self.L.st_force_rhs(self.h))

    def force_var(self):
        capi.check(self.L.st_force_var(self.h))

    def reset_forcing(self):
        capi.check(self.L.st_reset_forcing(self.h))

    def bcextr_var(self, var: int):
        capi.check(self.L.st_bcextr_var(self.h, var))

    def bcswap_var(self, var: int):
        capi.check(self.L.st_bcswap_var(self.h, var))

    def recyc_exchange(self):
        capi.check(self.L.st_recyc_exchange(self.h))

    def zero_flux(self):
        capi.check(self.L.st_zero_flux(self.h))

    def bc_nr(self):
        capi.check(self.L.st_bc_nr(self.h))

    def set_recycling(self, r):
        """Recycling data of streams_mi355x.cases.recycling_setup."""
        self._rc = [np.ascontiguousarray(a, dtype=np.float64) for a in
                    (r.wmean, r.weta, r.yplus_inflow, r.eta_inflow, r.yplus_recyc, r.eta_recyc)]
        self._rcm = [np.ascontiguousarray(a, dtype=np.int32) for a in (r.map_inn, r.map_out)]
        ip = C.POINTER(C.c_int32)
        capi.check(self.L.st_set_recycling(self.h, r.i_recyc, r.betarecyc, r.glund1, _dp(self._rc[0]),
                                           _dp(self._rc[1]), self._rcm[0].ctypes.data_as(ip),
                                           self._rcm[1].ctypes.data_as(ip), *[_dp(a) for a in self._rc[2:]]))

    def set_wmean(self, wmean):
        """Inflow mean planes wmean(1-ng:0, 1:ny, 1:4) (laminar inflow, tag 9)."""
        self._wm = np.ascontiguousarray(wmean, dtype=np.float64)
        capi.check(self.L.st_set_wmean(self.h, _dp(self._wm)))

    def init_run(self):
        capi.check(self.L.st_init_run(self.h))

    def compute_dt(self) -> float:
        dt = C.c_double()
        capi.check(self.L.st_compute_dt(self.h, C.byref(dt)))
        return dt.value

    def rk_sync(self):
        capi.check(self.L.st_rk_sync(self.h))

    def advance(self, nsteps: int):
        capi.check(self.L.st_advance(self.h, nsteps))

    def compute_residual(self) -> dict:
        s = StScalars()
        capi.check(self.L.st_compute_residual(self.h, C.byref(s)))
        return s.as_dict()

    def compute_stats(self, itav: int):
        """Fold the current state into the device's running statistics as
        the itav-th sample (compute_stats)."""
        capi.check(self.L.st_compute_stats(self.h, int(itav)))

    def get_stats(self) -> np.ndarray:
        """w_stat(1:70, 1:nx, 1:ny), Fortran order."""
        out = np.zeros((NV_STAT, self.p.nx, self.p.ny), order="F")
        capi.check(self.L.st_get_stats(self.h, _dp(out)))
        return out

    def set_stats(self, w_stat: np.ndarray):
        a = np.asfortranarray(w_stat, dtype=np.float64)
        assert a.shape == (NV_STAT, self.p.nx, self.p.ny)
        capi.check(self.L.st_set_stats(self.h, _dp(a)))

    def count_weno(self) -> np.ndarray:
        out = np.zeros(3)
        capi.check(self.L.st_count_weno(self.h, _dp(out)))
        return out

    def kernel_times(self, nsteps: int) -> dict:
        """Mean ms per launch of each kernel family over nsteps timed steps."""
        ms = (C.c_double * len(capi.KERNEL_FAMILIES))()
        cnt = (C.c_int32 * len(capi.KERNEL_FAMILIES))()
        capi.check(self.L.st_kernel_times(self.h, nsteps, ms, cnt))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(capi.KERNEL_FAMILIES) if cnt[i]}

    def synchronize(self):
        capi.check(self.L.st_synchronize(self.h))


PROGRESS_COLUMNS = ("icyc", "dt", "time", "residual_rhou", "-dpdx/dt", "rhobulk", "ubulk", "tbulk",
                    "rhomin", "rhomax", "tmin", "tmax", "pmin", "pmax")


def progress_row(s: dict, p: StParams) -> list:
    """The channel row of progress.out (print_progress,
    singleideal_cpu.F90:1695-1732, grid_dim 1): residual and dpdx scaled by
    l0 / (rho0 u0^2), dpdx divided by dt."""
    res = s["residual_rhou"] / (p.u0 ** 2) / p.rho0 * p.l0
    dpdx = s["dpdx"] / (p.u0 ** 2) / p.rho0 * p.l0
    return [s["icyc"], s["dt"], s["time"], res, -dpdx / s["dt"], s["rhobulk"], s["ubulk"], s["tbulk"],
            s["rhomin"], s["rhomax"], s["tmin"], s["tmax"], s["pmin"], s["pmax"]]


def progress_text(row: list) -> list:
    """Columns as the reference prints them: '(1i10,20es20.10)'."""
    return [str(int(row[0]))] + ["%.10E" % v for v in row[1:]]
