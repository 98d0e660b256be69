"""Seeded synthetic flow cases (grid + parameters + initial field)."""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from .grid import Grid, channel_grid, channel_initial_field
from .params import FlowCase, StParams, channel_params


@dataclass
class ChannelCase:
    case: FlowCase
    grid: Grid
    params: StParams
    w0: np.ndarray          # initial conserved field, Fortran layout [m,k,j,i] with ghosts


def make_channel(nx=32, ny=24, nz=16, mach=1.5, retau=220.0, noise=0.10, seed=0,
                 sensor_threshold=0.9, cfl=0.8, uniform_y=False, noise_source="counter", **kw) -> ChannelCase:
    """Isothermal-wall channel: periodic x/z, staggered walls in y.

    Domain 4pi x 2 x 4pi/3, tanh-type y stretching (generate_grid_cha),
    laminar Poiseuille profile plus ``noise`` (peak-to-peak fraction of u0)
    of seeded uniform velocity perturbations."""
    case = FlowCase(nx=nx, ny=ny, nz=nz, mach=mach, reynolds=retau, sensor_threshold=sensor_threshold,
                    cfl=cfl, **kw)
    grid = channel_grid(nx, ny, nz, ng=case.ng, lx=case.lx, ly=case.ly, lz=case.lz, retau=retau,
                        jbgrid=case.jbgrid, dyptarget=case.dyptarget, nsmooy=case.ysmoosteps,
                        metrics_order=case.ep_order, uniform_y=uniform_y)
    params = channel_params(case, grid.yn)
    e_ref = params.cv_coeff[1] * params.t0  # calorically perfect e(T=t0)
    w0 = channel_initial_field(grid, params.u0, noise=noise, seed=seed, rho0=params.rho0, t0=params.t0,
                               rgas0=params.rgas0, e_ref=e_ref, noise_source=noise_source)
    return ChannelCase(case, grid, params, w0)


# ------------------------------------------------ the reference's own decks
RAND_MAX = 2147483647.0


def crandom_stream(seed: int, n: int) -> np.ndarray:
    """n draws of get_crandom after init_crandom(seed).

    code/third-party/crandom/crandom.c: srand(seed), then rand()/RAND_MAX
    (scaled by 0 when seed == 0).  The C library's generator is called, as
    the reference does."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand.argtypes = [ctypes.c_uint]
    libc.rand.restype = ctypes.c_int
    libc.srand(seed)
    out = np.fromiter((libc.rand() for _ in range(n)), dtype=np.float64, count=n)
    rescale = 0.0 if seed == 0 else 1.0
    return rescale * out / RAND_MAX


def init_channel_field(grid: Grid, params: StParams, lz: float, splits=(1, 1, 1),
                       rand_type: int = 2) -> np.ndarray:
    """Initial channel field of init_channel (singleideal.F90:4344-4419).

    Laminar u = 1.5 u0 (1 - y^2) plus 2 % of u0 random fluctuations per
    component (rr - 0.5, three draws per cell) and streamwise rolls of
    amplitude 5 % u0 in v and w.  The random draws follow the MPI
    decomposition of the reference run: block (cx, cy, cz) of the
    x_split x y_split x z_split Cartesian topology (row-major rank order,
    field.F90:245-246) seeds with rank+1 (singleideal.F90:314-323, rand_type
    > 0) and draws over its local k, j, i loops.  rand_type == 0 disables
    the fluctuations (seed 0).  Returns w in Fortran layout [m, k, j, i]
    with zero ghosts.
    """
    nx, ny, nz, ng = grid.nx, grid.ny, grid.nz, grid.ng
    sxs, sys_, szs = splits
    if nx % sxs or ny % sys_ or nz % szs:
        raise ValueError("grid not divisible by the MPI splits")
    nxl, nyl, nzl = nx // sxs, ny // sys_, nz // szs
    rho0, u0, l0, t0, rgas0 = params.rho0, params.u0, params.l0, params.t0, params.rgas0
    p0 = rho0 * rgas0 * t0
    pi = math.acos(-1.0)
    yn_top = grid.yn[ny]
    nroll = int(lz / yn_top)
    u0_02 = 0.02 * u0
    if rand_type == 0:
        u0_02 = 0.0
    u0_05 = 0.05 * u0
    y = grid.y[ng:ng + ny]
    z = grid.z[ng:ng + nz]
    # random draws rr(1:3) - 0.5 of every interior cell, in the reference's order
    rr = np.zeros((3, nz, ny, nx))
    for cx in range(sxs):
        for cy in range(sys_):
            for cz in range(szs):
                rank = (cx * sys_ + cy) * szs + cz
                seed = 0 if rand_type == 0 else rank + 1
                r = crandom_stream(seed, 3 * nxl * nyl * nzl).reshape(nzl, nyl, nxl, 3)
                rr[:, cz * nzl:(cz + 1) * nzl, cy * nyl:(cy + 1) * nyl, cx * nxl:(cx + 1) * nxl] = \
                    np.moveaxis(r, 3, 0) - 0.5
    rho = rho0
    ubase = np.array([1.5 * rho0 * u0 * (1.0 - (yj / l0) * (yj / l0)) for yj in y]) / rho
    sy = np.array([math.sin(0.5 * pi * yj / l0) for yj in y])
    cz_ = np.array([math.cos(2.0 * pi * zk / lz * nroll) for zk in z])
    sz_ = np.array([math.sin(2.0 * pi * zk / lz * nroll) for zk in z])
    ufluc = u0_02 * rr[0]
    vfluc = u0_02 * rr[1] + (u0_05 * sy)[None, :, None] * cz_[:, None, None]
    wfluc = u0_02 * rr[2] + (u0_05 * sy)[None, :, None] * sz_[:, None, None]
    uu = ubase[None, :, None] + ufluc
    vv = 0.0 / rho + vfluc
    ww = 0.0 / rho + wfluc
    rhouu, rhovv, rhoww = rho * uu, rho * vv, rho * ww
    tt = p0 / rgas0 / rho
    # calorically perfect e(T) of get_e_from_temperature (kernels_cpu.F90:7155)
    ee = t0 * (params.cv_coeff[1] + params.cv_coeff[0] * (tt / t0 - 1.0))
    w = np.zeros((5, nz + 2 * ng, ny + 2 * ng, nx + 2 * ng))
    sl = (slice(ng, ng + nz), slice(ng, ng + ny), slice(ng, ng + nx))
    w[0][sl] = rho
    w[1][sl] = rhouu
    w[2][sl] = rhovv
    w[3][sl] = rhoww
    w[4][sl] = rho * ee + 0.5 * (rhouu ** 2 + rhovv ** 2 + rhoww ** 2) / rho
    return w


def make_subsonic_channel(nx=80, ny=60, nz=80) -> ChannelCase:
    """examples/subsonic_channel/singleideal.ini, resized as tests/tests.ini
    does for its [subsonic_channel] golden row (80 x 60 x 80, 2x1x2 ranks).

    Mach 0.2 at T_wall, Re_tau 180, theta_wall -2.5 (bulk temperature free),
    Sutherland viscosity (T_ref 220 K, S 110.4 K), conservative viscous form,
    sensor threshold 1.01, CFL 0.8, domain 12 x 2 x 6, CHA grid with
    dy+ 0.7, jbgrid 32, 50 smoothing sweeps."""
    case = FlowCase(nx=nx, ny=ny, nz=nz, lx=12.0, ly=2.0, lz=6.0, mach=0.2, reynolds=180.0,
                    theta_wall=-2.5, visc_model=2, s_suth=110.4, t_ref=220.0, conservative_viscous=1,
                    ep_order=6, nkeep=0, weno_scheme=3, weno_version=0, sensor_threshold=1.01, cfl=0.8,
                    dyptarget=0.7, jbgrid=32.0, ysmoosteps=50)
    grid = channel_grid(nx, ny, nz, ng=case.ng, lx=case.lx, ly=case.ly, lz=case.lz, retau=case.reynolds,
                        jbgrid=case.jbgrid, dyptarget=case.dyptarget, nsmooy=case.ysmoosteps,
                        metrics_order=6)
    params = channel_params(case, grid.yn)
    w0 = init_channel_field(grid, params, case.lz, splits=(2, 1, 2), rand_type=2)
    return ChannelCase(case, grid, params, w0)
