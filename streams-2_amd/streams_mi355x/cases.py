"""Seeded synthetic flow cases (grid + parameters + initial field)."""
from __future__ import annotations

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
                 sensor_threshold=0.9, cfl=0.8, uniform_y=False, **kw) -> ChannelCase:
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
                               rgas0=params.rgas0, e_ref=e_ref)
    return ChannelCase(case, grid, params, w0)
