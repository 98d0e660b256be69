"""Seeded synthetic flow cases (grid + parameters + initial field)."""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from .grid import Grid, channel_grid, channel_initial_field
from .params import FlowCase, StParams, channel_params, set_fluid_prop


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
    # e(T = t0): every term of get_e_from_temperature but the constant vanishes
    e_ref = params.cv_coeff[params.indx_cp_r + 1 - params.indx_cp_l] * params.t0
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
    params.rand_type = 2   # the deck's [numerics] rand_type (tests.ini's patch leaves it)
    w0 = init_channel_field(grid, params, case.lz, splits=(2, 1, 2), rand_type=2)
    return ChannelCase(case, grid, params, w0)


def bl_params(nx, ny, nz, mach=2.0, reynolds=200.0, t_wall=1.6, bctags=(1, 2, 6, 2, 0, 0),
              bctags_nr=(0, 1, 0, 1, 0, 0), sensor_threshold=0.2, cfl=0.5, conservative_viscous=1,
              shock_angle_deg=None, xshock_imp=None, tanhfacs=None, precision=0, rand_type=0,
              aoa_deg=0.0) -> StParams:
    """Parameters of a boundary-layer box (flow_init = 1 family): node-centred
    grid, free-stream reference state rho0 = t0 = rgas0 = l0 = 1, u0 = M c0,
    mu0 = u0 rho0 l0 / Re (set_flow_params, singleideal.F90:2798-2826), an
    isothermal wall at T_wall, no forcing.  winf is the free-stream
    conserved state; with shock_angle_deg the post-shock state of the
    oblique-shock relations (set_oblique_shock, singleideal.F90:2955-3017)
    fills winf_past_shock.  ttot0 and ptot0 are the total conditions of the
    free stream (singleideal.F90:2813-2814), aoa its angle of attack."""
    case = FlowCase(nx=nx, ny=ny, nz=nz, mach=mach, reynolds=reynolds, sensor_threshold=sensor_threshold, cfl=cfl,
                    conservative_viscous=conservative_viscous, precision=precision)
    gam, gm1, rfac, icl, icr, cp, cv = set_fluid_prop(case)
    rgas0 = t0 = rho0 = l0 = 1.0
    c0 = math.sqrt(gam * rgas0 * t0)
    u0 = mach * c0
    p = StParams()
    p.nx, p.ny, p.nz, p.ng = nx, ny, nz, 3
    p.ep_order = p.visc_order = 6
    p.nkeep = 0
    p.weno_scheme = 3
    p.weno_version = 0
    p.flux_splitting = 0
    p.conservative_viscous = conservative_viscous
    p.visc_model = 2
    p.calorically_perfect = 1
    p.indx_cp_l, p.indx_cp_r = icl, icr
    p.bctags[:] = list(bctags)
    p.bctags_nr[:] = list(bctags_nr)
    p.is_y_staggered = 0
    p.channel_case = 0
    p.rand_type = rand_type
    p.aoa = aoa_deg * math.pi / 180.0
    p.ttot0 = t0 * (1.0 + 0.5 * gm1 * mach ** 2)
    p.ptot0 = rho0 * rgas0 * t0 * (1.0 + 0.5 * gm1 * mach ** 2) ** (gam / gm1)
    p.correct_bound_ord = 1
    p.iter_dt_recompute = 1
    p.precision = precision
    p.rgas0, p.prandtl, p.t0, p.rho0, p.u0, p.l0 = rgas0, case.prandtl, t0, rho0, u0, l0
    p.mu0 = u0 * rho0 * l0 / reynolds
    p.sutherland_s, p.t_ref_dim, p.powerlaw_vtexp = 110.4, 220.0, 0.76
    p.sensor_threshold = sensor_threshold
    p.t_wall = t_wall
    p.theta_wall = 1.0
    p.volchan = 1.0
    p.cfl = cfl
    p.tol_iter_nr = 1e-12
    for i, v in enumerate(cp):
        p.cp_coeff[i] = v
    for i, v in enumerate(cv):
        p.cv_coeff[i] = v
    e0 = cv[1] * t0  # calorically perfect e(t0)
    winf = [rho0, rho0 * u0, 0.0, 0.0, rho0 * (e0 + 0.5 * u0 ** 2)]
    for m in range(5):
        p.winf[m] = winf[m]
        p.winf_past_shock[m] = winf[m]
    if shock_angle_deg is not None:
        beta = shock_angle_deg * math.pi / 180.0
        sins, coss = math.sin(beta), math.cos(beta)
        mn2 = (mach * sins) ** 2
        mn_ps = math.sqrt((1.0 + 0.5 * gm1 * mn2) / (gam * mn2 - 0.5 * gm1))
        vel_ratio = (2.0 + gm1 * mn2) / ((gam + 1.0) * mn2)
        rho_ratio = 1.0 / vel_ratio
        p_ratio = 1.0 + 2.0 * gam / (gam + 1.0) * (mn2 - 1.0)
        t_ratio = p_ratio / rho_ratio
        tdef = 2.0 * coss / sins * (mach ** 2 * sins ** 2 - 1.0) / (2.0 + mach ** 2 * (gam + math.cos(2.0 * beta)))
        delta = math.atan(tdef)
        m_ps = mn_ps / math.sin(beta - delta)
        t_ps = t0 * t_ratio
        c_ps = math.sqrt(gam * rgas0 * t_ps)
        u_ps, v_ps = m_ps * c_ps * math.cos(delta), m_ps * c_ps * math.sin(delta)
        rho_ps = rho0 * rho_ratio
        e_ps = cv[1] * t0 + cv[0] * (t_ps - t0)
        wps = [rho_ps, rho_ps * u_ps, -rho_ps * v_ps, 0.0, rho_ps * (e_ps + 0.5 * (u_ps ** 2 + v_ps ** 2))]
        for m in range(5):
            p.winf_past_shock[m] = wps[m]
        p.shock_angle = beta
        p.xshock_imp = xshock_imp
        p.tanhfacs = tanhfacs
    return p


def bl_grid(nx, ny, nz, lx=8.0, ly=3.0, lz=2.0, beta_y=1.8) -> Grid:
    """Node-centred grid of a boundary-layer box: uniform x (non-periodic:
    nx-1 intervals) and periodic z, y from the wall y=0 with tanh
    clustering (y_j = ly tanh(beta s)/tanh(beta), s = (j-1)/(ny-1)), ghosts
    mirrored about the wall and extended linearly at the top; 6th-order
    metrics (grid.F90:71-141)."""
    from .grid import local_metrics, metrics, uniform_nodes

    ng = 3
    xg = uniform_nodes(nx, ng, lx, False)
    zg = uniform_nodes(nz, ng, lz, True)
    j = np.arange(1 - ng, ny + ng + 1, dtype=np.float64)
    s = (j - 1.0) / (ny - 1.0)
    yg = ly * np.tanh(beta_y * s) / math.tanh(beta_y)
    # mirror below the wall, linear continuation above the top
    for l in range(1, ng + 1):
        yg[ng - l] = -yg[ng + l]
        yg[ng + ny - 1 + l] = 2.0 * yg[ng + ny - 1] - yg[ng + ny - 1 - l]
    yg = yg[: ny + 2 * ng]
    yn = np.concatenate([[0.0], 0.5 * (yg[ng:ng + ny - 1] + yg[ng + 1:ng + ny]), [yg[ng + ny - 1]]])
    dx1, dx2 = metrics(xg, nx, ng, 6)
    dy1, dy2 = metrics(yg, ny, ng, 6)
    dz1, dz2 = metrics(zg, nz, ng, 6)
    return Grid(nx, ny, nz, ng, xg[: nx + 2 * ng].copy(), yg.copy(), zg[: nz + 2 * ng].copy(), yn,
                *local_metrics(dx1, dx2), *local_metrics(dy1, dy2), *local_metrics(dz1, dz2))


def bl_initial_field(grid: Grid, params: StParams, delta=0.6, noise=0.05, seed=0) -> np.ndarray:
    """Seeded compressible boundary-layer-like field: u = u0 tanh(y/delta),
    T from the Crocco-Busemann relation between T_wall and T_inf = t0 with
    recovery factor Pr^(1/3), uniform pressure, plus ``noise`` (peak-to-peak
    fraction of u0) of uniform velocity fluctuations damped toward the wall.
    Ghosts are zero (filled by update_ghost)."""
    nx, ny, nz, ng = grid.nx, grid.ny, grid.nz, grid.ng
    u0, t0, rgas0 = params.u0, params.t0, params.rgas0
    gam = params.cp_coeff[0] / params.cv_coeff[0]
    rfac = params.prandtl ** (1.0 / 3.0)
    tr = t0 * (1.0 + 0.5 * (gam - 1.0) * rfac * (u0 ** 2 / (gam * rgas0 * t0)))
    y = grid.y[ng:ng + ny]
    uo = np.tanh(np.maximum(y, 0.0) / delta)
    tt = params.t_wall + (tr - params.t_wall) * uo + (t0 - tr) * uo ** 2
    r = np.random.default_rng(seed).random((3, nz, ny, nx)) - 0.5
    damp = (uo * (1.0 - uo) * 4.0)[None, :, None]
    uu = u0 * uo[None, :, None] + noise * u0 * r[0] * damp
    vv = noise * u0 * r[1] * damp
    ww = noise * u0 * r[2] * damp
    rho = (params.rho0 * t0 / tt)[None, :, None] * np.ones((nz, 1, nx))
    ee = params.cv_coeff[1] * t0 + params.cv_coeff[0] * (tt - t0)
    w = np.zeros((5, nz + 2 * ng, ny + 2 * ng, nx + 2 * ng))
    sl = (slice(ng, ng + nz), slice(ng, ng + ny), slice(ng, ng + nx))
    w[0][sl] = rho
    w[1][sl] = rho * uu
    w[2][sl] = rho * vv
    w[3][sl] = rho * ww
    w[4][sl] = rho * ee[None, :, None] + 0.5 * rho * (uu ** 2 + vv ** 2 + ww ** 2)
    return w


def make_bl_box(nx=40, ny=24, nz=16, seed=0, **kw) -> ChannelCase:
    g = bl_grid(nx, ny, nz)
    p = bl_params(nx, ny, nz, **kw)
    return ChannelCase(FlowCase(nx=nx, ny=ny, nz=nz), g, p, bl_initial_field(g, p, seed=seed))


def locateval(xx: np.ndarray, x: float) -> int:
    """Bisection locate of parameters.F90:145-164 for increasing xx(1:n):
    the number of entries strictly below x (0..n)."""
    return int(np.searchsorted(np.asarray(xx), x, side="left"))


@dataclass
class Recycling:
    i_recyc: int
    betarecyc: float
    glund1: float
    wmean: np.ndarray        # wmean(1-ng:0, 1:ny, 1:4) in Fortran order: array [m, j, ng-i] holds plane 1-i
    weta: np.ndarray         # (ny,)
    map_inn: np.ndarray      # (ny,) int32
    map_out: np.ndarray
    yplus_inflow: np.ndarray  # (ny + 2 ng,) index j + ng - 1
    eta_inflow: np.ndarray
    yplus_recyc: np.ndarray
    eta_recyc: np.ndarray


def recycling_setup(grid: Grid, x_recyc: float, delta_in: float, deltav_in: float, delta_rc: float,
                    deltav_rc: float, wmean_planes: np.ndarray) -> Recycling:
    """recyc_prepare (singleideal.F90:1512-1588) for one block: the station
    i_recyc from the x grid, the rescaling factors beta = dv(rc)/dv(in) and
    (d(rc)/d(in))**0.13, the inner (y+) and outer (eta) coordinates of the
    inflow and of the station, the maps between them and the blending
    weight.  The boundary-layer and viscous thicknesses at the inflow and at
    the station are given (in the reference they come from the mean
    profile of init_bl).  ``wmean_planes`` is the inflow mean profile
    wmean(1-ng:0, 1:ny, 1:4) in Fortran order, an array [m, j, q] whose
    q = 0..ng-1 is plane 1-ng+q."""
    ng, nx, ny = grid.ng, grid.nx, grid.ny
    xg = grid.x[ng:ng + nx]
    ig = locateval(xg, x_recyc)
    i_recyc = max(ig, ng)
    bexp = 0.13
    betarecyc = deltav_rc / deltav_in
    glund1 = (delta_rc / delta_in) ** bexp
    y = grid.y
    yplus_in = y / deltav_in
    eta_in = y / delta_in
    yplus_rc = y / deltav_rc
    eta_rc = y / delta_rc
    map_inn = np.array([locateval(yplus_rc[ng:ng + ny], yplus_in[ng + j]) for j in range(ny)], dtype=np.int32)
    map_out = np.array([locateval(eta_rc[ng:ng + ny], eta_in[ng + j]) for j in range(ny)], dtype=np.int32)
    alfa, beta = 4.0, 0.2
    cb = 1.0 - 2.0 * beta
    e = eta_in[ng:ng + ny]
    weta = 0.5 * (1.0 + np.tanh((alfa * (e - beta)) / (beta + e * cb)) / math.tanh(alfa))
    return Recycling(i_recyc, betarecyc, glund1, np.ascontiguousarray(wmean_planes, dtype=np.float64), weta,
                     map_inn, map_out, yplus_in, eta_in, yplus_rc, eta_rc)


def bl_mean_planes(case: ChannelCase) -> np.ndarray:
    """Spanwise/first-column mean profile of the seeded boundary-layer field
    (rho, rho*u, rho*v, rho*w) for the ng inflow ghost planes:
    wmean(1-ng:0, 1:ny, 1:4) as an array [m, j, q], q = 0 the plane 1-ng."""
    ng, ny = case.grid.ng, case.grid.ny
    w = case.w0
    prof = w[:4, ng:-ng, ng:ng + ny, ng].mean(axis=1)  # [m, j] at i = 1
    prof[2:4] = 0.0
    return np.repeat(prof[:, :, None], ng, axis=2)


def oblique_shock_states(params: StParams, shock_angle_deg: float) -> dict:
    """Ratios across an oblique shock of angle beta in the free stream of
    ``params`` (Mach u0/c0): density, temperature, post-shock velocity
    components (v < 0, toward the wall) and the flow deflection, from the
    Rankine-Hugoniot relations for the normal Mach number M sin(beta)."""
    gam = params.cp_coeff[0] / params.cv_coeff[0]
    gm1 = gam - 1.0
    c0 = math.sqrt(gam * params.rgas0 * params.t0)
    mach = params.u0 / c0
    beta = shock_angle_deg * math.pi / 180.0
    mn2 = (mach * math.sin(beta)) ** 2
    rho_ratio = (gam + 1.0) * mn2 / (2.0 + gm1 * mn2)
    p_ratio = 1.0 + 2.0 * gam / (gam + 1.0) * (mn2 - 1.0)
    t_ratio = p_ratio / rho_ratio
    tdef = 2.0 / math.tan(beta) * (mn2 - 1.0) / (2.0 + mach ** 2 * (gam + math.cos(2.0 * beta)))
    delta = math.atan(tdef)
    mn_ps = math.sqrt((1.0 + 0.5 * gm1 * mn2) / (gam * mn2 - 0.5 * gm1))
    m_ps = mn_ps / math.sin(beta - delta)
    q_ps = m_ps * math.sqrt(gam * params.rgas0 * params.t0 * t_ratio)
    return {"rho": rho_ratio, "t": t_ratio, "p": p_ratio, "u": q_ps * math.cos(delta), "v": -q_ps * math.sin(delta),
            "deflection": delta, "beta": beta}


def shock_angle_from_deflection(mach: float, theta_deg: float, gam: float = 1.4) -> float:
    """Weak-solution shock angle (degrees) of a wedge of ``theta_deg`` in a
    stream of Mach ``mach``: the root of the theta-beta-M relation
    tan(theta) = 2 cot(beta) (M^2 sin^2 beta - 1) / (M^2 (gam + cos 2 beta) + 2)
    between the Mach angle and the angle of largest deflection (bisection)."""
    def tdef(b):
        return 2.0 / math.tan(b) * ((mach * math.sin(b)) ** 2 - 1.0) / (mach ** 2 * (gam + math.cos(2.0 * b)) + 2.0)
    lo = math.asin(1.0 / mach)
    # the deflection rises from zero at the Mach angle to its maximum, then falls
    grid = np.linspace(lo, 0.5 * math.pi, 2001)
    hi = float(grid[int(np.argmax([tdef(b) for b in grid]))])
    target = math.tan(theta_deg * math.pi / 180.0)
    if tdef(hi) < target:
        raise ValueError("deflection beyond the attached-shock limit")
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if tdef(mid) < target:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi) * 180.0 / math.pi


def impose_oblique_shock(grid: Grid, params: StParams, w: np.ndarray, shock_angle_deg: float, x_imp: float,
                         width: float = 2.0) -> np.ndarray:
    """The incident oblique shock already across the field: behind the line
    x = x_imp - y / tan(beta) (the shock that the top boundary imposes and
    that meets the wall at x_imp) density and temperature take the jump
    ratios, u its post-shock ratio and v the post-shock downward velocity,
    both scaled by the local u/u0 so that the wall stays at rest; a tanh
    ramp of ``width`` x-spacings smooths the jump.  Interior cells only;
    total energy from the new state (calorically perfect gas)."""
    ng, nx, ny, nz = grid.ng, grid.nx, grid.ny, grid.nz
    s = oblique_shock_states(params, shock_angle_deg)
    x = grid.x[ng:ng + nx]
    y = grid.y[ng:ng + ny]
    dx = float(x[1] - x[0])
    xs = x_imp - np.maximum(y, 0.0) / math.tan(s["beta"])
    f = 0.5 * (1.0 + np.tanh((x[None, :] - xs[:, None]) / (width * dx)))[None, :, :]  # [1, j, i]
    out = w.copy()
    sl = (slice(ng, ng + nz), slice(ng, ng + ny), slice(ng, ng + nx))
    rho = w[0][sl]
    u, v, ww = w[1][sl] / rho, w[2][sl] / rho, w[3][sl] / rho
    cv0, cv1, t0 = params.cv_coeff[0], params.cv_coeff[1], params.t0
    tt = (w[4][sl] / rho - 0.5 * (u * u + v * v + ww * ww) - cv1 * t0) / cv0 + t0
    uu0 = u / params.u0
    rho2 = rho * (1.0 + (s["rho"] - 1.0) * f)
    t2 = tt * (1.0 + (s["t"] - 1.0) * f)
    u2 = u * (1.0 + (s["u"] / params.u0 - 1.0) * f)
    v2 = v + f * s["v"] * uu0
    e2 = cv1 * t0 + cv0 * (t2 - t0)
    out[0][sl] = rho2
    out[1][sl] = rho2 * u2
    out[2][sl] = rho2 * v2
    out[3][sl] = rho2 * ww
    out[4][sl] = rho2 * (e2 + 0.5 * (u2 * u2 + v2 * v2 + ww * ww))
    return out


def bl_recycling_case(nx=1024, ny=256, nz=192, lx=70.0, ly=12.0, lz=6.5, beta_y=3.0, mach=2.0, reynolds=200.0,
                      shock_angle_deg=None, xshock_imp=40.0, developed_shock=False, sensor_threshold=0.2,
                      conservative_viscous=1, rand_type=2, precision=0, noise=0.05, seed=0, x_recyc=20.0,
                      cfl=0.5):
    """Supersonic boundary layer over lx x ly x lz (inflow thickness delta = 1)
    with the boundary set of the reference's boundary-layer and SBLI decks:
    recycling-rescaling inflow at xmin (station x_recyc, random inflow plane
    per ``rand_type``), supersonic outflow with the non-reflecting treatment
    at xmax, isothermal wall at the recovery temperature, periodic span; the
    top is an extrapolated non-reflecting boundary, or, with
    ``shock_angle_deg``, the imposed oblique shock meeting the wall at
    ``xshock_imp``.  ``developed_shock`` starts from a field that the incident
    shock already crosses.  Initial field: the seeded compressible profile
    plus ``noise`` velocity fluctuations.  Returns (case, recycling)."""
    g0 = bl_grid(nx, ny, nz, lx=lx, ly=ly, lz=lz, beta_y=beta_y)
    gam, pr = 1.4, 0.72
    t_rec = 1.0 + 0.5 * (gam - 1.0) * pr ** (1.0 / 3.0) * mach ** 2
    if shock_angle_deg is None:
        tags, tags_nr, kw = (10, 2, 6, 2, 0, 0), (0, 1, 0, 1, 0, 0), {}
    else:
        tags, tags_nr = (10, 2, 6, 7, 0, 0), (0, 1, 0, 1, 0, 0)
        kw = dict(shock_angle_deg=shock_angle_deg, xshock_imp=xshock_imp, tanhfacs=lx / (nx - 1))
    p = bl_params(nx, ny, nz, mach=mach, reynolds=reynolds, t_wall=t_rec, bctags=tags, bctags_nr=tags_nr,
                  sensor_threshold=sensor_threshold, conservative_viscous=conservative_viscous, precision=precision,
                  rand_type=rand_type, cfl=cfl, **kw)
    w0 = bl_initial_field(g0, p, delta=1.0, noise=noise, seed=seed)
    c = ChannelCase(FlowCase(nx=nx, ny=ny, nz=nz), g0, p, w0)
    rc = recycling_setup(c.grid, x_recyc, 1.0, 0.01, 1.1, 0.011, bl_mean_planes(c))
    if shock_angle_deg is not None and developed_shock:
        c.w0 = impose_oblique_shock(g0, p, w0, shock_angle_deg, xshock_imp)
    return c, rc
