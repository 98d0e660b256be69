"""Scalar parameters of the hot path and the setup that produces them.

The ctypes struct mirrors ``st_params`` of include/streams_mi355x.h.  The
setup functions restate the parts of the reference equation module that turn
an input deck into those scalars:

* ``set_fluid_prop``  -- code/src/singleideal/singleideal.F90:2685-2774
* ``set_flow_params`` -- code/src/singleideal/singleideal.F90:2776-2856
* ``set_chan_prop``   -- code/src/singleideal/singleideal.F90:2894-2952
* ``get_reynolds_cha``/``velpiros_channel`` -- singleideal.F90:3052-3173, 3699-3725

Reference values are non-dimensional: rgas0 = l0 = t0 = rho0 = 1
(singleideal.F90:326-329).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

MAX_CP = 16
PREC_FP64 = 0
PREC_MIXED = 1


class StParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("ng", C.c_int32),
        ("ep_order", C.c_int32), ("visc_order", C.c_int32), ("nkeep", C.c_int32),
        ("weno_scheme", C.c_int32), ("weno_version", C.c_int32), ("flux_splitting", C.c_int32),
        ("conservative_viscous", C.c_int32), ("visc_model", C.c_int32),
        ("calorically_perfect", C.c_int32), ("indx_cp_l", C.c_int32), ("indx_cp_r", C.c_int32),
        ("bctags", C.c_int32 * 6), ("bctags_nr", C.c_int32 * 6),
        ("is_y_staggered", C.c_int32), ("channel_case", C.c_int32),
        ("correct_bound_ord", C.c_int32), ("iter_dt_recompute", C.c_int32),
        ("precision", C.c_int32), ("rand_type", C.c_int32), ("pad1", C.c_int32),
        ("rgas0", C.c_double), ("prandtl", C.c_double), ("t0", C.c_double), ("rho0", C.c_double),
        ("u0", C.c_double), ("l0", C.c_double), ("mu0", C.c_double),
        ("sutherland_s", C.c_double), ("t_ref_dim", C.c_double), ("powerlaw_vtexp", C.c_double),
        ("sensor_threshold", C.c_double), ("t_wall", C.c_double), ("theta_wall", C.c_double),
        ("t_bulk_target", C.c_double), ("volchan", C.c_double), ("cfl", C.c_double),
        ("tol_iter_nr", C.c_double),
        ("cp_coeff", C.c_double * MAX_CP), ("cv_coeff", C.c_double * MAX_CP),
        ("winf", C.c_double * 5), ("winf_past_shock", C.c_double * 5),
        ("xshock_imp", C.c_double), ("shock_angle", C.c_double), ("tanhfacs", C.c_double),
        ("aoa", C.c_double), ("ptot0", C.c_double), ("ttot0", C.c_double),
    ]


class StScalars(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("time", C.c_double), ("dpdx", C.c_double),
        ("rhobulk", C.c_double), ("ubulk", C.c_double), ("tbulk", C.c_double),
        ("residual_rhou", C.c_double),
        ("rhomin", C.c_double), ("rhomax", C.c_double), ("tmin", C.c_double),
        ("tmax", C.c_double), ("pmin", C.c_double), ("pmax", C.c_double),
        ("icyc", C.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


@dataclass
class FlowCase:
    """Input-deck values of one flow case (subset of singleideal.ini)."""
    nx: int
    ny: int
    nz: int
    ng: int = 3
    lx: float = 4.0 * math.pi
    ly: float = 2.0
    lz: float = 4.0 * math.pi / 3.0
    mach: float = 1.5
    reynolds: float = 220.0          # friction Reynolds number for the channel
    theta_wall: float = -2.0         # < -1: bulk temperature free (Mach based on T_wall)
    gam: float = 1.4
    prandtl: float = 0.72
    visc_model: int = 1              # 1 power law, 2 Sutherland
    calorically_perfect: int = 1     # 0: cp(T) polynomial (indx_cp_l..indx_cp_r, cp_coeff)
    indx_cp_l: int = 0
    indx_cp_r: int = 0
    cp_coeff: tuple = ()             # cp/R coefficients of T_dim^l, l = indx_cp_l..indx_cp_r, then h*/R
    flux_splitting: int = 0          # WENO faces: 0 Roe characteristic, 1 Rusanov
    conservative_viscous: int = 0    # 1: second-order conservative viscous form
    vt_exp: float = 0.76
    s_suth: float = 110.4
    t_ref: float = 160.0
    ep_order: int = 6
    nkeep: int = 0
    weno_scheme: int = 3
    weno_version: int = 0
    sensor_threshold: float = 0.9
    cfl: float = 0.8
    iter_dt_recompute: int = 1
    bctags: tuple = (0, 0, 6, 6, 0, 0)
    bctags_nr: tuple = (0, 0, 0, 0, 0, 0)
    correct_bound_ord: int = 1
    dyptarget: float = 0.8
    jbgrid: float = 32.0
    ysmoosteps: int = 50
    precision: int = PREC_FP64
    extra: dict = field(default_factory=dict)


def set_fluid_prop(case: FlowCase):
    """set_fluid_prop (singleideal.F90:2685-2774), both branches.

    Calorically perfect: cv = R/(gam-1), cp = gam cv.  Thermally perfect
    (calorically_perfect = 0): cp/R is the polynomial sum_l a_l T_dim^l,
    l = indx_cp_l..indx_cp_r, with h*/R as the last coefficient; the
    coefficients are made non-dimensional with T_ref, gamma is that of
    T_ref, and cv_l = cp_l except cv_0 = cp_0 - 1.  Returns (gam, gm1, rfac,
    indx_cp_l, indx_cp_r, cp_coeff, cv_coeff) with the coefficient arrays
    indexed l - indx_cp_l for l = indx_cp_l..indx_cp_r+1.
    """
    rgas0 = 1.0
    gam = case.gam
    rfac = case.prandtl ** (1.0 / 3.0)
    if case.calorically_perfect == 1:
        indx_cp_l = indx_cp_r = 0
    else:
        indx_cp_l, indx_cp_r = case.indx_cp_l, case.indx_cp_r
    n = indx_cp_r - indx_cp_l + 2
    cv = np.zeros(n)
    cp = np.zeros(n)
    if case.calorically_perfect == 0:
        if len(case.cp_coeff) != n:
            raise ValueError(f"Error! Check number of cp coefficients ({len(case.cp_coeff)} given, {n} needed)")
        cp[:] = case.cp_coeff
        tref = case.t_ref
        hstar = cp[n - 1]
        for l in range(indx_cp_l, indx_cp_r + 1):
            a = cp[l - indx_cp_l]
            hstar = hstar + (a * math.log(tref) if l == -1 else a / (l + 1) * tref ** (l + 1))
        bcoeff = hstar / tref
        cp_tref = sum(cp[l - indx_cp_l] * tref ** l for l in range(indx_cp_l, indx_cp_r + 1)) * rgas0
        cv_tref = cp_tref - rgas0
        gam = cp_tref / cv_tref
        for l in range(indx_cp_l, indx_cp_r + 1):
            cp[l - indx_cp_l] = cp[l - indx_cp_l] * tref ** l
            cv[l - indx_cp_l] = cp[l - indx_cp_l]
        cv[0 - indx_cp_l] = cp[0 - indx_cp_l] - 1.0
    gm1 = gam - 1.0
    gm = 1.0 / gm1
    if case.calorically_perfect == 1:
        cv[0] = gm
        cp[0] = gm * gam
        bcoeff = cv[0] + 1.0
    cp[n - 1] = bcoeff
    cv[n - 1] = bcoeff - 1.0
    cv *= rgas0
    cp *= rgas0
    return gam, gm1, rfac, indx_cp_l, indx_cp_r, cp, cv


def velpiros_channel(yplus: float, retau: float) -> float:
    """Composite channel velocity profile, singleideal.F90:3699-3725."""
    eta = min(1.0, yplus / retau)
    yp = eta * retau
    c = 5.3
    vkc = 0.39
    vkci = 1.0 / vkc
    c1 = -1.0 / vkc * math.log(vkc) + c
    etastar = 0.275
    b_rich = 0.33
    eta_rich = 11.0
    uinn = vkci * math.log(1.0 + vkc * yp) + c1 * (1.0 - math.exp(-yp / eta_rich) - yp / eta_rich * math.exp(-b_rich * yp))
    uout = c + vkci * math.log(retau) + vkci * math.log(etastar)
    uout = uout + (1 - etastar) / (2 * vkc * etastar) * (1 - (1 - eta) ** 2 / (1 - etastar) ** 2)
    return uinn if eta <= etastar else uout


def get_reynolds_cha(retau, rm, trat, s2tinf, vtexp, visc_exp, gam, rfac, ny, yn):
    """Bulk Reynolds number from Re_tau, singleideal.F90:3052-3173 (itransf = 0)."""
    tol_iter = 1e-9
    gm1h = 0.5 * (gam - 1.0)
    tr = 1.0 + gm1h * rfac * rm ** 2
    tw = trat * tr
    n = ny + 1
    ycplus = np.zeros(n)
    for j in range(1, n):
        ycplus[j] = (yn[j] - yn[0]) / (yn[ny] - yn[0]) * retau
    yplus = ycplus.copy()
    yplusold = yplus.copy()
    for _outer in range(200):
        uplus = np.zeros(n)
        for j in range(1, n):
            uplus[j] = velpiros_channel(yplus[j], retau)
        ucplus = uplus.copy()
        for _inner in range(200):
            ucold = ucplus.copy()
            uu = ucplus / ucplus[ny - 1]
            temperature = tw + (tr - tw) * uu + (1.0 - tr) * uu ** 2
            density = 1.0 / temperature
            if visc_exp:
                viscosity = temperature ** vtexp
            else:
                viscosity = temperature ** 1.5 * (1.0 + s2tinf) / (temperature + s2tinf)
            for j in range(1, n):
                du = uplus[j] - uplus[j - 1]
                ucij = math.sqrt(density[0] / density[j]) * math.sqrt(viscosity[j] / viscosity[0])
                ucijm = math.sqrt(density[0] / density[j - 1]) * math.sqrt(viscosity[j - 1] / viscosity[0])
                ucplus[j] = ucplus[j - 1] + 0.5 * (ucij + ucijm) * du
            res = np.sum(np.abs(ucplus - ucold)) / n
            if res < tol_iter:
                break
        yplus = np.zeros(n)
        for j in range(1, n):
            dy = ycplus[j] - ycplus[j - 1]
            ycij = 1.0 / (math.sqrt(density[0] / density[j]) * math.sqrt((viscosity[j] / viscosity[0]) ** 3))
            ycijm = 1.0 / (math.sqrt(density[0] / density[j - 1]) * math.sqrt((viscosity[j - 1] / viscosity[0]) ** 3))
            yplus[j] = yplus[j - 1] + 0.5 * (ycij + ycijm) * dy
        res = abs(yplus[n - 1] - yplusold[n - 1]) / n
        if res < tol_iter:
            break
        yplusold = yplus.copy()
    rhosum = rhousum = ysum = 0.0
    for j in range(ny):
        dy = yn[j + 1] - yn[j]
        ysum += dy
        rhosum += 0.5 * (density[j] + density[j + 1]) * dy
        rhousum += 0.5 * (density[j] * ucplus[j] + density[j + 1] * ucplus[j + 1]) * dy
    rhosum /= ysum
    rhousum /= ysum
    ubulk = rhousum / rhosum
    return retau * ubulk * rhosum / density[0]


def channel_params(case: FlowCase, yn: np.ndarray, nxmax=None, nzmax=None) -> StParams:
    """Fill st_params for the isothermal-wall channel (flow_init = 0).

    Follows set_fluid_prop, set_flow_params and set_chan_prop with
    rgas0 = t0 = rho0 = l0 = 1.  ``yn`` are the global y faces yn(1:ny+1).
    """
    nxmax = case.nx if nxmax is None else nxmax
    nzmax = case.nz if nzmax is None else nzmax
    rgas0 = t0 = rho0 = l0 = 1.0
    gam, gm1, rfac, icl, icr, cp, cv = set_fluid_prop(case)
    mach = case.mach
    theta_wall = case.theta_wall
    t_wall = t0
    t_bulk_target = 0.0
    if theta_wall >= -1.0:
        # constant bulk temperature: Mach given on T_bulk, converted to T_wall
        rmb = mach
        trec_over_tb = 1.0 + gm1 / 2.0 * rfac * rmb ** 2
        tw_over_tb = 1.0 + theta_wall * (trec_over_tb - 1.0)
        tb_over_tw = 1.0 / tw_over_tb
        tw_over_tr = 1.0 / (trec_over_tb * tb_over_tw)
        t_bulk_target = tb_over_tw * t_wall
        rmtw = rmb * math.sqrt(tb_over_tw)
        mach = rmtw
    else:
        thtmp = -1.0
        rmtw = mach
        rmb = math.sqrt(rmtw ** 2 / (1.0 - gm1 / 2.0 * rfac * thtmp * rmtw ** 2))
        trec_over_tb = 1.0 + gm1 / 2.0 * rfac * rmb ** 2
        tw_over_tb = 1.0 + thtmp * (trec_over_tb - 1.0)
        tb_over_tw = 1.0 / tw_over_tb
        tw_over_tr = 1.0 / (trec_over_tb * tb_over_tw)
    s2tinf = case.s_suth / case.t_ref
    ny = len(yn) - 1
    re_bulk = get_reynolds_cha(case.reynolds, rmb, tw_over_tr, s2tinf, case.vt_exp,
                               case.visc_model == 1, gam, rfac, ny // 2, yn[: ny // 2 + 1])
    c0 = math.sqrt(gam * rgas0 * t0)
    u0 = mach * c0
    mu0 = u0 * rho0 * l0 / re_bulk
    volchan = (yn[ny] - yn[0]) * nxmax * nzmax

    p = StParams()
    p.nx, p.ny, p.nz, p.ng = case.nx, case.ny, case.nz, case.ng
    p.ep_order = case.ep_order
    p.visc_order = case.ep_order
    p.nkeep = case.nkeep
    p.weno_scheme = case.weno_scheme
    p.weno_version = case.weno_version
    p.flux_splitting = case.flux_splitting
    p.conservative_viscous = case.conservative_viscous
    p.visc_model = case.visc_model
    p.calorically_perfect = case.calorically_perfect
    p.indx_cp_l, p.indx_cp_r = icl, icr
    p.bctags[:] = list(case.bctags)
    p.bctags_nr[:] = list(case.bctags_nr)
    p.is_y_staggered = 1
    p.channel_case = 1
    p.correct_bound_ord = case.correct_bound_ord
    p.iter_dt_recompute = case.iter_dt_recompute
    p.precision = case.precision
    p.rgas0, p.prandtl, p.t0, p.rho0, p.u0, p.l0, p.mu0 = rgas0, case.prandtl, t0, rho0, u0, l0, mu0
    p.sutherland_s, p.t_ref_dim, p.powerlaw_vtexp = case.s_suth, case.t_ref, case.vt_exp
    p.sensor_threshold = case.sensor_threshold
    p.t_wall = t_wall
    p.theta_wall = theta_wall
    p.t_bulk_target = t_bulk_target
    p.volchan = volchan
    p.cfl = case.cfl
    p.tol_iter_nr = 1e-12
    for i, v in enumerate(cp):
        p.cp_coeff[i] = v
    for i, v in enumerate(cv):
        p.cv_coeff[i] = v
    return p


def params_copy(p: StParams) -> StParams:
    q = StParams()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(StParams))
    return q
