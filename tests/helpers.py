"""Shared helpers of the parity tests."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num


def interior(w, ng=3):
    return w[..., ng:-ng, ng:-ng, ng:-ng]


def fill_ghosts_wrap(w, ng=3):
    """Give every ghost point (edges and corners too) the periodic image of
    the interior, so no ghost holds a zero density."""
    out = w.copy()
    inner = w[:, ng:-ng, ng:-ng, ng:-ng]
    out[:] = np.pad(inner, ((0, 0), (ng, ng), (ng, ng), (ng, ng)), mode="wrap")
    return out


def box_case(nx=16, ny=12, nz=10, L=(2 * np.pi, 2 * np.pi, 2 * np.pi), mu0=1e-2, threshold=2.0,
             nkeep=0, weno_version=0, visc_model=1, ng=3, ep_order=6, weno_scheme=3):
    """Triply periodic box on a uniform grid (no walls, no forcing):
    StParams + Grid with exact uniform metrics (dcsidx2 = 0)."""
    from streams_mi355x.grid import Grid
    from streams_mi355x.params import FlowCase, StParams, set_fluid_prop

    case = FlowCase(nx=nx, ny=ny, nz=nz)
    gam, gm1, rfac, icl, icr, cp, cv = set_fluid_prop(case)
    p = StParams()
    p.nx, p.ny, p.nz, p.ng = nx, ny, nz, ng
    p.ep_order = p.visc_order = ep_order
    p.nkeep = nkeep
    p.weno_scheme = weno_scheme
    p.weno_version = weno_version
    p.visc_model = visc_model
    p.calorically_perfect = 1
    p.indx_cp_l = p.indx_cp_r = 0
    p.bctags[:] = [0] * 6
    p.bctags_nr[:] = [0] * 6
    p.is_y_staggered = 0
    p.channel_case = 0
    p.correct_bound_ord = 0
    p.iter_dt_recompute = 1
    p.rgas0 = p.t0 = p.rho0 = p.l0 = 1.0
    p.prandtl = 0.72
    p.u0 = 1.0
    p.mu0 = mu0
    p.sutherland_s, p.t_ref_dim, p.powerlaw_vtexp = 110.4, 160.0, 0.76
    p.sensor_threshold = threshold
    p.t_wall = 1.0
    p.theta_wall = -2.0
    p.volchan = 1.0
    p.cfl = 0.5
    p.tol_iter_nr = 1e-12
    for i, v in enumerate(cp):
        p.cp_coeff[i] = v
    for i, v in enumerate(cv):
        p.cv_coeff[i] = v
    d = [L[0] / nx, L[1] / ny, L[2] / nz]
    co = [(np.arange(-ng, n + ng) + 0.5) * h for n, h in zip((nx, ny, nz), d)]
    met = []
    for n, h in zip((nx, ny, nz), d):
        met += [np.full(n, 1.0 / h), np.full(n, 1.0 / h ** 2), np.zeros(n)]
    yn = (np.arange(ny + 1)) * d[1]
    g = Grid(nx, ny, nz, ng, co[0], co[1], co[2], yn, *met)
    return p, g, d, co


def conserved_from_primitive(rho, u, v, w, T, cv=2.5, rgas=1.0):
    """w(1:5) from rho, velocity and temperature (calorically perfect gas)."""
    E = cv * T + 0.5 * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, rho * E])
