"""Boundary conditions of the boundary-layer family on the GPU against the
oracle: the physical ghost faces of update_ghost, the characteristic
treatment bc_nr, and short trajectories of a boundary-layer box."""
import numpy as np
import pytest

from helpers import interior, rel_l2
from oracle import Oracle
from streams_mi355x.cases import make_bl_box
from streams_mi355x.solver import ChannelSolver

pytestmark = pytest.mark.gpu

NG = 3
TRAJ_TOL = 1e-10

BC_SETS = {
    "inflow-outflow-wall-nr": dict(bctags=(1, 2, 6, 2, 0, 0), bctags_nr=(0, 1, 0, 1, 0, 0)),
    "shock-top": dict(bctags=(1, 2, 6, 7, 0, 0), bctags_nr=(0, 1, 0, 0, 0, 0), shock_angle_deg=35.0,
                      xshock_imp=5.0, tanhfacs=8.0 / 39),
    "symmetry-relaxation": dict(bctags=(1, 2, 5, 1, 0, 0), bctags_nr=(3, 2, 0, 3, 0, 0)),
    "reflective-wall": dict(bctags=(2, 2, 6, 2, 0, 0), bctags_nr=(1, 1, 6, 1, 0, 0)),
    "subsonic-total-conditions": dict(bctags=(3, 4, 6, 3, 0, 0), bctags_nr=(0, 0, 0, 0, 0, 0), mach=0.5,
                                      aoa_deg=2.0),
    "extrapolation-p0-no-op-faces": dict(bctags=(4, 4, 6, 4, 0, 0), bctags_nr=(0, 1, 0, 0, 0, 0)),
}


def _pair(name, **kw):
    c = make_bl_box(**BC_SETS[name], **kw)
    return c, Oracle(c.params, c.grid), ChannelSolver(c.params, c.grid)


@pytest.mark.parametrize("name", list(BC_SETS))
def test_update_ghost_parity(name):
    c, o, g = _pair(name)
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    o.compute_aux(1, 1)
    g.compute_aux(1, 1)
    o.update_ghost(1)
    g.update_ghost(1)
    wg = g.copy_to_field()
    # every point the boundary routines write: faces (no edges or corners)
    nz, ny, nx = c.params.nz, c.params.ny, c.params.nx
    for m in range(5):
        a, b = wg[m], o.w[m]
        assert rel_l2(a[NG:-NG, NG:-NG, :], b[NG:-NG, NG:-NG, :]) <= 1e-14, (name, m)
        assert rel_l2(a[NG:-NG, :, NG:-NG], b[NG:-NG, :, NG:-NG]) <= 1e-14, (name, m)
        assert rel_l2(a[:, NG:-NG, NG:-NG], b[:, NG:-NG, NG:-NG]) <= 1e-14, (name, m)
    g.close()


@pytest.mark.parametrize("name", ["inflow-outflow-wall-nr", "symmetry-relaxation", "reflective-wall"])
def test_bc_nr_parity(name):
    c, o, g = _pair(name)
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    o.update_ghost(1)
    g.update_ghost(1)
    o.compute_aux(1, 1)
    g.compute_aux(1, 1)
    o.zero_flux()
    g.zero_flux()
    o.bc_nr()
    g.bc_nr()
    fo, fg = o.fl, g.get_fl()
    assert np.abs(fo).max() > 0.0
    for m in range(5):
        assert rel_l2(fg[m], fo[m]) <= 1e-12, (name, m)
    g.close()


@pytest.mark.parametrize("name", list(BC_SETS))
def test_bl_trajectory_20_steps(name):
    c, o, g = _pair(name)
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    o.init_run()
    g.init_run()
    o.advance(20)
    g.advance(20)
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    print(name, "20-step rel L2 per variable:", errs)
    assert max(errs) <= TRAJ_TOL, (name, errs)
    so, sg = o.scalars(), g.scalars()
    assert abs(sg["dt"] - so["dt"]) <= 1e-12 * so["dt"]
    g.close()


@pytest.mark.parametrize("name", ["inflow-outflow-wall-nr", "shock-top"])
def test_bl_trajectory_laplacian_viscous_form(name):
    """The Laplacian viscous form (visflx modes 4 + 1) outside the channel:
    the viscous divergence rides on the z march without force_rhs partials."""
    c, o, g = _pair(name, conservative_viscous=0)
    assert c.params.conservative_viscous == 0 and c.params.channel_case == 0
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    o.init_run()
    g.init_run()
    o.advance(20)
    g.advance(20)
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    assert max(errs) <= TRAJ_TOL, (name, errs)
    g.close()


def _distinct_planes(c):
    """The inflow mean profile with its ng planes scaled apart (plane
    1-ng+q by 1 + 0.01 (q+1)), so the plane order of wmean shows."""
    from streams_mi355x.cases import bl_mean_planes

    wm = bl_mean_planes(c).copy()
    for q in range(NG):
        wm[:, :, q] *= 1.0 + 0.01 * (q + 1)
    return wm


def _recycling_pair(seed=5, rand_type=0):
    from streams_mi355x.cases import recycling_setup

    c = make_bl_box(bctags=(10, 2, 6, 2, 0, 0), bctags_nr=(0, 1, 0, 1, 0, 0), seed=seed, rand_type=rand_type)
    r = recycling_setup(c.grid, 5.0, 0.6, 0.01, 0.7, 0.011, _distinct_planes(c))
    w = c.w0.copy()
    w *= 1.0 + 0.04 * (np.random.default_rng(seed).random(w.shape[1:]) - 0.5)
    o, g = Oracle(c.params, c.grid), ChannelSolver(c.params, c.grid)
    o.w[:] = w
    g.copy_from_field(w)
    o.set_recycling(r)
    g.set_recycling(r)
    return c, o, g


def test_recycling_inflow_parity():
    c, o, g = _recycling_pair()
    o.compute_aux(1, 1)
    g.compute_aux(1, 1)
    o.update_ghost(1)
    g.update_ghost(1)
    wg = g.copy_to_field()
    for m in range(5):    # the inflow ghost planes
        assert rel_l2(wg[m][NG:-NG, NG:-NG, :NG], o.w[m][NG:-NG, NG:-NG, :NG]) <= 1e-13, m
    g.close()


@pytest.mark.parametrize("rand_type", [0, 2])
def test_recycling_trajectory_20_steps(rand_type):
    """20 steps of a recycling box against the oracle; with rand_type 2 the
    random inflow plane is redrawn every stage on the device and must
    follow the C library's rand() stream the oracle calls."""
    c, o, g = _recycling_pair(rand_type=rand_type)
    o.init_run()
    g.init_run()
    o.advance(20)
    g.advance(20)
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    print("recycling 20-step rel L2 per variable (rand_type %d):" % rand_type, errs)
    assert max(errs) <= TRAJ_TOL, errs
    # the inflow ghosts carry the random term
    for m in range(5):
        assert rel_l2(wg[m][NG:-NG, NG:-NG, :NG], o.w[m][NG:-NG, NG:-NG, :NG]) <= TRAJ_TOL, m
    g.close()


def test_random_plane_changes_the_inflow():
    """rand_type 2 against rand_type 0 from the same state: the inflow
    velocity differs by at most 1 % of u0 after one step."""
    out = []
    for rt in (0, 2):
        c, o, g = _recycling_pair(rand_type=rt)
        g.init_run()
        g.advance(1)
        out.append(g.copy_to_field())
        g.close()
    du = out[1][1:4, NG:-NG, NG:-NG, :NG] / out[1][0, NG:-NG, NG:-NG, :NG] - \
        out[0][1:4, NG:-NG, NG:-NG, :NG] / out[0][0, NG:-NG, NG:-NG, :NG]
    assert 0.0 < np.abs(du).max() <= 0.011 * c.params.u0


def test_laminar_inflow_from_wmean():
    """Tag 9 on xmin: the ghost planes of the laminar inflow from wmean."""
    c = make_bl_box(bctags=(9, 2, 6, 2, 0, 0), bctags_nr=(0, 1, 0, 1, 0, 0))
    o, g = Oracle(c.params, c.grid), ChannelSolver(c.params, c.grid)
    wm = _distinct_planes(c)
    o.set_wmean(wm)
    g.set_wmean(wm)
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    o.init_run()
    g.init_run()
    wg = g.copy_to_field()
    for m in range(5):
        assert rel_l2(wg[m][NG:-NG, NG:-NG, :NG], o.w[m][NG:-NG, NG:-NG, :NG]) <= 1e-14, m
    o.advance(10)
    g.advance(10)
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    assert max(errs) <= TRAJ_TOL, errs
    g.close()


def _c4_like(precision=0, seed=7):
    """The boundary set of the SBLI deck (configs C3/C4): recycling inflow,
    extrapolated outflow with the characteristic treatment, isothermal wall,
    imposed oblique shock on top, periodic span."""
    from streams_mi355x.cases import bl_mean_planes, recycling_setup

    c = make_bl_box(nx=48, ny=32, nz=16, bctags=(10, 2, 6, 7, 0, 0), bctags_nr=(0, 1, 0, 0, 0, 0),
                    shock_angle_deg=35.0, xshock_imp=5.0, tanhfacs=8.0 / 47, seed=seed, precision=precision)
    r = recycling_setup(c.grid, 5.0, 0.6, 0.01, 0.7, 0.011, bl_mean_planes(c))
    return c, r


def test_sbli_box_trajectory_20_steps():
    c, r = _c4_like()
    o, g = Oracle(c.params, c.grid), ChannelSolver(c.params, c.grid)
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    o.set_recycling(r)
    g.set_recycling(r)
    o.init_run()
    g.init_run()
    o.advance(20)
    g.advance(20)
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    assert max(errs) <= TRAJ_TOL, errs
    g.close()


def test_sbli_box_mixed_form():
    from streams_mi355x import PREC_MIXED

    out = []
    for prec in (0, PREC_MIXED):
        c, r = _c4_like(prec)
        with ChannelSolver(c.params, c.grid) as g:
            g.copy_from_field(c.w0)
            g.set_recycling(r)
            g.init_run()
            g.advance(20)
            out.append(g.copy_to_field())
    errs = [rel_l2(interior(out[1])[m], interior(out[0])[m]) for m in range(5)]
    assert max(errs) <= 1e-5, errs
