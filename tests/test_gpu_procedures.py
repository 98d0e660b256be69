"""The reference's stage procedure by procedure through the C ABI: the
range arguments of euler_x/y/z (singleideal_cpu.F90:674-843), init_flux /
update_flux / update_field, force_rhs, force_var, bcextr_var, bcswap_var and
recyc_exchange, against the oracle's rk_sync."""
import numpy as np
import pytest

from helpers import interior, rel_l2
from oracle import Oracle, lib as olib
from streams_mi355x import make_channel
from streams_mi355x.cases import make_bl_box
from streams_mi355x.solver import ChannelSolver

pytestmark = pytest.mark.gpu

NG = 3
RHORK = (0.0, -17.0 / 60.0, -5.0 / 12.0)
GAMRK = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)


def step_by_procedure(g, p, icyc0=0, split_x=True):
    """One step of run(): dt at the step start, then rk_sync as the
    reference calls its procedures (singleideal_cpu.F90:352-406), euler_x in
    two ranges whose face fluxes the second call's update combines."""
    sc = g.scalars()
    idt = max(p.iter_dt_recompute, 1)
    dt = g.compute_dt() if p.cfl > 0 and (sc["icyc"] + 1 - icyc0) % idt == 0 else sc["dt"]
    if p.channel_case:
        g.reset_forcing()
    rg = g.eul_range()
    imid = (rg[0] + rg[1]) // 2
    for istep in range(3):
        g.init_flux(RHORK[istep] * dt)
        g.bcswap()
        g.compute_aux(1, 1)
        if split_x:
            g.euler_x(rg[0], imid, rg[0] - 1, imid, False)
            g.euler_x(imid + 1, rg[1], imid, rg[1], True)
        else:
            g.euler_x()
        if p.conservative_viscous == 1:
            g.visflx(6)
            g.visflx(5)
        else:
            g.visflx(4)
        g.bcextr_var(10)
        g.bcswap_var(10)
        g.euler_y(rg[2], rg[3])
        g.euler_z(rg[4], rg[5])
        if istep == 2:
            g.visflx(3)
            g.bcswap_var(8)
        g.visflx(7 if p.conservative_viscous == 1 else 1)
        g.bc_nr()
        g.update_flux(GAMRK[istep] * dt)
        if p.channel_case:
            g.force_rhs()
        g.update_field()
        if p.channel_case:
            g.force_var()
        g.update_ghost(0)
    g.set_time(sc["time"] + dt, dt, sc["icyc"] + 1)


def _run_pair(c, nsteps, recycling=None, split_x=True):
    p = c.params
    o = Oracle(p, c.grid)
    g = ChannelSolver(p, c.grid)
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    if recycling is not None:
        o.set_recycling(recycling)
        g.set_recycling(recycling)
    o.init_run()
    g.init_run()
    for _ in range(nsteps):
        step_by_procedure(g, p, split_x=split_x)
    o.advance(nsteps)
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    return o, g, errs


@pytest.mark.parametrize("theta_wall", [-2.0, 1.0])
def test_channel_by_procedure_matches_oracle(theta_wall):
    """Bulk forcing (force_rhs) with and without bulk-temperature control
    (force_var), WENO faces present, 6 steps."""
    c = make_channel(nx=32, ny=24, nz=16, sensor_threshold=0.3, theta_wall=theta_wall)
    o, g, errs = _run_pair(c, 6)
    print("channel by procedure, theta_wall", theta_wall, errs)
    assert max(errs) <= 1e-10, errs
    so, sg = o.scalars(), g.scalars()
    for key in ("dpdx", "rhobulk", "ubulk", "tbulk", "time"):
        assert abs(sg[key] - so[key]) <= 1e-10 * max(abs(so[key]), 1e-30), (key, sg[key], so[key])
    res = g.compute_residual()
    o.compute_residual()
    assert abs(res["residual_rhou"] - o.scalars()["residual_rhou"]) <= 1e-8 * o.scalars()["residual_rhou"]
    g.close()


def test_boundary_layer_by_procedure_with_random_inflow():
    """Recycling inflow with the random plane (recyc_exchange inside
    update_ghost, draws in euler_z), characteristic outflow and top, the
    conservative viscous form, 6 steps."""
    from streams_mi355x.cases import bl_mean_planes, recycling_setup

    c = make_bl_box(bctags=(10, 2, 6, 2, 0, 0), bctags_nr=(0, 1, 0, 1, 0, 0), rand_type=2)
    r = recycling_setup(c.grid, 5.0, 0.6, 0.01, 0.7, 0.011, bl_mean_planes(c))
    o, g, errs = _run_pair(c, 6, recycling=r)
    print("boundary layer by procedure:", errs)
    assert max(errs) <= 1e-10, errs
    g.close()


@pytest.mark.parametrize("d", [0, 1, 2])
def test_sweep_ranges_match_oracle(d):
    """One sweep over a range other than bc_preproc's, through the face
    kernels, against the oracle's flux + update over the same range."""
    c = make_channel(nx=24, ny=20, nz=16, sensor_threshold=0.2)
    p = c.params
    o = Oracle(p, c.grid)
    g = ChannelSolver(p, c.grid)
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    for s in (o, g):
        s.update_ghost(1)
        s.compute_aux(1, 1)
    o.visflx(0)
    g.visflx(0)  # the Ducros sensor: some faces take WENO
    o.bcswap_var(8)
    g.bcswap_var(8)
    n = (p.nx, p.ny, p.nz)[d]
    lo, hi = 3, n - 4
    o.zero_flux()
    g.zero_flux()
    o.fhat[:] = 0.0  # the device face fluxes start at zero too
    olib().or_euler_fluxes(o.h, d, lo - 1, hi)
    if d == 0:
        # euler_x updates eul_imin..eul_imax whatever its face range: the
        # faces outside lo-1..hi keep their (zero) fluxes
        olib().or_euler_update(o.h, d, 1, p.nx)
        g.euler_x(lo, hi, lo - 1, hi, True)
    else:
        olib().or_euler_update(o.h, d, lo, hi)
        if d == 1:
            g.euler_y(lo, hi)
        else:
            g.euler_z(lo, hi)
    fg = g.get_fl()
    for m in range(5):
        assert rel_l2(fg[m], o.fl[m]) <= 1e-12, (d, m)
    g.close()


def test_split_euler_x_equals_the_fused_sweep():
    """euler_x over two face ranges with one update equals the fused sweep
    of the default range (round-off: the pair sums are grouped differently)."""
    c = make_channel(nx=32, ny=24, nz=16, sensor_threshold=0.3)
    g = ChannelSolver(c.params, c.grid)
    g.copy_from_field(c.w0)
    g.update_ghost(1)
    g.compute_aux(1, 1)
    g.visflx(0)
    g.bcswap_var(8)
    g.zero_flux()
    g.euler_x()
    fused = g.get_fl()
    g.zero_flux()
    g.euler_x(1, 16, 0, 16, False)
    g.euler_x(17, 32, 16, 32, True)
    split = g.get_fl()
    for m in range(5):
        assert rel_l2(split[m], fused[m]) <= 1e-13, m
    g.close()
