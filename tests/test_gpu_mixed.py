"""The mixed-precision form (ST_PREC_MIXED) against the FP64 oracle.

The mixed form keeps w (conserved variables) in FP64, recovers T and p from
E in FP64, runs the WENO branch, the RK accumulation and every reduction in
FP64, and stores w_aux and the right-hand sides in FP32, with the central
fluxes and viscous stresses evaluated in FP32.  Its bar: relative L2 error
per conserved variable <= 1e-5 after 100 steps against the FP64 oracle.
"""
import numpy as np
import pytest

from helpers import interior, rel_l2
from oracle import Oracle
from streams_mi355x import PREC_MIXED, make_channel, make_subsonic_channel
from streams_mi355x.params import params_copy
from streams_mi355x.solver import ChannelSolver

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-5


def _mixed_vs_oracle(c, nsteps):
    pm = params_copy(c.params)
    pm.precision = PREC_MIXED
    o = Oracle(c.params, c.grid)
    o.w[:] = c.w0
    o.init_run()
    o.advance(nsteps)
    with ChannelSolver(pm, c.grid) as g:
        g.copy_from_field(c.w0)
        g.init_run()
        g.advance(nsteps)
        wg = g.copy_to_field()
        sg = g.scalars()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    return errs, sg, o.scalars()


def test_mixed_aux_storage():
    c = make_channel()
    pm = params_copy(c.params)
    pm.precision = PREC_MIXED
    o = Oracle(c.params, c.grid)
    o.w[:] = c.w0
    o.update_ghost(1)
    o.compute_aux(1, 1)
    with ChannelSolver(pm, c.grid) as g:
        g.copy_from_field(c.w0)
        g.update_ghost(1)
        g.compute_aux(1, 1)
        for v in range(1, 8):
            a = interior(g.get_aux(v)[None])[0]
            b = interior(o.w_aux[v - 1][None])[0]
            assert rel_l2(a, b) <= 1e-7, v  # one FP32 rounding of an FP64 value


def test_mixed_c1_100_steps():
    """128x96x64 channel, M_b 1.5, Re_tau 220, sensor 0.9, 100 steps."""
    c = make_channel(nx=128, ny=96, nz=64)
    errs, sg, so = _mixed_vs_oracle(c, 100)
    print("mixed C1 100-step rel L2 per variable:", errs)
    assert max(errs) <= MIXED_TOL, errs
    assert abs(sg["time"] - so["time"]) <= 1e-6 * so["time"]


def test_mixed_subsonic_channel_100_steps():
    """The golden-row deck (M 0.2, conservative viscous form), 100 steps."""
    c = make_subsonic_channel()
    errs, sg, so = _mixed_vs_oracle(c, 100)
    print("mixed subsonic channel 100-step rel L2 per variable:", errs)
    assert max(errs) <= MIXED_TOL, errs


def test_mixed_all_weno_20_steps():
    c = make_channel(sensor_threshold=-1.0)
    errs, _, _ = _mixed_vs_oracle(c, 20)
    print("mixed all-WENO 20-step rel L2 per variable:", errs)
    assert max(errs) <= MIXED_TOL, errs


def test_fp64_unchanged_by_the_mixed_build():
    """The FP64 form stays bitwise reproducible and oracle-close."""
    c = make_channel()
    outs = []
    for _ in range(2):
        with ChannelSolver(c.params, c.grid) as g:
            g.copy_from_field(c.w0)
            g.init_run()
            g.advance(5)
            outs.append(g.copy_to_field())
    assert np.array_equal(interior(outs[0]), interior(outs[1]))


def test_mixed_switch_decisions_follow_fp64():
    """With the shock sensor active on many faces, the mixed form evaluates
    the sensor in FP64 from w: its central/WENO decisions are the FP64
    form's, so the WENO face counts agree and the error stays within the
    tolerance."""
    c = make_channel(nx=128, ny=96, nz=64, seed=2, sensor_threshold=0.3)
    pm = params_copy(c.params)
    pm.precision = PREC_MIXED
    out = []
    for p in (c.params, pm):
        with ChannelSolver(p, c.grid) as g:
            g.copy_from_field(c.w0)
            g.init_run()
            g.advance(40)
            out.append((g.copy_to_field(), g.count_weno()))
    (wd, cd), (wm, cm) = out
    assert cd.max() > 1e-3                     # the switch is in play
    assert np.allclose(cm, cd, rtol=1e-3, atol=0.0), (cm, cd)
    errs = [rel_l2(interior(wm)[m], interior(wd)[m]) for m in range(5)]
    assert max(errs) <= MIXED_TOL, errs
