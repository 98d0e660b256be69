"""Four ghost layers on the GPU: WENO7 (weno_scheme 4) and the
eighth-order central and viscous schemes (ep_order = visc_order 8), run
procedure by procedure on the per-face kernels (the fused kernels cover
three ghost layers).  Each case follows the oracle for 20 steps to 1e-10
relative L2 per conserved variable, with the same WENO face counts.
"""
import numpy as np
import pytest

from helpers import rel_l2
from oracle import Oracle
from streams_mi355x import make_channel
from streams_mi355x.solver import ChannelSolver

pytestmark = pytest.mark.gpu

NG = 4
TOL = 1e-10


def _interior(w):
    return w[:, NG:-NG, NG:-NG, NG:-NG]


def _pair(c, steps=20):
    o = Oracle(c.params, c.grid)
    o.w[:] = c.w0
    o.init_run()
    o.advance(steps)
    with ChannelSolver(c.params, c.grid) as g:
        g.copy_from_field(c.w0)
        g.init_run()
        g.advance(steps)
        weno = g.count_weno()
        sc = g.compute_residual()
        w = g.copy_to_field()
    errs = [rel_l2(_interior(w)[m], _interior(o.w)[m]) for m in range(5)]
    return errs, weno, o.count_weno(), sc


@pytest.mark.parametrize("kw", [
    dict(weno_scheme=4, ep_order=8),
    dict(weno_scheme=4, ep_order=8, flux_splitting=1),
    dict(weno_scheme=3, ep_order=8),
    dict(weno_scheme=3, ep_order=8, weno_version=1),
    dict(weno_scheme=4, ep_order=8, nkeep=2),
    dict(weno_scheme=4, ep_order=8, conservative_viscous=1),
    dict(weno_scheme=4, ep_order=8, theta_wall=0.35),
    dict(weno_scheme=4, ep_order=6),
], ids=["weno7-ep8", "weno7-rusanov", "weno5-ep8", "wenoz-ep8", "keep2", "conservative", "force-var", "ep6-ng4"])
def test_four_ghost_layers_match_oracle(kw):
    c = make_channel(nx=32, ny=24, nz=16, ng=NG, sensor_threshold=0.2, **kw)
    errs, weno, weno_o, sc = _pair(c)
    print(kw, "rel L2", errs, "WENO fractions", list(weno))
    assert max(errs) <= TOL, errs
    assert np.array_equal(weno, weno_o)
    assert max(weno) > 0.0          # the WENO branch is exercised
    assert sc["icyc"] == 20


def test_weno7_eighth_order_channel_128x96x64_100_steps():
    """WENO7 and the eighth-order schemes on the 128 x 96 x 64 channel
    (M_b 1.5, Re_tau 220) over 100 steps, sensor threshold 0.5."""
    c = make_channel(nx=128, ny=96, nz=64, ng=NG, ep_order=8, weno_scheme=4, sensor_threshold=0.5)
    errs, weno, weno_o, sc = _pair(c, steps=100)
    print("WENO7 / ep_order 8, 128x96x64, 100 steps: rel L2", errs, "WENO fractions", list(weno))
    assert max(errs) <= TOL, errs
    assert np.array_equal(weno, weno_o)
    assert sc["icyc"] == 100
