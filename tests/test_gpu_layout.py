"""Pitched device rows.  The library pads every ghosted row of w and w_aux to
whole 128-byte lines of w_aux values (16 values in FP64, 32 in the mixed
form), with i = 1 opening a line, and re-spaces the rows on every host copy.
On grids whose ghosted rows need no padding, one value of it and nearly a
whole line of it: the host round trips are exact, and short trajectories
follow the oracle (the aux pass writes zeros on the padding of each row;
nothing reads it)."""
import numpy as np
import pytest

from helpers import interior, rel_l2
from oracle import Oracle
from streams_mi355x import PREC_MIXED, make_channel
from streams_mi355x.solver import ChannelSolver

pytestmark = pytest.mark.gpu

# nx + 6 against the pitch: FP64 rows of 16 values, mixed rows of 32
CASES = [(10, 0), (25, 0), (11, 0), (26, PREC_MIXED), (27, PREC_MIXED)]


@pytest.mark.parametrize("nx,precision", CASES)
def test_host_round_trips_are_exact(nx, precision):
    c = make_channel(nx=nx, ny=12, nz=8, precision=precision)
    g = ChannelSolver(c.params, c.grid)
    rng = np.random.default_rng(nx)
    w = rng.uniform(0.5, 2.0, size=(5,) + g.shape_g)
    g.copy_from_field(w)
    assert np.array_equal(g.copy_to_field(), w)
    a = rng.uniform(-1.0, 1.0, size=g.shape_g)
    for var in (1, 6, 10):
        g.set_aux(var, a * var)
        expect = a * var if precision == 0 else (a * var).astype(np.float32).astype(np.float64)
        assert np.array_equal(g.get_aux(var), expect), var
    g.close()


@pytest.mark.parametrize("nx,precision", [(10, 0), (25, 0), (11, 0)])
def test_trajectory_on_padded_rows(nx, precision):
    c = make_channel(nx=nx, ny=12, nz=8, precision=precision, sensor_threshold=0.5, seed=nx)
    o = Oracle(c.params, c.grid)
    g = ChannelSolver(c.params, c.grid)
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    o.init_run()
    g.init_run()
    o.advance(10)
    g.advance(10)
    wg = g.copy_to_field()
    for m in range(5):
        err = rel_l2(interior(wg)[m], interior(o.w)[m])
        assert err <= 1e-10, (nx, m, err)
    g.close()
