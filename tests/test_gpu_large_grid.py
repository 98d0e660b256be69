"""A grid beyond 32-bit offsets on the GPU.

1024 x 512 x 448 cells hold 2.4e9 w_aux values (10 ghosted variables of
1030 x 518 x 454 points), so the offsets of the last w_aux variables pass
2^31.  The field is the 512 x 512 x 448 channel's field tiled twice along
the periodic x direction of a domain twice as long: the discrete scheme is
invariant under a shift by the small domain's length, so after 3 steps each
half of the large field equals the small field, up to the grouping of the
bulk sums of the forcing (round-off).  About 50 GB of HBM.
"""
import math

import numpy as np
import pytest

from helpers import interior, rel_l2
from streams_mi355x import make_channel
from streams_mi355x.solver import ChannelSolver

pytestmark = pytest.mark.gpu


def test_channel_beyond_32_bit_offsets_equals_its_tiled_half():
    small = make_channel(nx=512, ny=512, nz=448, retau=550.0, seed=3)
    big = make_channel(nx=1024, ny=512, nz=448, retau=550.0, lx=8.0 * math.pi, seed=3)
    assert 10 * (1024 + 6) * (512 + 6) * (448 + 6) > 2 ** 31
    ng = 3
    w = big.w0
    ws = small.w0[:, ng:-ng, ng:-ng, ng:-ng]
    w[:, ng:-ng, ng:-ng, ng:ng + 512] = ws
    w[:, ng:-ng, ng:-ng, ng + 512:ng + 1024] = ws
    out = []
    for c in (small, big):
        with ChannelSolver(c.params, c.grid) as g:
            g.copy_from_field(c.w0)
            g.init_run()
            g.advance(3)
            sc = g.compute_residual()
            out.append((interior(g.copy_to_field()), sc))
    (a, sa), (b, sb) = out
    assert np.isfinite(b).all()
    for half in (b[..., :512], b[..., 512:]):
        errs = [rel_l2(half[m], a[m]) for m in range(5)]
        print("half of the 1024 x 512 x 448 field vs 512 x 512 x 448:", errs)
        assert max(errs) <= 1e-12, errs
    assert abs(sb["dt"] - sa["dt"]) <= 1e-13 * sa["dt"]
    assert abs(sb["rhobulk"] - sa["rhobulk"]) <= 1e-13
