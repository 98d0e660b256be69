"""MI355X (gfx950) per-RK-stage hot path of the STREAmS-2 solver.

Host-side mirror of the reference backend interface over the C ABI of
libstreams_mi355x.so (include/streams_mi355x.h).
"""
from .params import FlowCase, StParams, StScalars, channel_params, PREC_FP64, PREC_MIXED  # noqa: F401
from .grid import Grid, channel_grid, channel_initial_field  # noqa: F401
from .cases import ChannelCase, make_channel, make_subsonic_channel  # noqa: F401

__all__ = ["FlowCase", "StParams", "StScalars", "channel_params", "Grid", "channel_grid",
           "channel_initial_field", "ChannelCase", "make_channel", "make_subsonic_channel", "PREC_FP64", "PREC_MIXED"]
