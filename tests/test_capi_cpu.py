"""CPU checks of the C ABI: the library loads and exports every entry point
declared in include/streams_mi355x.h (no compute calls without a GPU)."""
import ctypes as C
import os
import re

import pytest

from streams_mi355x import capi
from streams_mi355x.params import StParams, StScalars

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "streams_mi355x.h")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z0-9_]+\s*\*?\s*(st_[a-z0-9_]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(capi.exported_symbols())


def test_struct_sizes_match_header():
    # st_params: 22 int32 + 6+6 int32 arrays = 34 int32 (136 B), 17 doubles, 2 x 16 doubles
    assert C.sizeof(StParams) == 34 * 4 + 17 * 8 + 32 * 8 + 13 * 8
    assert C.sizeof(StScalars) == 13 * 8 + 8


def test_invalid_parameters_rejected_without_gpu():
    # validation runs before any device call
    from streams_mi355x import make_channel
    c = make_channel()
    p = c.params
    p.ng = 4
    h = C.c_void_p()
    rc = capi.lib().st_solver_create(C.byref(p), C.byref(h))
    assert rc == capi.ST_ERR_ARG
    assert b"ng" in capi.lib().st_last_error()
    assert not h.value


def test_footprint_grows_with_grid():
    from streams_mi355x import make_channel
    a = make_channel(nx=32, ny=24, nz=16).params
    b = make_channel(nx=64, ny=24, nz=16).params
    fa = capi.lib().st_solver_footprint(C.byref(a))
    fb = capi.lib().st_solver_footprint(C.byref(b))
    assert 0 < fa < fb
    # ~ (15 ghosted + 10 interior) doubles per cell
    assert fa > 8 * 25 * 32 * 24 * 16


def test_kernel_families_follow_the_header():
    """st_kernel_times fills ST_NKFAM slots indexed by the ST_K_* values; the
    Python names must sit at those indices."""
    text = open(HEADER).read()
    nfam = int(re.search(r"#define\s+ST_NKFAM\s+(\d+)", text).group(1))
    idx = {m.group(1).lower(): int(m.group(2)) for m in re.finditer(r"#define\s+ST_K_([A-Z_]+)\s+(\d+)", text)}
    assert len(capi.KERNEL_FAMILIES) == nfam == len(idx)
    for name, i in idx.items():
        assert capi.KERNEL_FAMILIES[i] == name, (name, i)
