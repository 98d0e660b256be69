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


def _c_layout(struct, fields):
    """sizeof and offsetof of the header's struct, from a C compiler."""
    import os
    import subprocess
    import tempfile

    inc = os.path.join(ROOT, "include")
    body = "\n".join(f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields)
    src = (f'#include <stdio.h>\n#include <stddef.h>\n#include "streams_mi355x.h"\n'
           f'int main(void) {{ printf("%zu\\n", sizeof({struct})); {body} return 0; }}\n')
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "l.c"), "w").write(src)
        subprocess.run(["gcc", "-I", inc, "-o", os.path.join(d, "l"), os.path.join(d, "l.c")], check=True)
        out = subprocess.run([os.path.join(d, "l")], check=True, capture_output=True, text=True).stdout.split()
    return int(out[0]), [int(v) for v in out[1:]]


@pytest.mark.parametrize("cls,name", [(StParams, "st_params"), (StScalars, "st_scalars")])
def test_struct_layouts_match_header(cls, name):
    fields = [f for f, _ in cls._fields_]
    size, offs = _c_layout(name, fields)
    assert C.sizeof(cls) == size
    assert [getattr(cls, f).offset for f in fields] == offs


@pytest.mark.parametrize("icl,icr,cp", [(1, 2, 1), (-3, -2, 1), (2, 1, 0)])
def test_cp_exponent_range_rejected_without_gpu(icl, icr, cp):
    from streams_mi355x import make_channel
    p = make_channel().params
    p.indx_cp_l, p.indx_cp_r, p.calorically_perfect = icl, icr, cp
    h = C.c_void_p()
    assert capi.lib().st_solver_create(C.byref(p), C.byref(h)) == capi.ST_ERR_ARG
    assert b"indx_cp" in capi.lib().st_last_error()


def test_unknown_boundary_tag_rejected_without_gpu():
    from streams_mi355x import make_channel
    p = make_channel().params
    p.bctags[2] = 11
    h = C.c_void_p()
    assert capi.lib().st_solver_create(C.byref(p), C.byref(h)) == capi.ST_ERR_ARG
    assert b"boundary tag" in capi.lib().st_last_error()


def test_invalid_parameters_rejected_without_gpu():
    # validation runs before any device call
    from streams_mi355x import make_channel
    c = make_channel()
    p = c.params
    p.ng = 2
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


def test_grid_limit_is_2_to_the_31_points_per_variable():
    """A point's index within one ghosted variable is 32-bit and offsets
    across variables are 64-bit: 1300^3 cells (2.2e9 ghosted points) are
    rejected before any device call; the footprint of 1024 x 512 x 448 (2.4e9
    w_aux values, beyond a 32-bit offset) is reported, 64-bit."""
    from streams_mi355x import make_channel
    p = make_channel().params
    p.nx = p.ny = p.nz = 1300
    h = C.c_void_p()
    assert capi.lib().st_solver_create(C.byref(p), C.byref(h)) == capi.ST_ERR_ARG
    assert b"2^31 points" in capi.lib().st_last_error()
    p.nx, p.ny, p.nz = 1024, 512, 448
    foot = capi.lib().st_solver_footprint(C.byref(p))
    assert foot > 10 * 1030 * 518 * 454 * 8 > 2 ** 31 * 8


@pytest.mark.parametrize("change,msg", [
    (dict(precision=1), b"FP64 only"),
    (dict(bctags=(1, 2, 6, 2, 0, 0)), b"periodic faces and staggered"),
    (dict(ng=3), b"ep_order must be even and at most 2 ng"),
    (dict(weno_version=1), b"weno_version 1 supported only for weno_scheme = 3"),
    (dict(ng=5), b"ng must be 3, or 4"),
])
def test_four_ghost_layers_are_validated(change, msg):
    """ng = 4 (WENO7 / ep_order 8, singleideal.F90:497-509) runs the channel
    procedure by procedure in FP64; every other combination is refused
    before any device call."""
    from streams_mi355x import make_channel
    p = make_channel(nx=16, ny=12, nz=8, ng=4, ep_order=8, weno_scheme=4).params
    for k, v in change.items():
        if k == "bctags":
            p.bctags[:] = list(v)
        else:
            setattr(p, k, v)
    h = C.c_void_p()
    assert capi.lib().st_solver_create(C.byref(p), C.byref(h)) == capi.ST_ERR_ARG
    assert msg in capi.lib().st_last_error(), capi.lib().st_last_error()


def test_footprint_counts_pitched_rows():
    """w and w_aux rows are padded to whole 128-byte lines of w_aux values:
    16 FP64 values, so nx + 6 = 16 needs no padding and nx + 6 = 17 pads to
    32; the 15 extra values of each of the 15 ghosted variables show in the
    footprint.  The mixed form pads to 32 values."""
    from streams_mi355x import PREC_MIXED, make_channel
    ny, nz = 12, 8
    planes = (ny + 6) * (nz + 6)
    f10 = capi.lib().st_solver_footprint(C.byref(make_channel(nx=10, ny=ny, nz=nz).params))
    f11 = capi.lib().st_solver_footprint(C.byref(make_channel(nx=11, ny=ny, nz=nz).params))
    # w (5) and w_aux (10) grow by 16 values a row; everything else by a few cells
    assert f11 - f10 >= 15 * 16 * planes * 8
    m26 = capi.lib().st_solver_footprint(C.byref(make_channel(nx=26, ny=ny, nz=nz, precision=PREC_MIXED).params))
    m27 = capi.lib().st_solver_footprint(C.byref(make_channel(nx=27, ny=ny, nz=nz, precision=PREC_MIXED).params))
    # w (5 doubles) and w_aux (11 floats) grow by 32 values a row
    assert m27 - m26 >= (5 * 8 + 11 * 4) * 32 * planes
