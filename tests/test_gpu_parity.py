"""Parity of the HIP path with the CPU oracle on identical seeded inputs.

Unit level: each procedure of the backend object on the same input state,
compared with the oracle's restatement of the reference routine (relative L2
per variable <= 1e-12: the two differ only by rounding, e.g. fused
multiply-adds on the GPU).  Trajectory level: init + N Wray RK3 steps,
relative L2 per conserved variable <= 1e-10 (the FP64 bar of the task).
"""
import numpy as np
import pytest

from helpers import fill_ghosts_wrap, interior, rel_l2
from oracle import Oracle
from streams_mi355x import make_channel
from streams_mi355x.params import params_copy
from streams_mi355x.solver import ChannelSolver

pytestmark = pytest.mark.gpu

UNIT_TOL = 1e-12
TRAJ_TOL = 1e-10


def _pair(nx=32, ny=24, nz=16, threshold=0.9, seed=0, **kw):
    c = make_channel(nx=nx, ny=ny, nz=nz, sensor_threshold=threshold, seed=seed, **kw)
    o = Oracle(c.params, c.grid)
    g = ChannelSolver(c.params, c.grid)
    return c, o, g


def _load_state(c, o, g, sensor_seed=None, energy_noise=0.0):
    """Same w (ghosts filled everywhere) and, optionally, a random sensor.
    ``energy_noise`` perturbs the total energy (hence T and mu) by a seeded
    relative amount, so that viscosity gradients are not zero."""
    w0 = c.w0.copy()
    if energy_noise:
        rng = np.random.default_rng(7)
        w0[4] *= 1.0 + energy_noise * rng.uniform(-1.0, 1.0, size=w0[4].shape)
    w = fill_ghosts_wrap(w0)
    o.w[:] = w
    g.copy_from_field(w)
    o.compute_aux(1, 1)
    g.compute_aux(1, 1)
    if sensor_seed is not None:
        rng = np.random.default_rng(sensor_seed)
        s = rng.uniform(0.0, 1.0, size=o.w_aux[7].shape)
        o.w_aux[7][:] = s
        g.set_aux(8, s)
    return w


def test_aux_parity():
    c, o, g = _pair()
    _load_state(c, o, g)
    for v in range(1, 8):
        assert rel_l2(g.get_aux(v), o.w_aux[v - 1]) <= UNIT_TOL, v


@pytest.mark.parametrize("sensor_seed", [None, 1])
@pytest.mark.parametrize("direction", ["x", "y", "z"])
def test_euler_parity(direction, sensor_seed):
    # sensor_seed None: sensor 0 everywhere -> all central faces;
    # sensor_seed 1: uniform random sensor with threshold 0.9 -> mixed faces
    c, o, g = _pair()
    _load_state(c, o, g, sensor_seed)
    o.zero_flux()
    g.zero_flux()
    getattr(o, "euler_" + direction)()
    getattr(g, "euler_" + direction)()
    fo, fg = o.fl, g.get_fl()
    for m in range(5):
        assert rel_l2(fg[m], fo[m]) <= UNIT_TOL, (direction, m, rel_l2(fg[m], fo[m]))


def test_euler_full_weno_parity():
    c, o, g = _pair(threshold=-1.0)
    _load_state(c, o, g)
    for d in ("x", "y", "z"):
        o.zero_flux()
        g.zero_flux()
        getattr(o, "euler_" + d)()
        getattr(g, "euler_" + d)()
        fo, fg = o.fl, g.get_fl()
        for m in range(5):
            assert rel_l2(fg[m], fo[m]) <= UNIT_TOL, (d, m, rel_l2(fg[m], fo[m]))


@pytest.mark.parametrize("mode", [4, 0])
def test_visflx_parity(mode):
    c, o, g = _pair()
    _load_state(c, o, g)
    o.zero_flux()
    g.zero_flux()
    o.visflx(mode)
    g.visflx(mode)
    fo, fg = o.fl, g.get_fl()
    for m in range(1, 5):
        assert rel_l2(fg[m], fo[m]) <= UNIT_TOL, (mode, m)
    ng = 3
    for v in ([8, 9, 10] if mode == 0 else [9, 10]):
        a = interior(g.get_aux(v)[None])[0]
        b = interior(o.w_aux[v - 1][None])[0]
        assert rel_l2(a, b) <= UNIT_TOL, (mode, v)


@pytest.mark.parametrize("mode", [4, 0])
def test_visflx_parity_banded_tiles(mode):
    """64x128x16: 16 rows of viscous tiles, so the workgroups are renumbered
    into per-XCD bands of two tile rows each; results must not change."""
    c, o, g = _pair(nx=64, ny=128, nz=16)
    _load_state(c, o, g, energy_noise=0.05)
    o.zero_flux()
    g.zero_flux()
    o.visflx(mode)
    g.visflx(mode)
    fo, fg = o.fl, g.get_fl()
    for m in range(1, 5):
        assert rel_l2(fg[m], fo[m]) <= UNIT_TOL, (mode, m)
    for v in ([8, 9, 10] if mode == 0 else [9, 10]):
        a = interior(g.get_aux(v)[None])[0]
        b = interior(o.w_aux[v - 1][None])[0]
        assert rel_l2(a, b) <= UNIT_TOL, (mode, v)
    g.close()


def test_trajectory_banded_tiles():
    c, o, g = _run_pair(5, nx=64, ny=128, nz=16, threshold=0.5)
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    assert max(errs) <= TRAJ_TOL, errs
    g.close()


def test_visflx_div_and_sensor_parity():
    c, o, g = _pair()
    _load_state(c, o, g)
    o.zero_flux()
    o.visflx(4)
    o.bcextr_var(10)
    o.bcswap_var(10)
    g.set_aux(9, o.w_aux[8])
    g.set_aux(10, o.w_aux[9])
    o.zero_flux()
    g.zero_flux()
    o.visflx(1)
    g.visflx(1)
    fo, fg = o.fl, g.get_fl()
    for m in range(1, 5):
        assert rel_l2(fg[m], fo[m]) <= UNIT_TOL, m
    o.visflx(3)
    g.visflx(3)
    a = interior(g.get_aux(8)[None])[0]
    b = interior(o.w_aux[7][None])[0]
    assert np.max(np.abs(a - b)) <= 1e-13


def test_ghost_fill_parity():
    c, o, g = _pair()
    w = c.w0.copy()
    o.w[:] = w
    g.copy_from_field(w)
    o.compute_aux(1, 1)
    g.compute_aux(1, 1)
    o.update_ghost(1)
    g.update_ghost(1)
    wg = g.copy_to_field()
    ng = 3
    # every ghost point the reference writes: x/z periodic faces, y wall faces
    for m in range(5):
        a, b = wg[m], o.w[m]
        assert rel_l2(a[ng:-ng, ng:-ng, :], b[ng:-ng, ng:-ng, :]) <= UNIT_TOL
        assert rel_l2(a[ng:-ng, :, ng:-ng], b[ng:-ng, :, ng:-ng]) <= UNIT_TOL
        assert rel_l2(a[:, ng:-ng, ng:-ng], b[:, ng:-ng, ng:-ng]) <= UNIT_TOL


def test_dt_parity():
    c, o, g = _pair()
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    o.init_run()
    g.init_run()
    dto = o.scalars()["dt"]
    dtg = g.compute_dt()
    assert abs(dtg - dto) <= 1e-14 * dto


def _run_pair(nsteps, **kw):
    c, o, g = _pair(**kw)
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    o.init_run()
    g.init_run()
    o.advance(nsteps)
    g.advance(nsteps)
    return c, o, g


@pytest.mark.parametrize("threshold", [0.9, 0.0, -1.0])
def test_trajectory_small(threshold):
    c, o, g = _run_pair(20, threshold=threshold)
    wg = g.copy_to_field()
    for m in range(5):
        err = rel_l2(interior(wg)[m], interior(o.w)[m])
        assert err <= TRAJ_TOL, (threshold, m, err)
    so, sg = o.scalars(), g.scalars()
    for key in ("dt", "time", "rhobulk", "ubulk", "tbulk"):
        assert abs(sg[key] - so[key]) <= 1e-12 * abs(so[key]), key
    assert abs(sg["dpdx"] - so["dpdx"]) <= 1e-9 * abs(so["dpdx"]) + 1e-15
    assert sg["icyc"] == so["icyc"] == 20


def test_trajectory_bulk_temperature_control():
    """theta_wall >= -1: force_var runs after every stage's update."""
    c, o, g = _run_pair(20, theta_wall=1.0)
    assert c.params.theta_wall >= -1.0
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    assert max(errs) <= TRAJ_TOL, errs
    so, sg = o.scalars(), g.scalars()
    for key in ("rhobulk", "ubulk", "tbulk"):
        assert abs(sg[key] - so[key]) <= 1e-12 * abs(so[key]), key
    g.close()


@pytest.mark.parametrize("visc_chunk,march_chunk", [(8, 8), (1, 3), (16, 32), (3, 5)])
def test_results_independent_of_launch_shape(visc_chunk, march_chunk):
    """The z-chunk of the viscous march and the chunk of the y/z flux
    marches change the launch only: results are bitwise those of the
    automatic shape."""
    from streams_mi355x import capi

    c = make_channel(nx=32, ny=24, nz=16, mach=0.8, retau=150.0, seed=0, sensor_threshold=0.5)
    out = []
    for vc, mc in ((0, 0), (visc_chunk, march_chunk)):
        with ChannelSolver(c.params, c.grid) as g:
            capi.check(g.L.st_set_tuning(g.h, vc, mc))
            g.copy_from_field(c.w0)
            g.init_run()
            g.advance(12)
            out.append(g.copy_to_field())
    assert np.array_equal(interior(out[0]), interior(out[1]))


def test_trajectory_ragged_cell_count():
    """29 x 21 x 11 cells (6699, not a multiple of 8): the last 64-bit word of
    the per-cell WENO flags also holds the sweeps' any-dirty bytes, which the
    fix-up's compaction must neither list nor clear.  Threshold 0.2 keeps
    WENO faces in every direction over the 10 steps."""
    c, o, g = _run_pair(10, nx=29, ny=21, nz=11, threshold=0.2, seed=2)
    assert min(o.count_weno()) > 0.005
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    assert max(errs) <= TRAJ_TOL, errs
    assert np.allclose(g.count_weno(), o.count_weno(), rtol=0, atol=0)
    g.close()


def test_residual_and_weno_count():
    c, o, g = _run_pair(5, threshold=0.0)
    o.compute_residual()
    sg = g.compute_residual()
    so = o.scalars()
    assert abs(sg["residual_rhou"] - so["residual_rhou"]) <= 1e-10 * so["residual_rhou"]
    for key in ("rhomin", "rhomax", "tmin", "tmax", "pmin", "pmax"):
        assert abs(sg[key] - so[key]) <= 1e-13 * abs(so[key]), key
    assert np.allclose(g.count_weno(), o.count_weno(), rtol=0, atol=0)


def test_trajectory_c1_100_steps():
    """Config C1 (128x96x64 channel, M_b 1.5, Re_tau 220) over 100 steps."""
    c, o, g = _run_pair(100, nx=128, ny=96, nz=64)
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    print("C1 100-step rel L2 per variable:", errs)
    assert max(errs) <= TRAJ_TOL, errs


def test_bitwise_reproducible():
    c = make_channel()
    outs = []
    for _ in range(2):
        g = ChannelSolver(c.params, c.grid)
        g.copy_from_field(c.w0)
        g.init_run()
        g.advance(10)
        outs.append(g.copy_to_field())
        g.close()
    assert np.array_equal(interior(outs[0]), interior(outs[1]))


def test_bitwise_reproducible_all_weno():
    """Threshold 0: every face is WENO, so every cell goes through the
    fix-up's segmented lists (several segments at 64x48x32).  The lists are
    a function of the flags alone, so two runs agree bit for bit."""
    c = make_channel(nx=64, ny=48, nz=32, sensor_threshold=0.0, seed=3)
    outs = []
    for _ in range(2):
        with ChannelSolver(c.params, c.grid) as g:
            g.copy_from_field(c.w0)
            g.init_run()
            g.advance(4)
            outs.append(g.copy_to_field())
    assert np.array_equal(interior(outs[0]), interior(outs[1]))


# ------------------------------------------------ conservative viscous form
@pytest.mark.parametrize("mode", [6, 5, 2])
def test_visflx_conservative_parity(mode):
    c, o, g = _pair(conservative_viscous=1, visc_model=2)
    _load_state(c, o, g, energy_noise=0.05)
    o.zero_flux()
    g.zero_flux()
    o.visflx(mode)
    g.visflx(mode)
    fo, fg = o.fl, g.get_fl()
    for m in range(1, 5):
        assert rel_l2(fg[m], fo[m]) <= UNIT_TOL, (mode, m, rel_l2(fg[m], fo[m]))
    if mode != 5:
        for v in ([8, 9, 10] if mode == 2 else [9, 10]):
            a = interior(g.get_aux(v)[None])[0]
            b = interior(o.w_aux[v - 1][None])[0]
            assert rel_l2(a, b) <= UNIT_TOL, (mode, v)


def test_visflx_div_ord2_parity():
    c, o, g = _pair(conservative_viscous=1, visc_model=2)
    _load_state(c, o, g, energy_noise=0.05)
    o.zero_flux()
    o.visflx(6)
    o.bcextr_var(10)
    o.bcswap_var(10)
    g.set_aux(10, o.w_aux[9])
    o.zero_flux()
    g.zero_flux()
    o.visflx(7)
    g.visflx(7)
    fo, fg = o.fl, g.get_fl()
    for m in range(1, 5):
        assert rel_l2(fg[m], fo[m]) <= UNIT_TOL, m


@pytest.mark.parametrize("threshold", [0.9, -1.0])
def test_trajectory_conservative_viscous(threshold):
    c, o, g = _run_pair(20, threshold=threshold, conservative_viscous=1, visc_model=2)
    wg = g.copy_to_field()
    for m in range(5):
        err = rel_l2(interior(wg)[m], interior(o.w)[m])
        assert err <= TRAJ_TOL, (threshold, m, err)
    so, sg = o.scalars(), g.scalars()
    for key in ("dt", "time", "rhobulk", "ubulk", "tbulk"):
        assert abs(sg[key] - so[key]) <= 1e-12 * abs(so[key]), key


def test_trajectory_subsonic_channel_100_steps():
    """examples/subsonic_channel at 80x60x80 (the golden-row deck) over 100
    steps: conservative viscous form, Sutherland viscosity, rank-seeded init."""
    from streams_mi355x import make_subsonic_channel

    c = make_subsonic_channel()
    o = Oracle(c.params, c.grid)
    g = ChannelSolver(c.params, c.grid)
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    o.init_run()
    g.init_run()
    o.advance(100)
    g.advance(100)
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    print("subsonic channel 100-step rel L2 per variable:", errs)
    assert max(errs) <= TRAJ_TOL, errs


def _pair_params(threshold, **fields):
    """A channel pair whose StParams get the given fields."""
    c = make_channel(nx=32, ny=24, nz=16, sensor_threshold=threshold, seed=0)
    p = params_copy(c.params)
    for k, v in fields.items():
        setattr(p, k, v)
    return c, Oracle(p, c.grid), ChannelSolver(p, c.grid)


@pytest.mark.parametrize("fields", [dict(flux_splitting=1), dict(flux_splitting=1, weno_scheme=2),
                                    dict(flux_splitting=0, weno_scheme=2), dict(flux_splitting=1, weno_version=1)],
                         ids=["rusanov", "rusanov-weno3", "roe-weno3", "rusanov-wenoz"])
def test_euler_all_weno_split_variants(fields):
    """Every face WENO (threshold -1): both flux splittings, WENO3 and WENO-Z."""
    c, o, g = _pair_params(-1.0, **fields)
    _load_state(c, o, g)
    for d in ("x", "y", "z"):
        o.zero_flux()
        g.zero_flux()
        getattr(o, "euler_" + d)()
        getattr(g, "euler_" + d)()
        fo, fg = o.fl, g.get_fl()
        for m in range(5):
            assert rel_l2(fg[m], fo[m]) <= UNIT_TOL, (fields, d, m, rel_l2(fg[m], fo[m]))
    g.close()


def test_trajectory_rusanov_20_steps():
    c, o, g = _pair_params(0.5, flux_splitting=1)
    o.w[:] = c.w0
    g.copy_from_field(c.w0)
    o.init_run()
    g.init_run()
    o.advance(20)
    g.advance(20)
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    assert max(errs) <= TRAJ_TOL, errs
    g.close()


NASA_AIR = (1.009950160e+04, -1.968275610e+02, 5.009155110e+00, -5.761013730e-03, 1.066859930e-05,
            -7.940297970e-09, 2.185231910e-12, -1.767962850e+02)   # cp/R of T, l = -2..4, then h*/R


@pytest.mark.parametrize("kw", [
    dict(calorically_perfect=0, indx_cp_l=-2, indx_cp_r=4, cp_coeff=NASA_AIR, t_ref=220.0),
    dict(calorically_perfect=0, indx_cp_l=-2, indx_cp_r=4, cp_coeff=NASA_AIR, t_ref=220.0, visc_model=2),
    dict(nkeep=2), dict(nkeep=-1), dict(weno_version=1), dict(ep_order=4), dict(cfl=-0.004),
    dict(iter_dt_recompute=2)],
    ids=["cp(T)", "cp(T)-sutherland", "keep-n2", "non-keep", "weno-z", "central-4th", "fixed-dt", "dt-every-2"])
def test_trajectory_variants(kw):
    """20 steps of the channel with the sensor at 0.5 (WENO faces present)
    in each variant of the scheme the reference's input deck selects."""
    c, o, g = _run_pair(20, threshold=0.5, **kw)
    wg = g.copy_to_field()
    errs = [rel_l2(interior(wg)[m], interior(o.w)[m]) for m in range(5)]
    assert max(errs) <= TRAJ_TOL, (kw, errs)
    so, sg = o.scalars(), g.scalars()
    assert abs(sg["time"] - so["time"]) <= 1e-12 * so["time"]
    g.close()
