"""CPU pins of the oracle's hybrid switch, WENO-Z weights and the KEEP(n) /
non-KEEP central splits (test infrastructure: the oracle against
mathematics that does not share its code).

- switch window (kernels_cpu.F90:466-470 and the y/z twins): a face takes
  the WENO branch exactly when a cell of its 2*weno_scheme-point window
  i-weno_scheme+1 .. i+weno_scheme carries a sensor above the threshold;
- count_weno (kernels_cpu.F90:52-113) against a numpy brute-force count;
- WENO-Z (kernels_cpu.F90:2628-2667, Borges et al., J. Comput. Phys. 227,
  2008): tau5 = |beta0 - beta2|, so mirror-symmetric stencil data gives the
  ideal weights and the linear fifth-order stencil even across a spike;
- KEEP(n) (kernels_cpu.F90:505-531): the density and internal-energy series
  tend to the logarithmic means of Kuya & Kawai (entropy-conserving limit);
- non-KEEP split (kernels_cpu.F90:556-596): the telescoped flux equals the
  expanded Kennedy-Gruber cubic split form with central derivatives.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import box_case, conserved_from_primitive
from oracle import Oracle, lib


def _random_box(o, p, seed, amp=0.3):
    rng = np.random.default_rng(seed)
    sh = (p.nz, p.ny, p.nx)
    wrap = lambda a: np.pad(a, 3, mode="wrap")
    rho = wrap(rng.uniform(1.0 - amp, 1.0 + amp, sh))
    u, v, w = (wrap(rng.uniform(-amp, amp, sh) + c) for c in (0.4, -0.1, 0.2))
    T = wrap(rng.uniform(1.0 - amp, 1.0 + amp, sh))
    o.w[:] = conserved_from_primitive(rho, u, v, w, T)
    o.compute_aux(1, 1)
    return rho, u, v, w, T


def _fhat(o, d, n):
    lib().or_euler_fluxes(o.h, d, 0, n)
    return o.fhat.copy()


# ---------------------------------------------------------------- switch window
@pytest.mark.parametrize("ws", [1, 2, 3])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_switch_window_is_the_faces_whose_stencil_covers_the_flagged_cell(d, ws):
    p, g, _, _ = box_case(nx=14, ny=13, nz=12)
    p.weno_scheme = ws
    p.sensor_threshold = 0.5
    o = Oracle(p, g)
    _random_box(o, p, seed=3 + d)
    n = (p.nx, p.ny, p.nz)[d]
    central = _fhat(o, d, n)  # sensor 0 everywhere: every face central
    c0 = [7, 6, 5]  # flagged interior cell (i, j, k)
    o.w_aux[7][c0[2] + 2, c0[1] + 2, c0[0] + 2] = 1.0
    hybrid = _fhat(o, d, n)
    diff = np.any(hybrid != central, axis=0)  # [k, j, i] over the ghosted box
    got = sorted(map(tuple, np.argwhere(diff)))
    want = []
    for f in range(c0[d] - ws, c0[d] + ws):
        idx = list(c0)
        idx[d] = f
        want.append((idx[2] + 2, idx[1] + 2, idx[0] + 2))
    assert got == sorted(want), (got, want)


@pytest.mark.parametrize("ws", [1, 2, 3])
def test_count_weno_brute_force(ws):
    p, g, _, _ = box_case(nx=12, ny=11, nz=10)
    p.weno_scheme = ws
    p.sensor_threshold = 0.8
    o = Oracle(p, g)
    rng = np.random.default_rng(11 + ws)
    o.w_aux[7][:] = rng.uniform(0.0, 1.0, o.w_aux[7].shape) ** 6  # about 26 % above 0.8, ghosts included
    got = o.count_weno()
    s = o.w_aux[7] > p.sensor_threshold
    for d, n in enumerate((p.nx, p.ny, p.nz)):
        axis = 2 - d
        cnt = 0
        for f in range(1, n + 1):  # the periodic box: eul range 1..n
            win = np.take(s, np.arange(f - ws + 1, f + ws + 1) + 2, axis=axis)
            inner = [slice(3, -3)] * 3
            inner[axis] = slice(None)
            cnt += int(np.any(win[tuple(inner)], axis=axis).sum())
        assert got[d] == cnt / (p.nx * p.ny * p.nz), (d, got[d] * p.nx * p.ny * p.nz, cnt)


# ----------------------------------------------------------------------- WENO-Z
def _wenorec(gp, gm, version):
    vp, vm = np.zeros((5, 9)), np.zeros((5, 9))
    vp[:, 1:7], vm[:, 1:7] = gp, gm
    out = np.zeros(5)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    lib().or_wenorec_1d(dp(vp), dp(vm), dp(out), 3, 3, version, 1.0, 1.0)
    return out


def _ideal(gp, gm):
    return ((2 * gp[:, 0] - 13 * gp[:, 1] + 47 * gp[:, 2] + 27 * gp[:, 3] - 3 * gp[:, 4]) / 60
            + (2 * gm[:, 5] - 13 * gm[:, 4] + 47 * gm[:, 3] + 27 * gm[:, 2] - 3 * gm[:, 1]) / 60)


def test_wenoz_symmetric_data_gives_ideal_weights_across_a_spike():
    """For data mirror-symmetric about the upwind point, beta0 = beta2, so
    tau5 = 0 and the WENO-Z weights are exactly (0.1, 0.6, 0.3) whatever
    the data: the linear fifth-order stencil, also across a spike where the
    Jiang-Shu weights move far from it."""
    rng = np.random.default_rng(5)
    # g+ uses points l = 1..5 around l = 3; g- uses l = 2..6 around l = 4
    gp = np.zeros((5, 6))
    gm = np.zeros((5, 6))
    for m in range(5):
        a, b, c = rng.uniform(-1, 1, 3) * (1 + 10 * m)
        gp[m, :5] = [a, b, c + 5.0, b, a]  # spike at the centre
        gp[m, 5] = rng.uniform()
        gm[m, 1:] = [a, b, c - 3.0, b, a]
        gm[m, 0] = rng.uniform()
    z = _wenorec(gp, gm, 1)
    js = _wenorec(gp, gm, 0)
    ideal = _ideal(gp, gm)
    assert np.allclose(z, ideal, rtol=1e-13, atol=1e-13)
    assert np.max(np.abs(js - ideal)) > 1e-2


def test_wenoz_discards_the_stencil_holding_a_jump():
    """A unit jump inside the left stencil only (tau5 large): the WENO-Z
    reconstruction of g+ lies within 1e-6 of the value the two smooth
    stencils give, like Jiang-Shu (ENO property)."""
    gp = np.array([[0.0, 1.0, 1.0, 1.0, 1.0, 1.0]] * 5)
    gm = np.zeros((5, 6))
    assert np.all(np.abs(_wenorec(gp, gm, 1) - 1.0) < 1e-6)
    assert np.all(np.abs(_wenorec(gp, gm, 0) - 1.0) < 1e-6)


# -------------------------------------------------------------- KEEP(n) limit
def _second_order_face_fluxes(nkeep, seed):
    p, g, _, _ = box_case(nx=10, ny=6, nz=6)
    p.ep_order = p.visc_order = 2
    p.nkeep = nkeep
    o = Oracle(p, g)
    rho, u, v, w, T = _random_box(o, p, seed, amp=0.25)
    fh = _fhat(o, 0, p.nx)
    return o, fh


def _log_mean(a, b):
    return np.where(np.abs(a - b) > 1e-14 * a, (a - b) / (np.log(a) - np.log(b)), a)


def test_keep_n_tends_to_the_logarithmic_means():
    """KEEP(n) with ep_order 2 (one pair per face): as n grows the face
    fluxes tend to rho_ln u_n for mass, rho_ln u_n u + p_bar for momentum and
    rho_ln u_n e+ e- / e_ln + rho_ln u_n (u+.u-)/2 + (u_n+ p- + u_n- p+)/2 for
    energy, where rho_ln and e_ln are logarithmic means of the two points:
    the density series 1 + xi^2/3 + xi^4/5 + ... sums to atanh(xi)/xi and the
    energy series to the matching closed forms.  The error falls
    geometrically with n."""
    errs = []
    for nk in (1, 2, 4, 60):
        o, fh = _second_order_face_fluxes(nk, seed=21)
        wa = o.w_aux
        a = [wa[m][3:-3, 3:-3, 2:-4] for m in range(6)]   # point i (faces 0..nx-1 -> i = face)
        b = [wa[m][3:-3, 3:-3, 3:-3] for m in range(6)]   # point i+1
        # faces 0..nx-1 sit between a and b
        ra, ua, va, wa_, ha, ta = a
        rb, ub, vb, wb, hb, tb = b
        pa, pb = ra * ta, rb * tb
        ea = ha - pa / ra - 0.5 * (ua ** 2 + va ** 2 + wa_ ** 2)
        eb = hb - pb / rb - 0.5 * (ub ** 2 + vb ** 2 + wb ** 2)
        rln = _log_mean(ra, rb)
        eln = _log_mean(ea, eb)
        un = 0.5 * (ua + ub)
        want = [rln * un, rln * un * 0.5 * (ua + ub) + 0.5 * (pa + pb), rln * un * 0.5 * (va + vb),
                rln * un * 0.5 * (wa_ + wb),
                rln * un * ea * eb / eln + rln * un * 0.5 * (ua * ub + va * vb + wa_ * wb)
                + 0.5 * (ua * pb + ub * pa)]
        got = [o.fhat[m][3:-3, 3:-3, 2:-4] for m in range(5)]
        errs.append(max(np.max(np.abs(got[m] - want[m])) / np.max(np.abs(want[m])) for m in (0, 1, 4)))
    assert errs[-1] < 1e-13, errs
    assert errs[0] > errs[1] > errs[2] > errs[3], errs
    assert errs[2] < 1e-3 * errs[0], errs


def test_keep_n_equals_keep_when_pair_density_and_energy_agree():
    """Equal density and internal energy at the two points of every pair
    (uniform rho and T, random velocity): every series term vanishes and
    KEEP(n) is KEEP bit for bit."""
    out = []
    for nk in (0, 3):
        p, g, _, co = box_case(nx=10, ny=6, nz=6)
        p.ep_order = p.visc_order = 2
        p.nkeep = nk
        o = Oracle(p, g)
        rng = np.random.default_rng(2)
        sh = (p.nz, p.ny, p.nx)
        u, v, w = (np.pad(rng.uniform(-1, 1, sh), 3, mode="wrap") for _ in range(3))
        one = np.ones_like(u)
        o.w[:] = conserved_from_primitive(1.2 * one, u, v, w, 0.9 * one)
        o.compute_aux(1, 1)
        out.append(_fhat(o, 0, p.nx))
    assert np.array_equal(out[0], out[1])


# --------------------------------------------------------- non-KEEP cubic split
@pytest.mark.parametrize("d", [0, 1, 2])
def test_non_keep_split_is_the_kennedy_gruber_cubic_form(d):
    """nkeep < 0: the flux-form sweep equals the expanded Kennedy-Gruber
    split with alpha = beta = kappa = 1/4 (Pirozzoli, J. Comput. Phys. 229,
    2010): (1/4)[D(rho u phi) + rho u D(phi) + rho phi D(u) + u phi D(rho)
    + rho D(u phi) + u D(rho phi) + phi D(rho u)] + D(p) on the normal
    momentum, phi = 1, u, v, w, H, with D the sixth-order central
    derivative.  Uniform data gives zero (free stream)."""
    p, g, dx, co = box_case(nx=12, ny=11, nz=10)
    p.nkeep = -1
    o = Oracle(p, g)
    rho, u, v, w, T = _random_box(o, p, seed=40 + d)
    o.zero_flux()
    getattr(o, "euler_" + "xyz"[d])()
    fl = o.fl.copy()
    axis = 2 - d
    a = [0.75, -0.15, 1.0 / 60.0]
    inner = (slice(3, -3),) * 3

    def D(f):
        return sum(a[l - 1] * (np.roll(f, -l, axis=axis) - np.roll(f, l, axis=axis)) for l in range(1, 4))[inner] / dx[d]

    un = (u, v, w)[d]
    H = 2.5 * T + T + 0.5 * (u * u + v * v + w * w)
    r_, u_ = rho[inner], un[inner]
    for m, phi in enumerate((np.ones_like(rho), u, v, w, H)):
        ph = phi[inner]
        kg = 0.25 * (D(rho * un * phi) + r_ * u_ * D(phi) + r_ * ph * D(un) + u_ * ph * D(rho)
                     + r_ * D(un * phi) + u_ * D(rho * phi) + ph * D(rho * un))
        if m == 1 + d:
            kg = kg + D(rho * T)
        assert np.allclose(fl[m], kg, rtol=1e-11, atol=1e-12 * np.abs(kg).max()), (d, m)
    # free stream
    one = np.ones_like(rho)
    o.w[:] = conserved_from_primitive(1.3 * one, 0.7 * one, -0.4 * one, 0.25 * one, 1.1 * one)
    o.compute_aux(1, 1)
    o.zero_flux()
    getattr(o, "euler_" + "xyz"[d])()
    assert np.max(np.abs(o.fl)) < 1e-13
