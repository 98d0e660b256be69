"""Pins of the oracle's compute_stats (singleideal.F90:2003-2179) and of the
stat.bin layout (write_stats, singleideal.F90:1779-1822).

The field is manufactured: density and temperature vary smoothly, the
velocity is linear in x, y and z, so every sixth-order first derivative of
it is exact and the velocity-gradient tensor is the constant G.  Each of the
70 statistics is then evaluated from the analytic field, per cell, and
averaged over z, independently of the oracle's code path.
"""
import numpy as np
import pytest

from helpers import box_case
from oracle import NV_STAT, Oracle

NG = 3
G = np.array([[0.20, 0.10, 0.05],     # du/dx, du/dy, du/dz
              [-0.05, 0.12, 0.02],    # dv/dx, ...
              [0.03, -0.04, 0.10]])
U0 = np.array([0.3, 0.1, 0.05])


def _field(p, co, phase=0.0):
    """Conserved w (all ghosts analytic) and T, mu on the ghosted box."""
    x, y, z = co
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rho = 1.0 + 0.1 * np.cos(X + phase)
    t = 1.0 + 0.2 * np.sin(Y - phase)
    mu = 0.01 * (1.0 + 0.1 * X)
    vel = [U0[a] + G[a, 0] * X + G[a, 1] * Y + G[a, 2] * Z for a in range(3)]
    w = np.empty((5,) + rho.shape)
    w[0] = rho
    for a in range(3):
        w[1 + a] = rho * vel[a]
    w[4] = rho * (p.cv_coeff[0] * t)  # not read by compute_stats
    return w, t, mu, (rho, vel, t, mu)


def _expected(p, fields):
    """The 70 statistics per interior cell from the analytic field, z-mean."""
    rho, (u, v, ww), t, mu = fields
    sl = (slice(NG, -NG),) * 3
    rho, u, v, ww, t, mu = rho[sl], u[sl], v[sl], ww[sl], t[sl], mu[sl]
    R = p.rgas0
    gam = p.cp_coeff[0] / p.cv_coeff[0]
    pp = rho * R * t
    qq = 0.5 * (u * u + v * v + ww * ww)
    c = np.sqrt(gam * R * t)
    mach = np.sqrt(2 * qq) / c
    ttot = t + qq * (gam - 1) / gam
    (ux, uy, uz), (vx, vy, vz), (wx, wy, wz) = G
    om = (wy - vz, uz - wx, vx - uy)
    div = ux + vy + wz
    S = mu[None, None] * np.array([[2 * (ux - div / 3), uy + vx, uz + wx],
                                   [uy + vx, 2 * (vy - div / 3), vz + wy],
                                   [uz + wx, vz + wy, 2 * (wz - div / 3)]])[:, :, None, None, None]
    ru, rv, rw = rho * u, rho * v, rho * ww
    one = np.ones_like(rho)
    st = [rho, u, v, ww, pp, t, rho ** 2, u ** 2, v ** 2, ww ** 2,
          pp ** 2, t ** 2, ru, rv, rw, ru * u, rv * v, rw * ww, ru * v, mu,
          mu / rho, om[0] ** 2 * one, om[1] ** 2 * one, om[2] ** 2 * one, rho * t, rho * t ** 2, ttot, rho * ttot,
          ttot ** 2, ru * t,
          rv * t, rw * t, mach, mach ** 2, ru * u ** 2, rv * u ** 2, ru * v ** 2, rv * v ** 2, ru * ww ** 2,
          rv * ww ** 2,
          pp * u, pp * v, S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2], S[0, 0] * u, S[0, 1] * u,
          S[1, 0] * v, S[1, 1] * v, S[2, 0] * ww, S[2, 1] * ww, S[0, 0] * v + S[1, 0] * u, S[0, 1] * v + S[1, 1] * u,
          S[0, 0] * ux + S[0, 1] * uy + S[0, 2] * uz, S[1, 0] * vx + S[1, 1] * vy + S[1, 2] * vz,
          S[2, 0] * wx + S[2, 1] * wy + S[2, 2] * wz,
          S[0, 0] * vx + S[0, 1] * (ux + vy) + S[1, 1] * uy + S[0, 2] * vz + S[1, 2] * uz,
          pp * ux, pp * vy, pp * wz, pp * (uy + vx), div ** 2 * one, rho * t ** 3, rho * t ** 4, ru * u ** 3,
          gam / (gam - 1) * R * one, gam * one]
    assert len(st) == NV_STAT
    # [k, j, i] -> mean over k -> (q, i, j)
    return np.stack([np.broadcast_to(a, rho.shape).mean(axis=0).T for a in st])


def _oracle_stats(p, g, co, samples):
    o = Oracle(p, g)
    ws = np.zeros((NV_STAT, p.nx, p.ny), order="F")
    for itav, phase in enumerate(samples):
        w, t, mu, _ = _field(p, co, phase)
        o.w[:] = w
        o.w_aux[5] = t
        o.w_aux[6] = mu
        o.compute_stats(itav, ws)
    return ws


def test_compute_stats_matches_the_analytic_field():
    p, g, d, co = box_case(nx=12, ny=10, nz=8, L=(3.0, 2.0, 1.5))
    ws = _oracle_stats(p, g, co, [0.0])
    want = _expected(p, _field(p, co, 0.0)[3])
    for q in range(NV_STAT):
        scale = max(np.abs(want[q]).max(), 1e-30)
        assert np.allclose(ws[q], want[q], rtol=1e-11, atol=1e-12 * scale), (q + 1, ws[q].ravel()[:3], want[q].ravel()[:3])


def test_compute_stats_time_average_of_two_samples():
    p, g, d, co = box_case(nx=12, ny=10, nz=8, L=(3.0, 2.0, 1.5))
    ws = _oracle_stats(p, g, co, [0.0, 0.7])
    want = 0.5 * (_expected(p, _field(p, co, 0.0)[3]) + _expected(p, _field(p, co, 0.7)[3]))
    for q in range(NV_STAT):
        scale = max(np.abs(want[q]).max(), 1e-30)
        assert np.allclose(ws[q], want[q], rtol=1e-11, atol=1e-12 * scale), q + 1


def test_stat_bin_layout(tmp_path):
    """write_stats: one nx*ny block per statistic, x fastest, no header."""
    from streams_mi355x.io import read_stats, write_stats

    nx, ny = 5, 4
    ws = np.asfortranarray(np.arange(NV_STAT * nx * ny, dtype=np.float64).reshape((NV_STAT, nx, ny), order="F"))
    path = tmp_path / "stat.bin"
    write_stats(path, ws)
    raw = np.fromfile(path, dtype="<f8")
    assert raw.size == NV_STAT * nx * ny
    for q in (0, 1, 37, NV_STAT - 1):
        block = raw[q * nx * ny:(q + 1) * nx * ny]
        assert np.array_equal(block, ws[q].ravel(order="F"))   # i fastest, then j
    assert np.array_equal(read_stats(path, nx, ny), ws)


def test_rst_bin_layout_and_round_trip(tmp_path):
    """write_field: one nx*ny*nz block per conserved variable, x fastest,
    interior cells only, no header; read_field reads it back."""
    from streams_mi355x.io import read_restart, write_restart

    nx, ny, nz, ng = 5, 4, 3, 3
    w = np.random.default_rng(2).normal(size=(5, nz + 2 * ng, ny + 2 * ng, nx + 2 * ng))
    path = tmp_path / "rst.bin"
    write_restart(path, w, ng)
    raw = np.fromfile(path, dtype="<f8")
    assert raw.size == 5 * nx * ny * nz
    n = nx * ny * nz
    for m in range(5):
        block = raw[m * n:(m + 1) * n]
        # Fortran w(1:nx,1:ny,1:nz,m): i fastest, then j, then k
        want = np.array([w[m, ng + k, ng + j, ng + i] for k in range(nz) for j in range(ny) for i in range(nx)])
        assert np.array_equal(block, want)
    assert np.array_equal(read_restart(path, nx, ny, nz), w[:, ng:-ng, ng:-ng, ng:-ng])


def test_field_info_round_trip(tmp_path):
    from streams_mi355x.io import FIELD_INFO_KEYS, read_field_info, write_field_info

    info = {"icyc0": 1234, "time0": 0.123456789012345678, "itav": 7, "istore": 3,
            "time_from_last_stat": 1.0 / 3.0}
    path = tmp_path / "field_info0.dat"
    write_field_info(path, info)
    text = path.read_text().splitlines()
    assert text[0] == "[field_info]"
    assert [ln.split(" = ")[0] for ln in text[1:]] == list(FIELD_INFO_KEYS)
    back = read_field_info(path)
    for k, v in info.items():
        assert back[k] == v
    # the reference's list-directed spelling (Fortran exponent letter) reads too
    path.write_text("[field_info]\nicyc0 = 10\ntime0 = 0.25D+01\nitav = 2\n")
    assert read_field_info(path)["time0"] == 2.5
