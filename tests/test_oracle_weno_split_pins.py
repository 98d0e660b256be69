"""Pins of the oracle's WENO faces for both flux splittings.

With weno_scheme = 1 every WENO face is first order: the split fluxes of
the two face points are combined without reconstruction, and the face flux
takes a closed form.
* Rusanov split (*_hybrid_rusanov_kernel, kernels_cpu.F90:692-941): the local
  Lax-Friedrichs flux 1/2 (F_L + F_R) - 1/2 lam (U_R - U_L), lam the largest
  |u_n -+ c| of the two points.
* Roe split (euler_x_fluxes_hybrid_kernel, :425-689): the same with one bound
  per wave, 1/2 (F_L + F_R) - 1/2 R diag(lam_m) L (U_R - U_L), R and L the
  eigenvectors of the Jacobian at the Roe average and lam_m the largest
  |u_n - c|, |u_n|, |u_n + c| of the two points.
The Jacobian and its eigenvectors are taken from finite differences and
numpy, independently of the oracle's eigenvector formulas.
"""
import numpy as np
import pytest

from helpers import box_case, conserved_from_primitive
from oracle import Oracle

NG = 3


def _case(split, seed=3):
    p, g, d, co = box_case(nx=10, ny=8, nz=6, threshold=-1.0)
    p.weno_scheme = 1
    p.flux_splitting = split
    rng = np.random.default_rng(seed)
    shape = (p.nz + 2 * NG, p.ny + 2 * NG, p.nx + 2 * NG)
    rho = 1.0 + 0.3 * rng.random(shape)
    u = [0.4 * (rng.random(shape) - 0.5) for _ in range(3)]
    t = 1.0 + 0.3 * rng.random(shape)
    o = Oracle(p, g)
    o.w[:] = conserved_from_primitive(rho, u[0], u[1], u[2], t, cv=p.cv_coeff[0], rgas=p.rgas0)
    o.compute_aux(1, 1)
    return p, o


def _state(o, k, j, i):
    return o.w[:, k, j, i].copy()


def _flux(wv, d, gam):
    rho, mom, E = wv[0], wv[1:4], wv[4]
    u = mom / rho
    pres = (gam - 1.0) * (E - 0.5 * rho * np.dot(u, u))
    f = np.empty(5)
    f[0] = mom[d]
    f[1:4] = mom * u[d]
    f[1 + d] += pres
    f[4] = (E + pres) * u[d]
    return f, u[d], np.sqrt(gam * pres / rho)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_rusanov_first_order_face_is_local_lax_friedrichs(d):
    p, o = _case(1)
    gam = p.cp_coeff[0] / p.cv_coeff[0]
    n = [p.nx, p.ny, p.nz][d]
    o.euler_fluxes(d, 0, n)
    for (k, j, i) in [(4, 5, 6), (5, 4, 3), (6, 7, 8)]:
        L = (k, j, i)
        R = [k, j, i]
        R[2 - d] += 1          # [k, j, i] index of the right point along d
        wl, wr = _state(o, *L), _state(o, *R)
        fl, ul, cl = _flux(wl, d, gam)
        fr, ur, cr = _flux(wr, d, gam)
        lam = max(abs(ul) + cl, abs(ur) + cr)
        want = 0.5 * (fl + fr) - 0.5 * lam * (wr - wl)
        assert np.allclose(o.fhat[:, k, j, i], want, rtol=1e-12, atol=1e-13), (d, o.fhat[:, k, j, i], want)


def _jacobian(wv, d, gam):
    A = np.empty((5, 5))
    for q in range(5):
        h = 1e-6 * max(1.0, abs(wv[q]))
        a, b = wv.copy(), wv.copy()
        a[q] += h
        b[q] -= h
        A[:, q] = (_flux(a, d, gam)[0] - _flux(b, d, gam)[0]) / (2 * h)
    return A


def _roe_state(wl, wr, gam):
    """Conserved state of the Roe average (sqrt(rho) weighting of u and H)."""
    def prim(wv):
        rho = wv[0]
        u = wv[1:4] / rho
        pres = (gam - 1.0) * (wv[4] - 0.5 * rho * np.dot(u, u))
        return rho, u, (wv[4] + pres) / rho
    rl, ul, hl = prim(wl)
    rr, ur, hr = prim(wr)
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    u = (sl * ul + sr * ur) / (sl + sr)
    h = (sl * hl + sr * hr) / (sl + sr)
    rho = sl * sr
    pres = (gam - 1.0) / gam * rho * (h - 0.5 * np.dot(u, u))
    return np.array([rho, *(rho * u), pres / (gam - 1.0) + 0.5 * rho * np.dot(u, u)])


@pytest.mark.parametrize("d", [0, 1, 2])
def test_roe_first_order_face_is_wavewise_lax_friedrichs(d):
    p, o = _case(0)
    gam = p.cp_coeff[0] / p.cv_coeff[0]
    n = [p.nx, p.ny, p.nz][d]
    o.euler_fluxes(d, 0, n)
    for (k, j, i) in [(4, 5, 6), (5, 4, 3)]:
        R_ = [k, j, i]
        R_[2 - d] += 1
        wl, wr = _state(o, k, j, i), _state(o, *R_)
        fl, ul, cl = _flux(wl, d, gam)
        fr, ur, cr = _flux(wr, d, gam)
        A = _jacobian(_roe_state(wl, wr, gam), d, gam)
        ev, Rm = np.linalg.eig(A)
        ev, Rm = ev.real, Rm.real
        Lm = np.linalg.inv(Rm)
        # the bound of each wave: |u-c| for the slow acoustic, |u| for the
        # entropy/shear waves, |u+c| for the fast acoustic (eigenvalue order)
        un = _flux(_roe_state(wl, wr, gam), d, gam)[1]
        lam = np.empty(5)
        for m in range(5):
            if abs(ev[m] - un) < 1e-6:
                lam[m] = max(abs(ul), abs(ur))
            elif ev[m] < un:
                lam[m] = max(abs(ul - cl), abs(ur - cr))
            else:
                lam[m] = max(abs(ul + cl), abs(ur + cr))
        want = 0.5 * (fl + fr) - 0.5 * Rm @ (lam * (Lm @ (wr - wl)))
        assert np.allclose(o.fhat[:, k, j, i], want, rtol=1e-6, atol=1e-8), (d, o.fhat[:, k, j, i], want)


def test_rusanov_and_roe_agree_on_uniform_data_and_differ_otherwise():
    """Both splittings reproduce the physical flux of a uniform state; on
    varying data the two WENO fluxes differ."""
    p, o = _case(1)
    q, o2 = _case(0)
    for oo in (o, o2):
        oo.euler_fluxes(0, 0, p.nx)
    assert np.abs(o.fhat[:, 4, 5, 6] - o2.fhat[:, 4, 5, 6]).max() > 1e-4
    gam = p.cp_coeff[0] / p.cv_coeff[0]
    for split in (0, 1):
        pp, oo = _case(split)
        oo.w[:] = oo.w[:, 4:5, 5:6, 6:7]     # one state everywhere
        oo.compute_aux(1, 1)
        oo.euler_fluxes(0, 0, pp.nx)
        want = _flux(oo.w[:, 4, 5, 6], 0, gam)[0]
        assert np.allclose(oo.fhat[:, 4, 5, 6], want, rtol=1e-13, atol=1e-14), split


def _wenorec(fn, vp, vm, rho0, u0, version=0):
    import ctypes as C
    from oracle import lib
    vp = np.ascontiguousarray(vp, dtype=np.float64)
    vm = np.ascontiguousarray(vm, dtype=np.float64)
    out = np.zeros(5)
    dp = C.POINTER(C.c_double)
    getattr(lib(), fn)(vp.ctypes.data_as(dp), vm.ctypes.data_as(dp), out.ctypes.data_as(dp), 3, 3, version,
                       rho0, u0)
    return out


@pytest.mark.parametrize("fn,expo", [("or_wenorec_1d", (2, 2, 4, 4, 2)),
                                     ("or_wenorec_1d_rusanov", (2, 4, 4, 4, 6))])
@pytest.mark.parametrize("version", [0, 1])
def test_weno_smoothness_scaling_per_component(fn, expo, version):
    """The smoothness indicators of component m are scaled by
    1/(rho0^2 u0^expo_m).  WENO5 is invariant under v -> a v with the scale
    -> scale/a^2, so the reconstruction at (rho0, u0) = (1, 2) equals the one
    at (1, 1) of the data times sqrt(scale_m), divided by sqrt(scale_m).
    Data of amplitude ~1e-3 keep beta*scale comparable to the epsilon of the
    weights, where a wrong exponent shows."""
    rng = np.random.default_rng(9)
    vp = 1e-3 * rng.normal(size=(5, 9))
    vm = 1e-3 * rng.normal(size=(5, 9))
    got = _wenorec(fn, vp, vm, 1.0, 2.0, version)
    for m in range(5):
        sq = np.sqrt(1.0 / 2.0 ** expo[m])
        ref = _wenorec(fn, vp * sq, vm * sq, 1.0, 1.0, version)[m] / sq
        assert abs(got[m] - ref) <= 1e-12 * max(abs(ref), 1e-12), (fn, m, got[m], ref)
    # with WENO-JS (epsilon 1e-6) the exponents matter at this amplitude:
    # another exponent gives another value (WENO-Z's epsilon is 1e-35)
    if version == 0:
        alt = _wenorec(fn, vp * 2.0, vm * 2.0, 1.0, 1.0, version) / 2.0
        assert np.abs(alt - got).max() > 1e-9
