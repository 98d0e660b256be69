"""Pins of the oracle's boundary conditions of the boundary-layer family.

Ghost faces are checked against index formulas and closed forms; the
characteristic (non-reflecting) treatment against the flux Jacobian built
by finite differences of the Euler flux, with left eigenvectors from
numpy's eigen-decomposition of that Jacobian, so none of the oracle's
eigenvector formulas is retyped here.
"""
import numpy as np
import pytest

from oracle import Oracle
from streams_mi355x.cases import bl_params, make_bl_box
from streams_mi355x.params import params_copy

NG = 3


def _oracle(**kw):
    c = make_bl_box(**kw)
    o = Oracle(c.params, c.grid)
    rng = np.random.default_rng(3)
    w = c.w0.copy()
    w[:, :, :, :] += 0.0
    inner = (slice(NG, -NG),) * 3
    for m in range(5):
        w[m][inner] *= 1.0 + 0.05 * rng.uniform(-1, 1, w[m][inner].shape)
    o.w[:] = w
    o.compute_aux(1, 0)
    return c, o


def test_free_stream_extrapolation_and_symmetry_faces():
    c, o = _oracle(bctags=(1, 2, 5, 2, 0, 0), bctags_nr=(0, 0, 0, 0, 0, 0))
    w0 = o.w.copy()
    o.update_ghost(0)
    w = o.w
    winf = np.array(c.params.winf[:])
    ny, nx = c.params.ny, c.params.nx
    jj, kk = slice(NG, NG + ny), slice(NG, NG + c.params.nz)
    for l in range(1, NG + 1):
        # xmin free stream
        assert np.all(w[:, kk, jj, NG - l] == winf[:, None, None])
        # xmax extrapolation of the last cell
        assert np.array_equal(w[:, kk, jj, NG + nx - 1 + l], w0[:, kk, jj, NG + nx - 1])
        # ymin symmetry: w(1-l) = w(1+l), all five variables, no sign change
        assert np.array_equal(w[:, kk, NG - l, NG:NG + nx], w0[:, kk, NG + l, NG:NG + nx])
    assert np.array_equal(w[:, kk, jj, NG:NG + nx], w0[:, kk, jj, NG:NG + nx])  # interior untouched


def test_node_wall_face():
    c, o = _oracle(bctags=(2, 2, 6, 2, 0, 0), bctags_nr=(0, 0, 0, 0, 0, 0))
    p = c.params
    w0 = o.w.copy()
    o.update_ghost(0)
    w = o.w
    cv0, cv1, tw = p.cv_coeff[0], p.cv_coeff[1], p.t_wall
    e_of = lambda t: cv1 * p.t0 + cv0 * (t - p.t0)  # calorically perfect e(T)
    kk, ii = slice(NG, NG + p.nz), slice(NG, NG + p.nx)
    row1 = w[:, kk, NG, ii]
    assert np.all(row1[1:4] == 0.0)
    assert np.allclose(row1[4], w0[0, kk, NG, ii] * e_of(tw), rtol=1e-15)
    for l in range(1, NG + 1):
        src = w0[:, kk, NG + l, ii]
        rho, u = src[0], src[1:4] / src[0]
        q = 0.5 * np.sum(u * u, axis=0)
        t = (src[4] / rho - q - cv1 * p.t0) / cv0 + p.t0
        pres = rho * t * p.rgas0
        tg = 2.0 * tw - t
        rg = pres / tg / p.rgas0
        g = w[:, kk, NG - l, ii]
        assert np.allclose(g[0], rg, rtol=1e-13)
        assert np.allclose(g[1:4], -rg * u, rtol=1e-13, atol=1e-14)
        assert np.allclose(g[4], rg * (e_of(tg) + q), rtol=1e-13)


def test_oblique_shock_face():
    dx = 8.0 / 39
    c, o = _oracle(bctags=(1, 2, 6, 7, 0, 0), bctags_nr=(0, 0, 0, 0, 0, 0), shock_angle_deg=35.0,
                   xshock_imp=5.0, tanhfacs=dx)
    p, gr = c.params, c.grid
    w0 = o.w.copy()
    o.update_ghost(0)
    w = o.w
    ny, nx = p.ny, p.nx
    winf, wps = np.array(p.winf[:]), np.array(p.winf_past_shock[:])
    x = gr.x[NG:NG + nx]
    kk = slice(NG, NG + p.nz)
    near_any = False
    for l in range(0, NG + 1):
        xsh = p.xshock_imp - gr.y[NG + ny - 1 + l] / np.tan(p.shock_angle)
        xx = x - xsh
        near = np.abs(xx) <= 8.0 * p.tanhfacs
        near_any |= near.any()
        blend = winf[:, None] + (wps - winf)[:, None] * (0.5 * (1.0 + np.tanh(xx / p.tanhfacs)))[None, :]
        got = w[:, kk, NG + ny - 1 + l, NG:NG + nx]
        # far from the trace: row ny (possibly already reset at l = 0)
        far_ref = w[:, kk, NG + ny - 1, NG:NG + nx]
        assert np.allclose(got[:, :, near], blend[:, None, near], rtol=1e-14, atol=1e-14)
        assert np.array_equal(got[:, :, ~near], far_ref[:, :, ~near])
    assert near_any
    # a shock state that actually differs from the free stream
    assert np.abs(wps - winf).max() > 0.1


# ----------------------------------------------------- characteristic faces
def _euler_flux(wv, d, gam, rgas=1.0):
    rho, mom, E = wv[0], wv[1:4], wv[4]
    u = mom / rho
    pres = (gam - 1.0) * (E - 0.5 * rho * np.dot(u, u))
    f = np.empty(5)
    f[0] = mom[d]
    f[1:4] = mom * u[d]
    f[1 + d] += pres
    f[4] = (E + pres) * u[d]
    return f


def _jacobian(wv, d, gam):
    A = np.empty((5, 5))
    for q in range(5):
        h = 1e-6 * max(1.0, abs(wv[q]))
        a, b = wv.copy(), wv.copy()
        a[q] += h
        b[q] -= h
        A[:, q] = (_euler_flux(a, d, gam) - _euler_flux(b, d, gam)) / (2 * h)
    return A


def _nr_case(face, nr_type, mach, **kw):
    """Oracle on a box whose w is linear along the normal of ``face`` near
    it (so the one-sided second-order derivative is exact); returns the
    boundary states, their normal derivatives and the fl increment."""
    nr = [0] * 6
    nr[face] = nr_type
    c, o = _oracle(bctags=(1, 2, 6, 2, 0, 0), bctags_nr=tuple(nr), mach=mach, **kw)
    p, gr = c.params, c.grid
    d = face // 2
    rng = np.random.default_rng(7)
    # free-stream state with its velocity turned along the face normal
    winf = np.array(p.winf[:])
    base = winf.copy()
    base[1:4] = 0.0
    base[1 + d] = winf[1]
    coords = [gr.x, gr.y, gr.z][d]
    span = coords.max() - coords.min()
    slope = 0.02 * rng.uniform(-1, 1, 5) * (np.abs(base) + 0.1 * np.abs(base).max()) / span
    nax = 3 - d  # array axis of direction d in [m, k, j, i]
    shape = [1, 1, 1, 1]
    shape[nax] = -1
    wlin = base[:, None, None, None] + slope[:, None, None, None] * coords.reshape(shape)
    o.w[:] = np.broadcast_to(wlin, o.w.shape)
    # small transverse variation so that every boundary point differs
    o.w[:] *= 1.0 + 0.01 * rng.uniform(-1, 1, o.w.shape)
    o.w[:] = o.w  # keep the view
    o.compute_aux(1, 1)
    o.zero_flux()
    o.bc_nr()
    return c, o, slope, d


@pytest.mark.parametrize("face", [1, 3])
def test_nr_supersonic_outflow_is_the_flux_derivative(face):
    """Type 1 at an outflow where every wave leaves: the correction equals
    A(w_b) dw/dn of the one-sided derivative (all waves kept)."""
    c, o, slope, d = _nr_case(face, 1, mach=3.0)
    p = c.params
    gam = p.cp_coeff[0] / p.cv_coeff[0]
    n = [p.nx, p.ny, p.nz][d]
    metric = [c.grid.dcsidx, c.grid.detady, c.grid.dzitdz][d]
    checked = 0
    for (k, j, i) in [(2, 5, 7), (5, 9, 11), (10, 3, 20)]:
        idx = [i, j, k]
        idx[d] = n  # boundary cell (1-based), end face
        ii, jj, kk = idx
        wb = o.w[:, NG + kk - 1, NG + jj - 1, NG + ii - 1]
        ub = wb[1 + d] / wb[0]
        cs = np.sqrt(gam * (gam - 1.0) * (wb[4] / wb[0] - 0.5 * np.sum((wb[1:4] / wb[0]) ** 2)))
        if ub - cs <= 0:
            continue
        # one-sided derivative in index space, then by the metric
        line = [o.w[:, NG + kk - 1, NG + jj - 1, NG + ii - 1 - q] if d == 0 else
                (o.w[:, NG + kk - 1, NG + jj - 1 - q, NG + ii - 1] if d == 1 else
                 o.w[:, NG + kk - 1 - q, NG + jj - 1, NG + ii - 1]) for q in range(3)]
        dwdn = 1.5 * line[0] - 2.0 * line[1] + 0.5 * line[2]
        ref = _jacobian(wb.copy(), d, gam) @ dwdn * metric[n - 1]
        got = o.fl[:, kk - 1, jj - 1, ii - 1]
        assert np.allclose(got, ref, rtol=1e-6, atol=1e-9 * np.abs(ref).max()), (got, ref)
        checked += 1
    assert checked > 0


def _char_components(c, o, d, face, ii, jj, kk):
    p = c.params
    gam = p.cp_coeff[0] / p.cv_coeff[0]
    wb = o.w[:, NG + kk - 1, NG + jj - 1, NG + ii - 1]
    A = _jacobian(wb.copy(), d, gam)
    lam, V = np.linalg.eig(A)
    lam, V = lam.real, V.real
    L = np.linalg.inv(V)
    start = face % 2 == 0
    sgn = 1 if start else -1
    st = [(0, 0, sgn), (0, sgn, 0), (sgn, 0, 0)][d]
    pts = [o.w[:, NG + kk - 1 + q * st[0], NG + jj - 1 + q * st[1], NG + ii - 1 + q * st[2]] for q in range(3)]
    dwdn = sgn * (-1.5 * pts[0] + 2.0 * pts[1] - 0.5 * pts[2])
    metric = [c.grid.dcsidx, c.grid.detady, c.grid.dzitdz][d]
    n = [p.nx, p.ny, p.nz][d]
    df = o.fl[:, kk - 1, jj - 1, ii - 1] / metric[(1 if start else n) - 1]
    return lam, L, V, dwdn, df, sgn


@pytest.mark.parametrize("face", [0, 1, 3])
def test_nr_subsonic_keeps_only_outgoing_waves(face):
    """Type 1 at a subsonic boundary: each outgoing characteristic component
    of the correction is lambda times that of dw/dn, each entering one is
    zero."""
    c, o, slope, d = _nr_case(face, 1, mach=0.5)
    p = c.params
    n = [p.nx, p.ny, p.nz][d]
    for (k, j, i) in [(3, 6, 9), (8, 10, 15)]:
        idx = [i, j, k]
        idx[d] = 1 if face % 2 == 0 else n
        lam, L, V, dwdn, df, sgn = _char_components(c, o, d, face, *idx)
        a_df, a_dw = L @ df, L @ dwdn
        scale = np.abs(lam).max() * np.abs(a_dw).max()
        for q in range(5):
            if sgn * lam[q] < 0:   # leaves the domain through this face
                assert abs(a_df[q] - lam[q] * a_dw[q]) <= 1e-6 * scale
            else:
                assert abs(a_df[q]) <= 1e-6 * scale


def test_nr_relaxation_uses_the_outer_difference_for_entering_waves():
    """Type 3 (relaxation to winf) at a subsonic outflow: the entering wave
    takes lambda times the characteristic component of sgn*(w - winf)."""
    face = 1
    c, o, slope, d = _nr_case(face, 3, mach=0.5)
    p = c.params
    winf = np.array(p.winf[:])
    lam, L, V, dwdn, df, sgn = _char_components(c, o, d, face, p.nx, 7, 4)
    wb = o.w[:, NG + 4 - 1, NG + 7 - 1, NG + p.nx - 1]
    outer = sgn * (wb - winf)
    a_df, a_dw, a_out = L @ df, L @ dwdn, L @ outer
    scale = np.abs(lam).max() * max(np.abs(a_dw).max(), np.abs(a_out).max())
    for q in range(5):
        want = lam[q] * (a_out[q] if sgn * lam[q] > 0 else a_dw[q])
        assert abs(a_df[q] - want) <= 1e-6 * scale


def test_nr_reflective_wall_keeps_only_acoustic_waves():
    """Type 6 on ymin: the correction lies in the span of the two acoustic
    right eigenvectors."""
    face = 2
    c, o, slope, d = _nr_case(face, 6, mach=0.5)
    lam, L, V, dwdn, df, sgn = _char_components(c, o, d, face, 9, 1, 5)
    un = lam[np.argsort(lam)][2]
    a_df = L @ df
    for q in range(5):
        if abs(lam[q] - un) < 1e-8 * np.abs(lam).max():
            assert abs(a_df[q]) <= 1e-8 * np.abs(a_df).max()
    # the reflected acoustic wave has the amplitude of the incident one, so
    # the correction is parallel to r(v-c) + r(v+c) = 2 (1, u, v, w, H)
    wb = o.w[:, NG + 5 - 1, NG + 1 - 1, NG + 9 - 1]
    p = c.params
    gam = p.cp_coeff[0] / p.cv_coeff[0]
    u = wb[1:4] / wb[0]
    pres = (gam - 1.0) * (wb[4] - 0.5 * wb[0] * np.dot(u, u))
    direction = np.array([1.0, u[0], u[1], u[2], (wb[4] + pres) / wb[0]])
    assert np.abs(np.cross(df[[0, 2, 4]], direction[[0, 2, 4]])).max() <= 1e-8 * np.abs(df).max() * np.abs(direction).max()
    assert np.allclose(df / df[0], direction, rtol=1e-8, atol=1e-10)


# --------------------------------------------------- recycling-rescaling inflow
def _distinct_planes(c):
    """The mean profile with each of the ng inflow planes scaled apart
    (plane 1-ng+q by 1 + 0.01 (q+1)), so a mirrored plane order shows."""
    from streams_mi355x.cases import bl_mean_planes

    wm = bl_mean_planes(c).copy()
    for q in range(NG):
        wm[:, :, q] *= 1.0 + 0.01 * (q + 1)
    return wm


def _recycling_case(same_scaling=False, uniform=False, seed=5, beta=None, rand_type=0):
    from streams_mi355x.cases import recycling_setup

    c = make_bl_box(bctags=(10, 2, 6, 2, 0, 0), bctags_nr=(0, 1, 0, 1, 0, 0), seed=seed, rand_type=rand_type)
    if same_scaling:
        r = recycling_setup(c.grid, 5.0, 0.6, 0.01, 0.6, 0.01, _distinct_planes(c))
    else:
        r = recycling_setup(c.grid, 5.0, 0.6, 0.01, 0.7, 0.011, _distinct_planes(c))
    if beta is not None:
        r.betarecyc = beta
    o = Oracle(c.params, c.grid)
    w = c.w0.copy()
    if uniform:  # spanwise-uniform field
        w[:] = w[:, NG:NG + 1, :, :]
    else:  # density fluctuations at fixed velocity and temperature
        w *= 1.0 + 0.04 * (np.random.default_rng(seed).random(w.shape[1:]) - 0.5)
    o.w[:] = w
    o.compute_aux(1, 1)
    o.set_recycling(r)
    return c, r, o


def test_recycling_uniform_station_gives_the_mean_profile():
    c, r, o = _recycling_case(uniform=True)
    o.update_ghost(0)
    p = c.params
    p0 = p.rho0 * p.rgas0 * p.t0
    e_of = lambda t: p.cv_coeff[1] * p.t0 + p.cv_coeff[0] * (t - p.t0)
    for i in range(1, NG + 1):
        g = o.w[:, NG:-NG, NG:-NG, NG - i]          # ghost plane 1-i: [m, k, j]
        rho = r.wmean[0, :, NG - i]
        u = r.wmean[1:4, :, NG - i] / rho
        assert np.allclose(g[0], rho[None, :], rtol=1e-14)
        assert np.allclose(g[1:4], (rho * u)[:, None, :], rtol=1e-13, atol=1e-15)
        tt = p0 / rho / p.rgas0
        want_e = rho * e_of(tt) + 0.5 * rho * np.sum(u * u, axis=0)
        assert np.allclose(g[4], want_e[None, :], rtol=1e-13)


def test_recycling_density_fluctuation_has_zero_spanwise_mean():
    c, r, o = _recycling_case()
    o.update_ghost(0)
    for i in range(1, NG + 1):
        rho_g = o.w[0, NG:-NG, NG:-NG, NG - i]       # [k, j]
        assert np.allclose(rho_g.mean(axis=0), r.wmean[0, :, NG - i], rtol=1e-12)


def test_recycling_shifts_the_span_by_half():
    c, r, o = _recycling_case(same_scaling=True, uniform=True)
    p = c.params
    nz = p.nz
    kb, jb = 3, 9                                   # bump at station cell (i_recyc, jb, kb)
    o.w[2, NG + kb - 1, NG + jb - 1, NG + r.i_recyc - 1] += 0.05 * o.w[0, NG + kb - 1, NG + jb - 1, NG + r.i_recyc - 1]
    o.update_ghost(0)
    rv = o.w[2, NG:-NG, NG:-NG, NG - 1] / o.w[0, NG:-NG, NG:-NG, NG - 1]     # v of ghost plane 0, [k, j]
    dev = np.abs(rv - rv.mean(axis=0))
    k_hit = int(np.argmax(dev[:, jb - 1])) + 1
    assert k_hit == (kb + nz // 2 - 1) % nz + 1


def test_recycling_identical_scaling_copies_the_damped_fluctuation():
    c, r, o = _recycling_case(same_scaling=True, beta=1.7)
    p = c.params
    assert np.array_equal(r.map_inn[1:], np.arange(1, p.ny))  # y maps onto itself
    w0 = o.w.copy()
    o.update_ghost(0)
    nz, ny = p.nz, p.ny
    eta = r.eta_inflow[NG:NG + ny]
    bdamp = 0.5 * (1.0 - np.tanh(4.0 * (eta - 2.0)))
    for i in range(1, NG + 1):
        st = w0[:, NG:-NG, NG:-NG, NG + r.i_recyc - i]                   # station plane [m, k, j]
        st = np.roll(st, nz // 2, axis=1)                                # half-span shift
        ufav = st[1].sum(axis=0) / st[0].sum(axis=0)
        uf = st[1] / st[0] - ufav[None, :]
        g = o.w[:, NG:-NG, NG:-NG, NG - i]
        rho_m = r.wmean[0, :, NG - i]
        u_g = g[1] / g[0]
        want = r.wmean[1, :, NG - i] / rho_m + uf * bdamp[None, :] * r.betarecyc
        jj = slice(1, ny - 1)   # rows with a full interpolation stencil
        assert np.allclose(u_g[:, jj], want[:, jj], rtol=1e-12, atol=1e-13)
        rf = st[0] - st[0].mean(axis=0)[None, :]                         # density: damped, not rescaled by beta
        want_rho = rho_m + np.clip(rf * bdamp[None, :], -0.75 * rho_m, 4.0 * rho_m)
        assert np.allclose(g[0][:, jj], want_rho[:, jj], rtol=1e-13)


def test_recycling_inflow_has_the_reference_pressure():
    """The inflow ghosts carry T = p0/(rho R), i.e. the pressure p0."""
    c, r, o = _recycling_case()
    o.update_ghost(0)
    p = c.params
    gam = p.cp_coeff[0] / p.cv_coeff[0]
    g = o.w[:, NG:-NG, NG:-NG, 0:NG]
    pres = (gam - 1.0) * (g[4] - 0.5 * (g[1] ** 2 + g[2] ** 2 + g[3] ** 2) / g[0])
    assert np.allclose(pres, p.rho0 * p.rgas0 * p.t0, rtol=1e-12)


# ------------------------------------------------ random inflow plane (rand_type)
def _jbl(c):
    """locateval(y(1:ny), ny, l0): rows strictly below l0."""
    y = c.grid.y[NG:NG + c.params.ny]
    return int(np.sum(y < c.params.l0))


def test_random_plane_follows_the_c_library_stream_in_get_crandom_order():
    """After run()'s seeding (seed rank+1 = 1) the first z sweep redraws
    inflow_random_plane(2:jbl, 1:nz, 1:3) from rand()/RAND_MAX with j
    outermost and the component innermost (get_crandom_f_array3d); rows 1
    and above jbl keep 0.5.  The stream is the C library's srand/rand,
    called directly here."""
    from streams_mi355x.cases import crandom_stream

    c, r, o = _recycling_case(rand_type=2)
    o.init_run()
    plane = o.inflow_plane()
    assert np.all(plane == 0.5)
    jbl = _jbl(c)
    nz = c.params.nz
    assert 2 <= jbl < c.params.ny
    o.euler_z()
    want = crandom_stream(1, 3 * (jbl - 1) * nz).reshape(jbl - 1, nz, 3)   # [j, k, c]
    got = o.inflow_plane()[:, :, 1:jbl].transpose(2, 1, 0)                  # [j, k, c]
    assert np.array_equal(got, want)
    assert np.all(o.inflow_plane()[:, :, 0] == 0.5)
    assert np.all(o.inflow_plane()[:, :, jbl:] == 0.5)
    o.euler_z()  # the next sweep continues the same stream
    want2 = crandom_stream(1, 6 * (jbl - 1) * nz)[3 * (jbl - 1) * nz:].reshape(jbl - 1, nz, 3)
    assert np.array_equal(o.inflow_plane()[:, :, 1:jbl].transpose(2, 1, 0), want2)


@pytest.mark.parametrize("uniform", [True, False])
def test_random_plane_adds_two_percent_of_u0_after_the_rescaling(uniform):
    """With identical inner/outer scalings the inflow velocity is the mean,
    plus beta*bdamp times the station fluctuation, plus
    0.02 u0 (r - 0.5) min(eta, 1) with r the plane's draw: the random term
    is added after the rescaling, is centred (r - 0.5) and its amplitude is
    at most 1 % of u0."""
    c, r, o = _recycling_case(same_scaling=True, uniform=uniform, beta=1.7, rand_type=2)
    p = c.params
    nz, ny = p.nz, p.ny
    o.init_run()
    o.euler_z()
    w0 = o.w.copy()
    o.update_ghost(0)
    plane = o.inflow_plane()                       # [c, k, j]
    eta = r.eta_inflow[NG:NG + ny]
    bdamp = 0.5 * (1.0 - np.tanh(4.0 * (eta - 2.0)))
    jj = slice(1, ny - 1)
    for i in range(1, NG + 1):
        st = np.roll(w0[:, NG:-NG, NG:-NG, NG + r.i_recyc - i], nz // 2, axis=1)
        g = o.w[:, NG:-NG, NG:-NG, NG - i]
        rho_m = r.wmean[0, :, NG - i]
        for m in range(3):
            fav = st[1 + m].sum(axis=0) / st[0].sum(axis=0)
            uf = st[1 + m] / st[0] - fav[None, :]
            rnd = 0.02 * p.u0 * (plane[m] - 0.5) * np.minimum(eta, 1.0)[None, :]
            want = r.wmean[1 + m, :, NG - i] / rho_m + uf * bdamp[None, :] * r.betarecyc + rnd
            assert np.allclose((g[1 + m] / g[0])[:, jj], want[:, jj], rtol=1e-12, atol=1e-13), (i, m)
            assert np.max(np.abs(rnd)) <= 0.01 * p.u0
    assert abs(np.mean(plane[:, :, 1:_jbl(c)]) - 0.5) < 0.02


# ----------------------------------------------------- ghost tags 3, 4, 8, 9
def test_subsonic_inflow_from_total_conditions():
    """Tag 3 on xmin and ymax (bcfree_sub): the ghosts hold the boundary
    cell's pressure, the total temperature and pressure of the free stream
    through the isentropic relations, and a velocity along the angle of
    attack; w = 0 on the span."""
    c, o = _oracle(bctags=(3, 2, 6, 3, 0, 0), bctags_nr=(0, 0, 0, 0, 0, 0), aoa_deg=3.0, mach=0.5)
    p = c.params
    gam = p.cp_coeff[0] / p.cv_coeff[0]
    w0 = o.w.copy()  # the wall face, processed after xmin, changes the row j = 1 afterwards
    o.update_ghost(0)
    for face in ("xmin", "ymax"):
        if face == "xmin":
            g = o.w[:, NG:-NG, NG:-NG, 0:NG]
            b = w0[:, NG:-NG, NG:-NG, NG:NG + 1]
        else:
            g = o.w[:, NG:-NG, -NG:, NG:-NG]
            b = w0[:, NG:-NG, -NG - 1:-NG, NG:-NG]
        pres_b = (gam - 1.0) * (b[4] - 0.5 * (b[1] ** 2 + b[2] ** 2 + b[3] ** 2) / b[0])
        u, v, w = g[1] / g[0], g[2] / g[0], g[3] / g[0]
        pres = (gam - 1.0) * (g[4] - 0.5 * g[0] * (u * u + v * v + w * w))
        T = pres / (g[0] * p.rgas0)
        mach2 = (u * u + v * v) / (gam * p.rgas0 * T)
        fac = 1.0 + 0.5 * (gam - 1.0) * mach2
        assert np.allclose(pres, np.broadcast_to(pres_b, pres.shape), rtol=1e-12), face
        assert np.allclose(T * fac, p.ttot0, rtol=1e-12), face
        assert np.allclose(pres * fac ** (gam / (gam - 1.0)), p.ptot0, rtol=1e-12), face
        assert np.allclose(np.arctan2(v, u), p.aoa, rtol=0, atol=1e-12), face
        assert np.all(w == 0.0)


def test_extrapolation_at_the_reference_pressure_and_untouched_ymin():
    """Tag 4 (bcextr_sub): density and momentum of the boundary cell and the
    pressure p0 = rho0 R t0 on every face but ymin, whose ghosts it leaves
    alone."""
    c, o = _oracle(bctags=(4, 4, 4, 4, 0, 0), bctags_nr=(0, 0, 0, 0, 0, 0))
    p = c.params
    gam = p.cp_coeff[0] / p.cv_coeff[0]
    p0 = p.rho0 * p.rgas0 * p.t0
    o.w[:, :, 0:NG, :] = 7.0
    o.update_ghost(0)
    w = o.w
    assert np.all(w[:, NG:-NG, 0:NG, NG:-NG] == 7.0)                # ymin untouched
    for sl_g, sl_b in (((slice(NG, -NG), slice(NG, -NG), slice(0, NG)), (slice(NG, -NG), slice(NG, -NG), slice(NG, NG + 1))),
                       ((slice(NG, -NG), slice(NG, -NG), slice(-NG, None)), (slice(NG, -NG), slice(NG, -NG), slice(-NG - 1, -NG))),
                       ((slice(NG, -NG), slice(-NG, None), slice(NG, -NG)), (slice(NG, -NG), slice(-NG - 1, -NG), slice(NG, -NG)))):
        g = w[(slice(None),) + sl_g]
        b = w[(slice(None),) + sl_b]
        for m in range(4):
            assert np.array_equal(g[m], np.broadcast_to(b[m], g[m].shape))
        pres = (gam - 1.0) * (g[4] - 0.5 * (g[1] ** 2 + g[2] ** 2 + g[3] ** 2) / g[0])
        assert np.allclose(pres, p0, rtol=1e-12)


def test_laminar_inflow_from_wmean_and_no_op_faces():
    """Tag 9 on xmin (bclam): ghost plane 1-l takes wmean(1-l, j, 1:4) at
    pressure p0.  Tag 8, and tag 9 on another face, leave the ghosts alone."""
    c, o = _oracle(bctags=(9, 8, 6, 9, 0, 0), bctags_nr=(0, 0, 0, 0, 0, 0))
    p = c.params
    gam = p.cp_coeff[0] / p.cv_coeff[0]
    wm = _distinct_planes(c)
    o.set_wmean(wm)
    o.w[:, :, :, -NG:] = 5.0
    o.w[:, :, -NG:, NG:-NG] = 6.0
    o.update_ghost(0)
    assert np.all(o.w[:, :, :, -NG:] == 5.0)
    assert np.all(o.w[:, NG:-NG, -NG:, NG:-NG] == 6.0)
    for l in range(1, NG + 1):
        g = o.w[:, NG:-NG, NG:-NG, NG - l]                          # [m, k, j]
        for m in range(4):
            assert np.array_equal(g[m], np.broadcast_to(wm[m, :, NG - l][None, :], g[m].shape)), (l, m)
        pres = (gam - 1.0) * (g[4] - 0.5 * (g[1] ** 2 + g[2] ** 2 + g[3] ** 2) / g[0])
        assert np.allclose(pres, p.rho0 * p.rgas0 * p.t0, rtol=1e-12)
