"""Cartesian grid, 1-D metrics and seeded initial fields.

Restates the grid setup the hot path consumes:

* uniform x/z nodes -- code/src/grid.F90:796-813 (generate_grid_uniform_x/z)
* channel y grid    -- code/src/grid.F90:431-485 (generate_grid_cha)
* derivative weights-- code/src/grid.F90:143-190 (get_deriv_coeffs)
* metrics           -- code/src/grid.F90:71-141 (compute_metrics)
* local 1-D metrics -- code/src/field.F90:456-516 (set_local_grid, one block)

All arrays are 0-based numpy views of the Fortran arrays: a ghosted
coordinate x(1-ng:n+ng) is stored at index i+ng-1.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def deriv_coeffs(order1: int, order2: int):
    """First/second-derivative central weights, grid.F90:143-190.

    Returns (c1[1..4] as array of 4, c2[0..4] as array of 5)."""
    c1 = np.zeros(4)
    c2 = np.zeros(5)
    h1 = order1 // 2
    if h1 == 1:
        c1[0] = 0.5
    elif h1 == 2:
        c1[:2] = [2.0 / 3.0, -1.0 / 12.0]
    elif h1 == 3:
        c1[:3] = [0.75, -0.15, 1.0 / 60.0]
    elif h1 == 4:
        c1[:4] = [4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0]
    h2 = order2 // 2
    if h2 == 1:
        c2[:2] = [-2.0, 1.0]
    elif h2 == 2:
        c2[:3] = [-2.5, 4.0 / 3.0, -1.0 / 12.0]
    elif h2 == 3:
        c2[:4] = [-245.0 / 90.0, 1.5, -0.15, 1.0 / 90.0]
    elif h2 == 4:
        c2[:5] = [-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0]
    return c1, c2


def uniform_nodes(n: int, ng: int, length: float, periodic: bool) -> np.ndarray:
    """xg(i) = (i-1)*dx for i = 1-ng..n+ng+1 (grid.F90:796-813)."""
    dx = length / n if periodic else length / (n - 1)
    i = np.arange(1 - ng, n + ng + 2, dtype=np.float64)
    return (i - 1.0) * dx


def channel_y(nymax: int, ng: int, jbgrid: float, dyptarget: float, nsmooy: int, retaucha: float):
    """generate_grid_cha (grid.F90:431-485).

    Returns (yn[1..nymax+1], yg[1-ng..nymax+ng]) as 0-based arrays."""
    half = nymax // 2
    yn = np.zeros(nymax + 1)
    alfc = 4.0 / 3.0 * retaucha ** 0.75 * 2.0 / nymax
    for j in range(1, half + 2):
        rj = float(j - 1)
        v = 1.0 / (1.0 + (rj / jbgrid) ** 2)
        v = v * (rj * dyptarget + (0.75 * alfc * rj) ** (4.0 / 3.0) * (rj / jbgrid) ** 2)
        yn[j - 1] = v
    ynf = yn[half]
    yn[: half + 1] /= ynf
    for _ in range(nsmooy):
        ynold = yn[: half + 1].copy()
        for j in range(2, half + 1):
            fsm = 6.0 - 2.0 * (np.tanh((yn[j - 1] - 0.75) / 0.15) + 1.0) * 0.5
            fsl = 0.5 * (6.0 - fsm)
            v = fsl * ynold[j - 2] + fsm * ynold[j - 1] + fsl * ynold[j]
            yn[j - 1] = v / 6.0
    yn[: half + 1] -= 1.0
    for j in range(1, half + 1):
        yn[nymax + 2 - j - 1] = -yn[j - 1]
    yg = np.zeros(nymax + 2 * ng)
    for j in range(1, nymax + 1):
        yg[j + ng - 1] = 0.5 * (yn[j - 1] + yn[j])
    for j in range(1, ng + 1):
        yg[nymax + j + ng - 1] = 2.0 - yg[nymax + 1 - j + ng - 1]
        yg[1 - j + ng - 1] = -2.0 - yg[j + ng - 1]
    return yn, yg


def metrics(xg: np.ndarray, n: int, ng: int, order: int):
    """dxg, d2xg for i = 1..n (grid.F90:71-141); xg is ghosted (index i+ng-1)."""
    c, cc = deriv_coeffs(order, order)
    mm = order // 2
    d1 = np.zeros(n)
    d2 = np.zeros(n)
    for i in range(1, n + 1):
        ii = i + ng - 1
        acc = 0.0
        for l in range(1, mm + 1):
            acc = acc + c[l - 1] * (xg[ii + l] - xg[ii - l])
        d1[i - 1] = acc
        acc2 = cc[0] * xg[ii]
        for l in range(1, mm + 1):
            acc2 = acc2 + cc[l] * (xg[ii + l] + xg[ii - l])
        d2[i - 1] = acc2
    return d1, d2


@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    ng: int
    x: np.ndarray        # x(1-ng:nx+ng)
    y: np.ndarray
    z: np.ndarray
    yn: np.ndarray       # yn(1:ny+1)
    dcsidx: np.ndarray
    dcsidxs: np.ndarray
    dcsidx2: np.ndarray
    detady: np.ndarray
    detadys: np.ndarray
    detady2: np.ndarray
    dzitdz: np.ndarray
    dzitdzs: np.ndarray
    dzitdz2: np.ndarray

    def arrays(self):
        return (self.x, self.y, self.z, self.yn, self.dcsidx, self.dcsidxs, self.dcsidx2,
                self.detady, self.detadys, self.detady2, self.dzitdz, self.dzitdzs, self.dzitdz2)


def local_metrics(d1, d2):
    """set_local_grid (field.F90:456-516): 1/dxg, its square, -d2xg/dxg^2."""
    a = 1.0 / d1
    s = a * a
    return a, s, -d2 * s


def channel_grid(nx, ny, nz, ng=3, lx=4 * np.pi, ly=2.0, lz=4 * np.pi / 3, retau=220.0,
                 jbgrid=32.0, dyptarget=0.8, nsmooy=50, metrics_order=6, uniform_y=False) -> Grid:
    """Channel grid: uniform periodic x/z, cell-centred stretched y."""
    xg = uniform_nodes(nx, ng, lx, True)
    zg = uniform_nodes(nz, ng, lz, True)
    if uniform_y:
        yn = np.linspace(-ly / 2, ly / 2, ny + 1)
        yg = np.zeros(ny + 2 * ng)
        yg[ng:ng + ny] = 0.5 * (yn[:-1] + yn[1:])
        for j in range(1, ng + 1):
            yg[ny + j + ng - 1] = ly - yg[ny + 1 - j + ng - 1]
            yg[1 - j + ng - 1] = -ly - yg[j + ng - 1]
    else:
        yn, yg = channel_y(ny, ng, jbgrid, dyptarget, nsmooy, retau)
    dx1, dx2 = metrics(xg, nx, ng, metrics_order)
    dy1, dy2 = metrics(yg, ny, ng, metrics_order)
    dz1, dz2 = metrics(zg, nz, ng, metrics_order)
    a = local_metrics(dx1, dx2)
    b = local_metrics(dy1, dy2)
    c = local_metrics(dz1, dz2)
    return Grid(nx, ny, nz, ng, xg[: nx + 2 * ng].copy(), yg.copy(), zg[: nz + 2 * ng].copy(), yn.copy(),
                *a, *b, *c)


# ----------------------------------------------------------------- inputs
def _splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014), vectorised on uint64."""
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def counter_uniform(seed: int, index: np.ndarray, stream: int) -> np.ndarray:
    """Uniform [0,1) from a counter-based hash of (seed, global index, stream).

    The value of a cell depends only on its global index, never on the
    decomposition or the traversal order."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0x100000001B3)) ^ (np.uint64(stream) << np.uint64(56))
        h = _splitmix64(index.astype(np.uint64) ^ key)
        h = _splitmix64(h + np.uint64(stream))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def channel_initial_field(grid: Grid, u0: float, noise: float = 0.10, seed: int = 0,
                          rho0: float = 1.0, t0: float = 1.0, rgas0: float = 1.0, cv0: float = 2.5,
                          e_ref: float = None, noise_source: str = "counter") -> np.ndarray:
    """Laminar Poiseuille profile u = 1.5*u_b*(1-y^2) plus seeded noise.

    ``noise`` is the peak-to-peak velocity perturbation as a fraction of u0:
    each component gets noise*u0*(r-0.5), r uniform in [0,1) from
    counter_uniform keyed by (seed, global cell index, component).  Density
    and temperature are uniform (rho0, t0); total energy follows the
    calorically perfect e(T) of get_e_from_temperature (kernels_cpu.F90:7155).
    ``noise_source="pcg"`` draws the noise from numpy's PCG64 stream seeded
    with ``seed`` instead (faster; used for the serving workload).
    Returns w(1-ng:nx+ng, 1-ng:ny+ng, 1-ng:nz+ng, 5) in Fortran order with
    zero ghosts.
    """
    nx, ny, nz, ng = grid.nx, grid.ny, grid.nz, grid.ng
    sx, sy, sz = nx + 2 * ng, ny + 2 * ng, nz + 2 * ng
    w = np.zeros((5, sz, sy, sx))
    yc = grid.y[ng:ng + ny]
    i = np.arange(nx, dtype=np.int64)
    j = np.arange(ny, dtype=np.int64)
    k = np.arange(nz, dtype=np.int64)
    gidx = (k[:, None, None] * ny + j[None, :, None]) * nx + i[None, None, :]
    ub = 1.5 * u0 * (1.0 - yc ** 2)
    if noise_source == "pcg":
        r = np.random.default_rng(seed).random((3, nz, ny, nx))
    else:
        r = [counter_uniform(seed, gidx, m) for m in range(3)]
    uu = np.broadcast_to(ub[None, :, None], (nz, ny, nx)) + noise * u0 * (r[0] - 0.5)
    vv = noise * u0 * (r[1] - 0.5)
    ww = noise * u0 * (r[2] - 0.5)
    rho = rho0
    if e_ref is None:
        # calorically perfect: e = t0*(cv(1) + cv(0)*(T/t0 - 1)), cv(1) = rgas0*gm
        e_ref = cv0 * t0
    ee = e_ref
    sl = (slice(ng, ng + nz), slice(ng, ng + ny), slice(ng, ng + nx))
    w[0][sl] = rho
    w[1][sl] = rho * uu
    w[2][sl] = rho * vv
    w[3][sl] = rho * ww
    w[4][sl] = rho * ee + 0.5 * ((rho * uu) ** 2 + (rho * vv) ** 2 + (rho * ww) ** 2) / rho
    return w
