"""The measured workloads at their full sizes on the GPU.

* Channel, 512 x 256 x 256, Re_tau 550, bulk Mach 1.5.
* Supersonic boundary layer with recycling-rescaling inflow, 1024 x 256 x 192.
* Shock / boundary-layer interaction, 1536 x 320 x 192, 8 deg wedge shock.
* Serving stream: 64 seeded channel solves, 64^3 to 256 x 128 x 128, 200 steps each,
  served plain and with batched launches; and 128 small solves (32-64 cells a
  side) served with batched launches.

The oracle cannot run these sizes in a test, so each workload is checked
in two parts:
* The same physics and boundary set on a reduced grid, against the oracle
  for 100 steps (200 for the served requests), to the FP64 tolerance of
  1e-10 relative L2 per conserved variable.
* The full size on the GPU alone, through properties that do not depend on
  size: finite fields, bitwise reproducibility, the bulk density that the
  forcing holds, the WENO region of the shock, served = plain.

The mixed form's error against FP64 after 100 steps is stated for every
workload. It is bounded by the switched scheme's own sensitivity: the FP64
form started from a field perturbed by 1e-7 (relative, seeded), the size of
an FP32 rounding, and run for the same 100 steps. Each conserved variable's
mixed error must stay within 4 times that spread or within 1e-5, whichever
is larger.
"""
import numpy as np
import pytest

from helpers import interior, rel_l2
from oracle import Oracle
from streams_mi355x import PREC_MIXED, make_channel
from streams_mi355x.cases import bl_recycling_case, shock_angle_from_deflection
from streams_mi355x.params import params_copy
from streams_mi355x.solver import ChannelSolver

pytestmark = pytest.mark.gpu

TRAJ_TOL = 1e-10
MIXED_FLOOR = 1e-5
MIXED_SPREAD_FACTOR = 4.0
WEDGE_DEG = 8.0


def _perturbed(w0, seed=1):
    return w0 * (1.0 + 1e-7 * (np.random.default_rng(seed).random(w0.shape) - 0.5))


def _run(params, grid, w0, steps, recycling=None, precision=None):
    p = params
    if precision is not None:
        p = params_copy(params)
        p.precision = precision
    with ChannelSolver(p, grid) as g:
        if recycling is not None:
            g.set_recycling(recycling)
        g.copy_from_field(w0)
        g.init_run()
        g.advance(steps)
        sc = g.compute_residual()
        weno = g.count_weno()
        w = interior(g.copy_to_field())
    return w, sc, weno


def _mixed_vs_spread(case, steps=100, recycling=None, label=""):
    """FP64, FP64 from the perturbed field, and mixed, all from the same
    initial field; returns (fp64 field, fp64 scalars, WENO fractions,
    mixed errors, spread)."""
    w64, sc, weno = _run(case.params, case.grid, case.w0, steps, recycling, precision=0)
    wp, _, _ = _run(case.params, case.grid, _perturbed(case.w0), steps, recycling, precision=0)
    spread = [rel_l2(wp[m], w64[m]) for m in range(5)]
    del wp
    wm, _, _ = _run(case.params, case.grid, case.w0, steps, recycling, precision=PREC_MIXED)
    err = [rel_l2(wm[m], w64[m]) for m in range(5)]
    del wm
    print(f"{label}: mixed rel L2 vs FP64 after {steps} steps {err}; FP64 spread under 1e-7 {spread}; "
          f"WENO fractions {list(weno)}")
    for m in range(5):
        assert err[m] <= max(MIXED_FLOOR, MIXED_SPREAD_FACTOR * spread[m]), (label, m, err, spread)
    return w64, sc, weno, err, spread


def _oracle_pair(case, steps, recycling=None):
    o = Oracle(case.params, case.grid)
    o.w[:] = case.w0
    if recycling is not None:
        o.set_recycling(recycling)
    o.init_run()
    o.advance(steps)
    wg, _, weno = _run(case.params, case.grid, case.w0, steps, recycling)
    errs = [rel_l2(wg[m], interior(o.w)[m]) for m in range(5)]
    return errs, weno, o.count_weno()


# ---------------------------------------------------------------- channel

def test_channel_retau550_reduced_grid_matches_oracle_100_steps():
    c = make_channel(nx=64, ny=64, nz=32, retau=550.0)
    errs, weno, weno_o = _oracle_pair(c, 100)
    print("channel Re_tau 550, 64x64x32, 100 steps: rel L2", errs)
    assert max(errs) <= TRAJ_TOL, errs
    assert np.array_equal(weno, weno_o)


def test_channel_512x256x256_full_size():
    c = make_channel(nx=512, ny=256, nz=256, retau=550.0)
    w64, sc, weno, err, spread = _mixed_vs_spread(c, 100, label="channel 512x256x256")
    assert np.isfinite(w64).all()
    # the constant-flow-rate forcing holds the bulk density at its initial value
    assert abs(sc["rhobulk"] - 1.0) <= 1e-12, sc["rhobulk"]
    # bitwise reproducible at full size
    w2, sc2, _ = _run(c.params, c.grid, c.w0, 100)
    assert np.array_equal(w2, w64)
    assert sc2["residual_rhou"] == sc["residual_rhou"]


# --------------------------------------------------------- boundary layer

def test_boundary_layer_reduced_grid_matches_oracle_100_steps():
    """The boundary set of the full-size box (recycling inflow with the
    random inflow plane, outflow and top with the characteristic treatment,
    isothermal wall) on 128 x 48 x 16 over 8 x 3 x 1."""
    c, rc = bl_recycling_case(128, 48, 16, lx=8.0, ly=3.0, lz=1.0, x_recyc=2.5)
    errs, _, _ = _oracle_pair(c, 100, rc)
    print("boundary layer 128x48x16, 100 steps: rel L2", errs)
    assert max(errs) <= TRAJ_TOL, errs


def test_boundary_layer_1024x256x192_full_size():
    c, rc = bl_recycling_case(1024, 256, 192)
    w64, sc, weno, err, spread = _mixed_vs_spread(c, 100, rc, label="boundary layer 1024x256x192")
    assert np.isfinite(w64).all()
    w2, _, _ = _run(c.params, c.grid, c.w0, 100, rc)
    assert np.array_equal(w2, w64)


def test_boundary_layer_mean_profiles_2000_steps():
    """Mean profiles of the 1024 x 256 x 192 box after 2000 steps: the
    spanwise and time averages of rho, u and T (compute_stats on the device,
    one sample every 20 steps over the last 1000 steps), mixed against FP64,
    within 4 times the spread of FP64 runs whose initial fields differ by 1e-7,
    or within 1e-4."""
    c, rc = bl_recycling_case(1024, 256, 192)
    first, every, nsamp = 1000, 20, 50
    means = []
    for prec, w0 in ((0, c.w0), (0, _perturbed(c.w0)), (PREC_MIXED, c.w0)):
        p = params_copy(c.params)
        p.precision = prec
        with ChannelSolver(p, c.grid) as g:
            g.set_recycling(rc)
            g.copy_from_field(w0)
            g.init_run()
            g.advance(first)
            for n in range(nsamp):
                g.advance(every)
                g.compute_stats(n)
            st = g.get_stats()
            assert np.isfinite(st).all()
            means.append(st[[0, 1, 5]].copy())   # rho, u, T over (x, y)
    names = ("rho", "u", "T")
    spread = [rel_l2(means[1][q], means[0][q]) for q in range(3)]
    err = [rel_l2(means[2][q], means[0][q]) for q in range(3)]
    print("boundary layer mean profiles after 2000 steps, mixed vs FP64:", dict(zip(names, err)),
          "FP64 spread under 1e-7:", dict(zip(names, spread)))
    for q in range(3):
        assert err[q] <= max(1e-4, MIXED_SPREAD_FACTOR * spread[q]), (names[q], err, spread)


# ----------------------------------------------- shock / boundary layer

def test_sbli_reduced_grid_matches_oracle_100_steps():
    """The shock / boundary-layer boundary set with the incident shock of an
    8 deg wedge already across the field, on 160 x 48 x 16 over 8 x 3 x 1:
    a shock-dominated WENO region (several per cent of the faces)."""
    beta = shock_angle_from_deflection(2.0, WEDGE_DEG)
    c, rc = bl_recycling_case(160, 48, 16, lx=8.0, ly=3.0, lz=1.0, x_recyc=2.5, shock_angle_deg=beta,
                              xshock_imp=5.0, developed_shock=True)
    errs, weno, weno_o = _oracle_pair(c, 100, rc)
    print("SBLI 160x48x16, 100 steps: rel L2", errs, "WENO fractions", list(weno))
    assert max(errs) <= TRAJ_TOL, errs
    assert np.array_equal(weno, weno_o)
    assert min(weno) >= 0.01, weno


def test_sbli_1536x320x192_full_size():
    beta = shock_angle_from_deflection(2.0, WEDGE_DEG)
    c, rc = bl_recycling_case(1536, 320, 192, shock_angle_deg=beta, xshock_imp=40.0, developed_shock=True)
    w64, sc, weno, err, spread = _mixed_vs_spread(c, 100, rc, label="SBLI 1536x320x192")
    assert np.isfinite(w64).all()
    # the shock crosses every x-y plane: WENO faces in all three directions
    assert min(weno) > 0.0 and max(weno) >= 1e-3, weno


# ---------------------------------------------------------------- serving

def test_serving_stream_64_requests():
    """The seeded 64-request stream (grids uniform in [64,256] x [64,128] x
    [64,128], bulk Mach 0.3-3.0, Re_tau 100-500, Poisson arrivals, 200 steps
    each) served in FP64: every completed request equals its plain FP64 solve
    bit for bit, and a request fails in the mix exactly when it fails alone.
    The two smallest completed requests match the oracle after their 200
    steps.  The same stream served in the mixed form: each request's error
    against its FP64 result is stated; the median must stay within 1e-4."""
    from streams_mi355x.serve import run_sequential, run_serving, serving_workload

    reqs = serving_workload(64, 200, rate=16.0, seed=0)
    srv = run_serving(reqs, hbm_bytes=64 << 30, max_slots=8, steps_per_graph=10, autotune=True, want_fields=True)
    seq = run_sequential(reqs, want_fields=True)
    ok = []
    for n, (a, b) in enumerate(zip(seq["results"], srv["results"])):
        assert (a["error"] is None) == (b["error"] is None), n
        if a["error"] is None:
            assert np.array_equal(interior(a["w"]), interior(b["w"])), n
            ok.append(n)
    print(f"served {len(ok)} of {len(reqs)} requests; p50 {srv['p50_s']:.2f} s, p99 {srv['p99_s']:.2f} s")
    assert len(ok) >= 56
    # the same stream with batched launches (2 lanes, solves up to 2^20 cells
    # batched).  Its small solves seldom share a lane while their class is
    # tuned, so how many chunks go out batched depends on the timing (none,
    # in some runs); the batched path itself is covered by test_gpu_serving.
    srv_b = run_serving(reqs, hbm_bytes=64 << 30, max_slots=8, steps_per_graph=10, autotune=True, want_fields=True,
                        batch=2, batch_max_cells=1 << 20)
    print("batched serving:", {k: v for k, v in srv_b["stats"].items() if k.startswith("batch")})
    for n, (a, b) in enumerate(zip(seq["results"], srv_b["results"])):
        assert (a["error"] is None) == (b["error"] is None), n
        if a["error"] is None:
            assert np.array_equal(interior(a["w"]), interior(b["w"])), n
    del srv_b
    by_size = sorted(ok, key=lambda n: reqs[n].case.params.nx * reqs[n].case.params.ny * reqs[n].case.params.nz)
    for n in by_size[:2]:
        r = reqs[n]
        o = Oracle(r.case.params, r.case.grid)
        o.w[:] = r.case.w0
        o.init_run()
        o.advance(r.nsteps)
        errs = [rel_l2(interior(srv["results"][n]["w"])[m], interior(o.w)[m]) for m in range(5)]
        print(f"request {n} ({r.case.params.nx}x{r.case.params.ny}x{r.case.params.nz}) vs oracle:", errs)
        assert max(errs) <= TRAJ_TOL, (n, errs)
    del seq
    reqs_m = serving_workload(64, 200, rate=16.0, seed=0, precision=PREC_MIXED)
    srv_m = run_serving(reqs_m, hbm_bytes=64 << 30, max_slots=8, steps_per_graph=10, autotune=True,
                        want_fields=True)
    errs = []
    for n in ok:
        b = srv_m["results"][n]
        if b["error"] is None:
            wa, wb = interior(srv["results"][n]["w"]), interior(b["w"])
            errs.append(max(rel_l2(wb[m], wa[m]) for m in range(5)))
    errs = np.array(errs)
    print(f"mixed served vs FP64 per request (worst variable): median {np.median(errs):.2e}, "
          f"p90 {np.percentile(errs, 90):.2e}, max {errs.max():.2e}, over {len(errs)} requests")
    assert len(errs) >= len(ok) - 2
    assert np.median(errs) <= 1e-4


def test_small_grid_serving_stream_batched():
    """128 seeded channel solves of 32-64 cells a side (200 steps each),
    served with batched launches (2 lanes): every request equals its plain
    FP64 solve bit for bit, and the two smallest match the oracle."""
    from streams_mi355x.serve import run_sequential, run_serving, serving_workload

    reqs = serving_workload(128, 200, rate=64.0, seed=1, nx_range=(32, 64), nyz_range=(32, 64))
    srv = run_serving(reqs, hbm_bytes=16 << 30, max_slots=32, steps_per_graph=10, autotune=True, want_fields=True,
                      batch=2)
    st = srv["stats"]
    print(f"small-grid stream batched: {srv['gcups']:.3f} Gcell-updates/s, {st['batched_chunks']} batched chunks, "
          f"{st['batched_solves'] / max(1, st['batched_chunks']):.1f} solves per chunk, p99 {srv['p99_s']:.2f} s")
    assert st["batched_chunks"] > 0 and srv["failed"] == 0
    seq = run_sequential(reqs, want_fields=True)
    for n, (a, b) in enumerate(zip(seq["results"], srv["results"])):
        assert a["error"] is None and np.array_equal(interior(a["w"]), interior(b["w"])), n
    by_size = sorted(range(len(reqs)), key=lambda n: reqs[n].case.params.nx * reqs[n].case.params.ny *
                     reqs[n].case.params.nz)
    for n in by_size[:2]:
        r = reqs[n]
        o = Oracle(r.case.params, r.case.grid)
        o.w[:] = r.case.w0
        o.init_run()
        o.advance(r.nsteps)
        errs = [rel_l2(interior(srv["results"][n]["w"])[m], interior(o.w)[m]) for m in range(5)]
        assert max(errs) <= TRAJ_TOL, (n, errs)
