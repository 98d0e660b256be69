// st_capi.hip -- extern "C" boundary of libstreams_mi355x (declared in
// include/streams_mi355x.h) and the host-side ordering of the hot loop.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <new>
#include <string>
#include <vector>

#include "st_solver.h"

namespace {
thread_local std::string g_err;
}  // namespace

namespace st {
int set_error(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
}  // namespace st

namespace {

int fail(int code, const std::string &msg) { return st::set_error(code, msg); }

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail(ST_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define NEED(s)                                                                \
    do {                                                                       \
        if (!(s)) return fail(ST_ERR_ARG, "null solver handle");               \
        if (!(s)->grid_set) return fail(ST_ERR_STATE, "st_set_grid not called"); \
    } while (0)

void deriv_coeffs(int order1, int order2, double d1[4], double d2[5]) {
    // central-difference weights, get_deriv_coeffs (grid.F90:143-190)
    for (int l = 0; l < 4; l++) d1[l] = 0.0;
    for (int l = 0; l < 5; l++) d2[l] = 0.0;
    switch (order1 / 2) {
    case 1: d1[0] = 0.5; break;
    case 2: d1[0] = 2.0 / 3.0; d1[1] = -1.0 / 12.0; break;
    case 3: d1[0] = 0.75; d1[1] = -0.15; d1[2] = 1.0 / 60.0; break;
    case 4: d1[0] = 4.0 / 5.0; d1[1] = -1.0 / 5.0; d1[2] = 4.0 / 105.0; d1[3] = -1.0 / 280.0; break;
    }
    switch (order2 / 2) {
    case 1: d2[0] = -2.0; d2[1] = 1.0; break;
    case 2: d2[0] = -2.5; d2[1] = 4.0 / 3.0; d2[2] = -1.0 / 12.0; break;
    case 3: d2[0] = -245.0 / 90.0; d2[1] = 1.5; d2[2] = -0.15; d2[3] = 1.0 / 90.0; break;
    case 4: d2[0] = -205.0 / 72.0; d2[1] = 8.0 / 5.0; d2[2] = -1.0 / 5.0; d2[3] = 8.0 / 315.0; d2[4] = -1.0 / 560.0; break;
    }
}

int validate(const st_params *p) {
    if (!p) return fail(ST_ERR_ARG, "null params");
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(ST_ERR_ARG, "grid sizes must be positive");
    // three ghost layers run the fused kernels; four (WENO7, ep_order 8)
    // run the stage procedure by procedure on the per-face kernels
    if (p->ng != 3 && p->ng != 4) return fail(ST_ERR_ARG, "ng must be 3, or 4 for weno_scheme 4 / ep_order 8");
    if (p->nx < p->ng || p->ny < p->ng || p->nz < p->ng)
        return fail(ST_ERR_ARG, "each grid size must be >= ng (singleideal.F90:578)");
    if (p->ep_order < 2 || p->ep_order > 2 * p->ng || (p->ep_order & 1))
        return fail(ST_ERR_ARG, "ep_order must be even and at most 2 ng (6 with ng = 3, 8 with ng = 4)");
    if (p->visc_order != p->ep_order) return fail(ST_ERR_ARG, "viscous order must be equal to ep_order");
    if (p->weno_scheme < 1 || p->weno_scheme > p->ng)
        return fail(ST_ERR_ARG, "weno_scheme must be 1..ng (3 with ng = 3, 4 with ng = 4)");
    if (p->weno_version != 0 && p->weno_version != 1) return fail(ST_ERR_ARG, "weno_version must be 0 or 1");
    if (p->weno_version == 1 && p->weno_scheme != 3)
        return fail(ST_ERR_ARG, "weno_version 1 supported only for weno_scheme = 3");
    if (p->flux_splitting != 0 && p->flux_splitting != 1)
        return fail(ST_ERR_ARG, "flux_splitting must be 0 (Roe) or 1 (Rusanov)");
    if (p->conservative_viscous != 0 && p->conservative_viscous != 1)
        return fail(ST_ERR_ARG, "conservative_viscous must be 0 or 1");
    if (p->calorically_perfect != 0 && p->calorically_perfect != 1)
        return fail(ST_ERR_ARG, "calorically_perfect must be 0 or 1");
    if (p->indx_cp_r < p->indx_cp_l) return fail(ST_ERR_ARG, "indx_cp_r must be >= indx_cp_l");
    if (p->indx_cp_r - p->indx_cp_l + 2 > ST_MAX_CP_COEFF) return fail(ST_ERR_ARG, "too many cp coefficients");
    // the constant-cp branch reads the coefficient of exponent 0
    if (p->calorically_perfect == 1 && (p->indx_cp_l > 0 || p->indx_cp_r + 1 < 0))
        return fail(ST_ERR_ARG, "calorically_perfect = 1 needs indx_cp_l <= 0 <= indx_cp_r + 1");
    if (p->precision != ST_PREC_FP64 && p->precision != ST_PREC_MIXED)
        return fail(ST_ERR_ARG, "precision must be ST_PREC_FP64 or ST_PREC_MIXED");
    if (p->ng == 4) {
        if (p->precision != ST_PREC_FP64) return fail(ST_ERR_ARG, "four ghost layers run in FP64 only");
        for (int l = 0; l < 6; l++)
            if ((p->bctags[l] != 0 && !(p->bctags[l] == 6 && p->is_y_staggered)) || p->bctags_nr[l] != 0)
                return fail(ST_ERR_ARG, "four ghost layers take periodic faces and staggered isothermal walls only");
    }
    bool any_nr = false;
    for (int l = 0; l < 6; l++) {
        const int t = p->bctags[l];
        // each tag acts on the faces update_ghost gives it and leaves the
        // ghosts of the other faces alone, as the reference does; tag 11
        // (airfoil wall) belongs to the curvilinear grids
        if (t < 0 || t > 10)
            return fail(ST_ERR_ARG, "boundary tag " + std::to_string(t) + " on face " + std::to_string(l + 1) +
                                        " is not supported (0..10)");
        const int nr = p->bctags_nr[l];
        if (nr != 0 && nr != 1 && nr != 2 && nr != 3 && nr != 6)
            return fail(ST_ERR_ARG, "characteristic tag " + std::to_string(nr) + " is not supported (0, 1, 2, 3, 6)");
        if (nr != 0 && !(t == 6 && p->is_y_staggered)) {
            if (t == 0) return fail(ST_ERR_ARG, "a characteristic boundary on a periodic face");
            any_nr = true;
        }
    }
    if (any_nr && p->channel_case)
        return fail(ST_ERR_ARG, "characteristic boundaries with the channel forcing are not supported");
    for (int d = 0; d < 3; d++)
        if ((p->bctags[2 * d] == 0) != (p->bctags[2 * d + 1] == 0))
            return fail(ST_ERR_ARG, "a periodic direction needs tag 0 on both faces");
    // a point's index within one variable is 32-bit (offsets across variables are 64-bit)
    const double rows = std::ceil(((double)p->nx + 2.0 * p->ng) / st::kRowAlign) * st::kRowAlign;  // pitched rows
    const double vol = rows * ((double)p->ny + 2.0 * p->ng) * ((double)p->nz + 2.0 * p->ng);
    if (vol > 2147483647.0 - 4096.0) return fail(ST_ERR_ARG, "grid too large: a ghosted variable must hold fewer than 2^31 points");
    return ST_OK;
}

// bc_preproc (singleideal.F90:4703-4750) and correct_bc_order
// (:946-993) for a single block per direction.
void build_tables(st_solver *s, int *eoc_host[3]) {
    const st_params &p = s->p;
    st::KPar &k = s->kp;
    int nr[6];
    for (int l = 0; l < 6; l++) {
        nr[l] = p.bctags_nr[l];
        k.fzf[l] = 0;
        if (p.bctags[l] == 6 && p.is_y_staggered) {
            k.fzf[l] = 1;
            nr[l] = 0;
        }
    }
    const int n[3] = {p.nx, p.ny, p.nz};
    for (int d = 0; d < 3; d++) {
        k.eul_lo[d] = 1 + (nr[2 * d] != 0 ? 1 : 0);
        k.eul_hi[d] = n[d] - (nr[2 * d + 1] != 0 ? 1 : 0);
        k.periodic[d] = p.bctags[2 * d] == 0 ? 1 : 0;
    }
    // the channel (periodic x/z, staggered walls in y) keeps its wall ghosts
    // fused into k_update; any other face set runs the per-face kernels
    s->general_bc = false;
    s->has_nr = false;
    s->recyc = p.bctags[0] == 10;
    for (int l = 0; l < 6; l++) {
        k.bctag[l] = p.bctags[l];
        k.nrtag[l] = nr[l];
        s->has_nr |= nr[l] != 0;
        if (p.bctags[l] != 0 && !(p.bctags[l] == 6 && p.is_y_staggered)) s->general_bc = true;
    }
    k.wall_stag[0] = !s->general_bc && p.bctags[2] == 6 && p.is_y_staggered;
    k.wall_stag[1] = !s->general_bc && p.bctags[3] == 6 && p.is_y_staggered;
    const int ss = p.ep_order / 2 > p.weno_scheme ? p.ep_order / 2 : p.weno_scheme;
    for (int d = 0; d < 3; d++) {
        int *e = eoc_host[d];
        for (int f = 0; f <= n[d]; f++) e[f] = 0;
        if (p.correct_bound_ord > 0 && (p.bctags[2 * d] != 0 || p.bctags[2 * d + 1] != 0)) {
            e[0] = -ss + 1;
            for (int f = 1; f <= ss - 1; f++) e[f] = -ss + f;
            e[n[d]] = -ss + 1;
            for (int f = 1; f <= ss - 1; f++) e[n[d] - f] = -ss + f;
        }
    }
}

}  // namespace

// The KPar fields that follow from the parameters alone (the boundary
// tables of build_tables come on top for a solver).
void st::params_to_kpar(const st_params &p, st::KPar &k) {
    k.lmax_base = p.ep_order / 2;
    k.visc_lmax = p.visc_order / 2;
    k.weno_scheme = p.weno_scheme;
    k.weno_version = p.weno_version;
    k.flux_splitting = p.flux_splitting;
    k.nkeep = p.nkeep;
    k.calorically_perfect = p.calorically_perfect;
    k.indx_cp_l = p.indx_cp_l;
    k.indx_cp_r = p.indx_cp_r;
    k.visc_model = p.visc_model;
    k.ndim = p.nz > 1 ? 3 : 2;
    k.rgas0 = p.rgas0; k.prandtl = p.prandtl; k.t0 = p.t0; k.rho0 = p.rho0; k.u0 = p.u0; k.l0 = p.l0;
    k.p_ref = p.rho0 * p.rgas0 * p.t0;
    for (int m = 0; m < 5; m++) {
        k.winf[m] = p.winf[m];
        k.winf_ps[m] = p.winf_past_shock[m];
    }
    k.xshock_imp = p.xshock_imp;
    k.shock_angle = p.shock_angle;
    k.tanhfacs = p.tanhfacs;
    k.aoa = p.aoa;
    k.ptot0 = p.ptot0;
    k.ttot0 = p.ttot0;
    k.mu0 = p.mu0; k.sutherland_s = p.sutherland_s; k.t_ref_dim = p.t_ref_dim; k.vtexp = p.powerlaw_vtexp;
    k.sensor_threshold = p.sensor_threshold; k.t_wall = p.t_wall; k.t_bulk_target = p.t_bulk_target;
    k.volchan = p.volchan; k.cfl = p.cfl; k.tol_iter_nr = p.tol_iter_nr;
    for (int l = 0; l < ST_MAX_CP_COEFF; l++) { k.cp[l] = p.cp_coeff[l]; k.cv[l] = p.cv_coeff[l]; }
    double d1[4], d2[5];
    deriv_coeffs(p.ep_order, p.ep_order, d1, d2);
    for (int l = 0; l < 4; l++) k.cd1[l] = d1[l];
    for (int lm = 1; lm <= 4; lm++) {
        double a[4], b[5];
        deriv_coeffs(2 * lm, 2 * lm, a, b);
        for (int l = 0; l < 4; l++) k.cd1tab[lm - 1][l] = a[l];
    }
    deriv_coeffs(p.visc_order, p.visc_order, d1, d2);
    for (int l = 0; l < 4; l++) k.vd1[l] = d1[l];
    for (int l = 0; l < 5; l++) k.vd2[l] = d2[l];
}

st::Geo st::params_to_geo(const st_params &p, bool solver) {
    st::Geo g;
    g.nx = p.nx; g.ny = p.ny; g.nz = p.nz; g.ng = p.ng;
    g.rx = p.nx + 2 * p.ng; g.sy = p.ny + 2 * p.ng; g.sz = p.nz + 2 * p.ng;
    const int a = st::row_align(p.precision);
    g.sx = solver ? (g.rx + a - 1) / a * a : g.rx;
    g.sxy = g.sx * g.sy;
    g.nvol = (long long)g.sxy * g.sz;
    g.nref = (long long)g.rx * g.sy * g.sz;
    g.nint = (long long)p.nx * p.ny * p.nz;
    return g;
}

int st::validate_params(const st_params *p) { return validate(p); }

namespace {

void fill_kpar(st_solver *s) { st::params_to_kpar(s->p, s->kp); }

// bytes of one w_aux / right-hand-side value
size_t rsize(const st_solver *s) { return s->mixed ? sizeof(float) : sizeof(double); }
// w_aux variables: the 10 of the reference, plus p - p_ref in the mixed form
size_t naux(const st_solver *s) { return s->mixed ? 11 : 10; }

template <class T>
hipError_t dalloc(T **p, size_t n) {
    return hipMalloc((void **)p, n * sizeof(T) > 0 ? n * sizeof(T) : 8);
}

// Frees what the solver allocated; arrays inside memory lent to it
// (create_solver with a SlotMem) and a lent scalar mirror stay with the lender.
void free_all(st_solver *s) {
    void *ptrs[] = {s->w_alloc ? s->w_alloc : (void *)s->w, s->w_aux_alloc ? s->w_aux_alloc : s->w_aux, s->F[0], s->F[1], s->wallprop, s->eoc[0], s->eoc[1], s->eoc[2], s->x, s->y,
                    s->z, s->yn, s->dcsidx, s->dcsidxs, s->dcsidx2, s->detady, s->detadys, s->detady2, s->dzitdz,
                    s->dzitdzs, s->dzitdz2, s->ds, s->partials, s->dirty, s->counters, s->fixlist, s->rc_real, s->rc_maps,
                    s->w_stat, s->rplane, s->rng_coef, s->rng_win, s->rng_ctl, s->wrecyc, s->fhat, s->fln};
    for (void *p : ptrs) {
        if (!p) continue;
        if ((char *)p >= s->arena_lo && (char *)p < s->arena_hi) continue;
        (void)hipFree(p);
    }
    if (s->hs && !s->borrowed_hs) (void)hipHostFree(s->hs);
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
}

int pull_scalars(st_solver *s) {
    HIPCHK(hipMemcpyAsync(s->hs, s->ds, sizeof(st::DevScalars), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return ST_OK;
}

// event pair around one launch when st_kernel_times is active
struct TimeScope {
    st_solver *s;
    int slot;
    TimeScope(st_solver *s_, int fam) : s(s_), slot(-1) {
        KTimer *t = s->timer;
        if (t && t->n < t->made) {
            slot = t->n++;
            t->fam[slot] = fam;
            (void)hipEventRecord(t->ev[slot][0], s->stream);
        }
    }
    ~TimeScope() {
        if (slot >= 0) (void)hipEventRecord(s->timer->ev[slot][1], s->stream);
    }
};
#define TIMED(fam, expr)           \
    do {                           \
        TimeScope ts_(s, fam);     \
        HIPCHK(expr);              \
    } while (0)

// The mixed form with the Laplacian viscous terms refines the Ducros sensor
// near the threshold in FP64 from w (its fluxes and stresses stay FP32).
bool fp64_sensor(const st_solver *s) { return s->mixed && s->p.conservative_viscous != 1; }

// one RK stage of rk_sync (singleideal_cpu.F90:352-406), fused
int stage(st_solver *s, int istep) {
    using namespace st;
    void *fprev = s->F[s->cur];   // RHS of the previous stage (reference: fl before init_flux)
    void *fcur = s->F[s->cur ^ 1];
    const Geo &g = s->g;
    // bcswap + compute_aux(ghost=1): aux over the box, periodic ghosts from
    // images; at the first stage of a step that recomputes dt, the same pass
    // takes the CFL reduction
    TIMED(ST_K_AUX, launch_aux(s, 1 - g.ng, g.nx + g.ng, 1 - g.ng, g.ny + g.ng, 1 - g.ng, g.nz + g.ng, true,
                               istep == 0 && s->dt_in_stage));
    if (istep == 0) s->dt_in_stage = false;
    // euler_x + euler_y in one pass; each sweep's WENO faces are then fixed
    // up (timed as their own family)
    const bool xy = s->fuse_xy;
    if (xy) {
        TIMED(ST_K_EULER_XY, launch_euler_xy(s, fcur, false));
        TIMED(ST_K_WENO_FIX, launch_weno_fix(s, fcur, 0, 1, 0));
        TIMED(ST_K_WENO_FIX, launch_weno_fix(s, fcur, 1, 2, 1));
    } else {
        TIMED(ST_K_EULER_X, launch_euler_x(s, fcur, true, false));
        TIMED(ST_K_WENO_FIX, launch_weno_fix(s, fcur, 0, 1, 0));
    }
    const bool cons = s->p.conservative_viscous == 1;
    // visflx(4), or visflx(6) + visflx(5); + bcextr_var / bcswap_var of w_aux(10)
    if (cons) TIMED(ST_K_VISC, launch_visc_cons(s, fcur, 56, true));
    else TIMED(ST_K_VISC, launch_visc(s, fcur, 4, true));
    if (!xy) {
        TIMED(ST_K_EULER_Y, launch_euler_y(s, fcur, false));
        TIMED(ST_K_WENO_FIX, launch_weno_fix(s, fcur, 1, 1, 0));
    }
    // euler_z, with the viscous divergence terms (visflx(1) or (7)) and the
    // force_rhs partials of the same cells folded into the march where the
    // z range spans all planes
    const bool zv = s->fuse_zv && s->kp.ndim == 3 && s->kp.eul_lo[2] == 1 && s->kp.eul_hi[2] == g.nz &&
                    (cons || s->kp.visc_lmax <= 3);
    if (zv) {
        TIMED(ST_K_EULER_ZV, launch_euler_zv(s, fcur, fprev, istep, s->p.channel_case != 0, cons, false));
        TIMED(ST_K_WENO_FIX, launch_euler_zv_fix(s, fcur, istep, s->p.channel_case != 0));
    } else if (s->kp.ndim == 3) {
        TIMED(ST_K_EULER_Z, launch_euler_z(s, fcur, false));
        TIMED(ST_K_WENO_FIX, launch_weno_fix(s, fcur, 2, 1, 0));
    }
    // visflx(3) + bcswap_var(w_aux 8); the mixed Laplacian form refines the
    // sensor near the threshold in FP64 so its switch decisions are the FP64 form's
    if (istep == 2) {
        TIMED(ST_K_SENSOR, launch_sensor(s, true));
        if (fp64_sensor(s)) TIMED(ST_K_SENSOR, launch_sensor_refine(s, false, true));
    }
    if (!zv) TIMED(ST_K_VISC_DIV, launch_visc_div(s, fcur, fprev, istep, s->p.channel_case != 0, cons));  // visflx(1) or (7)
    if (s->p.channel_case) TIMED(ST_K_FINISH, launch_force_finish(s, istep));
    if (s->has_nr) TIMED(ST_K_BC, launch_bc_nr(s, fcur));  // bc_nr
    TIMED(ST_K_UPDATE, launch_update(s, fprev, fcur, istep));
    if (s->p.channel_case && s->p.theta_wall >= -1.0) TIMED(ST_K_FORCE_VAR, launch_force_var(s));
    // get_crandom_f of euler_z: the random inflow plane of this stage
    if (s->recyc) TIMED(ST_K_BC, launch_inflow_draws(s));
    if (s->general_bc) TIMED(ST_K_BC, launch_physical_ghosts(s));  // update_ghost(do_swap=0)
    s->cur ^= 1;
    s->last_stage = istep;
    s->fln_current = false;
    return ST_OK;
}

int do_rk_sync(st_solver *s) {
    for (int istep = 0; istep < 3; istep++) {
        int rc = stage(s, istep);
        if (rc) return rc;
    }
    return ST_OK;
}

// The device arrays of one solve, in allocation order, and their sizes in
// bytes (st_solver_footprint sums them, each rounded up to 256 bytes).
constexpr int kNArrays = 26;
constexpr size_t kArenaAlign = 256;
size_t arena_align(size_t n) { return (n + kArenaAlign - 1) / kArenaAlign * kArenaAlign; }

int partial_slots(const st::Geo &g) {
    const int nblk = ((g.nx + 63) / 64) * ((g.ny + 3) / 4) * g.nz;
    // per-block partials (7 per cell block, or one per block of the aux
    // pass over the ghosted box, whichever is larger), the z fix-up's
    // 5 x 256 force_rhs partials, and the 128 x 8 second-level values at the end
    const int aux_blk = (int)((st::aux_lanes(g) + 255) / 256);
    return (7 * nblk > aux_blk ? 7 * nblk : aux_blk) + 256 * 5 + 128 * 8;
}

void array_sizes(const st_params &p, size_t n[kNArrays]) {
    const st::Geo g = st::params_to_geo(p);
    const size_t nvol = g.nvol, nint = g.nint;
    const size_t rs = p.precision == ST_PREC_MIXED ? sizeof(float) : sizeof(double);
    const size_t na = p.precision == ST_PREC_MIXED ? 11 : 10;
    const size_t d = sizeof(double);
    size_t q = 0;
    n[q++] = d * 5 * nvol + 128;           // w (+ the lead of its first point)
    n[q++] = rs * na * nvol + 128;         // w_aux
    n[q++] = rs * 5 * nint;                // F[0]
    n[q++] = rs * 5 * nint;                // F[1]
    n[q++] = d * 3 * g.sx * g.sz;          // wallprop
    n[q++] = sizeof(int) * (g.nx + 1);     // eoc x, y, z
    n[q++] = sizeof(int) * (g.ny + 1);
    n[q++] = sizeof(int) * (g.nz + 1);
    n[q++] = d * g.sx;                     // x, y, z, yn
    n[q++] = d * g.sy;
    n[q++] = d * g.sz;
    n[q++] = d * (g.ny + 1);
    for (int r = 0; r < 3; r++) n[q++] = d * g.nx;  // dcsidx, dcsidxs, dcsidx2
    for (int r = 0; r < 3; r++) n[q++] = d * g.ny;  // detady ...
    for (int r = 0; r < 3; r++) n[q++] = d * g.nz;  // dzitdz ...
    n[q++] = sizeof(st::DevScalars);       // ds
    n[q++] = d * partial_slots(g);         // partials
    n[q++] = nint + 64;                    // dirty flags + the any-dirty word
    n[q++] = sizeof(int) * (8 + 4096);     // counters (4096 = kFixMaxSeg)
    n[q++] = sizeof(int) * nint;           // fixlist
}

// st_solver_create, or with mem the same solver inside memory the caller
// lends (see st::SlotMem): no allocation, no synchronisation, every copy
// and fill ordered on mem->stream.
int create_solver(const st_params *p, const st::SlotMem *mem, st_solver **out) {
    if (!out) return fail(ST_ERR_ARG, "null output pointer");
    *out = nullptr;
    int rc = validate(p);
    if (rc) return rc;
    st_solver *s = new (std::nothrow) st_solver();
    if (!s) return fail(ST_ERR_NOMEM, "host allocation failed");
    std::memset((void *)s, 0, sizeof(st_solver));
    s->p = *p;
    s->fuse_xy = true;
    s->fuse_zv = true;
    s->g = st::params_to_geo(*p);
    st::Geo &g = s->g;
    s->rhork[0] = 0.0; s->rhork[1] = -17.0 / 60.0; s->rhork[2] = -5.0 / 12.0;   // Wray RK3,
    s->gamrk[0] = 8.0 / 15.0; s->gamrk[1] = 5.0 / 12.0; s->gamrk[2] = 3.0 / 4.0; // singleideal.F90:4780-4782
    s->last_stage = -1;
    s->mixed = p->precision == ST_PREC_MIXED;
    fill_kpar(s);
    hipError_t e = hipGetDevice(&s->device);
    if (mem) {
        s->stream = mem->stream;
        s->own_stream = false;
    } else {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        s->own_stream = true;
    }
    if (e != hipSuccess) {
        delete s;
        return fail(ST_ERR_HIP, std::string("stream creation: ") + hipGetErrorString(e));
    }
    const size_t nvol = g.nvol, nint = g.nint;
    s->npart = partial_slots(g);
    size_t n[kNArrays];
    array_sizes(*p, n);
    void **dst[kNArrays] = {(void **)&s->w, &s->w_aux, &s->F[0], &s->F[1], (void **)&s->wallprop,
                            (void **)&s->eoc[0], (void **)&s->eoc[1], (void **)&s->eoc[2], (void **)&s->x,
                            (void **)&s->y, (void **)&s->z, (void **)&s->yn, (void **)&s->dcsidx,
                            (void **)&s->dcsidxs, (void **)&s->dcsidx2, (void **)&s->detady, (void **)&s->detadys,
                            (void **)&s->detady2, (void **)&s->dzitdz, (void **)&s->dzitdzs, (void **)&s->dzitdz2,
                            (void **)&s->ds, (void **)&s->partials, (void **)&s->dirty, (void **)&s->counters,
                            (void **)&s->fixlist};
    hipError_t ea = hipSuccess;
    if (mem) {
        size_t off = 0;
        for (int q = 0; q < kNArrays; q++) {
            *dst[q] = mem->arena + off;
            off += arena_align(n[q]);
        }
        if (off > mem->bytes) {
            delete s;
            return fail(ST_ERR_NOMEM, "the lent memory is smaller than the solve's footprint");
        }
        s->arena_lo = mem->arena;
        s->arena_hi = mem->arena + off;
        s->hs = mem->hs;
        s->borrowed_hs = true;
    } else {
        for (int q = 0; q < kNArrays && ea == hipSuccess; q++) ea = hipMalloc(dst[q], n[q] > 0 ? n[q] : 8);
        if (ea == hipSuccess) ea = hipHostMalloc((void **)&s->hs, sizeof(st::DevScalars), hipHostMallocDefault);
        if (ea != hipSuccess) {
            free_all(s);
            delete s;
            return fail(ST_ERR_NOMEM, std::string("device allocation failed: ") + hipGetErrorString(ea));
        }
    }
    // the first points of w and w_aux sit lead_bytes into their allocations
    s->w_alloc = s->w;
    s->w_aux_alloc = s->w_aux;
    s->w = (double *)((char *)s->w_alloc + st::lead_bytes(g.ng, sizeof(double)));
    s->w_aux = (char *)s->w_aux_alloc + st::lead_bytes(g.ng, rsize(s));
    // tables
    int *eh[3];
    int *buf = mem ? mem->tabs : (int *)std::calloc((size_t)(g.nx + g.ny + g.nz + 3), sizeof(int));
    if (!buf) {
        free_all(s);
        delete s;
        return fail(ST_ERR_NOMEM, "host allocation failed");
    }
    if (mem) std::memset(buf, 0, sizeof(int) * (size_t)(g.nx + g.ny + g.nz + 3));
    eh[0] = buf; eh[1] = buf + g.nx + 1; eh[2] = buf + g.nx + g.ny + 2;
    build_tables(s, eh);
    // zero state (w_aux = 0 as in the reference alloc, singleideal.F90:1596; then
    // w_aux(6) = t0 as in initialize, singleideal.F90:3032)
    const size_t nz_bytes[7] = {sizeof(double) * 5 * nvol, rsize(s) * naux(s) * nvol, rsize(s) * 5 * nint,
                                rsize(s) * 5 * nint, sizeof(double) * 3 * g.sx * g.sz, sizeof(int) * (8 + 4096),
                                nint + 64};
    void *nz_ptr[7] = {s->w, s->w_aux, s->F[0], s->F[1], s->wallprop, s->counters, s->dirty};
    hipError_t ec = hipSuccess;
    std::memset(s->hs, 0, sizeof(st::DevScalars));
    if (p->cfl < 0) s->hs->dt = -p->cfl;
    if (mem) {
        s->hs->time = mem->time0;
        s->hs->icyc = mem->icyc0;
        s->icyc = s->icyc0 = mem->icyc0;
        // stream-ordered from pinned host memory the lender keeps alive
        const size_t tb[3] = {sizeof(int) * (g.nx + 1), sizeof(int) * (g.ny + 1), sizeof(int) * (g.nz + 1)};
        for (int d = 0; d < 3 && ec == hipSuccess; d++)
            ec = hipMemcpyAsync(s->eoc[d], eh[d], tb[d], hipMemcpyHostToDevice, s->stream);
        for (int q = 0; q < 7 && ec == hipSuccess; q++) ec = hipMemsetAsync(nz_ptr[q], 0, nz_bytes[q], s->stream);
        if (ec == hipSuccess)
            ec = hipMemcpyAsync(s->ds, s->hs, sizeof(st::DevScalars), hipMemcpyHostToDevice, s->stream);
        if (ec == hipSuccess) ec = st::launch_init_temperature(s);
    } else {
        if (ec == hipSuccess) ec = hipMemcpy(s->eoc[0], eh[0], sizeof(int) * (g.nx + 1), hipMemcpyHostToDevice);
        if (ec == hipSuccess) ec = hipMemcpy(s->eoc[1], eh[1], sizeof(int) * (g.ny + 1), hipMemcpyHostToDevice);
        if (ec == hipSuccess) ec = hipMemcpy(s->eoc[2], eh[2], sizeof(int) * (g.nz + 1), hipMemcpyHostToDevice);
        std::free(buf);
        for (int q = 0; q < 7 && ec == hipSuccess; q++) ec = hipMemset(nz_ptr[q], 0, nz_bytes[q]);
        if (ec == hipSuccess) ec = hipMemcpy(s->ds, s->hs, sizeof(st::DevScalars), hipMemcpyHostToDevice);
        if (ec == hipSuccess) ec = st::launch_init_temperature(s);
        if (ec == hipSuccess) ec = hipStreamSynchronize(s->stream);
    }
    if (ec != hipSuccess) {
        free_all(s);
        delete s;
        return fail(ST_ERR_HIP, std::string("initialisation: ") + hipGetErrorString(ec));
    }
    *out = s;
    return ST_OK;
}

// nvars ghosted variables of es-byte values between the solver's pitched
// layout (device) and the reference layout (host): each variable's rows are
// re-spaced in a host buffer of one variable and move in one copy on the
// solver's stream.  Returns once the copies have completed.
hipError_t copy_rows(st_solver *s, void *dst, const void *src, size_t es, int nvars, hipMemcpyKind kind) {
    const st::Geo &g = s->g;
    const size_t dpitch = (size_t)g.sx * es, hpitch = (size_t)g.rx * es;
    const size_t rows = (size_t)g.sy * g.sz;  // of one variable
    const size_t dvar = (size_t)g.nvol * es, hvar = (size_t)g.nref * es;
    std::vector<char> tmp;
    try {
        tmp.assign(rows * dpitch, 0);
    } catch (const std::bad_alloc &) {
        return hipErrorOutOfMemory;
    }
    hipError_t e = hipSuccess;
    for (int m = 0; m < nvars && e == hipSuccess; m++) {
        if (kind == hipMemcpyHostToDevice) {
            const char *h = (const char *)src + m * hvar;
            for (size_t r = 0; r < rows; r++) std::memcpy(tmp.data() + r * dpitch, h + r * hpitch, hpitch);
            e = hipMemcpyAsync((char *)dst + m * dvar, tmp.data(), dvar, kind, s->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        } else {
            e = hipMemcpyAsync(tmp.data(), (const char *)src + m * dvar, dvar, kind, s->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
            char *h = (char *)dst + m * hvar;
            if (e == hipSuccess)
                for (size_t r = 0; r < rows; r++) std::memcpy(h + r * hpitch, tmp.data() + r * dpitch, hpitch);
        }
    }
    return e;
}

}  // namespace

extern "C" {

const char *st_last_error(void) { return g_err.c_str(); }
const char *st_version(void) { return "streams_mi355x 0.1 (gfx950)"; }

int64_t st_solver_footprint(const st_params *p) {
    if (!p) return -1;
    size_t n[kNArrays];
    array_sizes(*p, n);
    int64_t tot = 0;
    for (int q = 0; q < kNArrays; q++) tot += (int64_t)arena_align(n[q]);
    return tot;
}

int st_solver_create(const st_params *p, st_solver **out) { return create_solver(p, nullptr, out); }

int st_solver_destroy(st_solver *s) {
    if (!s) return ST_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
    return ST_OK;
}

int st_set_grid(st_solver *s, const double *x, const double *y, const double *z, const double *yn,
                const double *dcsidx, const double *dcsidxs, const double *dcsidx2, const double *detady,
                const double *detadys, const double *detady2, const double *dzitdz, const double *dzitdzs,
                const double *dzitdz2) {
    if (!s) return fail(ST_ERR_ARG, "null solver handle");
    const double *src[13] = {x, y, z, yn, dcsidx, dcsidxs, dcsidx2, detady, detadys, detady2, dzitdz, dzitdzs, dzitdz2};
    double *dst[13] = {s->x, s->y, s->z, s->yn, s->dcsidx, s->dcsidxs, s->dcsidx2,
                       s->detady, s->detadys, s->detady2, s->dzitdz, s->dzitdzs, s->dzitdz2};
    const st::Geo &g = s->g;
    const size_t n[13] = {(size_t)g.rx, (size_t)g.sy, (size_t)g.sz, (size_t)g.ny + 1, (size_t)g.nx, (size_t)g.nx,
                          (size_t)g.nx, (size_t)g.ny, (size_t)g.ny, (size_t)g.ny, (size_t)g.nz, (size_t)g.nz,
                          (size_t)g.nz};
    for (int q = 0; q < 13; q++) {
        if (!src[q]) return fail(ST_ERR_ARG, "null grid array " + std::to_string(q));
        HIPCHK(hipMemcpyAsync(dst[q], src[q], n[q] * sizeof(double), hipMemcpyHostToDevice, s->stream));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    // jbl_inflow = locateval(y(1:ny), ny, l0) (singleideal.F90:3284): rows
    // below l0 in an increasing y
    s->jbl = 0;
    for (int j = 1; j <= g.ny; j++)
        if (s->p.l0 > y[j + g.ng - 1]) s->jbl = j;
    s->grid_set = true;
    return ST_OK;
}

int st_copy_from_field(st_solver *s, const double *w_host) {
    NEED(s);
    if (!w_host) return fail(ST_ERR_ARG, "null host array");
    HIPCHK(copy_rows(s, s->w, w_host, sizeof(double), 5, hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(s->stream));
    return ST_OK;
}

int st_copy_to_field(st_solver *s, double *w_host) {
    NEED(s);
    if (!w_host) return fail(ST_ERR_ARG, "null host array");
    HIPCHK(copy_rows(s, w_host, s->w, sizeof(double), 5, hipMemcpyDeviceToHost));
    HIPCHK(hipStreamSynchronize(s->stream));
    return ST_OK;
}

// Device array of n values (double, or float in the mixed form) to a host
// double array.
static int to_host(st_solver *s, double *host, const void *dev, size_t n) {
    if (!s->mixed) {
        HIPCHK(hipMemcpyAsync(host, dev, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return ST_OK;
    }
    std::vector<float> tmp(n);
    HIPCHK(hipMemcpyAsync(tmp.data(), dev, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (size_t t = 0; t < n; t++) host[t] = tmp[t];
    return ST_OK;
}

// One ghosted variable (double, or float in the mixed form) between the
// solver's pitched layout and a host array in the reference layout.
static int ghost_to_host(st_solver *s, double *host, const void *dev) {
    const size_t n = (size_t)s->g.nref;
    if (!s->mixed) {
        HIPCHK(copy_rows(s, host, dev, sizeof(double), 1, hipMemcpyDeviceToHost));
        HIPCHK(hipStreamSynchronize(s->stream));
        return ST_OK;
    }
    std::vector<float> tmp(n);
    HIPCHK(copy_rows(s, tmp.data(), dev, sizeof(float), 1, hipMemcpyDeviceToHost));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (size_t t = 0; t < n; t++) host[t] = tmp[t];
    return ST_OK;
}

static int ghost_from_host(st_solver *s, void *dev, const double *host) {
    const size_t n = (size_t)s->g.nref;
    if (!s->mixed) {
        HIPCHK(copy_rows(s, dev, host, sizeof(double), 1, hipMemcpyHostToDevice));
        HIPCHK(hipStreamSynchronize(s->stream));
        return ST_OK;
    }
    std::vector<float> tmp(host, host + n);
    HIPCHK(copy_rows(s, dev, tmp.data(), sizeof(float), 1, hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(s->stream));
    return ST_OK;
}

int st_get_aux(st_solver *s, int32_t var, double *host) {
    NEED(s);
    if (var < 1 || var > 10 || !host) return fail(ST_ERR_ARG, "var must be 1..10");
    const size_t nv = s->g.nvol;
    return ghost_to_host(s, host, (const char *)s->w_aux + rsize(s) * (size_t)(var - 1) * nv);
}

int st_set_aux(st_solver *s, int32_t var, const double *host) {
    NEED(s);
    if (var < 1 || var > 10 || !host) return fail(ST_ERR_ARG, "var must be 1..10");
    const size_t nv = s->g.nvol;
    return ghost_from_host(s, (char *)s->w_aux + rsize(s) * (size_t)(var - 1) * nv, host);
}

int st_get_fl(st_solver *s, double *host) {
    NEED(s);
    if (!host) return fail(ST_ERR_ARG, "null host array");
    return to_host(s, host, s->F[s->cur], 5 * (size_t)s->g.nint);
}

int st_get_scalars(st_solver *s, st_scalars *out) {
    if (!s || !out) return fail(ST_ERR_ARG, "null argument");
    int rc = pull_scalars(s);
    if (rc) return rc;
    const st::DevScalars &d = *s->hs;
    out->dt = d.dt; out->time = d.time; out->dpdx = d.dpdx;
    out->rhobulk = d.rhobulk; out->ubulk = d.ubulk; out->tbulk = d.tbulk;
    out->residual_rhou = d.residual;
    out->rhomin = d.rhomin; out->rhomax = d.rhomax; out->tmin = d.tmin; out->tmax = d.tmax;
    out->pmin = d.pmin; out->pmax = d.pmax;
    out->icyc = d.icyc;
    return ST_OK;
}

int st_set_time(st_solver *s, double time, double dt, int64_t icyc) {
    if (!s) return fail(ST_ERR_ARG, "null solver handle");
    int rc = pull_scalars(s);
    if (rc) return rc;
    s->hs->time = time;
    s->hs->dt = dt;
    s->hs->icyc = icyc;
    HIPCHK(hipMemcpyAsync(s->ds, s->hs, sizeof(st::DevScalars), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->icyc = icyc;
    s->icyc0 = icyc;
    return ST_OK;
}

namespace {

// The C library's rand() after srand(seed): r[0] = seed, r[i] = 16807 r[i-1]
// mod (2^31 - 1) for i < 31 (Schrage's form), r[31..33] = r[0..2], then
// r[i] = r[i-31] + r[i-3] (mod 2^32) with outputs r[i] >> 1 from i = 344 on.
// Fills the 31 words r[313..343] that precede the first output.
void crandom_seed_window(uint32_t seed, uint32_t win[31]) {
    std::vector<int64_t> r(344);
    r[0] = seed == 0 ? 1 : (int32_t)seed;
    for (int i = 1; i < 31; i++) {
        const int64_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
        int64_t word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        r[i] = word;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    for (int i = 34; i < 344; i++) r[i] = (int64_t)(uint32_t)((uint32_t)r[i - 31] + (uint32_t)r[i - 3]);
    for (int q = 0; q < 31; q++) win[q] = (uint32_t)r[313 + q];
}

// Set-up of the random inflow plane: seed, the jump coefficients of the
// draws of one z sweep and the device buffers (plane at 0.5,
// singleideal.F90:1543).
int setup_inflow_random(st_solver *s) {
    const int ny = s->g.ny, nz = s->g.nz;
    const int n = s->jbl >= 2 ? 3 * (s->jbl - 1) * nz : 0;
    if (s->rplane) return ST_OK;
    uint32_t seed = 1;  // init_crandom_f(rank+1) of one rank
    if (s->p.rand_type < 0) seed = (uint32_t)std::time(nullptr) + 1u;
    crandom_seed_window(seed, s->rng_seeded);
    // coef(t) = coef(t-31) + coef(t-3), coef(-31+q) = e_q
    const size_t tot = (size_t)(n + 31) * 31;
    std::vector<uint32_t> c(tot, 0u);
    for (int q = 0; q < 31; q++) c[(size_t)q * 31 + q] = 1u;
    for (int t = 0; t < n; t++) {
        uint32_t *dst = &c[(size_t)(t + 31) * 31];
        const uint32_t *a = &c[(size_t)t * 31], *b = &c[(size_t)(t + 28) * 31];
        for (int q = 0; q < 31; q++) dst[q] = a[q] + b[q];
    }
    std::vector<uint32_t> coef((size_t)31 * (n > 0 ? n : 1));
    for (int t = 0; t < n; t++)
        for (int q = 0; q < 31; q++) coef[(size_t)q * n + t] = c[(size_t)(t + 31) * 31 + q];
    std::vector<double> plane((size_t)ny * nz * 3, 0.5);
    HIPCHK(hipMalloc(&s->rplane, plane.size() * sizeof(double)));
    HIPCHK(hipMalloc(&s->rng_coef, coef.size() * sizeof(uint32_t)));
    HIPCHK(hipMalloc(&s->rng_win, 62 * sizeof(uint32_t)));
    HIPCHK(hipMalloc(&s->rng_ctl, 2 * sizeof(int)));
    HIPCHK(hipMemcpyAsync(s->rplane, plane.data(), plane.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->rng_coef, coef.data(), coef.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                          s->stream));
    s->rng_n = n;
    HIPCHK(st::launch_rng_reset(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return ST_OK;
}

int ensure_rc_real(st_solver *s) {
    const int ng = s->g.ng, ny = s->g.ny, ny2 = ny + 2 * ng;
    const size_t nreal = (size_t)ng * ny * 4 + ny + 4 * (size_t)ny2 + (size_t)ng * ny * 5;
    if (!s->rc_real) {
        HIPCHK(hipMalloc(&s->rc_real, nreal * sizeof(double)));
        HIPCHK(hipMalloc(&s->rc_maps, 2 * (size_t)ny * sizeof(int)));
    }
    return ST_OK;
}

}  // namespace

int st_set_wmean(st_solver *s, const double *wmean) {
    NEED(s);
    if (!wmean) return fail(ST_ERR_ARG, "null wmean");
    int rc = ensure_rc_real(s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(s->rc_real, wmean, sizeof(double) * (size_t)s->g.ng * s->g.ny * 4, hipMemcpyHostToDevice,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->wmean_set = true;
    return ST_OK;
}

int st_set_recycling(st_solver *s, int32_t i_recyc, double betarecyc, double glund1, const double *wmean_inflow,
                     const double *weta, const int32_t *map_j_inn, const int32_t *map_j_out,
                     const double *yplus_inflow, const double *eta_inflow, const double *yplus_recyc,
                     const double *eta_recyc) {
    NEED(s);
    if (!s->recyc) return fail(ST_ERR_STATE, "bctags[0] is not 10 (recycling)");
    const int ng = s->g.ng, ny = s->g.ny, ny2 = ny + 2 * ng;
    if (i_recyc < ng || i_recyc > s->g.nx) return fail(ST_ERR_ARG, "i_recyc must be in [ng, nx]");
    if (!wmean_inflow || !weta || !map_j_inn || !map_j_out || !yplus_inflow || !eta_inflow || !yplus_recyc ||
        !eta_recyc)
        return fail(ST_ERR_ARG, "null recycling array");
    for (int j = 2; j <= ny; j++) {
        const int a = map_j_inn[j - 1], b = map_j_out[j - 1];
        if (a < 0 || b < 0 || a > ny || b > ny || ((a < ny && b < ny) && (a < 1 || b < 1)))
            return fail(ST_ERR_ARG, "recycling y map out of range at j = " + std::to_string(j));
    }
    const size_t nreal = (size_t)ng * ny * 4 + ny + 4 * (size_t)ny2 + (size_t)ng * ny * 5;
    int rc = ensure_rc_real(s);
    if (rc) return rc;
    if (!s->wrecyc) HIPCHK(hipMalloc(&s->wrecyc, sizeof(double) * 5 * (size_t)ng * ny * s->g.nz));
    if (s->p.rand_type != 0 && (rc = setup_inflow_random(s))) return rc;
    std::vector<double> h(nreal, 0.0);
    size_t o = 0;
    std::memcpy(&h[o], wmean_inflow, sizeof(double) * ng * ny * 4); o += (size_t)ng * ny * 4;
    std::memcpy(&h[o], weta, sizeof(double) * ny); o += ny;
    std::memcpy(&h[o], yplus_inflow, sizeof(double) * ny2); o += ny2;
    std::memcpy(&h[o], eta_inflow, sizeof(double) * ny2); o += ny2;
    std::memcpy(&h[o], yplus_recyc, sizeof(double) * ny2); o += ny2;
    std::memcpy(&h[o], eta_recyc, sizeof(double) * ny2);
    std::vector<int> hm(2 * (size_t)ny);
    std::memcpy(hm.data(), map_j_inn, sizeof(int) * ny);
    std::memcpy(hm.data() + ny, map_j_out, sizeof(int) * ny);
    HIPCHK(hipMemcpyAsync(s->rc_real, h.data(), nreal * sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->rc_maps, hm.data(), 2 * (size_t)ny * sizeof(int), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->i_recyc = i_recyc;
    s->betarecyc = betarecyc;
    s->glund1 = glund1;
    s->recyc_set = true;
    s->wmean_set = true;
    return ST_OK;
}

int st_update_ghost(st_solver *s, int32_t do_swap) {
    NEED(s);
    if (s->recyc && !s->recyc_set) return fail(ST_ERR_STATE, "st_set_recycling not called");
    if (s->general_bc) HIPCHK(st::launch_physical_ghosts(s));
    else HIPCHK(st::launch_wall_ghosts(s));
    if (do_swap) HIPCHK(st::launch_fill_periodic(s, s->w, 5));
    return ST_OK;
}

int st_bcswap(st_solver *s) {
    NEED(s);
    HIPCHK(st::launch_fill_periodic(s, s->w, 5));
    return ST_OK;
}

int st_compute_aux(st_solver *s, int32_t central, int32_t ghost) {
    NEED(s);
    const st::Geo &g = s->g;
    const int nx = g.nx, ny = g.ny, nz = g.nz, ng = g.ng;
    if (central == 1 && ghost == 1) {
        HIPCHK(st::launch_aux(s, 1 - ng, nx + ng, 1 - ng, ny + ng, 1 - ng, nz + ng, false));
    } else if (central == 1 && ghost == 0) {
        HIPCHK(st::launch_aux(s, 1, nx, 1, ny, 1, nz, false));
    } else if (central == 0 && ghost == 1) {
        HIPCHK(st::launch_aux(s, 1 - ng, 0, 1 - ng, ny + ng, 1 - ng, nz + ng, false));
        HIPCHK(st::launch_aux(s, nx + 1, nx + ng, 1 - ng, ny + ng, 1 - ng, nz + ng, false));
        HIPCHK(st::launch_aux(s, 1 - ng, nx + ng, 1 - ng, 0, 1 - ng, nz + ng, false));
        HIPCHK(st::launch_aux(s, 1 - ng, nx + ng, ny + 1, ny + ng, 1 - ng, nz + ng, false));
        HIPCHK(st::launch_aux(s, 1 - ng, nx + ng, 1 - ng, ny + ng, 1 - ng, 0, false));
        HIPCHK(st::launch_aux(s, 1 - ng, nx + ng, 1 - ng, ny + ng, nz + 1, nz + ng, false));
    }
    return ST_OK;
}

int st_visflx(st_solver *s, int32_t mode) {
    NEED(s);
    void *fl = s->F[s->cur];
    switch (mode) {
    case 0:
        HIPCHK(st::launch_visc(s, fl, 0, false));
        if (fp64_sensor(s)) HIPCHK(st::launch_sensor_refine(s, true, false));
        break;
    case 1: HIPCHK(st::launch_visc_div(s, fl, s->F[s->cur ^ 1], -1, false, false)); break;
    case 2: HIPCHK(st::launch_visc_cons(s, fl, 2, false)); break;
    case 3:
        HIPCHK(st::launch_sensor(s, false));
        if (fp64_sensor(s)) HIPCHK(st::launch_sensor_refine(s, false, false));
        break;
    case 4: HIPCHK(st::launch_visc(s, fl, 4, false)); break;
    case 5: HIPCHK(st::launch_visc_cons(s, fl, 5, false)); break;
    case 6: HIPCHK(st::launch_visc_cons(s, fl, 6, false)); break;
    case 7: HIPCHK(st::launch_visc_div(s, fl, s->F[s->cur ^ 1], -1, false, true)); break;
    default: return fail(ST_ERR_ARG, "visflx mode " + std::to_string(mode) + " is not supported");
    }
    return ST_OK;
}

namespace {

int ensure_piece_buffers(st_solver *s, bool fhat, bool fln) {
    if (fhat && !s->fhat) {
        HIPCHK(hipMalloc(&s->fhat, sizeof(double) * 5 * (size_t)s->g.nvol));
        HIPCHK(st::launch_zero(s, s->fhat, 5 * (size_t)s->g.nvol));
    }
    if (fln && !s->fln) {
        HIPCHK(hipMalloc(&s->fln, sizeof(double) * 5 * (size_t)s->g.nint));
        HIPCHK(st::launch_zero(s, s->fln, 5 * (size_t)s->g.nint));
    }
    return ST_OK;
}

// One sweep over an explicit range: face fluxes of faces f0..f1 into fhat,
// then (update) the divergence of cells c0..c1.  The default range runs the
// fused sweep instead (fluxes and divergence in registers, WENO fix-up).
int sweep_range(st_solver *s, int dir, int f0, int f1, bool update, int c0, int c1, bool fused) {
    const int n = dir == 0 ? s->g.nx : (dir == 1 ? s->g.ny : s->g.nz);
    if (fused && s->g.ng == 3) {
        if (dir == 0) HIPCHK(st::launch_euler_x(s, s->F[s->cur], false));
        else if (dir == 1) HIPCHK(st::launch_euler_y(s, s->F[s->cur]));
        else HIPCHK(st::launch_euler_z(s, s->F[s->cur]));
        return ST_OK;
    }
    if (f0 < 0 || f1 > n || (update && (c0 < 1 || c1 > n)))
        return fail(ST_ERR_ARG, "sweep range outside faces 0..n / cells 1..n of direction " + std::to_string(dir + 1));
    int rc = ensure_piece_buffers(s, true, false);
    if (rc) return rc;
    if (f1 >= f0) HIPCHK(st::launch_face_range(s, dir, f0, f1));
    if (update && c1 >= c0) HIPCHK(st::launch_face_update(s, dir, s->F[s->cur], c0, c1));
    return ST_OK;
}

}  // namespace

int st_euler_x(st_solver *s, int32_t istart, int32_t iend, int32_t istart_face, int32_t iend_face,
               int32_t do_update) {
    NEED(s);
    if (iend - istart < 0) return ST_OK;
    const st::KPar &k = s->kp;
    const bool fused = istart == k.eul_lo[0] && iend == k.eul_hi[0] && istart_face == istart - 1 &&
                       iend_face == iend && do_update != 0;
    return sweep_range(s, 0, istart_face, iend_face, do_update != 0, k.eul_lo[0], k.eul_hi[0], fused);
}
int st_euler_y(st_solver *s, int32_t eul_jmin, int32_t eul_jmax) {
    NEED(s);
    const st::KPar &k = s->kp;
    const bool fused = eul_jmin == k.eul_lo[1] && eul_jmax == k.eul_hi[1];
    return sweep_range(s, 1, eul_jmin - 1, eul_jmax, true, eul_jmin, eul_jmax, fused);
}
int st_euler_z(st_solver *s, int32_t eul_kmin, int32_t eul_kmax) {
    NEED(s);
    const st::KPar &k = s->kp;
    const bool fused = eul_kmin == k.eul_lo[2] && eul_kmax == k.eul_hi[2];
    int rc = sweep_range(s, 2, eul_kmin - 1, eul_kmax, true, eul_kmin, eul_kmax, fused);
    if (rc) return rc;
    if (s->recyc) HIPCHK(st::launch_inflow_draws(s));  // get_crandom_f of euler_z
    return ST_OK;
}

int st_get_eul_range(st_solver *s, int32_t range[6]) {
    if (!s || !range) return fail(ST_ERR_ARG, "null argument");
    for (int d = 0; d < 3; d++) {
        range[2 * d] = s->kp.eul_lo[d];
        range[2 * d + 1] = s->kp.eul_hi[d];
    }
    return ST_OK;
}

int st_init_flux(st_solver *s, double rhodt) {
    NEED(s);
    int rc = ensure_piece_buffers(s, false, true);
    if (rc) return rc;
    HIPCHK(st::launch_flux_step(s, s->F[s->cur], rhodt, true));
    return ST_OK;
}

int st_update_flux(st_solver *s, double gamdt) {
    NEED(s);
    if (!s->fln) return fail(ST_ERR_STATE, "st_init_flux not called");
    HIPCHK(st::launch_flux_step(s, s->F[s->cur], gamdt, false));
    return ST_OK;
}

int st_update_field(st_solver *s) {
    NEED(s);
    if (!s->fln) return fail(ST_ERR_STATE, "st_init_flux not called");
    HIPCHK(st::launch_update_field(s));
    s->fln_current = true;
    return ST_OK;
}

int st_force_rhs(st_solver *s) {
    NEED(s);
    if (!s->fln) return fail(ST_ERR_STATE, "st_init_flux not called");
    HIPCHK(st::launch_force_rhs_piece(s));
    return ST_OK;
}

int st_force_var(st_solver *s) {
    NEED(s);
    if (s->p.theta_wall >= -1.0) HIPCHK(st::launch_force_var_piece(s));
    return ST_OK;
}

int st_reset_forcing(st_solver *s) {
    NEED(s);
    HIPCHK(st::launch_zero(s, &s->ds->dpdx, 1));
    return ST_OK;
}

int st_bcextr_var(st_solver *s, int32_t var) {
    NEED(s);
    if (var < 1 || var > 10) return fail(ST_ERR_ARG, "var must be 1..10");
    HIPCHK(st::launch_extrapolate_aux(s, var));
    return ST_OK;
}

int st_bcswap_var(st_solver *s, int32_t var) {
    NEED(s);
    if (var < 1 || var > 10) return fail(ST_ERR_ARG, "var must be 1..10");
    HIPCHK(st::launch_fill_periodic_aux(s, var));
    return ST_OK;
}

int st_recyc_exchange(st_solver *s) {
    NEED(s);
    if (!s->recyc || !s->recyc_set) return fail(ST_ERR_STATE, "no recycling inflow set (st_set_recycling)");
    HIPCHK(st::launch_recyc_exchange(s));
    return ST_OK;
}

int st_bc_nr(st_solver *s) {
    NEED(s);
    if (s->has_nr) HIPCHK(st::launch_bc_nr(s, s->F[s->cur]));
    return ST_OK;
}

int st_zero_flux(st_solver *s) {
    NEED(s);
    HIPCHK(st::launch_zero_rhs(s, s->F[0]));
    HIPCHK(st::launch_zero_rhs(s, s->F[1]));
    return ST_OK;
}

// Periodic ghost swap / extrapolation of one aux variable (bcswap_var,
// bcextr_var); used by st_init_run and exposed through the tests.
static int swap_var(st_solver *s, int var, bool extrapolate) {
    if (extrapolate) HIPCHK(st::launch_extrapolate_aux(s, var));
    HIPCHK(st::launch_fill_periodic_aux(s, var));
    return ST_OK;
}

int st_init_run(st_solver *s) {
    NEED(s);
    int rc;
    if ((rc = st_update_ghost(s, 1))) return rc;
    if ((rc = st_compute_aux(s, 1, 1))) return rc;
    if ((rc = st_zero_flux(s))) return rc;
    // visflx(2) for the conservative viscous form, else visflx(0)
    if (s->p.conservative_viscous == 1) HIPCHK(st::launch_visc_cons(s, s->F[s->cur], 2, false));
    else HIPCHK(st::launch_visc(s, s->F[s->cur], 0, false));
    if (fp64_sensor(s)) HIPCHK(st::launch_sensor_refine(s, true, false));
    if ((rc = swap_var(s, 8, false))) return rc;
    if ((rc = st_zero_flux(s))) return rc;
    s->last_stage = -1;
    if ((rc = st_compute_dt(s, nullptr))) return rc;
    // init_crandom_f before the integration loop (singleideal_cpu.F90:1585-1594)
    HIPCHK(st::launch_rng_reset(s));
    return ST_OK;
}

int st_compute_dt(st_solver *s, double *dt_out) {
    NEED(s);
    if (s->p.cfl >= 0) {
        HIPCHK(st::launch_dt(s));
    }
    if (dt_out) {
        int rc = pull_scalars(s);
        if (rc) return rc;
        *dt_out = s->hs->dt;
    }
    return ST_OK;
}

// rk_sync procedure by procedure (singleideal_cpu.F90:352-406) with the
// step's dt: the stage of four ghost layers (WENO7, ep_order 8), whose
// stencils the fused kernels do not cover.  fl is F[cur], with fln and fhat
// as in the reference.
static int rk_sync_pieces(st_solver *s, double dt) {
    const st::KPar &k = s->kp;
    const bool cons = s->p.conservative_viscous == 1;
    int rc;
#define ST_TRY(x)                     \
    do {                              \
        if ((rc = (x)) != ST_OK) return rc; \
    } while (0)
    for (int istep = 0; istep < 3; istep++) {
        ST_TRY(st_init_flux(s, s->rhork[istep] * dt));
        ST_TRY(st_bcswap(s));
        ST_TRY(st_compute_aux(s, 1, 1));
        ST_TRY(sweep_range(s, 0, k.eul_lo[0] - 1, k.eul_hi[0], true, k.eul_lo[0], k.eul_hi[0], false));
        if (cons) {
            ST_TRY(st_visflx(s, 6));
            ST_TRY(st_visflx(s, 5));
        } else {
            ST_TRY(st_visflx(s, 4));
        }
        ST_TRY(st_bcextr_var(s, 10));
        ST_TRY(st_bcswap_var(s, 10));
        ST_TRY(sweep_range(s, 1, k.eul_lo[1] - 1, k.eul_hi[1], true, k.eul_lo[1], k.eul_hi[1], false));
        ST_TRY(sweep_range(s, 2, k.eul_lo[2] - 1, k.eul_hi[2], true, k.eul_lo[2], k.eul_hi[2], false));
        if (istep == 2) {
            ST_TRY(st_visflx(s, 3));
            ST_TRY(st_bcswap_var(s, 8));
        }
        ST_TRY(st_visflx(s, cons ? 7 : 1));
        ST_TRY(st_bc_nr(s));
        ST_TRY(st_update_flux(s, s->gamrk[istep] * dt));
        if (s->p.channel_case) ST_TRY(st_force_rhs(s));
        ST_TRY(st_update_field(s));
        if (s->p.channel_case) ST_TRY(st_force_var(s));
        ST_TRY(st_update_ghost(s, 0));
    }
#undef ST_TRY
    s->last_stage = 2;
    return ST_OK;
}

int st_rk_sync(st_solver *s) {
    NEED(s);
    if (s->recyc && !s->recyc_set) return fail(ST_ERR_STATE, "st_set_recycling not called");
    if (s->g.ng == 4) {
        int rc = pull_scalars(s);
        if (rc) return rc;
        return rk_sync_pieces(s, s->hs->dt);
    }
    return do_rk_sync(s);
}

int st_advance(st_solver *s, int32_t nsteps) {
    NEED(s);
    if (s->recyc && !s->recyc_set) return fail(ST_ERR_STATE, "st_set_recycling not called");
    const int iter_dt = s->p.iter_dt_recompute > 0 ? s->p.iter_dt_recompute : 1;
    for (int n = 0; n < nsteps; n++) {
        s->icyc += 1;
        if (s->g.ng == 4) {
            // procedure by procedure: dt at the step start, the stages, the
            // step's time and counter (singleideal_cpu.F90:1599-1645)
            double dt;
            int rc;
            if (s->p.cfl > 0 && ((s->icyc - s->icyc0) % iter_dt) == 0) rc = st_compute_dt(s, &dt);
            else if ((rc = pull_scalars(s)) == ST_OK) dt = s->hs->dt;
            if (rc) return rc;
            if (s->p.channel_case && (rc = st_reset_forcing(s))) return rc;
            if ((rc = rk_sync_pieces(s, dt))) return rc;
            HIPCHK(st::launch_step_end(s));
            continue;
        }
        if (s->p.cfl > 0 && ((s->icyc - s->icyc0) % iter_dt) == 0) {
            // compute_aux(central) + compute_dt of the step start form the
            // same primitives of the same w as the first stage's aux, which
            // therefore takes the reduction
            s->dt_in_stage = true;
        }
        int rc = do_rk_sync(s);
        if (rc) return rc;
        TIMED(ST_K_STEP_END, st::launch_step_end(s));
    }
    return ST_OK;
}

int st_kernel_times(st_solver *s, int32_t nsteps, double ms[ST_NKFAM], int32_t count[ST_NKFAM]) {
    NEED(s);
    if (!ms || !count || nsteps < 1) return fail(ST_ERR_ARG, "bad arguments");
    if (s->g.ng == 4) return fail(ST_ERR_STATE, "kernel times cover the fused stage of three ghost layers");
    KTimer *t = new (std::nothrow) KTimer();
    if (!t) return fail(ST_ERR_NOMEM, "timer allocation failed");
    t->n = 0;
    const int need = nsteps * 64 < KTimer::kMax ? nsteps * 64 : KTimer::kMax;  // up to ~20 scopes a stage
    int made = 0;
    hipError_t e = hipSuccess;
    for (; made < need && e == hipSuccess; made++) {
        e = hipEventCreate(&t->ev[made][0]);
        if (e == hipSuccess) e = hipEventCreate(&t->ev[made][1]);
    }
    int rc = ST_OK;
    t->made = e == hipSuccess ? made : 0;
    if (e != hipSuccess) {
        rc = fail(ST_ERR_HIP, std::string("event creation: ") + hipGetErrorString(e));
    } else {
        s->timer = t;
        rc = st_advance(s, nsteps);
        s->timer = nullptr;
        if (rc == ST_OK) {
            e = hipStreamSynchronize(s->stream);
            if (e != hipSuccess) rc = fail(ST_ERR_HIP, hipGetErrorString(e));
        }
        if (rc == ST_OK) {
            for (int f = 0; f < ST_NKFAM; f++) { ms[f] = 0.0; count[f] = 0; }
            for (int q = 0; q < t->n && q < made; q++) {
                float el = 0.f;
                if (hipEventElapsedTime(&el, t->ev[q][0], t->ev[q][1]) == hipSuccess) {
                    ms[t->fam[q]] += el;
                    count[t->fam[q]] += 1;
                }
            }
            for (int f = 0; f < ST_NKFAM; f++)
                if (count[f]) ms[f] /= count[f];
        }
    }
    for (int q = 0; q < made; q++) {
        (void)hipEventDestroy(t->ev[q][0]);
        (void)hipEventDestroy(t->ev[q][1]);
    }
    s->timer = nullptr;
    delete t;
    return rc;
}

int st_compute_residual(st_solver *s, st_scalars *out) {
    NEED(s);
    HIPCHK(st::launch_residual(s));
    st_scalars tmp;
    int rc = st_get_scalars(s, out ? out : &tmp);
    if (rc) return rc;
    const double r = s->hs->residual;
    if (r != r) return fail(ST_ERR_NAN, "NaN residual (boom!!!)");
    return ST_OK;
}

int st_count_weno(st_solver *s, double frac_xyz[3]) {
    NEED(s);
    if (!frac_xyz) return fail(ST_ERR_ARG, "null output");
    HIPCHK(st::launch_zero(s, s->ds->cnt, 3));
    HIPCHK(st::launch_count_weno(s));
    int rc = pull_scalars(s);
    if (rc) return rc;
    const st::KPar &k = s->kp;
    const double nx = s->g.nx, ny = s->g.ny, nz = s->g.nz;
    frac_xyz[0] = s->hs->cnt[0] / ((double)(k.eul_hi[0] - k.eul_lo[0] + 1) * ny * nz);
    frac_xyz[1] = s->hs->cnt[1] / (nx * (double)(k.eul_hi[1] - k.eul_lo[1] + 1) * nz);
    frac_xyz[2] = s->hs->cnt[2] / (nx * ny * (double)(k.eul_hi[2] - k.eul_lo[2] + 1));
    return ST_OK;
}

static constexpr int kNvStat = 70;

static int ensure_stats(st_solver *s) {
    if (s->w_stat) return ST_OK;
    const size_t n = (size_t)kNvStat * s->g.nx * s->g.ny;
    HIPCHK(hipMalloc(&s->w_stat, n * sizeof(double)));
    HIPCHK(st::launch_zero(s, s->w_stat, n));
    return ST_OK;
}

int st_compute_stats(st_solver *s, int32_t itav) {
    NEED(s);
    if (itav < 0) return fail(ST_ERR_ARG, "itav must be >= 0");
    int rc = ensure_stats(s);
    if (rc) return rc;
    // the steps read periodic neighbours from their images and leave the
    // periodic ghost cells of w alone: refresh them for the gradients
    HIPCHK(st::launch_fill_periodic(s, s->w, 5));
    HIPCHK(st::launch_stats(s, itav));
    return ST_OK;
}

int st_get_stats(st_solver *s, double *w_stat_host) {
    NEED(s);
    if (!w_stat_host) return fail(ST_ERR_ARG, "null output");
    if (!s->w_stat) return fail(ST_ERR_STATE, "no statistics yet (st_compute_stats or st_set_stats first)");
    HIPCHK(hipMemcpyAsync(w_stat_host, s->w_stat, (size_t)kNvStat * s->g.nx * s->g.ny * sizeof(double),
                          hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return ST_OK;
}

int st_set_stats(st_solver *s, const double *w_stat_host) {
    NEED(s);
    if (!w_stat_host) return fail(ST_ERR_ARG, "null input");
    int rc = ensure_stats(s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(s->w_stat, w_stat_host, (size_t)kNvStat * s->g.nx * s->g.ny * sizeof(double),
                          hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return ST_OK;
}

int st_set_tuning(st_solver *s, int32_t visc_chunk, int32_t march_chunk) {
    NEED(s);
    if (visc_chunk < 0 || visc_chunk > 1024 || march_chunk < 0 || march_chunk > 1024)
        return fail(ST_ERR_ARG, "chunks must be 0 (automatic) or 1..1024");
    s->visc_chunk = visc_chunk;
    s->march_chunk = march_chunk;
    return ST_OK;
}

int st_set_stream(st_solver *s, void *hip_stream) {
    if (!s) return fail(ST_ERR_ARG, "null solver handle");
    HIPCHK(hipStreamSynchronize(s->stream));
    if (hip_stream) {
        if (s->own_stream) HIPCHK(hipStreamDestroy(s->stream));
        s->stream = (hipStream_t)hip_stream;
        s->own_stream = false;
    } else if (!s->own_stream) {
        HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->own_stream = true;
    }
    return ST_OK;
}

int st_synchronize(st_solver *s) {
    if (!s) return fail(ST_ERR_ARG, "null solver handle");
    HIPCHK(hipStreamSynchronize(s->stream));
    return ST_OK;
}

}  // extern "C"

namespace st {
int solver_create_in(const st_params *p, const SlotMem &mem, st_solver **out) { return create_solver(p, &mem, out); }
}  // namespace st
