/*
 * streams_oracle.c -- CPU oracle of the STREAmS-2 per-RK-stage hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see streams_oracle.h).  Every function restates
 * one routine of the reference CPU backend; the routine is cited as
 * file:line relative to the reference repository root.  The arithmetic keeps
 * the reference's operation order (Fortran evaluates a*b*c as (a*b)*c and
 * a+b+c as (a+b)+c) so that a correct HIP path differs from it only by
 * rounding.  Compile without floating-point contraction (-ffp-contract=off)
 * so the oracle itself is deterministic.
 *
 * Parity pinning of each function (tests/test_oracle_pins.py):
 *   or_get_deriv_coeffs       closed-form central-difference weights
 *   Wray RK3 coefficients     published values, sum(rho+gamma)=1
 *   or_eval_aux / EOS         closed forms on hand inputs, e(T(e)) round trip
 *   or_euler_fluxes (central) free-stream preservation, telescoping
 *                             conservation, KEEP kinetic-energy property,
 *                             brute-force split-form evaluation
 *   or_wenorec_1d             constants/linear data reproduced, linear weights
 *                             on smooth data
 *   or_eigenvectors/Roe       el*er = I, el*A*er = diag(u-c,u,u,u,u+c)
 *   or_visflx                 manufactured polynomial fields
 *   or_compute_dt, reductions brute force
 *   ghost fills               exact equality with index formulas
 *   conservative viscous      manufactured fields on which the second-order
 *   (modes 2, 5, 6, 7)        face/cell differences are exact
 *   whole channel path        the reference's golden progress row
 *                             (tests/tests.ini, [subsonic_channel])
 *   WENO faces, first order   the local Lax-Friedrichs flux of the Roe and
 *   (Roe and Rusanov split)   the Rusanov split in closed form
 *   or_force_var              bulk T equal to the target after the call
 *   compute_stats             a manufactured field (tests/test_stats_and_restart_files.py)
 * Boundary conditions (tests/test_oracle_bc_pins.py): free stream,
 * extrapolation, symmetry, node wall and shock faces by their definitions,
 * bc_nr against a finite-difference Jacobian and its eigen-decomposition,
 * the recycling inflow on stations with known fluctuations.
 */
#include "streams_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---------------------------------------------------------------- indexing */
static inline size_t G(const or_state *o, int i, int j, int k)
{
    return (size_t)(i + o->ng - 1) +
           (size_t)o->sx * ((size_t)(j + o->ng - 1) + (size_t)o->sy * (size_t)(k + o->ng - 1));
}
static inline size_t I(const or_state *o, int i, int j, int k)
{
    return (size_t)(i - 1) + (size_t)o->nx * ((size_t)(j - 1) + (size_t)o->ny * (size_t)(k - 1));
}
#define W(i, j, k, m)    o->w[((size_t)(m)-1) * o->nvol + G(o, i, j, k)]
#define WA(i, j, k, m)   o->w_aux[((size_t)(m)-1) * o->nvol + G(o, i, j, k)]
#define FL(i, j, k, m)   o->fl[((size_t)(m)-1) * o->nint + I(o, i, j, k)]
#define FLN(i, j, k, m)  o->fln[((size_t)(m)-1) * o->nint + I(o, i, j, k)]
#define FH(i, j, k, m)   o->fhat[((size_t)(m)-1) * o->nvol + G(o, i, j, k)]
#define EOC(i, j, k, d)  o->eoc[(size_t)(i) + (size_t)(o->nx + 1) * ((size_t)(j) + (size_t)(o->ny + 1) * (size_t)(k)) + \
                            (size_t)((d)-1) * (size_t)(o->nx + 1) * (o->ny + 1) * (o->nz + 1)]
#define CP(l)            o->p.cp_coeff[(l) - o->p.indx_cp_l]
#define XC(i)            o->x[(i) + o->ng - 1]
#define YC(j)            o->y[(j) + o->ng - 1]
#define ZC(k)            o->z[(k) + o->ng - 1]
#define CV(l)            o->p.cv_coeff[(l) - o->p.indx_cp_l]

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* x**n for integer n as binary exponentiation (gfortran's integer power) */
static double ipow(double x, int n)
{
    if (n < 0) return 1.0 / ipow(x, -n);
    double r = 1.0, b = x;
    while (n) {
        if (n & 1) r *= b;
        b *= b;
        n >>= 1;
    }
    return r;
}

/* --------------------------------------------------------------- lifecycle */
/* get_deriv_coeffs, code/src/grid.F90:143-190 */
void or_get_deriv_coeffs(double *d1, double *d2, int order1, int order2)
{
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

/* bc_preproc, code/src/singleideal/singleideal.F90:4703-4750 */
static void bc_preproc(or_state *o)
{
    int bctags_nr[6];
    memcpy(bctags_nr, o->p.bctags_nr, sizeof bctags_nr);
    for (int l = 0; l < 6; l++) o->force_zero_flux[l] = 0;
    o->eul_imin = 1; o->eul_imax = o->nx;
    o->eul_jmin = 1; o->eul_jmax = o->ny;
    o->eul_kmin = 1; o->eul_kmax = o->nz;
    for (int ilat = 1; ilat <= 6; ilat++) {
        if (o->p.bctags[ilat - 1] == 6 && o->p.is_y_staggered) {
            o->force_zero_flux[ilat - 1] = 1;
            bctags_nr[ilat - 1] = 0;
        }
        o->nr_eff[ilat - 1] = bctags_nr[ilat - 1];
        int off = bctags_nr[ilat - 1] != 0 ? 1 : 0;
        switch (ilat) {
        case 1: o->eul_imin += off; break;
        case 2: o->eul_imax -= off; break;
        case 3: o->eul_jmin += off; break;
        case 4: o->eul_jmax -= off; break;
        case 5: o->eul_kmin += off; break;
        case 6: o->eul_kmax -= off; break;
        }
    }
}

/* correct_bc_order, code/src/singleideal/singleideal.F90:946-993 (one block
 * per direction, so every block is both first and last) */
static void correct_bc_order(or_state *o)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    int ss = o->p.ep_order / 2 > o->p.weno_scheme ? o->p.ep_order / 2 : o->p.weno_scheme;
    int *t = o->p.bctags;
    if (t[0] != 0 || t[1] != 0) {
        for (int k = 0; k <= nz; k++) for (int j = 0; j <= ny; j++) {
            EOC(0, j, k, 1) = -ss + 1;
            for (int i = 1; i <= ss - 1; i++) EOC(i, j, k, 1) = -ss + i;
            EOC(nx, j, k, 1) = -ss + 1;
            for (int i = 1; i <= ss - 1; i++) EOC(nx - i, j, k, 1) = -ss + i;
        }
    }
    if (t[2] != 0 || t[3] != 0) {
        for (int k = 0; k <= nz; k++) for (int i = 0; i <= nx; i++) {
            EOC(i, 0, k, 2) = -ss + 1;
            for (int j = 1; j <= ss - 1; j++) EOC(i, j, k, 2) = -ss + j;
            EOC(i, ny, k, 2) = -ss + 1;
            for (int j = 1; j <= ss - 1; j++) EOC(i, ny - j, k, 2) = -ss + j;
        }
    }
    if (t[4] != 0 || t[5] != 0) {
        for (int j = 0; j <= ny; j++) for (int i = 0; i <= nx; i++) {
            EOC(i, j, 0, 3) = -ss + 1;
            for (int k = 1; k <= ss - 1; k++) EOC(i, j, k, 3) = -ss + k;
            EOC(i, j, nz, 3) = -ss + 1;
            for (int k = 1; k <= ss - 1; k++) EOC(i, j, nz - k, 3) = -ss + k;
        }
    }
}

or_state *or_create(const st_params *p)
{
    or_state *o = (or_state *)calloc(1, sizeof(or_state));
    if (!o) return NULL;
    o->p = *p;
    o->nx = p->nx; o->ny = p->ny; o->nz = p->nz; o->ng = p->ng;
    o->sx = p->nx + 2 * p->ng; o->sy = p->ny + 2 * p->ng; o->sz = p->nz + 2 * p->ng;
    o->nvol = (size_t)o->sx * o->sy * o->sz;
    o->nint = (size_t)p->nx * p->ny * p->nz;
    o->w = (double *)calloc(5 * o->nvol, sizeof(double));
    o->w_aux = (double *)calloc(10 * o->nvol, sizeof(double));
    /* initialize (singleideal.F90:3032): T = t0 as the first Newton start */
    for (size_t q = 0; q < o->nvol; q++) o->w_aux[5 * o->nvol + q] = p->t0;
    o->fl = (double *)calloc(5 * o->nint, sizeof(double));
    o->fln = (double *)calloc(5 * o->nint, sizeof(double));
    o->fhat = (double *)calloc(5 * o->nvol, sizeof(double));
    o->wallprop = (double *)calloc(3 * (size_t)o->sx * o->sz, sizeof(double));
    o->eoc = (int *)calloc(3 * (size_t)(p->nx + 1) * (p->ny + 1) * (p->nz + 1), sizeof(int));
    o->x = (double *)calloc(o->sx, sizeof(double));
    o->y = (double *)calloc(o->sy, sizeof(double));
    o->z = (double *)calloc(o->sz, sizeof(double));
    o->yn = (double *)calloc(p->ny + 1, sizeof(double));
    o->dcsidx = (double *)calloc(p->nx, sizeof(double));
    o->dcsidxs = (double *)calloc(p->nx, sizeof(double));
    o->dcsidx2 = (double *)calloc(p->nx, sizeof(double));
    o->detady = (double *)calloc(p->ny, sizeof(double));
    o->detadys = (double *)calloc(p->ny, sizeof(double));
    o->detady2 = (double *)calloc(p->ny, sizeof(double));
    o->dzitdz = (double *)calloc(p->nz, sizeof(double));
    o->dzitdzs = (double *)calloc(p->nz, sizeof(double));
    o->dzitdz2 = (double *)calloc(p->nz, sizeof(double));
    /* coefficient tables of orders 2,4,6,8 (singleideal.F90:518-525) */
    for (int i = 0; i < 4; i++) {
        double d1[4], d2[5];
        or_get_deriv_coeffs(d1, d2, 2 * (i + 1), 2 * (i + 1));
        for (int l = 0; l < 4; l++) o->cd1[i][l] = d1[l];
        for (int l = 0; l < 5; l++) o->cd2[i][l] = d2[l];
    }
    /* Wray RK3, runge_kutta_initialize, singleideal.F90:4773-4782 */
    o->rhork[0] = 0.0; o->rhork[1] = -17.0 / 60.0; o->rhork[2] = -5.0 / 12.0;
    o->gamrk[0] = 8.0 / 15.0; o->gamrk[1] = 5.0 / 12.0; o->gamrk[2] = 3.0 / 4.0;
    bc_preproc(o);
    if (p->correct_bound_ord > 0) correct_bc_order(o);
    o->print_control = 1;
    return o;
}

void or_destroy(or_state *o)
{
    if (!o) return;
    free(o->w); free(o->w_aux); free(o->fl); free(o->fln); free(o->fhat); free(o->wallprop);
    free(o->rc_wmean); free(o->rc_weta); free(o->rc_inn); free(o->rc_out); free(o->rc_y); free(o->rc_wrecyc);
    free(o->rc_plane);
    free(o->eoc); free(o->x); free(o->y); free(o->z); free(o->yn);
    free(o->dcsidx); free(o->dcsidxs); free(o->dcsidx2);
    free(o->detady); free(o->detadys); free(o->detady2);
    free(o->dzitdz); free(o->dzitdzs); free(o->dzitdz2);
    free(o);
}

int or_set_grid(or_state *o, const double *x, const double *y, const double *z, const double *yn,
                const double *dcsidx, const double *dcsidxs, const double *dcsidx2,
                const double *detady, const double *detadys, const double *detady2,
                const double *dzitdz, const double *dzitdzs, const double *dzitdz2)
{
    memcpy(o->x, x, sizeof(double) * o->sx);
    memcpy(o->y, y, sizeof(double) * o->sy);
    memcpy(o->z, z, sizeof(double) * o->sz);
    memcpy(o->yn, yn, sizeof(double) * (o->ny + 1));
    memcpy(o->dcsidx, dcsidx, sizeof(double) * o->nx);
    memcpy(o->dcsidxs, dcsidxs, sizeof(double) * o->nx);
    memcpy(o->dcsidx2, dcsidx2, sizeof(double) * o->nx);
    memcpy(o->detady, detady, sizeof(double) * o->ny);
    memcpy(o->detadys, detadys, sizeof(double) * o->ny);
    memcpy(o->detady2, detady2, sizeof(double) * o->ny);
    memcpy(o->dzitdz, dzitdz, sizeof(double) * o->nz);
    memcpy(o->dzitdzs, dzitdzs, sizeof(double) * o->nz);
    memcpy(o->dzitdz2, dzitdz2, sizeof(double) * o->nz);
    return 0;
}

double *or_w(or_state *o) { return o->w; }
double *or_w_aux(or_state *o) { return o->w_aux; }
double *or_fl(or_state *o) { return o->fl; }
double *or_fln(or_state *o) { return o->fln; }
double *or_fhat(or_state *o) { return o->fhat; }
double *or_wallprop(or_state *o) { return o->wallprop; }
int *or_eoc(or_state *o) { return o->eoc; }

void or_get_scalars(const or_state *o, st_scalars *s)
{
    s->dt = o->dt; s->time = o->time; s->dpdx = o->dpdx;
    s->rhobulk = o->rhobulk; s->ubulk = o->ubulk; s->tbulk = o->tbulk;
    s->residual_rhou = o->residual_rhou;
    s->rhomin = o->rhomin; s->rhomax = o->rhomax; s->tmin = o->tmin; s->tmax = o->tmax;
    s->pmin = o->pmin; s->pmax = o->pmax;
    s->icyc = o->icyc;
}

void or_rk_coeffs(const or_state *o, double rho[3], double gam[3])
{
    for (int q = 0; q < 3; q++) { rho[q] = o->rhork[q]; gam[q] = o->gamrk[q]; }
}

void or_set_time(or_state *o, double time, double dt, int64_t icyc)
{
    o->time = time; o->dt = dt; o->icyc = icyc; o->icyc0 = icyc;
}

/* ------------------------------------------------------------ thermodynamics */
/* get_gamloc_dev, kernels_cpu.F90:7096-7113 */
double or_get_gamloc(const or_state *o, double tt)
{
    double cploc;
    if (o->p.calorically_perfect == 1) {
        cploc = CP(0);
    } else {
        cploc = 0.0;
        for (int l = o->p.indx_cp_l; l <= o->p.indx_cp_r; l++) cploc = cploc + CP(l) * ipow(tt / o->p.t0, l);
    }
    return cploc / (cploc - o->p.rgas0);
}

/* get_temperature_from_e_dev, kernels_cpu.F90:7116-7152 */
double or_get_temperature_from_e(const or_state *o, double ee, double t_start)
{
    double t0 = o->p.t0, tt;
    double ebar = ee - CV(o->p.indx_cp_r + 1) * t0;
    if (o->p.calorically_perfect == 1) {
        tt = t0 + ebar / CV(0);
    } else {
        double t_old = t_start;
        tt = t_old;
        for (int iter = 1; iter <= 50; iter++) {
            double den = 0.0, num = 0.0;
            for (int l = o->p.indx_cp_l; l <= o->p.indx_cp_r; l++) {
                if (l == -1) {
                    double t_pow = ipow(t_old / t0, l);
                    den = den + CV(l) * t_pow;
                    num = num + CV(l) * log(t_old / t0);
                } else {
                    double t_pow = ipow(t_old / t0, l);
                    double t_powp = (t_old / t0) * t_pow;
                    den = den + CV(l) * t_pow;
                    num = num + CV(l) * (t_powp - 1.0) / (l + 1.0);
                }
            }
            num = num * t0;
            tt = t_old + (ebar - num) / den;
            if (fabs(tt - t_old) < o->p.tol_iter_nr) break;
            t_old = tt;
        }
    }
    return tt;
}

/* get_e_from_temperature_dev, kernels_cpu.F90:7155-7176 */
double or_get_e_from_temperature(const or_state *o, double tt)
{
    double t0 = o->p.t0;
    double ee = CV(o->p.indx_cp_r + 1);
    if (o->p.calorically_perfect == 1) {
        ee = ee + CV(0) * (tt / t0 - 1.0);
    } else {
        for (int l = o->p.indx_cp_l; l <= o->p.indx_cp_r; l++) {
            if (l == -1) ee = ee + CV(l) * log(tt / t0);
            else ee = ee + CV(l) / (l + 1.0) * (ipow(tt / t0, l + 1) - 1.0);
        }
    }
    return ee * t0;
}

/* ------------------------------------------------------------------ eval_aux */
/* eval_aux_subroutine, kernels_cpu.F90:6960-7011 */
void or_eval_aux(or_state *o, int i0, int i1, int j0, int j1, int k0, int k1)
{
    const st_params *p = &o->p;
#pragma omp parallel for schedule(static)
    for (int k = k0; k <= k1; k++)
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++) {
                double rho = W(i, j, k, 1), rhou = W(i, j, k, 2), rhov = W(i, j, k, 3);
                double rhow = W(i, j, k, 4), rhoe = W(i, j, k, 5);
                double ri = 1.0 / rho;
                double uu = rhou * ri, vv = rhov * ri, ww = rhow * ri;
                double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
                double ee = rhoe / rho - qq;
                double tt = or_get_temperature_from_e(o, ee, WA(i, j, k, 6));
                double pp = rho * tt * p->rgas0;
                WA(i, j, k, 1) = rho;
                WA(i, j, k, 2) = uu;
                WA(i, j, k, 3) = vv;
                WA(i, j, k, 4) = ww;
                WA(i, j, k, 5) = (rhoe + pp) / rho;
                WA(i, j, k, 6) = tt;
                double mu = 0.0;
                if (p->visc_model == 1)
                    mu = p->mu0 * pow(tt / p->t0, p->powerlaw_vtexp);
                else if (p->visc_model == 2)
                    mu = p->mu0 * pow(tt / p->t0, 1.5) * (1.0 + p->sutherland_s / p->t_ref_dim) /
                         (tt / p->t0 + p->sutherland_s / p->t_ref_dim);
                WA(i, j, k, 7) = mu;
            }
}

/* compute_aux, singleideal_cpu.F90:281-330 */
void or_compute_aux(or_state *o, int central, int ghost)
{
    int nx = o->nx, ny = o->ny, nz = o->nz, ng = o->ng;
    if (central == 1 && ghost == 1) {
        or_eval_aux(o, 1 - ng, nx + ng, 1 - ng, ny + ng, 1 - ng, nz + ng);
    } else if (central == 1 && ghost == 0) {
        or_eval_aux(o, 1, nx, 1, ny, 1, nz);
    } else if (central == 0 && ghost == 1) {
        or_eval_aux(o, 1 - ng, 0, 1 - ng, ny + ng, 1 - ng, nz + ng);
        or_eval_aux(o, nx + 1, nx + ng, 1 - ng, ny + ng, 1 - ng, nz + ng);
        or_eval_aux(o, 1 - ng, nx + ng, 1 - ng, 0, 1 - ng, nz + ng);
        or_eval_aux(o, 1 - ng, nx + ng, ny + 1, ny + ng, 1 - ng, nz + ng);
        or_eval_aux(o, 1 - ng, nx + ng, 1 - ng, ny + ng, 1 - ng, 0);
        or_eval_aux(o, 1 - ng, nx + ng, 1 - ng, ny + ng, nz + 1, nz + ng);
    }
}

/* ------------------------------------------------------------- ghost fills */
/* bcswap on one rank per direction: pack (base_cpu.F90:410-450), local copy
 * when the neighbour is self (:1382-1417), unpack (:669-745).  Non-periodic
 * directions have mpi_proc_null neighbours and are left untouched. */
static void swap_array(or_state *o, double *arr, int nv)
{
    int nx = o->nx, ny = o->ny, nz = o->nz, ng = o->ng;
    for (int m = 1; m <= nv; m++) {
        double *a = arr + (size_t)(m - 1) * o->nvol;
        if (o->p.bctags[0] == 0) /* x periodic */
            for (int k = 1; k <= nz; k++) for (int j = 1; j <= ny; j++) for (int i = 1; i <= ng; i++) {
                double left = a[G(o, i, j, k)], right = a[G(o, nx - ng + i, j, k)];
                a[G(o, i - ng, j, k)] = right;
                a[G(o, nx + i, j, k)] = left;
            }
        if (o->p.bctags[2] == 0) /* y periodic */
            for (int k = 1; k <= nz; k++) for (int j = 1; j <= ng; j++) for (int i = 1; i <= nx; i++) {
                double lo = a[G(o, i, j, k)], hi = a[G(o, i, ny - ng + j, k)];
                a[G(o, i, j - ng, k)] = hi;
                a[G(o, i, ny + j, k)] = lo;
            }
        if (o->p.bctags[4] == 0 && o->nz > 1) /* z periodic (ndim == 3) */
            for (int k = 1; k <= ng; k++) for (int j = 1; j <= ny; j++) for (int i = 1; i <= nx; i++) {
                double lo = a[G(o, i, j, k)], hi = a[G(o, i, j, nz - ng + k)];
                a[G(o, i, j, k - ng)] = hi;
                a[G(o, i, j, nz + k)] = lo;
            }
    }
}

/* bcswap, base_cpu.F90:1326-1435 */
void or_bcswap(or_state *o) { swap_array(o, o->w, 5); }

/* bcswap_var, base_cpu.F90:1835-1912 */
void or_bcswap_var(or_state *o, int var) { swap_array(o, o->w_aux + (size_t)(var - 1) * o->nvol, 1); }

/* bcextr_var_subroutine, kernels_cpu.F90:6199-6254 */
void or_bcextr_var(or_state *o, int var)
{
    int nx = o->nx, ny = o->ny, nz = o->nz, ng = o->ng;
    double *a = o->w_aux + (size_t)(var - 1) * o->nvol;
    for (int k = 1; k <= nz; k++) for (int j = 1; j <= ny; j++) for (int l = 1; l <= ng; l++)
        a[G(o, 1 - l, j, k)] = a[G(o, 1, j, k)];
    for (int k = 1; k <= nz; k++) for (int j = 1; j <= ny; j++) for (int l = 1; l <= ng; l++)
        a[G(o, nx + l, j, k)] = a[G(o, nx, j, k)];
    for (int k = 1; k <= nz; k++) for (int i = 1; i <= nx; i++) for (int l = 1; l <= ng; l++)
        a[G(o, i, 1 - l, k)] = a[G(o, i, 1, k)];
    for (int k = 1; k <= nz; k++) for (int i = 1; i <= nx; i++) for (int l = 1; l <= ng; l++)
        a[G(o, i, ny + l, k)] = a[G(o, i, ny, k)];
    for (int j = 1; j <= ny; j++) for (int i = 1; i <= nx; i++) for (int l = 1; l <= ng; l++)
        a[G(o, i, j, 1 - l)] = a[G(o, i, j, 1)];
    for (int j = 1; j <= ny; j++) for (int i = 1; i <= nx; i++) for (int l = 1; l <= ng; l++)
        a[G(o, i, j, nz + l)] = a[G(o, i, j, nz)];
}

/* bcwall_staggered_subroutine, kernels_cpu.F90:6612-6678 */
static void bcwall_staggered(or_state *o, int ilat)
{
    int nx = o->nx, ny = o->ny, nz = o->nz, ng = o->ng;
    const st_params *p = &o->p;
    if (ilat != 3 && ilat != 4) return;
    for (int k = 1; k <= nz; k++)
        for (int l = 1; l <= ng; l++)
            for (int i = 1; i <= nx; i++) {
                int js = ilat == 3 ? l : ny + 1 - l;   /* interior source */
                int jg = ilat == 3 ? 1 - l : ny + l;   /* ghost target    */
                double rho = W(i, js, k, 1);
                double uu = W(i, js, k, 2) / rho;
                double vv = W(i, js, k, 3) / rho;
                double ww = W(i, js, k, 4) / rho;
                double rhoe = W(i, js, k, 5);
                double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
                double ee = rhoe / rho - qq;
                double tt = or_get_temperature_from_e(o, ee, WA(i, js, k, 6));
                double pp = rho * tt * p->rgas0;
                tt = 2.0 * p->t_wall - tt;
                ee = or_get_e_from_temperature(o, tt);
                rho = pp / tt / p->rgas0;
                W(i, jg, k, 1) = rho;
                W(i, jg, k, 2) = -rho * uu;
                W(i, jg, k, 3) = -rho * vv;
                W(i, jg, k, 4) = -rho * ww;
                W(i, jg, k, 5) = rho * (ee + qq);
            }
}

/* Ghost point (a, b, l) of face ilat: the two tangential indices and the
 * depth l = 1..ng, and its mirror / boundary source. */
static size_t ghost_at(const or_state *o, int ilat, int a, int b, int l)
{
    switch (ilat) {
    case 1: return G(o, 1 - l, a, b);
    case 2: return G(o, o->nx + l, a, b);
    case 3: return G(o, a, 1 - l, b);
    case 4: return G(o, a, o->ny + l, b);
    case 5: return G(o, a, b, 1 - l);
    default: return G(o, a, b, o->nz + l);
    }
}
static size_t boundary_at(const or_state *o, int ilat, int a, int b)
{
    switch (ilat) {
    case 1: return G(o, 1, a, b);
    case 2: return G(o, o->nx, a, b);
    case 3: return G(o, a, 1, b);
    case 4: return G(o, a, o->ny, b);
    case 5: return G(o, a, b, 1);
    default: return G(o, a, b, o->nz);
    }
}
/* tangential extents (first, second index) of face ilat */
static void face_extent(const or_state *o, int ilat, int *na, int *nb)
{
    if (ilat <= 2) { *na = o->ny; *nb = o->nz; }
    else if (ilat <= 4) { *na = o->nx; *nb = o->nz; }
    else { *na = o->nx; *nb = o->ny; }
}

/* bcfree_subroutine (kernels_cpu.F90:5989-6056): ghosts = winf */
static void bcfree(or_state *o, int ilat)
{
    int na, nb;
    face_extent(o, ilat, &na, &nb);
    for (int b = 1; b <= nb; b++) for (int a = 1; a <= na; a++) for (int l = 1; l <= o->ng; l++)
        for (int m = 0; m < 5; m++) o->w[(size_t)m * o->nvol + ghost_at(o, ilat, a, b, l)] = o->p.winf[m];
}

/* bcextr_subroutine (kernels_cpu.F90:6323-6390): ghosts = boundary cell */
static void bcextr(or_state *o, int ilat)
{
    int na, nb;
    face_extent(o, ilat, &na, &nb);
    for (int b = 1; b <= nb; b++) for (int a = 1; a <= na; a++) for (int l = 1; l <= o->ng; l++)
        for (int m = 0; m < 5; m++)
            o->w[(size_t)m * o->nvol + ghost_at(o, ilat, a, b, l)] = o->w[(size_t)m * o->nvol + boundary_at(o, ilat, a, b)];
}

/* bcsym_subroutine (kernels_cpu.F90:6394-6418): ymin only, w(i,1-l) =
 * w(i,1+l) for all five variables */
static void bcsym(or_state *o, int ilat)
{
    if (ilat != 3) return;
    for (int k = 1; k <= o->nz; k++) for (int i = 1; i <= o->nx; i++) for (int l = 1; l <= o->ng; l++)
        for (int m = 1; m <= 5; m++) W(i, 1 - l, k, m) = W(i, 1 + l, k, m);
}

/* bcwall_subroutine (kernels_cpu.F90:6478-6525), node-centred isothermal
 * wall at ymin: the wall row j=1 is set to rest at T_wall, the ghosts mirror
 * rows 1+l with T_g = 2 T_wall - T, p_g = p, u_g = -u. */
static void bcwall_node(or_state *o, int ilat)
{
    const st_params *p = &o->p;
    if (ilat != 3) return;
    for (int k = 1; k <= o->nz; k++)
        for (int i = 1; i <= o->nx; i++) {
            W(i, 1, k, 2) = 0.0;
            W(i, 1, k, 3) = 0.0;
            W(i, 1, k, 4) = 0.0;
            double ee = or_get_e_from_temperature(o, p->t_wall);
            W(i, 1, k, 5) = W(i, 1, k, 1) * ee;
            for (int l = 1; l <= o->ng; l++) {
                double rho = W(i, 1 + l, k, 1);
                double uu = W(i, 1 + l, k, 2) / rho;
                double vv = W(i, 1 + l, k, 3) / rho;
                double ww = W(i, 1 + l, k, 4) / rho;
                double rhoe = W(i, 1 + l, k, 5);
                double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
                ee = rhoe / rho - qq;
                double tt = or_get_temperature_from_e(o, ee, WA(i, 1 + l, k, 6));
                double pp = rho * tt * p->rgas0;
                tt = 2.0 * p->t_wall - tt;
                ee = or_get_e_from_temperature(o, tt);
                rho = pp / tt / p->rgas0;
                W(i, 1 - l, k, 1) = rho;
                W(i, 1 - l, k, 2) = -rho * uu;
                W(i, 1 - l, k, 3) = -rho * vv;
                W(i, 1 - l, k, 4) = -rho * ww;
                W(i, 1 - l, k, 5) = rho * (ee + qq);
            }
        }
}

/* bcshock_subroutine (kernels_cpu.F90:6159-6196), ymax: rows ny..ny+ng
 * (the last interior row included) take the tanh blend of winf and the
 * post-shock state near the shock trace, the value of row ny elsewhere. */
static void bcshock(or_state *o, int ilat)
{
    const st_params *p = &o->p;
    if (ilat != 4) return;
    int ny = o->ny;
    double tanhlen = 8.0 * p->tanhfacs;
    for (int k = 1; k <= o->nz; k++)
        for (int i = 1; i <= o->nx; i++)
            for (int l = 0; l <= o->ng; l++) {
                double xsh = p->xshock_imp - YC(ny + l) / tan(p->shock_angle);
                double xx = XC(i) - xsh;
                if (fabs(xx) > tanhlen) {
                    for (int m = 1; m <= 5; m++) W(i, ny + l, k, m) = W(i, ny, k, m);
                } else {
                    double tanhf = 0.5 * (1.0 + tanh(xx / p->tanhfacs));
                    for (int m = 1; m <= 5; m++) {
                        double dwinf = p->winf_past_shock[m - 1] - p->winf[m - 1];
                        W(i, ny + l, k, m) = p->winf[m - 1] + dwinf * tanhf;
                    }
                }
            }
}

/* Recycling setup of recyc_prepare (singleideal.F90:1512-1588), given by the
 * caller: station i_recyc, rescaling factors, the inflow mean profile
 * wmean(1-ng:0, 1:ny, 1:4), blending weights, the y maps between inflow
 * and recycling station and the y+ / eta coordinates of both. */
/* wmean(1-ng:0, 1:ny, 1:4) of the inflow, Fortran order (offset 0 is plane
 * 1-ng): the data of bclam and of the recycling inflow. */
double *or_inflow_plane(or_state *o) { return o->rc_plane; }

int or_set_wmean(or_state *o, const double *wmean)
{
    const size_t n = (size_t)o->ng * o->ny * 4;
    if (!o->rc_wmean) o->rc_wmean = (double *)calloc(n, sizeof(double));
    if (!o->rc_wmean) return 1;
    memcpy(o->rc_wmean, wmean, sizeof(double) * n);
    return 0;
}

/* init_crandom_f (crandom_f.F90, crandom.c) as run() calls it before the
 * integration loop (singleideal_cpu.F90:1585-1594): seed rank+1 = 1 for one
 * rank when rand_type > 0, a clock-derived seed when rand_type < 0.  The C
 * library's generator is used through random_r with a private state, whose
 * sequence is that of srand(seed)/rand() (TYPE_3, 128-byte state). */
void or_seed_random(or_state *o)
{
    struct random_data *rd = (struct random_data *)o->rng_data;
    _Static_assert(sizeof(struct random_data) <= sizeof(o->rng_data), "random_data does not fit");
    memset(o->rng_data, 0, sizeof(o->rng_data));
    unsigned seed = 1;
    if (o->p.rand_type < 0) seed = (unsigned)time(NULL) + 1u;
    initstate_r(seed, o->rng_state, sizeof(o->rng_state), rd);
    srandom_r(seed, rd);
}

/* inflow_random_plane(2:jbl_inflow, 1:nz, 1:3) redrawn by get_crandom_f
 * after the z sweep (singleideal_cpu.F90:841): j outermost, component
 * innermost (get_crandom_f_array3d, crandom_f.F90), each rand()/RAND_MAX;
 * jbl_inflow = locateval(y(1:ny), ny, l0) (singleideal.F90:3284). */
static void draw_inflow_plane(or_state *o)
{
    if (o->p.rand_type == 0 || !o->rc_plane) return;  /* rescale 0: the term vanishes */
    int ny = o->ny, nz = o->nz, jbl = 0;
    for (int j = 1; j <= ny; j++) if (o->p.l0 > YC(j)) jbl = j;
    struct random_data *rd = (struct random_data *)o->rng_data;
    for (int j = 2; j <= jbl; j++)
        for (int k = 1; k <= nz; k++)
            for (int c = 1; c <= 3; c++) {
                int32_t r;
                random_r(rd, &r);
                o->rc_plane[(size_t)(j - 1) + (size_t)ny * ((size_t)(k - 1) + (size_t)nz * (c - 1))] =
                    (double)r / 2147483647.0;
            }
}

int or_set_recycling(or_state *o, int i_recyc, double betarecyc, double glund1, const double *wmean,
                     const double *weta, const int *map_inn, const int *map_out, const double *yplus_inflow,
                     const double *eta_inflow, const double *yplus_recyc, const double *eta_recyc)
{
    int ng = o->ng, ny = o->ny, nz = o->nz, ny2 = ny + 2 * ng;
    if (i_recyc < ng || i_recyc > o->nx) return 1;
    o->i_recyc = i_recyc;
    o->betarecyc = betarecyc;
    o->glund1 = glund1;
    if (!o->rc_plane) {
        o->rc_plane = (double *)malloc(sizeof(double) * (size_t)ny * nz * 3);
        for (size_t q = 0; q < (size_t)ny * nz * 3; q++) o->rc_plane[q] = 0.5;
        or_seed_random(o);
    }
    if (!o->rc_wmean) {
        o->rc_wmean = (double *)calloc((size_t)ng * ny * 4, sizeof(double));
        o->rc_weta = (double *)calloc(ny, sizeof(double));
        o->rc_inn = (int *)calloc(ny, sizeof(int));
        o->rc_out = (int *)calloc(ny, sizeof(int));
        o->rc_y = (double *)calloc(4 * (size_t)ny2, sizeof(double));
        o->rc_wrecyc = (double *)calloc((size_t)ng * ny * nz * 5, sizeof(double));
    }
    if (!o->rc_weta) {
        o->rc_weta = (double *)calloc(ny, sizeof(double));
        o->rc_inn = (int *)calloc(ny, sizeof(int));
        o->rc_out = (int *)calloc(ny, sizeof(int));
        o->rc_y = (double *)calloc(4 * (size_t)ny2, sizeof(double));
        o->rc_wrecyc = (double *)calloc((size_t)ng * ny * nz * 5, sizeof(double));
    }
    memcpy(o->rc_wmean, wmean, sizeof(double) * ng * ny * 4);
    memcpy(o->rc_weta, weta, sizeof(double) * ny);
    memcpy(o->rc_inn, map_inn, sizeof(int) * ny);
    memcpy(o->rc_out, map_out, sizeof(int) * ny);
    memcpy(o->rc_y, yplus_inflow, sizeof(double) * ny2);
    memcpy(o->rc_y + ny2, eta_inflow, sizeof(double) * ny2);
    memcpy(o->rc_y + 2 * ny2, yplus_recyc, sizeof(double) * ny2);
    memcpy(o->rc_y + 3 * ny2, eta_recyc, sizeof(double) * ny2);
    o->recyc_set = 1;
    return 0;
}

#define WRC(i, j, k, m) o->rc_wrecyc[(size_t)((i)-1) + (size_t)o->ng * ((size_t)((j)-1) + (size_t)o->ny * ((size_t)((k)-1) + (size_t)o->nz * ((m)-1)))]

/* recyc_exchange (singleideal_cpu.F90:935-999) for one block in x and z:
 * wrecyc(i,j,k) = w(i_recyc+1-i, j, k shifted by nzmax/2)
 * (recyc_exchange_subroutine_1/2/3, kernels_cpu.F90:5003-5060) */
static void recyc_exchange(or_state *o)
{
    int nz = o->nz, h = (nz / 2) % nz;
    for (int m = 1; m <= 5; m++)
        for (int k = 1; k <= nz; k++) {
            int ks = k > h ? k - h : k + nz - h;
            for (int j = 1; j <= o->ny; j++)
                for (int i = 1; i <= o->ng; i++) WRC(i, j, k, m) = W(o->i_recyc + 1 - i, j, ks, m);
        }
}

/* bcrecyc (singleideal_cpu.F90:1140-1176) with bcrecyc_subroutine_1/2/3
 * (kernels_cpu.F90:5810-5947): spanwise Favre averages of the recycled
 * planes, their fluctuations interpolated to the inflow y (inner y+ and
 * outer eta scaling blended by weta) and added, rescaled, to the inflow
 * mean profile; with rand_type != 0 plus 0.02*u0*(r - 0.5)*min(eta, 1) from
 * the random inflow plane on each velocity fluctuation. */
static void bcrecyc(or_state *o)
{
    const st_params *p = &o->p;
    int ng = o->ng, ny = o->ny, nz = o->nz, ny2 = ny + 2 * ng;
    double *av = (double *)calloc((size_t)ng * ny * 5, sizeof(double));
#define AV(i, j, m) av[(size_t)((i)-1) + (size_t)ng * ((size_t)((j)-1) + (size_t)ny * ((m)-1))]
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= ng; i++)
            for (int m = 1; m <= 5; m++) {
                double a = 0.0;
                for (int k = 1; k <= nz; k++) a = a + WRC(i, j, k, m);
                AV(i, j, m) = a;
            }
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= ng; i++) {
            double ufav = AV(i, j, 2) / AV(i, j, 1);
            double vfav = AV(i, j, 3) / AV(i, j, 1);
            double wfav = AV(i, j, 4) / AV(i, j, 1);
            double rhom = AV(i, j, 1) / nz;
            for (int k = 1; k <= nz; k++) {
                WRC(i, j, k, 2) = WRC(i, j, k, 2) / WRC(i, j, k, 1) - ufav;
                WRC(i, j, k, 3) = WRC(i, j, k, 3) / WRC(i, j, k, 1) - vfav;
                WRC(i, j, k, 4) = WRC(i, j, k, 4) / WRC(i, j, k, 1) - wfav;
                WRC(i, j, k, 1) = WRC(i, j, k, 1) - rhom;
            }
        }
#undef AV
    free(av);
    const double *ypi = o->rc_y, *eti = o->rc_y + ny2, *ypr = o->rc_y + 2 * ny2, *etr = o->rc_y + 3 * ny2;
#define YI(a, j) a[(j) + ng - 1]
    const double p0 = p->rho0 * p->rgas0 * p->t0;
    const double u0_02 = p->rand_type == 0 ? 0.0 : 0.02 * p->u0;
#define RP(j, k, c) o->rc_plane[(size_t)((j)-1) + (size_t)ny * ((size_t)((k)-1) + (size_t)nz * ((c)-1))]
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++) {
            double eta = YI(eti, j);
            double etamin = dmin(eta, 1.0);
            double weta = o->rc_weta[j - 1];
            double weta1 = 1.0 - weta;
            double bdamp = 0.5 * (1.0 - tanh(4.0 * (YI(eti, j) - 2.0)));
            int j_inn = o->rc_inn[j - 1], j_out = o->rc_out[j - 1];
            double disty_inn = 0.0, disty_out = 0.0;
            if (!(j == 1 || j_inn >= ny || j_out >= ny)) {
                disty_inn = (YI(ypi, j) - YI(ypr, j_inn)) / (YI(ypr, j_inn + 1) - YI(ypr, j_inn));
                disty_out = (eta - YI(etr, j_out)) / (YI(etr, j_out + 1) - YI(etr, j_out));
            }
            for (int i = 1; i <= ng; i++) {
                double rhomean = o->rc_wmean[(size_t)(ng - i) + (size_t)ng * ((size_t)(j - 1) + (size_t)ny * 0)];
                double uumean = o->rc_wmean[(size_t)(ng - i) + (size_t)ng * ((size_t)(j - 1) + (size_t)ny * 1)] / rhomean;
                double vvmean = o->rc_wmean[(size_t)(ng - i) + (size_t)ng * ((size_t)(j - 1) + (size_t)ny * 2)] / rhomean;
                double wwmean = o->rc_wmean[(size_t)(ng - i) + (size_t)ng * ((size_t)(j - 1) + (size_t)ny * 3)] / rhomean;
                double rhofluc, ufluc, vfluc, wfluc;
                if (j == 1 || j_inn >= ny || j_out >= ny) {
                    rhofluc = 0.0; ufluc = 0.0; vfluc = 0.0; wfluc = 0.0;
                } else {
                    double f_in[5], f_out[5];
                    for (int m = 1; m <= 4; m++) {
                        f_in[m] = WRC(i, j_inn, k, m) * (1.0 - disty_inn) + WRC(i, j_inn + 1, k, m) * disty_inn;
                        f_out[m] = WRC(i, j_out, k, m) * (1.0 - disty_out) + WRC(i, j_out + 1, k, m) * disty_out;
                    }
                    rhofluc = f_in[1] * weta1 + f_out[1] * weta;
                    ufluc = f_in[2] * weta1 + f_out[2] * weta;
                    vfluc = f_in[3] * weta1 + f_out[3] * weta;
                    wfluc = f_in[4] * weta1 + f_out[4] * weta;
                    rhofluc = rhofluc * bdamp;
                    ufluc = ufluc * bdamp * o->betarecyc;
                    vfluc = vfluc * bdamp * o->betarecyc;
                    wfluc = wfluc * bdamp * o->betarecyc;
                    ufluc = ufluc + u0_02 * (RP(j, k, 1) - 0.5) * etamin;
                    vfluc = vfluc + u0_02 * (RP(j, k, 2) - 0.5) * etamin;
                    wfluc = wfluc + u0_02 * (RP(j, k, 3) - 0.5) * etamin;
                }
                rhofluc = dmax(-0.75 * rhomean, rhofluc);
                rhofluc = dmin(4.0 * rhomean, rhofluc);
                double rho = rhomean + rhofluc;
                double tt = p0 / rho / p->rgas0;
                double uu = uumean + ufluc, vv = vvmean + vfluc, ww = wwmean + wfluc;
                double rhou = rho * uu, rhov = rho * vv, rhow = rho * ww;
                double ee = or_get_e_from_temperature(o, tt);
                W(1 - i, j, k, 1) = rho;
                W(1 - i, j, k, 2) = rhou;
                W(1 - i, j, k, 3) = rhov;
                W(1 - i, j, k, 4) = rhow;
                W(1 - i, j, k, 5) = rho * ee + 0.5 * (rhou * rhou + rhov * rhov + rhow * rhow) / rho;
            }
        }
#undef YI
#undef RP
}

/* bcfree_sub_subroutine (kernels_cpu.F90:6060-6155), xmin and ymax only:
 * the boundary cell's pressure with the total conditions ptot0, ttot0
 * gives the Mach number, T and the velocity along the angle of attack;
 * every ghost of the line takes that state. */
static void bcfree_sub(or_state *o, int ilat)
{
    const st_params *p = &o->p;
    if (ilat != 1 && ilat != 4) return;
    const double cosangle = cos(p->aoa), sinangle = sin(p->aoa);
    int na, nb;
    face_extent(o, ilat, &na, &nb);
    for (int b = 1; b <= nb; b++)
        for (int a = 1; a <= na; a++) {
            const size_t c = boundary_at(o, ilat, a, b);
            double rho = o->w[c], rhou = o->w[o->nvol + c], rhov = o->w[2 * o->nvol + c];
            double rhow = o->w[3 * o->nvol + c], rhoe = o->w[4 * o->nvol + c];
            double ri = 1.0 / rho;
            double uu = rhou * ri, vv = rhov * ri, ww = rhow * ri;
            double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
            double ee = rhoe / rho - qq;
            double tt = or_get_temperature_from_e(o, ee, o->w_aux[5 * o->nvol + c]);
            double pp = rho * tt * p->rgas0;
            double gamloc = or_get_gamloc(o, tt);
            double del = 0.5 * (gamloc - 1.0);
            double rmfac = pow(pp / p->ptot0, -(gamloc - 1.0) / gamloc);
            double rml = sqrt((rmfac - 1.0) / del);
            tt = p->ttot0 / rmfac;
            double vel_mod = rml * sqrt(gamloc * p->rgas0 * tt);
            uu = vel_mod * cosangle;
            vv = vel_mod * sinangle;
            ww = 0.0;
            rho = pp / tt / p->rgas0;
            rhou = rho * uu; rhov = rho * vv; rhow = rho * ww;
            ee = or_get_e_from_temperature(o, tt);
            for (int l = 1; l <= o->ng; l++) {
                const size_t d = ghost_at(o, ilat, a, b, l);
                o->w[d] = rho;
                o->w[o->nvol + d] = rhou;
                o->w[2 * o->nvol + d] = rhov;
                o->w[3 * o->nvol + d] = rhow;
                o->w[4 * o->nvol + d] = rho * ee + 0.5 * (rhou * rhou + rhov * rhov + rhow * rhow) / rho;
            }
        }
}

/* bcextr_sub_subroutine (kernels_cpu.F90:5064-5164), every face but ymin:
 * density and momentum of the boundary cell, temperature p0/(rho R). */
static void bcextr_sub(or_state *o, int ilat)
{
    const st_params *p = &o->p;
    if (ilat == 3) return;
    const double p0 = p->rho0 * p->rgas0 * p->t0;
    int na, nb;
    face_extent(o, ilat, &na, &nb);
    for (int b = 1; b <= nb; b++)
        for (int a = 1; a <= na; a++)
            for (int l = 1; l <= o->ng; l++) {
                const size_t c = boundary_at(o, ilat, a, b), d = ghost_at(o, ilat, a, b, l);
                double rho = o->w[c], rhou = o->w[o->nvol + c], rhov = o->w[2 * o->nvol + c];
                double rhow = o->w[3 * o->nvol + c];
                double tt = p0 / rho / p->rgas0;
                double ee = or_get_e_from_temperature(o, tt);
                o->w[d] = rho;
                o->w[o->nvol + d] = rhou;
                o->w[2 * o->nvol + d] = rhov;
                o->w[3 * o->nvol + d] = rhow;
                o->w[4 * o->nvol + d] = rho * ee + 0.5 * (rhou * rhou + rhov * rhov + rhow * rhow) / rho;
            }
}

/* bclam_subroutine (kernels_cpu.F90:5951-5985), xmin only: ghost plane 1-l
 * takes wmean(1-l, j, 1:4) with T = p0/(rho R). */
static void bclam(or_state *o, int ilat)
{
    const st_params *p = &o->p;
    if (ilat != 1 || !o->rc_wmean) return;
    const int ng = o->ng, ny = o->ny;
    const double p0 = p->rho0 * p->rgas0 * p->t0;
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int l = 1; l <= ng; l++) {
                const double *wm = o->rc_wmean + (size_t)(ng - l) + (size_t)ng * (size_t)(j - 1);
                const size_t st = (size_t)ng * ny;
                double rho = wm[0], rhou = wm[st], rhov = wm[2 * st], rhow = wm[3 * st];
                double tt = p0 / rho / p->rgas0;
                W(1 - l, j, k, 1) = rho;
                W(1 - l, j, k, 2) = rhou;
                W(1 - l, j, k, 3) = rhov;
                W(1 - l, j, k, 4) = rhow;
                double ee = or_get_e_from_temperature(o, tt);
                W(1 - l, j, k, 5) = rho * ee + 0.5 * (rhou * rhou + rhov * rhov + rhow * rhow) / rho;
            }
}

/* update_ghost, singleideal_cpu.F90:1064-1138: the recycled planes are
 * exchanged first, then the physical boundary of each face in the order
 * xmin..zmax, then the periodic swap. */
void or_update_ghost(or_state *o, int do_swap)
{
    if (o->p.bctags[0] == 10 && o->recyc_set) recyc_exchange(o);
    for (int ilat = 1; ilat <= 6; ilat++) {
        if (o->p.bctags[ilat - 1] == 10 && ilat == 1 && o->recyc_set) bcrecyc(o);
        switch (o->p.bctags[ilat - 1]) {
        case 1: bcfree(o, ilat); break;
        case 2: bcextr(o, ilat); break;
        case 3: bcfree_sub(o, ilat); break;
        case 4: bcextr_sub(o, ilat); break;
        case 9: bclam(o, ilat); break;
        case 5: bcsym(o, ilat); break;
        case 6:
            if (o->p.is_y_staggered) bcwall_staggered(o, ilat);
            else bcwall_node(o, ilat);
            break;
        case 7: bcshock(o, ilat); break;
        default: break;
        }
    }
    if (do_swap) or_bcswap(o);
}

/* bc_nr_lat_x/y/z_kernel (kernels_cpu.F90:5169-5283, 5420-5540,
 * 5688-5807) dispatched by bc_nr (singleideal_cpu.F90:1001-1062):
 * characteristic treatment of the boundary cell of face ilat.  The
 * one-sided second-order derivative of w along the normal is projected on
 * the left eigenvectors at the boundary state; the wave amplitudes are
 * multiplied by the eigenvalues (type 1 keeps only waves leaving the
 * domain; types 2 and 3 replace entering waves by the difference to the
 * neighbour or to winf; type 6 reflects the acoustic wave and drops the
 * others) and the result is added to fl through the right eigenvectors and
 * the 1-D metric. */
static void bc_nr_face(or_state *o, int ilat, int nr_type)
{
    const st_params *p = &o->p;
    const int dir = (ilat - 1) / 2, start = (ilat - 1) % 2 == 0;
    const int sgn = start ? 1 : -1;
    const double c_one[3] = {-1.5, 2.0, -0.5};
    int na, nb;
    face_extent(o, ilat, &na, &nb);
    const int n = dir == 0 ? o->nx : (dir == 1 ? o->ny : o->nz);
    const int ib = start ? 1 : n;
    const size_t st = dir == 0 ? 1 : (dir == 1 ? (size_t)o->sx : (size_t)o->sx * o->sy);
    const double *metric = dir == 0 ? o->dcsidx : (dir == 1 ? o->detady : o->dzitdz);
    for (int b = 1; b <= nb; b++)
        for (int a = 1; a <= na; a++) {
            int i, j, k;
            if (dir == 0) { i = ib; j = a; k = b; }
            else if (dir == 1) { i = a; j = ib; k = b; }
            else { i = a; j = b; k = ib; }
            const size_t c = G(o, i, j, k);
            double dw_dn[5], dw_dn_outer[5], dwc_dn[5], dwc_dn_outer[5], ev[5];
            for (int m = 0; m < 5; m++) {
                const double *wm = o->w + (size_t)m * o->nvol;
                dw_dn[m] = 0.0;
                for (int l = 1; l <= 3; l++)
                    dw_dn[m] = dw_dn[m] + sgn * c_one[l - 1] * wm[(long)c + (long)sgn * (l - 1) * (long)st];
                double w_target = 0.0;
                if (nr_type == 2) w_target = wm[(long)c - (long)sgn * (long)st];
                else if (nr_type == 3) w_target = p->winf[m];
                dw_dn_outer[m] = sgn * (wm[c] - w_target);
            }
            const double rho = o->w_aux[c], uu = o->w_aux[o->nvol + c], vv = o->w_aux[2 * o->nvol + c];
            const double ww = o->w_aux[3 * o->nvol + c], h = o->w_aux[4 * o->nvol + c];
            const double tt = o->w_aux[5 * o->nvol + c];
            const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
            const double gamloc = or_get_gamloc(o, tt);
            const double cc = gamloc * tt * p->rgas0;
            const double cs = sqrt(cc);
            const double ci = 1.0 / cs;
            const double p_rho = tt * p->rgas0;
            const double p_e = rho * (gamloc - 1.0);
            const double etot = h - tt * p->rgas0;
            const double b3 = etot - rho * p_rho / p_e;
            const double b2 = p_e / (rho * cc);
            const double b1 = p_rho / cc - b2 * (etot - 2.0 * qq);
            const double roe[9] = {b1, b2, b3, cs, ci, h, uu, vv, ww};
            double el[5][5], er[5][5];
            or_eigenvectors(dir, roe, el, er);
            for (int m = 0; m < 5; m++) {
                dwc_dn[m] = 0.0;
                dwc_dn_outer[m] = 0.0;
                for (int mm = 0; mm < 5; mm++) {
                    dwc_dn[m] = dwc_dn[m] + el[mm][m] * dw_dn[mm];
                    dwc_dn_outer[m] = dwc_dn_outer[m] + el[mm][m] * dw_dn_outer[mm];
                }
            }
            const double un = dir == 0 ? uu : (dir == 1 ? vv : ww);
            ev[0] = un - cs;
            ev[1] = un;
            ev[2] = ev[1];
            ev[3] = ev[1];
            ev[4] = un + cs;
            if (nr_type == 1)
                for (int l = 0; l < 5; l++) ev[l] = sgn * dmin(sgn * ev[l], 0.0);
            if (nr_type == 2 || nr_type == 3)
                for (int m = 0; m < 5; m++)
                    if (sgn * ev[m] > 0.0) dwc_dn[m] = dwc_dn_outer[m];
            for (int m = 0; m < 5; m++) dwc_dn[m] = ev[m] * dwc_dn[m];
            if (nr_type == 6) {
                dwc_dn[1] = 0.0;
                dwc_dn[2] = 0.0;
                dwc_dn[3] = 0.0;
                if (start) dwc_dn[4] = dwc_dn[0];
                else dwc_dn[0] = dwc_dn[4];
            }
            const int ic = dir == 0 ? i : (dir == 1 ? j : k);
            for (int m = 0; m < 5; m++) {
                double df = 0.0;
                for (int mm = 0; mm < 5; mm++) df = df + er[mm][m] * dwc_dn[mm];
                FL(i, j, k, m + 1) = FL(i, j, k, m + 1) + df * metric[ic - 1];
            }
        }
}

void or_bc_nr(or_state *o)
{
    for (int ilat = 1; ilat <= 6; ilat++) {
        if (ilat >= 5 && o->nz == 1) break;  /* 2*ndim faces */
        if (o->nr_eff[ilat - 1] > 0) bc_nr_face(o, ilat, o->nr_eff[ilat - 1]);
    }
}

/* ---------------------------------------------------------- RK arithmetic */
/* zero_flux_subroutine, kernels_cpu.F90:9-25 */
void or_zero_flux(or_state *o) { memset(o->fl, 0, sizeof(double) * 5 * o->nint); }

/* init_flux_subroutine, kernels_cpu.F90:29-47 */
void or_init_flux(or_state *o, double rhodt)
{
#pragma omp parallel for schedule(static)
    for (size_t n = 0; n < 5 * o->nint; n++) {
        o->fln[n] = -rhodt * o->fl[n];
        o->fl[n] = 0.0;
    }
}

/* update_flux_subroutine, kernels_cpu.F90:3400-3418 */
void or_update_flux(or_state *o, double gamdt)
{
#pragma omp parallel for schedule(static)
    for (size_t n = 0; n < 5 * o->nint; n++) o->fln[n] = o->fln[n] - gamdt * o->fl[n];
}

/* update_field_subroutine, kernels_cpu.F90:3422-3442 (fluid_mask == 0) */
void or_update_field(or_state *o)
{
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= o->nz; k++)
        for (int m = 1; m <= 5; m++)
            for (int j = 1; j <= o->ny; j++) for (int i = 1; i <= o->nx; i++)
                W(i, j, k, m) = W(i, j, k, m) + FLN(i, j, k, m);
}

/* -------------------------------------------------------- Roe + eigenvectors */
/* point of a stencil along direction dir */
static inline void stp(int dir, int i, int j, int k, int s, int *a, int *b, int *c)
{
    *a = i + (dir == 0 ? s : 0);
    *b = j + (dir == 1 ? s : 0);
    *c = k + (dir == 2 ? s : 0);
}

/* compute_roe_average, kernels_cpu.F90:3013-3089, between (i,j,k) and the
 * next point along dir.  out = b1,b2,b3,c,ci,h,uu,vv,ww */
void or_compute_roe_average(const or_state *o, int dir, int i, int j, int k, double out[9])
{
    const st_params *p = &o->p;
    int ip, jp, kp;
    stp(dir, i, j, k, 1, &ip, &jp, &kp);
    double uu = WA(i, j, k, 2), vv = WA(i, j, k, 3), ww = WA(i, j, k, 4);
    double h = WA(i, j, k, 5);
    double up = WA(ip, jp, kp, 2), vp = WA(ip, jp, kp, 3), wp = WA(ip, jp, kp, 4);
    double hp = WA(ip, jp, kp, 5);
    double r = WA(ip, jp, kp, 1) / WA(i, j, k, 1);
    r = sqrt(r);
    double rho = r * WA(i, j, k, 1);
    double rp1 = 1.0 / (r + 1.0);
    uu = (r * up + uu) * rp1;
    vv = (r * vp + vv) * rp1;
    ww = (r * wp + ww) * rp1;
    h = (r * hp + h) * rp1;
    double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
    double hbar = h - qq - CP(p->indx_cp_r + 1) * p->t0;
    double tt, gamloc;
    if (p->calorically_perfect == 1) {
        tt = p->t0 + hbar / CP(0);
        gamloc = CP(0) / (CP(0) - p->rgas0);
    } else {
        double t0 = p->t0;
        tt = WA(i, j, k, 6);
        double ttp = WA(ip, jp, kp, 6);
        tt = (r * ttp + tt) * rp1;
        double told = tt;
        for (int iter = 1; iter <= 50; iter++) {
            double num = 0.0, den = 0.0;
            for (int ll = p->indx_cp_l; ll <= p->indx_cp_r; ll++) {
                if (ll == -1) {
                    double tpow = ipow(told / t0, ll);
                    den = den + CP(ll) * tpow;
                    num = num + CP(ll) * log(told / t0);
                } else {
                    double tpow = ipow(told / t0, ll);
                    double tpowp = (told / t0) * tpow;
                    den = den + CP(ll) * tpow;
                    num = num + CP(ll) * (tpowp - 1.0) / (ll + 1.0);
                }
            }
            num = num * t0;
            tt = told + (hbar - num) / den;
            if (fabs(tt - told) < p->tol_iter_nr) break;
            told = tt;
        }
        double cploc = 0.0;
        for (int ll = p->indx_cp_l; ll <= p->indx_cp_r; ll++) cploc = cploc + CP(ll) * ipow(tt / t0, ll);
        gamloc = cploc / (cploc - p->rgas0);
    }
    double cc = gamloc * tt * p->rgas0;
    double gm1loc = gamloc - 1.0;
    double c = sqrt(cc);
    double ci = 1.0 / c;
    double p_rho = tt * p->rgas0;
    double p_e = rho * gm1loc;
    double etot = h - tt * p->rgas0;
    double b3 = etot - rho * p_rho / p_e;
    double b2 = p_e / (rho * cc);
    double b1 = p_rho / cc - b2 * (etot - 2.0 * qq);
    out[0] = b1; out[1] = b2; out[2] = b3; out[3] = c; out[4] = ci; out[5] = h;
    out[6] = uu; out[7] = vv; out[8] = ww;
}

/* eigenvectors_x/y/z, kernels_cpu.F90:2728-2782, 2842-2896, 2956-3010.
 * el[v][m] = el(v+1,m+1) (left eigenvector m, component v);
 * er[m][v] = er(m+1,v+1) (right eigenvector m, component v). */
void or_eigenvectors(int dir, const double roe[9], double el[5][5], double er[5][5])
{
    double b1 = roe[0], b2 = roe[1], b3 = roe[2], c = roe[3], ci = roe[4], h = roe[5];
    double uu = roe[6], vv = roe[7], ww = roe[8];
    if (dir == 0) {
        el[0][0] = 0.5 * (b1 + uu * ci); el[1][0] = -0.5 * (b2 * uu + ci); el[2][0] = -0.5 * (b2 * vv);
        el[3][0] = -0.5 * (b2 * ww); el[4][0] = 0.5 * b2;
        el[0][1] = 1.0 - b1; el[1][1] = b2 * uu; el[2][1] = b2 * vv; el[3][1] = b2 * ww; el[4][1] = -b2;
        el[0][2] = -vv; el[1][2] = 0.0; el[2][2] = 1.0; el[3][2] = 0.0; el[4][2] = 0.0;
        el[0][3] = -ww; el[1][3] = 0.0; el[2][3] = 0.0; el[3][3] = 1.0; el[4][3] = 0.0;
        el[0][4] = 0.5 * (b1 - uu * ci); el[1][4] = -0.5 * (b2 * uu - ci); el[2][4] = -0.5 * (b2 * vv);
        el[3][4] = -0.5 * (b2 * ww); el[4][4] = 0.5 * b2;
        er[0][0] = 1.0; er[1][0] = 1.0; er[2][0] = 0.0; er[3][0] = 0.0; er[4][0] = 1.0;
        er[0][1] = uu - c; er[1][1] = uu; er[2][1] = 0.0; er[3][1] = 0.0; er[4][1] = uu + c;
        er[0][2] = vv; er[1][2] = vv; er[2][2] = 1.0; er[3][2] = 0.0; er[4][2] = vv;
        er[0][3] = ww; er[1][3] = ww; er[2][3] = 0.0; er[3][3] = 1.0; er[4][3] = ww;
        er[0][4] = h - uu * c; er[1][4] = b3; er[2][4] = vv; er[3][4] = ww; er[4][4] = h + uu * c;
    } else if (dir == 1) {
        el[0][0] = 0.5 * (b1 + vv * ci); el[1][0] = -0.5 * (b2 * uu); el[2][0] = -0.5 * (b2 * vv + ci);
        el[3][0] = -0.5 * (b2 * ww); el[4][0] = 0.5 * b2;
        el[0][1] = 1.0 - b1; el[1][1] = b2 * uu; el[2][1] = b2 * vv; el[3][1] = b2 * ww; el[4][1] = -b2;
        el[0][2] = -uu; el[1][2] = 1.0; el[2][2] = 0.0; el[3][2] = 0.0; el[4][2] = 0.0;
        el[0][3] = -ww; el[1][3] = 0.0; el[2][3] = 0.0; el[3][3] = 1.0; el[4][3] = 0.0;
        el[0][4] = 0.5 * (b1 - vv * ci); el[1][4] = -0.5 * (b2 * uu); el[2][4] = -0.5 * (b2 * vv - ci);
        el[3][4] = -0.5 * (b2 * ww); el[4][4] = 0.5 * b2;
        er[0][0] = 1.0; er[1][0] = 1.0; er[2][0] = 0.0; er[3][0] = 0.0; er[4][0] = 1.0;
        er[0][1] = uu; er[1][1] = uu; er[2][1] = 1.0; er[3][1] = 0.0; er[4][1] = uu;
        er[0][2] = vv - c; er[1][2] = vv; er[2][2] = 0.0; er[3][2] = 0.0; er[4][2] = vv + c;
        er[0][3] = ww; er[1][3] = ww; er[2][3] = 0.0; er[3][3] = 1.0; er[4][3] = ww;
        er[0][4] = h - vv * c; er[1][4] = b3; er[2][4] = uu; er[3][4] = ww; er[4][4] = h + vv * c;
    } else {
        el[0][0] = 0.5 * (b1 + ww * ci); el[1][0] = -0.5 * (b2 * uu); el[2][0] = -0.5 * (b2 * vv);
        el[3][0] = -0.5 * (b2 * ww + ci); el[4][0] = 0.5 * b2;
        el[0][1] = 1.0 - b1; el[1][1] = b2 * uu; el[2][1] = b2 * vv; el[3][1] = b2 * ww; el[4][1] = -b2;
        el[0][2] = -uu; el[1][2] = 1.0; el[2][2] = 0.0; el[3][2] = 0.0; el[4][2] = 0.0;
        el[0][3] = -vv; el[1][3] = 0.0; el[2][3] = 1.0; el[3][3] = 0.0; el[4][3] = 0.0;
        el[0][4] = 0.5 * (b1 - ww * ci); el[1][4] = -0.5 * (b2 * uu); el[2][4] = -0.5 * (b2 * vv);
        el[3][4] = -0.5 * (b2 * ww - ci); el[4][4] = 0.5 * b2;
        er[0][0] = 1.0; er[1][0] = 1.0; er[2][0] = 0.0; er[3][0] = 0.0; er[4][0] = 1.0;
        er[0][1] = uu; er[1][1] = uu; er[2][1] = 1.0; er[3][1] = 0.0; er[4][1] = uu;
        er[0][2] = vv; er[1][2] = vv; er[2][2] = 0.0; er[3][2] = 1.0; er[4][2] = vv;
        er[0][3] = ww - c; er[1][3] = ww; er[2][3] = 0.0; er[3][3] = 0.0; er[4][3] = ww + c;
        er[0][4] = h - ww * c; er[1][4] = b3; er[2][4] = uu; er[3][4] = vv; er[4][4] = h + ww * c;
    }
}

/* ------------------------------------------------------------------- WENO */
static double sq(double a) { return a * a; }

/* wenorec_1d, kernels_cpu.F90:2526-2724.  gp/gm[m][l], l = 1..2*iweno. */
static void wenorec_scaled(const double vp[5][9], const double vm[5][9], double vhat[5], int iweno,
                           int wenorec_ord, int weno_version, const double betascale[5]);

/* wenorec_1d (kernels_cpu.F90:2526-2724): characteristic variables, scaled
 * by rho0^2 u0^2 (waves 1, 2, 5) and rho0^2 u0^4 (waves 3, 4). */
void or_wenorec_1d(const double vp[5][9], const double vm[5][9], double vhat[5], int iweno,
                   int wenorec_ord, int weno_version, double rho0, double u0)
{
    double u0_2 = u0 * u0;
    double rho0_2u0_2 = rho0 * rho0 * u0_2;
    double rho0_2u0_4 = rho0_2u0_2 * u0_2;
    double betascale[5];
    betascale[0] = 1.0 / rho0_2u0_2;
    betascale[1] = betascale[0];
    betascale[2] = 1.0 / rho0_2u0_4;
    betascale[3] = betascale[2];
    betascale[4] = betascale[0];
    wenorec_scaled(vp, vm, vhat, iweno, wenorec_ord, weno_version, betascale);
}

/* wenorec_1d_rusanov (kernels_cpu.F90:2325-2523): the same reconstruction
 * of the conserved components, scaled by rho0^2 u0^2 (mass), rho0^2 u0^4
 * (momenta) and rho0^2 u0^6 (energy). */
void or_wenorec_1d_rusanov(const double vp[5][9], const double vm[5][9], double vhat[5], int iweno,
                           int wenorec_ord, int weno_version, double rho0, double u0)
{
    double u0_2 = u0 * u0;
    double rho0_2u0_2 = rho0 * rho0 * u0_2;
    double rho0_2u0_4 = rho0_2u0_2 * u0_2;
    double betascale[5];
    betascale[0] = 1.0 / rho0_2u0_2;
    betascale[1] = 1.0 / rho0_2u0_4;
    betascale[2] = betascale[1];
    betascale[3] = betascale[1];
    betascale[4] = betascale[1] / u0_2;
    wenorec_scaled(vp, vm, vhat, iweno, wenorec_ord, weno_version, betascale);
}

static void wenorec_scaled(const double vp[5][9], const double vm[5][9], double vhat[5], int iweno,
                           int wenorec_ord, int weno_version, const double betascale[5])
{
    int i = iweno;
    if (wenorec_ord == 1) {
        for (int m = 0; m < 5; m++) vhat[m] = vp[m][i] + vm[m][i + 1];
    } else if (wenorec_ord == 2) {
        double dwe[2] = {1.0 / 3.0, 2.0 / 3.0};
        for (int m = 0; m < 5; m++) {
            double bp[2], bm[2];
            bp[0] = sq(vp[m][i] - vp[m][i - 1]);
            bp[1] = sq(vp[m][i + 1] - vp[m][i]);
            bp[1] = betascale[m] * bp[1];
            bp[0] = betascale[m] * bp[0];
            bm[0] = sq(vm[m][i + 2] - vm[m][i + 1]);
            bm[1] = sq(vm[m][i + 1] - vm[m][i]);
            bm[1] = betascale[m] * bm[1];
            bm[0] = betascale[m] * bm[0];
            double sump = 0.0, summ = 0.0;
            for (int l = 0; l < 2; l++) {
                bp[l] = dwe[l] / sq(0.000001 + bp[l]);
                bm[l] = dwe[l] / sq(0.000001 + bm[l]);
                sump = sump + bp[l];
                summ = summ + bm[l];
            }
            for (int l = 0; l < 2; l++) { bp[l] = bp[l] / sump; bm[l] = bm[l] / summ; }
            double vminus = bp[0] * (-vp[m][i - 1] + 3.0 * vp[m][i]) + bp[1] * (vp[m][i] + vp[m][i + 1]);
            double vplus = bm[0] * (-vm[m][i + 2] + 3.0 * vm[m][i + 1]) + bm[1] * (vm[m][i] + vm[m][i + 1]);
            vhat[m] = 0.5 * (vminus + vplus);
        }
    } else if (wenorec_ord == 3) {
        double dwe[3] = {1.0 / 10.0, 6.0 / 10.0, 3.0 / 10.0};
        double d0 = 13.0 / 12.0, d1 = 1.0 / 4.0;
        double c0 = 1.0 / 3.0, c1 = 5.0 / 6.0, c2 = -1.0 / 6.0, c3 = -7.0 / 6.0, c4 = 11.0 / 6.0;
        for (int m = 0; m < 5; m++) {
            double bp[3], bm[3];
            bp[2] = d0 * sq(vp[m][i] - 2.0 * vp[m][i + 1] + vp[m][i + 2]) + d1 * sq(3.0 * vp[m][i] - 4.0 * vp[m][i + 1] + vp[m][i + 2]);
            bp[1] = d0 * sq(vp[m][i - 1] - 2.0 * vp[m][i] + vp[m][i + 1]) + d1 * sq(vp[m][i - 1] - vp[m][i + 1]);
            bp[0] = d0 * sq(vp[m][i] - 2.0 * vp[m][i - 1] + vp[m][i - 2]) + d1 * sq(3.0 * vp[m][i] - 4.0 * vp[m][i - 1] + vp[m][i - 2]);
            bp[2] = betascale[m] * bp[2];
            bp[1] = betascale[m] * bp[1];
            bp[0] = betascale[m] * bp[0];
            bm[2] = d0 * sq(vm[m][i + 1] - 2.0 * vm[m][i] + vm[m][i - 1]) + d1 * sq(3.0 * vm[m][i + 1] - 4.0 * vm[m][i] + vm[m][i - 1]);
            bm[1] = d0 * sq(vm[m][i + 2] - 2.0 * vm[m][i + 1] + vm[m][i]) + d1 * sq(vm[m][i + 2] - vm[m][i]);
            bm[0] = d0 * sq(vm[m][i + 1] - 2.0 * vm[m][i + 2] + vm[m][i + 3]) + d1 * sq(3.0 * vm[m][i + 1] - 4.0 * vm[m][i + 2] + vm[m][i + 3]);
            bm[2] = betascale[m] * bm[2];
            bm[1] = betascale[m] * bm[1];
            bm[0] = betascale[m] * bm[0];
            if (weno_version == 0) {
                double sump = 0.0, summ = 0.0;
                for (int l = 0; l < 3; l++) {
                    bp[l] = dwe[l] / sq(0.000001 + bp[l]);
                    bm[l] = dwe[l] / sq(0.000001 + bm[l]);
                    sump = sump + bp[l];
                    summ = summ + bm[l];
                }
                for (int l = 0; l < 3; l++) { bp[l] = bp[l] / sump; bm[l] = bm[l] / summ; }
            } else {
                /* eps40 = 1.e-35 is a default-real (single) literal in the reference */
                double eps40 = (double)1.e-35f;
                double tau5p = fabs(bp[0] - bp[2]) + eps40;
                double tau5m = fabs(bm[0] - bm[2]) + eps40;
                for (int l = 0; l < 3; l++) {
                    bp[l] = (bp[l] + eps40) / (bp[l] + tau5p);
                    bm[l] = (bm[l] + eps40) / (bm[l] + tau5m);
                }
                double sump = 0.0, summ = 0.0;
                for (int l = 0; l < 3; l++) {
                    bp[l] = dwe[l] / bp[l];
                    bm[l] = dwe[l] / bm[l];
                    sump = sump + bp[l];
                    summ = summ + bm[l];
                }
                for (int l = 0; l < 3; l++) { bp[l] = bp[l] / sump; bm[l] = bm[l] / summ; }
            }
            double vminus = bp[2] * (c0 * vp[m][i] + c1 * vp[m][i + 1] + c2 * vp[m][i + 2]) +
                            bp[1] * (c2 * vp[m][i - 1] + c1 * vp[m][i] + c0 * vp[m][i + 1]) +
                            bp[0] * (c0 * vp[m][i - 2] + c3 * vp[m][i - 1] + c4 * vp[m][i]);
            double vplus = bm[2] * (c0 * vm[m][i + 1] + c1 * vm[m][i] + c2 * vm[m][i - 1]) +
                           bm[1] * (c2 * vm[m][i + 2] + c1 * vm[m][i + 1] + c0 * vm[m][i]) +
                           bm[0] * (c0 * vm[m][i + 3] + c3 * vm[m][i + 2] + c4 * vm[m][i + 1]);
            vhat[m] = vminus + vplus;
        }
    } else if (wenorec_ord == 4) {
        double dwe[4] = {1.0 / 35.0, 12.0 / 35.0, 18.0 / 35.0, 4.0 / 35.0};
        double d1 = 1.0 / 36.0, d2 = 13.0 / 12.0, d3 = 781.0 / 720.0;
        for (int m = 0; m < 5; m++) {
            const double *a = vp[m], *b = vm[m];
            double bp[4], bm[4];
            bp[3] = d1 * sq(-11 * a[i] + 18 * a[i + 1] - 9 * a[i + 2] + 2 * a[i + 3]) + d2 * sq(2 * a[i] - 5 * a[i + 1] + 4 * a[i + 2] - a[i + 3]) + d3 * sq(-a[i] + 3 * a[i + 1] - 3 * a[i + 2] + a[i + 3]);
            bp[2] = d1 * sq(-2 * a[i - 1] - 3 * a[i] + 6 * a[i + 1] - a[i + 2]) + d2 * sq(a[i - 1] - 2 * a[i] + a[i + 1]) + d3 * sq(-a[i - 1] + 3 * a[i] - 3 * a[i + 1] + a[i + 2]);
            bp[1] = d1 * sq(a[i - 2] - 6 * a[i - 1] + 3 * a[i] + 2 * a[i + 1]) + d2 * sq(a[i - 1] - 2 * a[i] + a[i + 1]) + d3 * sq(-a[i - 2] + 3 * a[i - 1] - 3 * a[i] + a[i + 1]);
            bp[0] = d1 * sq(-2 * a[i - 3] + 9 * a[i - 2] - 18 * a[i - 1] + 11 * a[i]) + d2 * sq(-a[i - 3] + 4 * a[i - 2] - 5 * a[i - 1] + 2 * a[i]) + d3 * sq(-a[i - 3] + 3 * a[i - 2] - 3 * a[i - 1] + a[i]);
            for (int l = 3; l >= 0; l--) bp[l] = betascale[m] * bp[l];
            bm[3] = d1 * sq(-11 * b[i + 1] + 18 * b[i] - 9 * b[i - 1] + 2 * b[i - 2]) + d2 * sq(2 * b[i + 1] - 5 * b[i] + 4 * b[i - 1] - b[i - 2]) + d3 * sq(-b[i + 1] + 3 * b[i] - 3 * b[i - 1] + b[i - 2]);
            bm[2] = d1 * sq(-2 * b[i + 2] - 3 * b[i + 1] + 6 * b[i] - b[i - 1]) + d2 * sq(b[i + 2] - 2 * b[i + 1] + b[i]) + d3 * sq(-b[i + 2] + 3 * b[i + 1] - 3 * b[i] + b[i - 1]);
            bm[1] = d1 * sq(b[i + 3] - 6 * b[i + 2] + 3 * b[i + 1] + 2 * b[i]) + d2 * sq(b[i + 2] - 2 * b[i + 1] + b[i]) + d3 * sq(-b[i + 3] + 3 * b[i + 2] - 3 * b[i + 1] + b[i]);
            bm[0] = d1 * sq(-2 * b[i + 4] + 9 * b[i + 3] - 18 * b[i + 2] + 11 * b[i + 1]) + d2 * sq(-b[i + 4] + 4 * b[i + 3] - 5 * b[i + 2] + 2 * b[i + 1]) + d3 * sq(-b[i + 4] + 3 * b[i + 3] - 3 * b[i + 2] + b[i + 1]);
            for (int l = 3; l >= 0; l--) bm[l] = betascale[m] * bm[l];
            double sump = 0.0, summ = 0.0;
            for (int l = 0; l < 4; l++) {
                bp[l] = dwe[l] / sq(0.000001 + bp[l]);
                bm[l] = dwe[l] / sq(0.000001 + bm[l]);
                sump = sump + bp[l];
                summ = summ + bm[l];
            }
            for (int l = 0; l < 4; l++) { bp[l] = bp[l] / sump; bm[l] = bm[l] / summ; }
            double vminus = bp[3] * (6 * a[i] + 26 * a[i + 1] - 10 * a[i + 2] + 2 * a[i + 3]) + bp[2] * (-2 * a[i - 1] + 14 * a[i] + 14 * a[i + 1] - 2 * a[i + 2]) + bp[1] * (2 * a[i - 2] - 10 * a[i - 1] + 26 * a[i] + 6 * a[i + 1]) + bp[0] * (-6 * a[i - 3] + 26 * a[i - 2] - 46 * a[i - 1] + 50 * a[i]);
            double vplus = bm[3] * (6 * b[i + 1] + 26 * b[i] - 10 * b[i - 1] + 2 * b[i - 2]) + bm[2] * (-2 * b[i + 2] + 14 * b[i + 1] + 14 * b[i] - 2 * b[i - 1]) + bm[1] * (2 * b[i + 3] - 10 * b[i + 2] + 26 * b[i + 1] + 6 * b[i]) + bm[0] * (-6 * b[i + 4] + 26 * b[i + 3] - 46 * b[i + 2] + 50 * b[i + 1]);
            vhat[m] = (vminus + vplus) / 24.0;
        }
    }
}

/* ------------------------------------------------------------ Euler fluxes */
/* euler_x_fluxes_hybrid_kernel, kernels_cpu.F90:425-689; the y and z
 * variants (:1808-2071, :1063-1325) differ only in the stencil direction,
 * the normal velocity and the momentum that carries the pressure.  Faces
 * fstart..fend along dir; j,k (or i,k / i,j) over the interior. */
void or_euler_fluxes(or_state *o, int dir, int fstart, int fend)
{
    const st_params *p = &o->p;
    int nx = o->nx, ny = o->ny, nz = o->nz;
    int lmax_base = p->ep_order / 2;
    int weno_scheme = p->weno_scheme, weno_size = 2 * weno_scheme;
    int nkeep = p->nkeep;
    double rgas0 = p->rgas0;
    int fzmin = o->force_zero_flux[2 * dir], fzmax = o->force_zero_flux[2 * dir + 1];
    int nlast = dir == 0 ? nx : (dir == 1 ? ny : nz);
    int k0 = 1, k1 = nz, j0 = 1, j1 = ny, i0 = 1, i1 = nx;
    if (dir == 0) { i0 = fstart; i1 = fend; }
    if (dir == 1) { j0 = fstart; j1 = fend; }
    if (dir == 2) { k0 = fstart; k1 = fend; }
#pragma omp parallel for schedule(static)
    for (int k = k0; k <= k1; k++)
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++) {
                int face = dir == 0 ? i : (dir == 1 ? j : k);
                int eoc = EOC(i, j, k, dir + 1);
                int ishk = 0;
                for (int s = -weno_scheme + 1; s <= weno_scheme; s++) {
                    int a, b, c;
                    stp(dir, i, j, k, s, &a, &b, &c);
                    if (WA(a, b, c, 8) > p->sensor_threshold) ishk = 1;
                }
                if (ishk == 0) {
                    double ft1 = 0, ft2 = 0, ft3 = 0, ft4 = 0, ft5 = 0, ft6 = 0;
                    int lmax = lmax_base + eoc;
                    if (lmax < 1) lmax = 1;
                    for (int l = 1; l <= lmax; l++) {
                        double uvs1 = 0, uvs2 = 0, uvs3 = 0, uvs4 = 0, uvs5 = 0, uvs6 = 0;
                        double uvs5_i = 0, uvs5_k = 0, uvs5_p = 0;
                        for (int m = 0; m <= l - 1; m++) {
                            int a, b, c, ap, bp, cp;
                            stp(dir, i, j, k, -m, &a, &b, &c);
                            stp(dir, i, j, k, -m + l, &ap, &bp, &cp);
                            double rhoi = WA(a, b, c, 1), uui = WA(a, b, c, 2), vvi = WA(a, b, c, 3);
                            double wwi = WA(a, b, c, 4), enti = WA(a, b, c, 5), tti = WA(a, b, c, 6);
                            double ppi = tti * rhoi * rgas0;
                            double rhoip = WA(ap, bp, cp, 1), uuip = WA(ap, bp, cp, 2), vvip = WA(ap, bp, cp, 3);
                            double wwip = WA(ap, bp, cp, 4), entip = WA(ap, bp, cp, 5), ttip = WA(ap, bp, cp, 6);
                            double ppip = ttip * rhoip * rgas0;
                            double uni = dir == 0 ? uui : (dir == 1 ? vvi : wwi);
                            double unip = dir == 0 ? uuip : (dir == 1 ? vvip : wwip);
                            double rhom = rhoi + rhoip;
                            if (nkeep >= 0) {
                                double eei = enti - ppi / rhoi - 0.5 * (uui * uui + vvi * vvi + wwi * wwi);
                                double eeip = entip - ppip / rhoip - 0.5 * (uuip * uuip + vvip * vvip + wwip * wwip);
                                double eem = eei + eeip;
                                double drhof, deef;
                                if (nkeep == 0) {
                                    drhof = 1.0;
                                    deef = 1.0;
                                } else {
                                    double sumnumrho = 1.0;
                                    double drho = 2.0 * (rhoip - rhoi) / rhom;
                                    double dee = 2.0 * (eeip - eei) / eem;
                                    double t_sumdenrho = (0.5 * drho) * (0.5 * drho);
                                    double t_sumdenee = (0.5 * dee) * (0.5 * dee);
                                    double t2_sumdenrho = t_sumdenrho, t2_sumdenee = t_sumdenee;
                                    double sumdenrho = 1.0 + t_sumdenrho / 3.0;
                                    double sumdenee = 1.0 + t_sumdenee;
                                    double sumnumee = 1.0 + t_sumdenee / 3.0;
                                    for (int n = 2; n <= nkeep; n++) {
                                        int n2 = 2 * n;
                                        t_sumdenrho = t2_sumdenrho * t_sumdenrho;
                                        t_sumdenee = t2_sumdenee * t_sumdenee;
                                        sumdenrho = sumdenrho + t_sumdenrho / (1.0 + n2);
                                        sumdenee = sumdenee + t_sumdenee;
                                        sumnumee = sumnumee + t_sumdenee / (1.0 + n2);
                                    }
                                    drhof = sumnumrho / sumdenrho;
                                    deef = sumnumee / sumdenee;
                                }
                                double uv_part = (uni + unip) * rhom * drhof;
                                uvs1 = uvs1 + uv_part * 2.0;
                                uvs2 = uvs2 + uv_part * (uui + uuip);
                                uvs3 = uvs3 + uv_part * (vvi + vvip);
                                uvs4 = uvs4 + uv_part * (wwi + wwip);
                                uvs5_i = uvs5_i + uv_part * eem * deef;
                                uvs5_k = uvs5_k + uv_part * (uui * uuip + vvi * vvip + wwi * wwip);
                                uvs5_p = uvs5_p + 4.0 * (uni * ppip + unip * ppi);
                                uvs6 = uvs6 + 2.0 * (ppi + ppip);
                            } else {
                                double uv_part = (uni + unip) * rhom;
                                uvs1 = uvs1 + uv_part * 2.0;
                                uvs2 = uvs2 + uv_part * (uui + uuip);
                                uvs3 = uvs3 + uv_part * (vvi + vvip);
                                uvs4 = uvs4 + uv_part * (wwi + wwip);
                                uvs5 = uvs5 + uv_part * (enti + entip);
                                uvs6 = uvs6 + 2.0 * (ppi + ppip);
                            }
                        }
                        double cl = o->cd1[lmax - 1][l - 1];
                        ft1 = ft1 + cl * uvs1;
                        ft2 = ft2 + cl * uvs2;
                        ft3 = ft3 + cl * uvs3;
                        ft4 = ft4 + cl * uvs4;
                        if (nkeep >= 0) ft5 = ft5 + cl * (uvs5_i + uvs5_k + uvs5_p);
                        else ft5 = ft5 + cl * uvs5;
                        ft6 = ft6 + cl * uvs6;
                    }
                    double fh[5];
                    fh[0] = 0.25 * ft1; fh[1] = 0.25 * ft2; fh[2] = 0.25 * ft3; fh[3] = 0.25 * ft4; fh[4] = 0.25 * ft5;
                    if ((face == 0 && fzmin == 1) || (face == nlast && fzmax == 1)) {
                        for (int m = 0; m < 5; m++) fh[m] = 0.0;
                    }
                    fh[1 + dir] = fh[1 + dir] + 0.5 * ft6;
                    for (int m = 0; m < 5; m++) FH(i, j, k, m + 1) = fh[m];
                } else if (p->flux_splitting == 1) {
                    /* Rusanov split of the WENO faces, the *_hybrid_rusanov_kernel
                     * branches (kernels_cpu.F90:887-934, 2268-2315, 1521-1568):
                     * one eigenvalue bound over the stencil, conserved components */
                    double evmax = -1.0;
                    for (int l = 1; l <= weno_size; l++) {
                        int a, b, c;
                        stp(dir, i, j, k, l - weno_scheme, &a, &b, &c);
                        double un = WA(a, b, c, 2 + dir);
                        double tt = WA(a, b, c, 6);
                        double gamloc = or_get_gamloc(o, tt);
                        double cs = sqrt(gamloc * rgas0 * tt);
                        double evm = dmax(fabs(un - cs), fabs(un + cs));
                        evmax = dmax(evm, evmax);
                    }
                    double gp[5][9], gm[5][9];
                    for (int l = 1; l <= weno_size; l++) {
                        int a, b, c;
                        stp(dir, i, j, k, l - weno_scheme, &a, &b, &c);
                        double rho = WA(a, b, c, 1), uu = WA(a, b, c, 2), vv = WA(a, b, c, 3);
                        double ww = WA(a, b, c, 4), h = WA(a, b, c, 5);
                        double un = dir == 0 ? uu : (dir == 1 ? vv : ww);
                        double rhoun = rho * un;
                        double pp = rho * WA(a, b, c, 6) * rgas0;
                        double rhoevm = rho * evmax;
                        double evm, cc;
                        evm = rhoun;
                        cc = 0.5 * (evm + rhoevm);
                        gp[0][l] = cc; gm[0][l] = evm - cc;
                        evm = uu * rhoun + (dir == 0 ? pp : 0.0);
                        cc = 0.5 * (evm + rhoevm * uu);
                        gp[1][l] = cc; gm[1][l] = evm - cc;
                        evm = vv * rhoun + (dir == 1 ? pp : 0.0);
                        cc = 0.5 * (evm + rhoevm * vv);
                        gp[2][l] = cc; gm[2][l] = evm - cc;
                        evm = ww * rhoun + (dir == 2 ? pp : 0.0);
                        cc = 0.5 * (evm + rhoevm * ww);
                        gp[3][l] = cc; gm[3][l] = evm - cc;
                        evm = h * rhoun;
                        cc = 0.5 * (evm + evmax * (rho * h - pp));
                        gp[4][l] = cc; gm[4][l] = evm - cc;
                    }
                    int wenorec_ord = weno_scheme + eoc;
                    if (wenorec_ord < 1) wenorec_ord = 1;
                    double fi[5];
                    or_wenorec_1d_rusanov(gp, gm, fi, weno_scheme, wenorec_ord, p->weno_version, p->rho0, p->u0);
                    for (int m = 0; m < 5; m++) FH(i, j, k, m + 1) = fi[m];
                } else {
                    double roe[9], el[5][5], er[5][5];
                    or_compute_roe_average(o, dir, i, j, k, roe);
                    or_eigenvectors(dir, roe, el, er);
                    double evmax[5] = {-1.0, -1.0, -1.0, -1.0, -1.0};
                    for (int l = 1; l <= weno_size; l++) {
                        int a, b, c;
                        stp(dir, i, j, k, l - weno_scheme, &a, &b, &c);
                        double un = WA(a, b, c, 2 + dir);
                        double tt = WA(a, b, c, 6);
                        double gamloc = or_get_gamloc(o, tt);
                        double cs = sqrt(gamloc * rgas0 * tt);
                        evmax[0] = dmax(fabs(un - cs), evmax[0]);
                        evmax[1] = dmax(fabs(un), evmax[1]);
                        evmax[2] = evmax[1];
                        evmax[3] = evmax[1];
                        evmax[4] = dmax(fabs(un + cs), evmax[4]);
                    }
                    double gp[5][9], gm[5][9];
                    for (int l = 1; l <= weno_size; l++) {
                        int a, b, c;
                        stp(dir, i, j, k, l - weno_scheme, &a, &b, &c);
                        double rho = WA(a, b, c, 1), uu = WA(a, b, c, 2), vv = WA(a, b, c, 3);
                        double ww = WA(a, b, c, 4), h = WA(a, b, c, 5);
                        double un = dir == 0 ? uu : (dir == 1 ? vv : ww);
                        double rhoun = rho * un;
                        double pp = rho * WA(a, b, c, 6) * rgas0;
                        double fi[5];
                        fi[0] = rhoun;
                        fi[1] = uu * rhoun;
                        fi[2] = vv * rhoun;
                        fi[3] = ww * rhoun;
                        fi[4] = h * rhoun;
                        fi[1 + dir] = fi[1 + dir] + pp;
                        for (int m = 0; m < 5; m++) {
                            double wc = 0.0, gc = 0.0;
                            wc = wc + el[0][m] * rho;
                            gc = gc + el[0][m] * fi[0];
                            wc = wc + el[1][m] * rho * uu;
                            gc = gc + el[1][m] * fi[1];
                            wc = wc + el[2][m] * rho * vv;
                            gc = gc + el[2][m] * fi[2];
                            wc = wc + el[3][m] * rho * ww;
                            gc = gc + el[3][m] * fi[3];
                            wc = wc + el[4][m] * (rho * h - pp);
                            gc = gc + el[4][m] * fi[4];
                            double cc = 0.5 * (gc + evmax[m] * wc);
                            gp[m][l] = cc;
                            gm[m][l] = gc - cc;
                        }
                    }
                    int wenorec_ord = weno_scheme + eoc;
                    if (wenorec_ord < 1) wenorec_ord = 1;
                    double fi[5];
                    or_wenorec_1d(gp, gm, fi, weno_scheme, wenorec_ord, p->weno_version, p->rho0, p->u0);
                    for (int m = 0; m < 5; m++) {
                        double f = 0.0;
                        for (int mm = 0; mm < 5; mm++) f = f + er[mm][m] * fi[mm];
                        FH(i, j, k, m + 1) = f;
                    }
                }
            }
}

/* euler_x/y/z_update_subroutine, kernels_cpu.F90:944-963, 991-1010, 1039-1058 */
void or_euler_update(or_state *o, int dir, int cstart, int cend)
{
    int i0 = 1, i1 = o->nx, j0 = 1, j1 = o->ny, k0 = 1, k1 = o->nz;
    if (dir == 0) { i0 = cstart; i1 = cend; }
    if (dir == 1) { j0 = cstart; j1 = cend; }
    if (dir == 2) { k0 = cstart; k1 = cend; }
#pragma omp parallel for schedule(static)
    for (int k = k0; k <= k1; k++)
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++)
                for (int iv = 1; iv <= 5; iv++) {
                    double d;
                    int a, b, c;
                    stp(dir, i, j, k, -1, &a, &b, &c);
                    if (dir == 0) d = o->dcsidx[i - 1];
                    else if (dir == 1) d = o->detady[j - 1];
                    else d = o->dzitdz[k - 1];
                    FL(i, j, k, iv) = FL(i, j, k, iv) + (FH(i, j, k, iv) - FH(a, b, c, iv)) * d;
                }
}

/* euler_x / euler_y / euler_z, singleideal_cpu.F90:674-843 */
void or_euler_x(or_state *o)
{
    if (o->eul_imax - o->eul_imin >= 0) {
        or_euler_fluxes(o, 0, o->eul_imin - 1, o->eul_imax);
        or_euler_update(o, 0, o->eul_imin, o->eul_imax);
    }
}
void or_euler_y(or_state *o)
{
    or_euler_fluxes(o, 1, o->eul_jmin - 1, o->eul_jmax);
    or_euler_update(o, 1, o->eul_jmin, o->eul_jmax);
}
void or_euler_z(or_state *o)
{
    or_euler_fluxes(o, 2, o->eul_kmin - 1, o->eul_kmax);
    or_euler_update(o, 2, o->eul_kmin, o->eul_kmax);
    if (o->p.bctags[0] == 10) draw_inflow_plane(o);
}

/* ----------------------------------------------------------------- viscous */
/* visflx_subroutine (mode 0, kernels_cpu.F90:3598-3777) and
 * visflx_nosensor_subroutine (mode 4, :3996-4180); with_sensor selects the
 * Ducros sensor store of mode 0 (:3755). */
static void visflx_lapl(or_state *o, int with_sensor)
{
    const st_params *p = &o->p;
    int lmax = p->visc_order / 2;
    const double *c1 = o->cd1[lmax - 1], *c2 = o->cd2[lmax - 1];
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) {
                double uu = WA(i, j, k, 2), vv = WA(i, j, k, 3), ww = WA(i, j, k, 4);
                double tt = WA(i, j, k, 6), mu = WA(i, j, k, 7);
                double ux = 0, vx = 0, wx = 0, tx = 0, mux = 0;
                double uy = 0, vy = 0, wy = 0, ty = 0, muy = 0;
                double uz = 0, vz = 0, wz = 0, tz = 0, muz = 0;
                double ulapx = c2[0] * uu, ulapy = ulapx, ulapz = ulapx;
                double vlapx = c2[0] * vv, vlapy = vlapx, vlapz = vlapx;
                double wlapx = c2[0] * ww, wlapy = wlapx, wlapz = wlapx;
                double tlapx = c2[0] * tt, tlapy = tlapx, tlapz = tlapx;
                for (int l = 1; l <= lmax; l++) {
                    double ccl = c1[l - 1], clapl = c2[l];
                    ux = ux + ccl * (WA(i + l, j, k, 2) - WA(i - l, j, k, 2));
                    ulapx = ulapx + clapl * (WA(i + l, j, k, 2) + WA(i - l, j, k, 2));
                    vx = vx + ccl * (WA(i + l, j, k, 3) - WA(i - l, j, k, 3));
                    vlapx = vlapx + clapl * (WA(i + l, j, k, 3) + WA(i - l, j, k, 3));
                    wx = wx + ccl * (WA(i + l, j, k, 4) - WA(i - l, j, k, 4));
                    wlapx = wlapx + clapl * (WA(i + l, j, k, 4) + WA(i - l, j, k, 4));
                    tx = tx + ccl * (WA(i + l, j, k, 6) - WA(i - l, j, k, 6));
                    tlapx = tlapx + clapl * (WA(i + l, j, k, 6) + WA(i - l, j, k, 6));
                    mux = mux + ccl * (WA(i + l, j, k, 7) - WA(i - l, j, k, 7));
                }
                for (int l = 1; l <= lmax; l++) {
                    double ccl = c1[l - 1], clapl = c2[l];
                    uy = uy + ccl * (WA(i, j + l, k, 2) - WA(i, j - l, k, 2));
                    ulapy = ulapy + clapl * (WA(i, j + l, k, 2) + WA(i, j - l, k, 2));
                    vy = vy + ccl * (WA(i, j + l, k, 3) - WA(i, j - l, k, 3));
                    vlapy = vlapy + clapl * (WA(i, j + l, k, 3) + WA(i, j - l, k, 3));
                    wy = wy + ccl * (WA(i, j + l, k, 4) - WA(i, j - l, k, 4));
                    wlapy = wlapy + clapl * (WA(i, j + l, k, 4) + WA(i, j - l, k, 4));
                    ty = ty + ccl * (WA(i, j + l, k, 6) - WA(i, j - l, k, 6));
                    tlapy = tlapy + clapl * (WA(i, j + l, k, 6) + WA(i, j - l, k, 6));
                    muy = muy + ccl * (WA(i, j + l, k, 7) - WA(i, j - l, k, 7));
                }
                for (int l = 1; l <= lmax; l++) {
                    double ccl = c1[l - 1], clapl = c2[l];
                    uz = uz + ccl * (WA(i, j, k + l, 2) - WA(i, j, k - l, 2));
                    ulapz = ulapz + clapl * (WA(i, j, k + l, 2) + WA(i, j, k - l, 2));
                    vz = vz + ccl * (WA(i, j, k + l, 3) - WA(i, j, k - l, 3));
                    vlapz = vlapz + clapl * (WA(i, j, k + l, 3) + WA(i, j, k - l, 3));
                    wz = wz + ccl * (WA(i, j, k + l, 4) - WA(i, j, k - l, 4));
                    wlapz = wlapz + clapl * (WA(i, j, k + l, 4) + WA(i, j, k - l, 4));
                    tz = tz + ccl * (WA(i, j, k + l, 6) - WA(i, j, k - l, 6));
                    tlapz = tlapz + clapl * (WA(i, j, k + l, 6) + WA(i, j, k - l, 6));
                    muz = muz + ccl * (WA(i, j, k + l, 7) - WA(i, j, k - l, 7));
                }
                double dx = o->dcsidx[i - 1], dy = o->detady[j - 1], dz = o->dzitdz[k - 1];
                ux = ux * dx; vx = vx * dx; wx = wx * dx; tx = tx * dx; mux = mux * dx;
                uy = uy * dy; vy = vy * dy; wy = wy * dy; ty = ty * dy; muy = muy * dy;
                uz = uz * dz; vz = vz * dz; wz = wz * dz; tz = tz * dz; muz = muz * dz;
                double dxs = o->dcsidxs[i - 1], dys = o->detadys[j - 1], dzs = o->dzitdzs[k - 1];
                double dx2 = o->dcsidx2[i - 1], dy2 = o->detady2[j - 1], dz2 = o->dzitdz2[k - 1];
                ulapx = ulapx * dxs + ux * dx2; vlapx = vlapx * dxs + vx * dx2;
                wlapx = wlapx * dxs + wx * dx2; tlapx = tlapx * dxs + tx * dx2;
                ulapy = ulapy * dys + uy * dy2; vlapy = vlapy * dys + vy * dy2;
                wlapy = wlapy * dys + wy * dy2; tlapy = tlapy * dys + ty * dy2;
                ulapz = ulapz * dzs + uz * dz2; vlapz = vlapz * dzs + vz * dz2;
                wlapz = wlapz * dzs + wz * dz2; tlapz = tlapz * dzs + tz * dz2;
                double ulap = ulapx + ulapy + ulapz;
                double vlap = vlapx + vlapy + vlapz;
                double wlap = wlapx + wlapy + wlapz;
                double tlap = tlapx + tlapy + tlapz;
                double div = ux + vy + wz;
                double div3l = div / 3.0;
                WA(i, j, k, 10) = div3l;
                double omegax = wy - vz, omegay = uz - wx, omegaz = vx - uy;
                double omod2 = omegax * omegax + omegay * omegay + omegaz * omegaz;
                WA(i, j, k, 9) = sqrt(omod2);
                if (with_sensor) {
                    double s = -div / sqrt(omod2 + div * div + sq(p->u0 / p->l0));
                    s = dmax(s, 0.0);
                    WA(i, j, k, 8) = s * s;
                }
                double cploc;
                if (p->calorically_perfect == 1) {
                    cploc = CP(0);
                } else {
                    cploc = 0.0;
                    for (int ll = p->indx_cp_l; ll <= p->indx_cp_r; ll++) cploc = cploc + CP(ll) * ipow(tt / p->t0, ll);
                }
                double sig11 = 2.0 * (ux - div3l), sig12 = uy + vx, sig13 = uz + wx;
                double sig22 = 2.0 * (vy - div3l), sig23 = vz + wy, sig33 = 2.0 * (wz - div3l);
                double sigx = mux * sig11 + muy * sig12 + muz * sig13 + mu * ulap;
                double sigy = mux * sig12 + muy * sig22 + muz * sig23 + mu * vlap;
                double sigz = mux * sig13 + muy * sig23 + muz * sig33 + mu * wlap;
                double sigqt = (mux * tx + muy * ty + muz * tz + mu * tlap) * cploc / p->prandtl;
                double sigah = (sig11 * ux + sig12 * uy + sig13 * uz + sig12 * vx + sig22 * vy + sig23 * vz +
                                sig13 * wx + sig23 * wy + sig33 * wz) * mu;
                double sigq = sigx * uu + sigy * vv + sigz * ww + sigah + sigqt;
                FL(i, j, k, 2) = FL(i, j, k, 2) - sigx;
                FL(i, j, k, 3) = FL(i, j, k, 3) - sigy;
                FL(i, j, k, 4) = FL(i, j, k, 4) - sigz;
                FL(i, j, k, 5) = FL(i, j, k, 5) - sigq;
                if (j == 1) {
                    size_t wp = (size_t)(i + o->ng - 1) + (size_t)o->sx * (size_t)(k + o->ng - 1);
                    size_t ws = (size_t)o->sx * o->sz;
                    o->wallprop[0 * ws + wp] = mu * uy;
                    o->wallprop[1 * ws + wp] = mu * wy;
                    o->wallprop[2 * ws + wp] = mu * ty * cploc / p->prandtl;
                }
            }
}

/* visflx_div_subroutine, kernels_cpu.F90:3484-3532 */
static void visflx_div(or_state *o)
{
    int lmax = o->p.visc_order / 2;
    const double *c1 = o->cd1[lmax - 1];
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) {
                double uu = WA(i, j, k, 2), vv = WA(i, j, k, 3), ww = WA(i, j, k, 4);
                double mu = WA(i, j, k, 7);
                double divx3l = 0, divy3l = 0, divz3l = 0;
                for (int l = 1; l <= lmax; l++) {
                    double ccl = c1[l - 1];
                    divx3l = divx3l + ccl * (WA(i + l, j, k, 10) - WA(i - l, j, k, 10));
                    divy3l = divy3l + ccl * (WA(i, j + l, k, 10) - WA(i, j - l, k, 10));
                    divz3l = divz3l + ccl * (WA(i, j, k + l, 10) - WA(i, j, k - l, 10));
                }
                divx3l = divx3l * o->dcsidx[i - 1];
                divy3l = divy3l * o->detady[j - 1];
                divz3l = divz3l * o->dzitdz[k - 1];
                double sigx = mu * divx3l, sigy = mu * divy3l, sigz = mu * divz3l;
                double sigq = sigx * uu + sigy * vv + sigz * ww;
                FL(i, j, k, 2) = FL(i, j, k, 2) - sigx;
                FL(i, j, k, 3) = FL(i, j, k, 3) - sigy;
                FL(i, j, k, 4) = FL(i, j, k, 4) - sigz;
                FL(i, j, k, 5) = FL(i, j, k, 5) - sigq;
            }
}

/* sensor_subroutine, kernels_cpu.F90:4726-4744 */
static void sensor(or_state *o)
{
    const st_params *p = &o->p;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) {
                double omod2 = sq(WA(i, j, k, 9));
                double div = 3.0 * WA(i, j, k, 10);
                double s = -div / sqrt(omod2 + div * div + sq(p->u0 / p->l0));
                s = dmax(s, 0.0);
                WA(i, j, k, 8) = s * s;
            }
}

/* ------------------------------------------- conservative viscous form */

static double cp_at(const or_state *o, double tt)
{
    const st_params *p = &o->p;
    if (p->calorically_perfect == 1) return CP(0);
    double cploc = 0.0;
    for (int ll = p->indx_cp_l; ll <= p->indx_cp_r; ll++) cploc = cploc + CP(ll) * ipow(tt / p->t0, ll);
    return cploc;
}

/* visflx_reduced_ord2_subroutine, kernels_cpu.F90:4623-4721 (visflx modes 2
 * and 6): second-order cell terms grad(mu).sigma and the dissipation, div/3
 * and |omega| into w_aux(10), w_aux(9), wall shear into wallprop and, with
 * update_sensor, the Ducros sensor. */
static void visflx_reduced_ord2(or_state *o, int update_sensor)
{
    const st_params *p = &o->p;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) {
                double uu = WA(i, j, k, 2), vv = WA(i, j, k, 3), ww = WA(i, j, k, 4);
                double tt = WA(i, j, k, 6), mu = WA(i, j, k, 7);
                double cploc = cp_at(o, tt);
                double dxl = 1.0 / (XC(i + 1) - XC(i - 1));
                double dyl = 1.0 / (YC(j + 1) - YC(j - 1));
                double dzl = 1.0 / (ZC(k + 1) - ZC(k - 1));
                double ux = dxl * (WA(i + 1, j, k, 2) - WA(i - 1, j, k, 2));
                double vx = dxl * (WA(i + 1, j, k, 3) - WA(i - 1, j, k, 3));
                double wx = dxl * (WA(i + 1, j, k, 4) - WA(i - 1, j, k, 4));
                double mux = dxl * (WA(i + 1, j, k, 7) - WA(i - 1, j, k, 7));
                double uy = dyl * (WA(i, j + 1, k, 2) - WA(i, j - 1, k, 2));
                double vy = dyl * (WA(i, j + 1, k, 3) - WA(i, j - 1, k, 3));
                double wy = dyl * (WA(i, j + 1, k, 4) - WA(i, j - 1, k, 4));
                double ty = dyl * (WA(i, j + 1, k, 6) - WA(i, j - 1, k, 6));
                double muy = dyl * (WA(i, j + 1, k, 7) - WA(i, j - 1, k, 7));
                double uz = dzl * (WA(i, j, k + 1, 2) - WA(i, j, k - 1, 2));
                double vz = dzl * (WA(i, j, k + 1, 3) - WA(i, j, k - 1, 3));
                double wz = dzl * (WA(i, j, k + 1, 4) - WA(i, j, k - 1, 4));
                double muz = dzl * (WA(i, j, k + 1, 7) - WA(i, j, k - 1, 7));
                if (j == 1) {
                    size_t wp = (size_t)(i + o->ng - 1) + (size_t)o->sx * (size_t)(k + o->ng - 1);
                    size_t ws = (size_t)o->sx * o->sz;
                    o->wallprop[0 * ws + wp] = mu * uy;
                    o->wallprop[1 * ws + wp] = mu * wy;
                    o->wallprop[2 * ws + wp] = mu * ty * cploc / p->prandtl;
                }
                double div = ux + vy + wz;
                double div3l = div / 3.0;
                WA(i, j, k, 10) = div3l;
                double omegax = wy - vz, omegay = uz - wx, omegaz = vx - uy;
                double omod2 = omegax * omegax + omegay * omegay + omegaz * omegaz;
                WA(i, j, k, 9) = sqrt(omod2);
                if (update_sensor == 1) {
                    double s = -div / sqrt(omod2 + div * div + sq(p->u0 / p->l0));
                    s = dmax(s, 0.0);
                    WA(i, j, k, 8) = s * s;
                }
                double sig11 = ux - 2.0 * div3l, sig12 = vx, sig13 = wx;
                double sig21 = uy, sig22 = vy - 2.0 * div3l, sig23 = wy;
                double sig31 = uz, sig32 = vz, sig33 = wz - 2.0 * div3l;
                double sigx = mux * sig11 + muy * sig12 + muz * sig13;
                double sigy = mux * sig21 + muy * sig22 + muz * sig23;
                double sigz = mux * sig31 + muy * sig32 + muz * sig33;
                double sigah = (sig11 * ux + sig12 * uy + sig13 * uz + sig21 * vx + sig22 * vy + sig23 * vz +
                                sig31 * wx + sig32 * wy + sig33 * wz) * mu;
                double sigq = sigx * uu + sigy * vv + sigz * ww + sigah;
                FL(i, j, k, 2) = FL(i, j, k, 2) - sigx;
                FL(i, j, k, 3) = FL(i, j, k, 3) - sigy;
                FL(i, j, k, 4) = FL(i, j, k, 4) - sigz;
                FL(i, j, k, 5) = FL(i, j, k, 5) - sigq;
            }
}

/* Face flux of the conservative Laplacian terms between points a (left) and
 * b (right) whose node distance is h: mu_face * d(u,v,w)/dn and
 * mu_face * (cp/Pr dT/dn + d(|u|^2/2)/dn); kernels_cpu.F90:4805-4836. */
static void cons_face(const or_state *o, size_t a, size_t b, double h, double out[4])
{
    const st_params *p = &o->p;
    const double *wa = o->w_aux;
    size_t nv = o->nvol;
    double uu = wa[1 * nv + a], uup = wa[1 * nv + b];
    double vv = wa[2 * nv + a], vvp = wa[2 * nv + b];
    double ww = wa[3 * nv + a], wwp = wa[3 * nv + b];
    double tt = wa[5 * nv + a], ttp = wa[5 * nv + b];
    double mu = wa[6 * nv + a], mup = wa[6 * nv + b];
    double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
    double qqp = 0.5 * (uup * uup + vvp * vvp + wwp * wwp);
    double cploc = p->calorically_perfect == 1 ? CP(0) : cp_at(o, 0.5 * (tt + ttp));
    double sigx = uup - uu, sigy = vvp - vv, sigz = wwp - ww;
    double sigq_tt = ttp - tt, sigq_qq = qqp - qq;
    double muf = mu + mup;
    muf = 0.5 * muf / h;
    out[0] = sigx * muf;
    out[1] = sigy * muf;
    out[2] = sigz * muf;
    out[3] = (sigq_tt * cploc / p->prandtl + sigq_qq) * muf;
}

/* visflx_x_subroutine, kernels_cpu.F90:4779-4852 (visflx mode 5, x part):
 * face values into fhat(i = 0..nx), then fl += dxhl*(fhat(i)-fhat(i-1)). */
static void visflx_cons_x(or_state *o)
{
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 0; i <= o->nx; i++) {
                double s[4];
                cons_face(o, G(o, i, j, k), G(o, i + 1, j, k), XC(i + 1) - XC(i), s);
                FH(i, j, k, 2) = -s[0];
                FH(i, j, k, 3) = -s[1];
                FH(i, j, k, 4) = -s[2];
                FH(i, j, k, 5) = -s[3];
            }
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) {
                double dxhl = 2.0 / (XC(i + 1) - XC(i - 1));
                for (int iv = 2; iv <= 5; iv++)
                    FL(i, j, k, iv) = FL(i, j, k, iv) + dxhl * (FH(i, j, k, iv) - FH(i - 1, j, k, iv));
            }
}

/* visflx_y_subroutine (kernels_cpu.F90:4856-4926) for dir 1 and
 * visflx_z_subroutine (:4929-4999) for dir 2: a march over the faces
 * l = 0..n carrying the previous face flux times the cell's dyhl, so that
 * fl(l) += (flux(l-1/2)*dyhl(l)) - flux(l+1/2)*dyhl(l). */
static void visflx_cons_march(or_state *o, int dir)
{
    int n = dir == 1 ? o->ny : o->nz;
    int na = dir == 1 ? o->nx : o->nx;
    int nb = dir == 1 ? o->nz : o->ny;
    const double *c = dir == 1 ? o->y : o->z;
#define CC(l) c[(l) + o->ng - 1]
#pragma omp parallel for schedule(static)
    for (int b = 1; b <= nb; b++)
        for (int i = 1; i <= na; i++) {
            double fo[4] = {0, 0, 0, 0}, dhl = 0.0;
            for (int l = 0; l <= n; l++) {
                size_t pa, pb;
                if (dir == 1) { pa = G(o, i, l, b); pb = G(o, i, l + 1, b); }
                else          { pa = G(o, i, b, l); pb = G(o, i, b, l + 1); }
                double s[4];
                cons_face(o, pa, pb, CC(l + 1) - CC(l), s);
                if (l > 0) {
                    for (int q = 0; q < 4; q++) {
                        double *f = dir == 1 ? &FL(i, l, b, q + 2) : &FL(i, b, l, q + 2);
                        *f = *f + fo[q] - s[q] * dhl;
                    }
                }
                if (l < n) {
                    dhl = 2.0 / (CC(l + 2) - CC(l));
                    for (int q = 0; q < 4; q++) fo[q] = s[q] * dhl;
                }
            }
        }
#undef CC
}

/* visflx_div_ord2_subroutine, kernels_cpu.F90:3446-3480 (visflx mode 7) */
static void visflx_div_ord2(or_state *o)
{
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) {
                double uu = WA(i, j, k, 2), vv = WA(i, j, k, 3), ww = WA(i, j, k, 4);
                double mu = WA(i, j, k, 7);
                double dxl = mu / (XC(i + 1) - XC(i - 1));
                double dyl = mu / (YC(j + 1) - YC(j - 1));
                double dzl = mu / (ZC(k + 1) - ZC(k - 1));
                double sigx = dxl * (WA(i + 1, j, k, 10) - WA(i - 1, j, k, 10));
                double sigy = dyl * (WA(i, j + 1, k, 10) - WA(i, j - 1, k, 10));
                double sigz = dzl * (WA(i, j, k + 1, 10) - WA(i, j, k - 1, 10));
                double sigq = sigx * uu + sigy * vv + sigz * ww;
                FL(i, j, k, 2) = FL(i, j, k, 2) - sigx;
                FL(i, j, k, 3) = FL(i, j, k, 3) - sigy;
                FL(i, j, k, 4) = FL(i, j, k, 4) - sigz;
                FL(i, j, k, 5) = FL(i, j, k, 5) - sigq;
            }
}

/* visflx(mode), singleideal_cpu.F90:186-279 (grid_dim == 1 branches) */
void or_visflx(or_state *o, int mode)
{
    if (mode == 0) visflx_lapl(o, 1);
    else if (mode == 1) visflx_div(o);
    else if (mode == 2) visflx_reduced_ord2(o, 1);
    else if (mode == 3) sensor(o);
    else if (mode == 4) visflx_lapl(o, 0);
    else if (mode == 5) {
        visflx_cons_x(o);
        visflx_cons_march(o, 1);
        visflx_cons_march(o, 2);
    }
    else if (mode == 6) visflx_reduced_ord2(o, 0);
    else if (mode == 7) visflx_div_ord2(o);
}

/* -------------------------------------------------------------- forcing */
/* force_rhs_1_subroutine (kernels_cpu.F90:3157-3193), the allreduce and
 * bulk update of force_rhs (singleideal_cpu.F90:845-889, grid_dim == 1) and
 * force_rhs_2_subroutine (kernels_cpu.F90:3092-3116) */
void or_force_rhs(or_state *o)
{
    double b1 = 0, b2 = 0, b3 = 0, b4 = 0, b5 = 0;
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) {
                double dy = o->yn[j] - o->yn[j - 1];
                b1 = b1 + FLN(i, j, k, 1) * dy;
                b2 = b2 + FLN(i, j, k, 2) * dy;
                b3 = b3 + W(i, j, k, 1) * dy;
                b4 = b4 + W(i, j, k, 2) * dy;
                b5 = b5 + W(i, j, k, 2) * dy * WA(i, j, k, 6);
            }
    double v = o->p.volchan;
    b1 = b1 / v; b2 = b2 / v; b3 = b3 / v; b4 = b4 / v; b5 = b5 / v;
    o->dpdx = o->dpdx + b2;
    o->rhobulk = b3;
    o->ubulk = b4 / o->rhobulk;
    o->tbulk = b5 / o->rhobulk / o->ubulk;
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) {
                double uu = WA(i, j, k, 2);
                FLN(i, j, k, 1) = FLN(i, j, k, 1) - b1;
                FLN(i, j, k, 2) = FLN(i, j, k, 2) - b2;
                FLN(i, j, k, 5) = FLN(i, j, k, 5) - uu * b2;
            }
}

/* force_var (singleideal_cpu.F90:890-933) with force_var_1/2_subroutine
 * (kernels_cpu.F90:3241-3314) */
void or_force_var(or_state *o)
{
    if (!(o->p.theta_wall >= -1.0)) return;
    double bulk_5 = 0.0;
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) {
                double dy = o->yn[j] - o->yn[j - 1];
                double rho = W(i, j, k, 1), rhou = W(i, j, k, 2), rhov = W(i, j, k, 3);
                double rhow = W(i, j, k, 4), rhoe = W(i, j, k, 5);
                double ri = 1.0 / rho;
                double uu = rhou * ri, vv = rhov * ri, ww = rhow * ri;
                double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
                double ee = rhoe / rho - qq;
                double tt = or_get_temperature_from_e(o, ee, WA(i, j, k, 6));
                WA(i, j, k, 6) = tt;
                bulk_5 = bulk_5 + rhou * tt * dy;
            }
    double bulkt = bulk_5 / o->p.volchan;
    bulkt = bulkt / o->rhobulk / o->ubulk;
    double tbdiff = o->p.t_bulk_target - bulkt;
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) {
                double rho = W(i, j, k, 1), rhou = W(i, j, k, 2), rhov = W(i, j, k, 3), rhow = W(i, j, k, 4);
                double tt = WA(i, j, k, 6);
                double ttnew = tt + tbdiff;
                double ee = or_get_e_from_temperature(o, ttnew);
                W(i, j, k, 5) = rho * ee + 0.5 * (rhou * rhou + rhov * rhov + rhow * rhow) / rho;
            }
}

/* ---------------------------------------------------------------- dt / RK */
/* compute_dt_subroutine (kernels_cpu.F90:6814-6880) and compute_dt
 * (singleideal_cpu.F90:1476-1517) */
void or_compute_dt(or_state *o)
{
    const st_params *p = &o->p;
    if (p->cfl < 0) {
        o->dt = -p->cfl;
        return;
    }
    double mx[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) {
                double rho = W(i, j, k, 1);
                double ri = 1.0 / rho;
                double uu = WA(i, j, k, 2), vv = WA(i, j, k, 3), ww = WA(i, j, k, 4);
                double tt = WA(i, j, k, 6), mu = WA(i, j, k, 7);
                double nu = ri * mu;
                double k_over_rhocp = nu / p->prandtl;
                double gamloc = or_get_gamloc(o, tt);
                double c = sqrt(gamloc * p->rgas0 * tt);
                double v[9];
                v[0] = (fabs(uu) + c) * o->dcsidx[i - 1];
                v[1] = (fabs(vv) + c) * o->detady[j - 1];
                v[2] = (fabs(ww) + c) * o->dzitdz[k - 1];
                v[3] = nu * o->dcsidxs[i - 1];
                v[4] = nu * o->detadys[j - 1];
                v[5] = nu * o->dzitdzs[k - 1];
                v[6] = k_over_rhocp * o->dcsidxs[i - 1];
                v[7] = k_over_rhocp * o->detadys[j - 1];
                v[8] = k_over_rhocp * o->dzitdzs[k - 1];
                for (int q = 0; q < 9; q++) mx[q] = dmax(mx[q], v[q]);
            }
    double dtinv_max = mx[0];
    for (int q = 1; q < 9; q++) dtinv_max = dmax(dtinv_max, mx[q]);
    double dt_min = 1.0 / dtinv_max;
    o->dt = p->cfl * dt_min;
}

/* rk_sync, singleideal_cpu.F90:332-409 (Wray RK3, grid_dim 1, no
 * limiter/sponge/tripping) */
void or_rk_sync(or_state *o)
{
    if (o->p.channel_case) o->dpdx = 0.0;
    for (int istep = 1; istep <= 3; istep++) {
        double rhodt = o->rhork[istep - 1] * o->dt;
        double gamdt = o->gamrk[istep - 1] * o->dt;
        or_init_flux(o, rhodt);
        or_bcswap(o);
        or_compute_aux(o, 1, 1);
        or_euler_x(o);
        if (o->p.conservative_viscous == 1) {
            or_visflx(o, 6);
            or_visflx(o, 5);
        } else {
            or_visflx(o, 4);
        }
        or_bcextr_var(o, 10);
        or_bcswap_var(o, 10);
        or_euler_y(o);
        if (o->nz > 1) or_euler_z(o);
        if (istep == 3) {
            or_visflx(o, 3);
            or_bcswap_var(o, 8);
        }
        or_visflx(o, o->p.conservative_viscous == 1 ? 7 : 1);
        or_bc_nr(o);
        or_update_flux(o, gamdt);
        if (o->p.channel_case) or_force_rhs(o);
        or_update_field(o);
        if (o->p.channel_case) or_force_var(o);
        or_update_ghost(o, 0);
    }
}

/* pre-loop sequence of run(), singleideal_cpu.F90:1547-1594 */
void or_init_run(or_state *o)
{
    or_update_ghost(o, 1);
    or_compute_aux(o, 1, 1);
    or_visflx(o, o->p.conservative_viscous == 1 ? 2 : 0);
    or_bcswap_var(o, 8);
    or_zero_flux(o);
    or_compute_dt(o);
    if (o->rc_plane) or_seed_random(o);
}

/* compute_residual_subroutine (kernels_cpu.F90:6682-6701),
 * compute_rho_t_p_minmax_subroutine (:6765-6810) and compute_residual
 * (singleideal_cpu.F90:1339-1364) */
void or_compute_residual(or_state *o)
{
    double r = 0.0;
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) r = r + sq(FLN(i, j, k, 2) / o->dt);
    double restemp = (double)o->nx * (double)o->ny * (double)o->nz;
    o->residual_rhou = sqrt(r / restemp);
    double rhomin = 1.7976931348623157e308, rhomax = -100.0;
    double tmin = rhomin, tmax = -100.0, pmin = rhomin, pmax = -100.0;
    for (int k = 1; k <= o->nz; k++)
        for (int j = 1; j <= o->ny; j++)
            for (int i = 1; i <= o->nx; i++) {
                double rho = WA(i, j, k, 1), tt = WA(i, j, k, 6);
                double pp = rho * tt * o->p.rgas0;
                rhomin = dmin(rhomin, rho); tmin = dmin(tmin, tt); pmin = dmin(pmin, pp);
                rhomax = dmax(rhomax, rho); tmax = dmax(tmax, tt); pmax = dmax(pmax, pp);
            }
    o->rhomin = rhomin; o->rhomax = rhomax; o->tmin = tmin; o->tmax = tmax; o->pmin = pmin; o->pmax = pmax;
}

/* integration loop body of run(), singleideal_cpu.F90:1599-1645 */
void or_advance(or_state *o, int nsteps)
{
    int iter_dt = o->p.iter_dt_recompute > 0 ? o->p.iter_dt_recompute : 1;
    for (int n = 0; n < nsteps; n++) {
        o->icyc = o->icyc + 1;
        if (o->p.cfl > 0 && ((o->icyc - o->icyc0) % iter_dt) == 0) {
            or_compute_aux(o, 1, 0);
            or_compute_dt(o);
        }
        or_rk_sync(o);
        o->time = o->time + o->dt;
    }
}

/* compute_stats (singleideal.F90:2003-2179, grid_dim 1): the 70 spanwise
 * sums of each (i, j), accumulated over k = 1..nz in order, divided by nz
 * and folded into the running time average
 * w_stat = (w_stat*itav + w_stat_z)/(itav + 1).  Velocity gradients come
 * from w (ghosts as they are), T and mu from w_aux(6), w_aux(7).
 * w_stat is (1:70, 1:nx, 1:ny) in Fortran order. */
void or_compute_stats(or_state *o, int itav, double *w_stat)
{
    const st_params *p = &o->p;
    const int nx = o->nx, ny = o->ny, nz = o->nz, nv = OR_NV_STAT;
    const int lmax = p->visc_order / 2;
    const double *c1 = o->cd1[lmax - 1];
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            double z[OR_NV_STAT];
            for (int q = 0; q < nv; q++) z[q] = 0.0;
            for (int k = 1; k <= nz; k++) {
                double ux = 0, vx = 0, wx = 0, uy = 0, vy = 0, wy = 0, uz = 0, vz = 0, wz = 0;
                for (int l = 1; l <= lmax; l++) {
                    const double ccl = c1[l - 1];
                    ux = ux + ccl * (W(i + l, j, k, 2) / W(i + l, j, k, 1) - W(i - l, j, k, 2) / W(i - l, j, k, 1));
                    vx = vx + ccl * (W(i + l, j, k, 3) / W(i + l, j, k, 1) - W(i - l, j, k, 3) / W(i - l, j, k, 1));
                    wx = wx + ccl * (W(i + l, j, k, 4) / W(i + l, j, k, 1) - W(i - l, j, k, 4) / W(i - l, j, k, 1));
                    uy = uy + ccl * (W(i, j + l, k, 2) / W(i, j + l, k, 1) - W(i, j - l, k, 2) / W(i, j - l, k, 1));
                    vy = vy + ccl * (W(i, j + l, k, 3) / W(i, j + l, k, 1) - W(i, j - l, k, 3) / W(i, j - l, k, 1));
                    wy = wy + ccl * (W(i, j + l, k, 4) / W(i, j + l, k, 1) - W(i, j - l, k, 4) / W(i, j - l, k, 1));
                    uz = uz + ccl * (W(i, j, k + l, 2) / W(i, j, k + l, 1) - W(i, j, k - l, 2) / W(i, j, k - l, 1));
                    vz = vz + ccl * (W(i, j, k + l, 3) / W(i, j, k + l, 1) - W(i, j, k - l, 3) / W(i, j, k - l, 1));
                    wz = wz + ccl * (W(i, j, k + l, 4) / W(i, j, k + l, 1) - W(i, j, k - l, 4) / W(i, j, k - l, 1));
                }
                const double dx = o->dcsidx[i - 1], dy = o->detady[j - 1], dz = o->dzitdz[k - 1];
                ux = ux * dx; vx = vx * dx; wx = wx * dx;
                uy = uy * dy; vy = vy * dy; wy = wy * dy;
                uz = uz * dz; vz = vz * dz; wz = wz * dz;
                const double omx = wy - vz, omy = uz - wx, omz = vx - uy;
                const double rho = W(i, j, k, 1), rhou = W(i, j, k, 2), rhov = W(i, j, k, 3), rhow = W(i, j, k, 4);
                const double ri = 1.0 / rho;
                const double uu = rhou * ri, vv = rhov * ri, ww = rhow * ri;
                const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
                const double tt = WA(i, j, k, 6);
                const double pp = rho * p->rgas0 * tt;
                const double mu = WA(i, j, k, 7);
                const double nu = mu / rho;
                const double gamloc = or_get_gamloc(o, tt);
                const double cploc = gamloc / (gamloc - 1.0) * p->rgas0;
                const double c = sqrt(gamloc * p->rgas0 * tt);
                const double machlocal = sqrt(2.0 * qq) / c;
                const double t_tot = tt + qq * (gamloc - 1.0) / gamloc;
                const double rho2 = rho * rho, uu2 = uu * uu, vv2 = vv * vv, ww2 = ww * ww, uv = uu * vv;
                const double pp2 = pp * pp, tt2 = tt * tt;
                const double divl = ux + vy + wz;
                const double div3l = divl / 3.0;
                double s11 = 2.0 * (ux - div3l), s12 = uy + vx, s13 = uz + wx;
                double s22 = 2.0 * (vy - div3l), s23 = vz + wy, s33 = 2.0 * (wz - div3l);
                s11 = s11 * mu; s12 = s12 * mu; s13 = s13 * mu;
                s22 = s22 * mu; s23 = s23 * mu; s33 = s33 * mu;
                const double s21 = s12, s31 = s13, s32 = s23;
                (void)uv;
                const double v[OR_NV_STAT] = {
                    rho, uu, vv, ww, pp, tt, rho2, uu2, vv2, ww2,
                    pp2, tt2, rhou, rhov, rhow, rhou * uu, rhov * vv, rhow * ww, rhou * vv, mu,
                    nu, omx * omx, omy * omy, omz * omz, rho * tt, rho * tt2, t_tot, rho * t_tot, t_tot * t_tot,
                    rhou * tt,
                    rhov * tt, rhow * tt, machlocal, machlocal * machlocal, rhou * uu2, rhov * uu2, rhou * vv2,
                    rhov * vv2, rhou * ww2, rhov * ww2,
                    pp * uu, pp * vv, s11, s12, s13, s22, s23, s33, s11 * uu, s12 * uu,
                    s21 * vv, s22 * vv, s31 * ww, s32 * ww, s11 * vv + s21 * uu, s12 * vv + s22 * uu,
                    s11 * ux + s12 * uy + s13 * uz, s21 * vx + s22 * vy + s23 * vz, s31 * wx + s32 * wy + s33 * wz,
                    s11 * vx + s12 * (ux + vy) + s22 * uy + s13 * vz + s23 * uz,
                    pp * ux, pp * vy, pp * wz, pp * (uy + vx), divl * divl, rho * tt2 * tt, rho * tt2 * tt2,
                    rhou * uu2 * uu, cploc, gamloc};
                for (int q = 0; q < nv; q++) z[q] = z[q] + v[q];
            }
            double *ws = w_stat + (size_t)nv * ((size_t)(i - 1) + (size_t)nx * (size_t)(j - 1));
            for (int q = 0; q < nv; q++) {
                const double zq = z[q] / (double)nz;
                ws[q] = (ws[q] * (double)itav + zq) / (double)(itav + 1);
            }
        }
}

/* count_weno_subroutine, kernels_cpu.F90:52-113, and the fractions of
 * count_weno (singleideal_cpu.F90:1366-1395) */
void or_count_weno(or_state *o, double frac[3])
{
    int ws = o->p.weno_scheme;
    double thr = o->p.sensor_threshold;
    double cx = 0, cy = 0, cz = 0;
    for (int k = 1; k <= o->nz; k++) for (int j = 1; j <= o->ny; j++) for (int i = o->eul_imin; i <= o->eul_imax; i++) {
        int ishk = 0;
        for (int ii = i - ws + 1; ii <= i + ws; ii++) if (WA(ii, j, k, 8) > thr) ishk = 1;
        if (ishk) cx = cx + 1;
    }
    for (int k = 1; k <= o->nz; k++) for (int j = o->eul_jmin; j <= o->eul_jmax; j++) for (int i = 1; i <= o->nx; i++) {
        int ishk = 0;
        for (int jj = j - ws + 1; jj <= j + ws; jj++) if (WA(i, jj, k, 8) > thr) ishk = 1;
        if (ishk) cy = cy + 1;
    }
    for (int k = o->eul_kmin; k <= o->eul_kmax; k++) for (int j = 1; j <= o->ny; j++) for (int i = 1; i <= o->nx; i++) {
        int ishk = 0;
        for (int kk = k - ws + 1; kk <= k + ws; kk++) if (WA(i, j, kk, 8) > thr) ishk = 1;
        if (ishk) cz = cz + 1;
    }
    frac[0] = cx / ((double)(o->eul_imax - o->eul_imin + 1) * o->ny * o->nz);
    frac[1] = cy / ((double)o->nx * (o->eul_jmax - o->eul_jmin + 1) * o->nz);
    frac[2] = cz / ((double)o->nx * o->ny * (o->eul_kmax - o->eul_kmin + 1));
}
