/*
 * streams_oracle.h -- CPU oracle of the STREAmS-2 per-RK-stage hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU backend (code/src/singleideal/kernels_cpu.F90 and the stage ordering of
 * singleideal_cpu.F90).  It is the checker the HIP path is compared with; it
 * is never linked into, called by or shipped with the product library.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Arrays use the reference layout exactly (Fortran column-major, lower bound
 * 1-ng for ghosted arrays); see streams_mi355x.h for the list.
 */
#ifndef STREAMS_ORACLE_H
#define STREAMS_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/streams_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_state {
    st_params p;
    int nx, ny, nz, ng;
    int sx, sy, sz;          /* ghosted extents                          */
    size_t nvol, nint;       /* points per ghosted / interior variable   */
    double *w;               /* 5  * nvol                                */
    double *w_aux;           /* 10 * nvol                                */
    double *fl, *fln;        /* 5  * nint                                */
    double *fhat;            /* 5  * nvol                                */
    double *wallprop;        /* 3  * sx * sz  (index 2..4)               */
    int *eoc;                /* ep_ord_change(0:nx,0:ny,0:nz,1:3)        */
    double *x, *y, *z, *yn;
    double *dcsidx, *dcsidxs, *dcsidx2;
    double *detady, *detadys, *detady2;
    double *dzitdz, *dzitdzs, *dzitdz2;
    double cd1[4][4];        /* coeff_deriv1(l,lmax) = cd1[lmax-1][l-1]   */
    double cd2[4][5];        /* coeff_deriv2(l,lmax) = cd2[lmax-1][l]     */
    int eul_imin, eul_imax, eul_jmin, eul_jmax, eul_kmin, eul_kmax;
    int force_zero_flux[6];
    int nr_eff[6];           /* bctags_nr after bc_preproc               */
    /* recycling-rescaling inflow (bctags(1) == 10) */
    int recyc_set, i_recyc;
    double betarecyc, glund1;
    double *rc_wmean;        /* wmean(1-i, j, m): [(ng-i) + ng*((j-1) + ny*(m-1))] */
    double *rc_plane;        /* inflow_random_plane(1:ny, 1:nz, 1:3), starts at 0.5 */
    char rng_state[128];     /* C library random_r state of get_crandom (TYPE_3) */
    double rng_data[8];      /* struct random_data of the C library, kept opaque here */
    double *rc_weta;         /* weta_inflow(1:ny)                         */
    int *rc_inn, *rc_out;    /* map_j_inn / map_j_out (1:ny)              */
    double *rc_y;            /* yplus_inflow, eta_inflow, yplus_recyc, eta_recyc (1-ng:ny+ng) */
    double *rc_wrecyc;       /* wrecyc(1:ng, 1:ny, 1:nz, 1:5)             */
    double rhork[3], gamrk[3];
    double dt, time, dpdx, rhobulk, ubulk, tbulk;
    double residual_rhou, rhomin, rhomax, tmin, tmax, pmin, pmax;
    int64_t icyc, icyc0;
    int print_control;
} or_state;

or_state *or_create(const st_params *p);
void or_destroy(or_state *o);
int or_set_grid(or_state *o, const double *x, const double *y, const double *z, const double *yn,
                const double *dcsidx, const double *dcsidxs, const double *dcsidx2,
                const double *detady, const double *detadys, const double *detady2,
                const double *dzitdz, const double *dzitdzs, const double *dzitdz2);

/* pointers to the state arrays, for the ctypes wrapper */
double *or_w(or_state *o);
double *or_w_aux(or_state *o);
double *or_fl(or_state *o);
double *or_fln(or_state *o);
double *or_fhat(or_state *o);
double *or_wallprop(or_state *o);
int *or_eoc(or_state *o);
double *or_inflow_plane(or_state *o);  /* inflow_random_plane(1:ny,1:nz,1:3), NULL before or_set_recycling */
void or_get_scalars(const or_state *o, st_scalars *out);
void or_set_time(or_state *o, double time, double dt, int64_t icyc);
void or_rk_coeffs(const or_state *o, double rho[3], double gam[3]);

/* building blocks, each following one reference routine */
void or_get_deriv_coeffs(double *d1, double *d2, int order1, int order2);
double or_get_gamloc(const or_state *o, double tt);
double or_get_temperature_from_e(const or_state *o, double ee, double t_start);
double or_get_e_from_temperature(const or_state *o, double tt);
void or_compute_roe_average(const or_state *o, int dir, int i, int j, int k, double out[9]);
void or_eigenvectors(int dir, const double roe[9], double el[5][5], double er[5][5]);
void or_wenorec_1d(const double gp[5][9], const double gm[5][9], double vhat[5], int iweno,
                   int wenorec_ord, int weno_version, double rho0, double u0);
void or_wenorec_1d_rusanov(const double gp[5][9], const double gm[5][9], double vhat[5], int iweno,
                           int wenorec_ord, int weno_version, double rho0, double u0);

void or_eval_aux(or_state *o, int i0, int i1, int j0, int j1, int k0, int k1);
void or_compute_aux(or_state *o, int central, int ghost);
void or_bcswap(or_state *o);
void or_bcswap_var(or_state *o, int var);
void or_bcextr_var(or_state *o, int var);
void or_update_ghost(or_state *o, int do_swap);
void or_bc_nr(or_state *o);
int or_set_wmean(or_state *o, const double *wmean);
/* init_crandom_f of run() (singleideal_cpu.F90:1585-1594): seed the draws of
 * the random inflow plane (rand_type > 0: seed 1 = rank+1 of one rank) */
void or_seed_random(or_state *o);
int or_set_recycling(or_state *o, int i_recyc, double betarecyc, double glund1, const double *wmean,
                     const double *weta, const int *map_inn, const int *map_out, const double *yplus_inflow,
                     const double *eta_inflow, const double *yplus_recyc, const double *eta_recyc);
void or_zero_flux(or_state *o);
void or_init_flux(or_state *o, double rhodt);
void or_euler_fluxes(or_state *o, int dir, int fstart, int fend);
void or_euler_update(or_state *o, int dir, int cstart, int cend);
void or_euler_x(or_state *o);
void or_euler_y(or_state *o);
void or_euler_z(or_state *o);
void or_visflx(or_state *o, int mode);
void or_update_flux(or_state *o, double gamdt);
void or_force_rhs(or_state *o);
void or_update_field(or_state *o);
void or_force_var(or_state *o);
void or_compute_dt(or_state *o);
void or_rk_sync(or_state *o);
void or_init_run(or_state *o);
void or_advance(or_state *o, int nsteps);
void or_compute_residual(or_state *o);
void or_count_weno(or_state *o, double frac[3]);
#define OR_NV_STAT 70
void or_compute_stats(or_state *o, int itav, double *w_stat);

#ifdef __cplusplus
}
#endif

#endif
