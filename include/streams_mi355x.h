/*
 * streams_mi355x.h -- C ABI of the MI355X (gfx950) per-RK-stage hot path of the
 * STREAmS-2 single-ideal-gas solver.
 *
 * The library owns every device array of one flow solve behind an opaque
 * handle.  Each entry point below replaces one type-bound procedure of the
 * reference backend object `equation_singleideal_cpu_object`
 * (code/src/singleideal/singleideal_cpu.F90:29-138) or one device-side
 * helper of the AMD backend (code/src/singleideal/kernels_amd.F90 bind(C)
 * interfaces <-> kernels_amd_cpp.cpp extern "C" launchers).  A Fortran
 * backend binds them with `bind(c, name="...")` interfaces; plain C types
 * only: int32 scalars by value, double scalars by value, host arrays as
 * pointers, results through pointers.
 *
 * Array conventions (identical to the reference, Fortran column-major):
 *   w      (1-ng:nx+ng, 1-ng:ny+ng, 1-ng:nz+ng, 1:5)   conserved variables
 *   w_aux  (1-ng:nx+ng, 1-ng:ny+ng, 1-ng:nz+ng, 1:10)  rho,u,v,w,H,T,mu,sensor,|omega|,div/3
 *   fl     (1:nx, 1:ny, 1:nz, 1:5)                     right-hand side
 *   x,y,z  (1-ng:n+ng) node coordinates, yn (1:ny+1) cell faces in y
 *   metric arrays dcsidx, dcsidxs, dcsidx2 (1:nx), detady.. (1:ny), dzitdz.. (1:nz)
 * Host arrays passed to and from the library have these layouts.  The
 * library's own device copies of w and w_aux pad each row to whole 128-byte
 * lines (an internal layout; the copies re-space the rows).
 *
 * Every function returns 0 on success and a non-zero ST_ERR_* code on
 * failure; st_last_error() returns a message for the calling thread.  The
 * Fortran side turns a non-zero status into fail_input_any/mpi_abort, as the
 * reference does for its own input errors (parameters.F90:166-171).
 */
#ifndef STREAMS_MI355X_H
#define STREAMS_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ST_OK               0
#define ST_ERR_ARG          1   /* invalid argument / unsupported option     */
#define ST_ERR_HIP          2   /* HIP runtime error                          */
#define ST_ERR_NOMEM        3   /* device allocation failed                   */
#define ST_ERR_STATE        4   /* call order violated (e.g. no grid set)     */
#define ST_ERR_NAN          5   /* NaN residual (reference 'boom!!!' path)    */

#define ST_MAX_CP_COEFF     16

/* Precision of the flux/stress evaluation.  FP64 everywhere is the plain
 * form; MIXED evaluates the convective reconstructions and viscous stresses
 * in FP32 while the state, RK accumulation, T/p recovery, WENO smoothness
 * indicators and all reductions stay FP64. */
#define ST_PREC_FP64        0
#define ST_PREC_MIXED       1

/* Scalar parameters of the path: the fields of equation_singleideal_object
 * that the hot loop reads (singleideal.F90:58-125) plus the grid sizes. */
typedef struct st_params {
    int32_t nx, ny, nz, ng;          /* interior cells per direction, ghost width: 3, or
                                        4 for weno_scheme 4 / ep_order 8 (FP64, periodic
                                        faces and staggered walls, stage run procedure by
                                        procedure)                                          */
    int32_t ep_order;                /* central convective order (2,4,6,8)                */
    int32_t visc_order;              /* viscous order (== ep_order)                       */
    int32_t nkeep;                   /* KEEP(n): 0 KEEP, >0 KEEP(n), <0 non-KEEP split    */
    int32_t weno_scheme;             /* 1..4, WENO order 2*weno_scheme-1 (3 = WENO5)      */
    int32_t weno_version;            /* 0 Jiang-Shu, 1 WENO-Z                             */
    int32_t flux_splitting;          /* WENO faces: 0 Roe characteristic, 1 Rusanov      */
    int32_t conservative_viscous;    /* 0 Laplacian form, 1 conservative form             */
    int32_t visc_model;              /* 0 inviscid, 1 power law, 2 Sutherland             */
    int32_t calorically_perfect;     /* 1 constant cp, 0 cp(T) polynomial                 */
    int32_t indx_cp_l, indx_cp_r;    /* exponent range of the cp polynomial               */
    int32_t bctags[6];               /* xmin,xmax,ymin,ymax,zmin,zmax (update_ghost,
                                        singleideal_cpu.F90:1064-1138): 0 periodic, 1 free stream,
                                        2 extrapolation, 3 subsonic free stream from total
                                        conditions (xmin, ymax), 4 extrapolation at p0 (all but
                                        ymin), 5 symmetry (ymin), 6 isothermal wall (ymin; ymax
                                        too when staggered), 7 oblique shock (ymax), 8 none,
                                        9 laminar inflow from wmean (xmin), 10 recycling-rescaling
                                        inflow (xmin).  On the faces not listed a tag leaves the
                                        ghosts alone, as in the reference. */
    int32_t bctags_nr[6];            /* characteristic BC: 0 none, 1 non-reflecting, 2 relaxation
                                        to the neighbour, 3 relaxation to winf, 6 reflective wall */
    int32_t is_y_staggered;          /* 1 for the channel (cell-centred y)                */
    int32_t channel_case;            /* 1 enables force_rhs / force_var                   */
    int32_t correct_bound_ord;       /* 1 lowers the stencil order at non-periodic faces  */
    int32_t iter_dt_recompute;       /* CFL dt recomputed every this many steps           */
    int32_t precision;               /* ST_PREC_FP64 or ST_PREC_MIXED                     */
    int32_t rand_type;               /* random inflow plane of the recycling inflow: 0 off,
                                        > 0 reproducible (C library rand() seeded 1, the
                                        single-rank seed rank+1), < 0 seeded from the clock
                                        (singleideal_cpu.F90:1585-1594) */
    int32_t pad1;                    /* explicit: the record has no implicit padding      */
    double rgas0, prandtl, t0, rho0, u0, l0, mu0;
    double sutherland_s, t_ref_dim, powerlaw_vtexp;
    double sensor_threshold;
    double t_wall;                   /* isothermal wall temperature                        */
    double theta_wall;               /* >= -1 activates bulk-temperature control          */
    double t_bulk_target;
    double volchan;                  /* channel volume (yn(ny+1)-yn(1))*nxmax*nzmax        */
    double cfl;                      /* > 0 CFL number, < 0 fixed dt = -cfl               */
    double tol_iter_nr;              /* Newton tolerance for T(e) (1e-12 in FP64)          */
    double cp_coeff[ST_MAX_CP_COEFF];/* cp_coeff(indx_cp_l:indx_cp_r+1), index l-indx_cp_l */
    double cv_coeff[ST_MAX_CP_COEFF];
    /* free-stream and oblique-shock boundary data (bctags 1, 7, bctags_nr 3):
     * winf and winf_past_shock are conserved states, xshock_imp the impinging
     * point, shock_angle in radians, tanhfacs the smoothing length
     * (singleideal.F90:2798-2826 set_flow_params, :2955-3017 set_oblique_shock) */
    double winf[5], winf_past_shock[5];
    double xshock_imp, shock_angle, tanhfacs;
    /* total conditions of the subsonic free-stream boundary (bctags 3):
     * angle of attack in radians, ptot0 and ttot0 (set_flow_params,
     * singleideal.F90:2807-2814) */
    double aoa, ptot0, ttot0;
} st_params;

/* Device-resident scalars of the run, copied to the host on request. */
typedef struct st_scalars {
    double dt;          /* current time step                                   */
    double time;        /* simulated time                                      */
    double dpdx;        /* accumulated forcing of the current step             */
    double rhobulk, ubulk, tbulk;
    double residual_rhou;
    double rhomin, rhomax, tmin, tmax, pmin, pmax;
    int64_t icyc;       /* completed steps                                     */
} st_scalars;

typedef struct st_solver st_solver;

/* ---- lifecycle: device allocation owned by the library -------------------
 * Replaces the hipfort hipMalloc of the backend arrays
 * (base_amd.F90:1211-1413, singleideal_amd.F90:1268-1327) and the setup of
 * bc_preproc / correct_bc_order (singleideal.F90:4703-4750, 946-993). */
int st_solver_create(const st_params *p, st_solver **out);
int st_solver_destroy(st_solver *s);
/* Bytes of HBM one solve of these sizes occupies (admission control). */
int64_t st_solver_footprint(const st_params *p);

/* Grid and 1-D metrics (field.F90:456-516 set_local_grid output). */
int st_set_grid(st_solver *s,
                const double *x, const double *y, const double *z, const double *yn,
                const double *dcsidx, const double *dcsidxs, const double *dcsidx2,
                const double *detady, const double *detadys, const double *detady2,
                const double *dzitdz, const double *dzitdzs, const double *dzitdz2);

/* copy_from_field / copy_to_field (base_cpu.F90:1914-1929). */
int st_copy_from_field(st_solver *s, const double *w_host);
int st_copy_to_field(st_solver *s, double *w_host);
/* Read / write one w_aux variable (1..10) over the ghosted box. */
int st_get_aux(st_solver *s, int32_t var, double *host);
int st_set_aux(st_solver *s, int32_t var, const double *host);
/* Read the current right-hand side fl (interior, 5 variables). */
int st_get_fl(st_solver *s, double *host);
int st_get_scalars(st_solver *s, st_scalars *out);
int st_set_time(st_solver *s, double time, double dt, int64_t icyc);

/* ---- procedures of the backend object ------------------------------------ */
/* update_ghost (singleideal_cpu.F90:1064-1138): physical BC ghosts, then the
 * periodic ghost copy of bcswap when do_swap != 0. */
int st_update_ghost(st_solver *s, int32_t do_swap);
/* bcswap (base_cpu.F90:1326-1435), single-GPU periodic copy. */
int st_bcswap(st_solver *s);
/* compute_aux (singleideal_cpu.F90:281-330). */
int st_compute_aux(st_solver *s, int32_t central, int32_t ghost);
/* visflx(mode) (singleideal_cpu.F90:186-279), accumulating into fl:
 *   0 Laplacian form + sensor, 1 div/3 terms, 3 sensor, 4 Laplacian form;
 *   conservative form (conservative_viscous=1): 2 cell terms + sensor,
 *   5 face terms, 6 cell terms, 7 second-order div/3 terms. */
int st_visflx(st_solver *s, int32_t mode);
/* euler_x(istart, iend, istart_face, iend_face, do_update)
 * (singleideal_cpu.F90:674-747): when iend >= istart, the hybrid fluxes of
 * faces istart_face..iend_face (the reference's defaults are istart-1 and
 * iend) and, with do_update != 0, their divergence into fl over cells
 * eul_imin..eul_imax.  Face fluxes persist between calls, so a sweep split
 * into ranges (as rk_async splits it) updates from all of them.
 * euler_y(eul_jmin, eul_jmax) / euler_z(eul_kmin, eul_kmax) (:749-843):
 * faces eul_min-1..eul_max and cells eul_min..eul_max; euler_z also redraws
 * the random inflow plane of a recycling inflow.  The ranges of bc_preproc
 * (st_get_eul_range) run the fused sweeps; any other range runs the face
 * kernels of the reference's decomposition. */
int st_euler_x(st_solver *s, int32_t istart, int32_t iend, int32_t istart_face, int32_t iend_face,
               int32_t do_update);
int st_euler_y(st_solver *s, int32_t eul_jmin, int32_t eul_jmax);
int st_euler_z(st_solver *s, int32_t eul_kmin, int32_t eul_kmax);
/* eul_imin, eul_imax, eul_jmin, eul_jmax, eul_kmin, eul_kmax of bc_preproc
 * (singleideal.F90:4703-4750). */
int st_get_eul_range(st_solver *s, int32_t range[6]);

/* The remaining procedures of one rk_sync stage (singleideal_cpu.F90:
 * 352-406), for a backend that calls them one by one instead of st_rk_sync:
 * init_flux (fln = -rhodt*fl, fl = 0; kernels_cpu.F90:29-47), update_flux
 * (fln -= gamdt*fl; :3400-3418), update_field (w += fln; :3422-3442),
 * force_rhs (bulk sums of fln, w and T -> dpdx, rho_b, u_b, T_b, then the
 * forcing of fln; singleideal_cpu.F90:845-889), force_var (bulk-temperature
 * control when theta_wall >= -1; :890-933), the dpdx = 0 of rk_sync's step
 * start (:347-350), bcextr_var and bcswap_var of one w_aux variable
 * (:411-426, base_cpu.F90:1835-1912) and recyc_exchange (the recycling
 * station planes with the half-span shift; :935-999).  rhodt and gamdt are
 * rhork*dt and gamrk*dt of the stage.  fl is the right-hand side the sweeps
 * and visflx accumulate into. */
int st_init_flux(st_solver *s, double rhodt);
int st_update_flux(st_solver *s, double gamdt);
int st_update_field(st_solver *s);
int st_force_rhs(st_solver *s);
int st_force_var(st_solver *s);
int st_reset_forcing(st_solver *s);
int st_bcextr_var(st_solver *s, int32_t var);
int st_bcswap_var(st_solver *s, int32_t var);
int st_recyc_exchange(st_solver *s);
/* Recycling-rescaling inflow of face xmin (bctags[0] == 10): the data of
 * recyc_prepare (singleideal.F90:1512-1588).  i_recyc is the station (cell
 * index >= ng), betarecyc / glund1 the rescaling factors, wmean_inflow the
 * inflow mean profile wmean(1-ng:0, 1:ny, 1:4) in Fortran order, weta the
 * blending weight (1:ny), map_j_inn / map_j_out the locateval maps (1:ny;
 * 1..ny-1 where used), the four coordinate arrays y+ and eta of the inflow
 * and of the station over 1-ng:ny+ng.  With rand_type != 0 the random
 * inflow plane (rows 2..jbl_inflow, jbl_inflow = locateval(y(1:ny), l0))
 * is redrawn on the device after every z sweep in the order of
 * get_crandom_f (j, then k, then component), bit-identical to the C
 * library's rand() seeded as the reference seeds it, and adds
 * 0.02*u0*(r - 0.5)*min(eta, 1) to the recycled velocity fluctuations
 * (kernels_cpu.F90:5879-5923).  Required before the first step of such a
 * solve. */
/* The inflow mean planes wmean(1-ng:0, 1:ny, 1:4) (rho, rho*u, rho*v,
 * rho*w) in Fortran order, offset 0 = plane 1-ng: the data of the laminar
 * inflow (bctags[0] == 9, bclam_subroutine, kernels_cpu.F90:5951-5985).
 * st_set_recycling sets them too. */
int st_set_wmean(st_solver *s, const double *wmean);
int st_set_recycling(st_solver *s, int32_t i_recyc, double betarecyc, double glund1, const double *wmean_inflow,
                     const double *weta, const int32_t *map_j_inn, const int32_t *map_j_out,
                     const double *yplus_inflow, const double *eta_inflow, const double *yplus_recyc,
                     const double *eta_recyc);
/* bc_nr (singleideal_cpu.F90:1001-1062): characteristic boundary treatment
 * of the faces with bctags_nr != 0, added to fl. */
int st_bc_nr(st_solver *s);
/* zero_flux_subroutine (kernels_cpu.F90:9-25). */
int st_zero_flux(st_solver *s);
/* Start-of-run sequence of run() (singleideal_cpu.F90:1547-1594):
 * update_ghost, compute_aux, visflx(0) (visflx(2) for the conservative
 * form), sensor ghost swap, zero_flux, compute_dt. */
int st_init_run(st_solver *s);
/* compute_dt (singleideal_cpu.F90:1476-1517); dt_out may be NULL, in which
 * case dt stays on the device and no host synchronisation happens. */
int st_compute_dt(st_solver *s, double *dt_out);
/* rk_sync (singleideal_cpu.F90:332-409): one full Wray RK3 time step. */
int st_rk_sync(st_solver *s);
/* The integration loop body of run() (singleideal_cpu.F90:1599-1645) for
 * nsteps steps: dt every iter_dt_recompute steps, rk_sync, time += dt.  No
 * host synchronisation inside. */
int st_advance(st_solver *s, int32_t nsteps);
/* compute_residual (singleideal_cpu.F90:1339-1364): residual of rho*u and
 * rho/T/p extrema into the scalars; returns ST_ERR_NAN on a NaN residual. */
int st_compute_residual(st_solver *s, st_scalars *out);
/* count_weno (singleideal_cpu.F90:1366-1395): fraction of WENO faces. */
int st_count_weno(st_solver *s, double frac_xyz[3]);
/* compute_stats (singleideal.F90:2003-2179, called from run() at
 * singleideal_cpu.F90:2020-2041): the 70 spanwise means of the current
 * state folded into the running time average held on the device,
 * w_stat = (w_stat*itav + mean_z)/(itav + 1), as the itav-th sample
 * (itav = 0 starts a new average).  Each spanwise sum runs over k = 1..nz in
 * order.  Velocity gradients use the periodic ghosts of w refreshed from the
 * current interior; the reference's run() holds those of the last stage's
 * start at that point.  No host copy of the field is made. */
int st_compute_stats(st_solver *s, int32_t itav);
/* The running statistics w_stat(1:70, 1:nx, 1:ny) in Fortran order
 * (write_stats, singleideal.F90:1779-1822, writes them to stat.bin);
 * st_set_stats loads them back (restart_type 2, read_stats). */
int st_get_stats(st_solver *s, double *w_stat);
int st_set_stats(st_solver *s, const double *w_stat);

/* No counterpart in the reference: compute_dt (singleideal_cpu.F90:1476-1517)
 * without a solver handle, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It allocates nothing: the block maxima live in a static
 * device array of 257 KiB per GPU, shared by every call on that GPU, so
 * calls that overlap on one GPU race.  All arrays are device pointers of the
 * solve described by p, in the reference layouts above (ghosted rows of
 * nx + 2 ng values, no padding): w (5 ghosted variables), w_aux (the solve's w_aux:
 * double, or float with 11 variables when p->precision is ST_PREC_MIXED;
 * variable 6, T, is the start of the T(e) Newton iteration and receives T
 * when calorically_perfect = 0, as compute_aux does), the six metric arrays
 * of st_set_grid, and dt_out (one double, written as cfl / max of the
 * spectral radii; untouched when cfl < 0).  Work is enqueued on hip_stream
 * (NULL = the null stream) and not waited for.  Limit: ceil(nx/64) *
 * ceil(ny/4) <= 32768 (ST_ERR_ARG beyond). */
int st_compute_dt_noalloc(const st_params *p, void *hip_stream, const double *w, void *w_aux,
                          const double *dcsidx, const double *detady, const double *dzitdz,
                          const double *dcsidxs, const double *detadys, const double *dzitdzs, double *dt_out);

/* Measurement helper, no counterpart in the reference (its only timer is the
 * 'averaged timing' of run(), singleideal_cpu.F90:1661-1666): runs nsteps
 * steps of st_advance with HIP events recorded on the solver's stream around
 * every kernel and returns, per kernel family ST_K_*, the mean duration of
 * one launch in ms (ms[]) and the number of launches (count[]). */
#define ST_K_AUX        0
#define ST_K_EULER_X    1
#define ST_K_VISC       2
#define ST_K_EULER_Y    3
#define ST_K_EULER_Z    4
#define ST_K_SENSOR     5
#define ST_K_VISC_DIV   6
#define ST_K_FINISH     7
#define ST_K_UPDATE     8
#define ST_K_FORCE_VAR  9
#define ST_K_DT         10
#define ST_K_STEP_END   11
#define ST_K_BC         12   /* physical ghost faces and characteristic boundaries */
#define ST_K_EULER_XY   13   /* x and y Euler fluxes in one pass (y range spanning all rows) */
#define ST_K_WENO_FIX   14   /* WENO faces of the Euler sweeps (list compaction + fix-up) */
#define ST_K_EULER_ZV   15   /* z Euler fluxes + viscous divergence terms + force_rhs partials (its z fix-up counts under ST_K_WENO_FIX) */
#define ST_NKFAM        16
int st_kernel_times(st_solver *s, int32_t nsteps, double ms[ST_NKFAM], int32_t count[ST_NKFAM]);

/* Stream used by every launch of this solver (NULL = the library's own). */
/* Launch shapes (no reference counterpart; tuning only): planes per tile
 * of the z-marching viscous kernel and points per thread of the y/z flux
 * marches, 0 = automatic.  Results do not depend on them. */
int st_set_tuning(st_solver *s, int32_t visc_chunk, int32_t march_chunk);
int st_set_stream(st_solver *s, void *hip_stream);
int st_synchronize(st_solver *s);

const char *st_last_error(void);
const char *st_version(void);

/* ---- serving loop (no reference counterpart) ----------------------------
 * Continuous batching of many independent solves on one GPU: requests are
 * admitted in arrival order as slots and HBM free up (the server holds
 * hbm_capacity_bytes as one device pool from its creation and carves each
 * solve's st_solver_footprint out of it: no allocation while solves run),
 * each runs on its own stream from a HIP graph of steps_per_graph steps
 * (0: no graph), and their work overlaps on the device.  autotune != 0
 * times launch shapes (st_set_tuning) on the first solves of each
 * grid-shape class, during their own steps, and keeps the fastest.  A
 * result is identical to the same solve run alone through st_advance. */
typedef struct st_server st_server;
int st_server_create(int64_t hbm_capacity_bytes, int32_t max_slots, int32_t steps_per_graph, int32_t autotune,
                     st_server **out);
int st_server_destroy(st_server *sv);
/* Queue one solve: parameters, the 13 grid arrays of st_set_grid, the
 * initial field w (5 ghosted variables) and the number of RK3 steps after
 * st_init_run.  Inputs are copied; *ticket identifies the request. */
int st_server_submit(st_server *sv, const st_params *p,
                     const double *x, const double *y, const double *z, const double *yn,
                     const double *dcsidx, const double *dcsidxs, const double *dcsidx2,
                     const double *detady, const double *detadys, const double *detady2,
                     const double *dzitdz, const double *dzitdzs, const double *dzitdz2,
                     const double *w0, int32_t nsteps, int64_t *ticket);
/* st_server_submit with a start (time0, icyc0: a request resumed from its
 * restart files, as st_set_time) and, with checkpoint_every > 0 (a multiple
 * of steps_per_graph, or of 10 without graphs), the request's restart files
 * rst.bin and field_info0.dat written to checkpoint_dir every
 * checkpoint_every steps and after its last step, in the formats of
 * write_field / write_field_info (field.F90:1853-1966,
 * singleideal.F90:2546-2659).  Requests with an inflow plane (xmin tags 9,
 * 10) are not taken. */
int st_server_submit2(st_server *sv, const st_params *p,
                      const double *x, const double *y, const double *z, const double *yn,
                      const double *dcsidx, const double *dcsidxs, const double *dcsidx2,
                      const double *detady, const double *detadys, const double *detady2,
                      const double *dzitdz, const double *dzitdzs, const double *dzitdz2,
                      const double *w0, int32_t nsteps, double time0, int64_t icyc0,
                      int32_t checkpoint_every, const char *checkpoint_dir, int64_t *ticket);
/* Wait (timeout_s < 0: no limit) for a request, copy its final field (may
 * be NULL), scalars after st_compute_residual (may be NULL) and its submit /
 * admit / done times in seconds since server creation (may be NULL), then
 * forget it.  ST_ERR_STATE on time-out. */
int st_server_wait(st_server *sv, int64_t ticket, double timeout_s, double *w_out, st_scalars *scalars,
                   double times[3]);
/* [0] queued, [1] running, [2] completed, [3] failed, [4] HBM bytes reserved,
 * [5] tuned shape classes. */
int st_server_stats(st_server *sv, int64_t out[6]);
/* Batched launches (call before the first submission; an even number of
 * steps per chunk: steps_per_graph, or 10 with steps_per_graph 0, in which
 * case no graph is captured and batched chunks are enqueued launch by
 * launch).  With lanes > 0 (at most 8) each running solve of at most
 * max_cells interior cells (0: any size) is bound to one of `lanes` batch
 * streams (the one holding the fewest cells).  While a lane holds two
 * solves or more, the chunks of its solves whose launch sequences agree
 * (same precision, scheme and boundary set, any grid) run as one sequence
 * of launches, each launch a single grid over a table of the solves: one
 * launch per kernel per stage and lane instead of one per solve.  A solve
 * alone in its lane, larger solves, solves whose steps have no batched
 * form, the chunks that tune a class's launch shape and a final chunk
 * shorter than steps_per_graph run on their own streams as without
 * batching.  lanes = 0 turns batching off.  Results are unchanged, bit for
 * bit. */
int st_server_set_batching(st_server *sv, int32_t lanes, int64_t max_cells);
/* [0] batched chunks, [1] batched launches, [2] solves summed over the
 * batched chunks, [3] table layouts (recorded chunks; a lane whose solves
 * are unchanged reuses its last layout), [4] microseconds the scheduler
 * spent recording and laying out, [5] microseconds enqueuing batched
 * launches, [6] batched chunks replayed from a HIP graph (a reused layout
 * is captured once per table buffer). */
int st_server_batch_stats(st_server *sv, int64_t out[7]);
const char *st_server_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* STREAMS_MI355X_H */
