// st_solver.h -- host-side solver state shared by the C ABI and the launchers.
#pragma once

#include <string>

#include "st_batch.h"
#include "st_internal.h"

struct st_solver {
    st_params p;
    st::Geo g;
    st::KPar kp;
    hipStream_t stream;
    bool own_stream;
    int device;
    // state
    double *w;          // 5 ghosted variables
    void *w_aux;        // 10 ghosted variables (double, or float in the mixed form)
    void *w_alloc, *w_aux_alloc;  // the allocations of w and w_aux (their first points sit lead_bytes in)
    void *F[2];         // right-hand sides of the current and the previous stage (same type)
    bool mixed;         // ST_PREC_MIXED: w_aux and F in float
    bool general_bc;    // boundary faces other than periodic / staggered wall: per-face ghost kernels
    bool has_nr;        // a characteristic boundary treatment is active
    // recycling-rescaling inflow (bctags[0] == 10)
    bool recyc, recyc_set;
    int i_recyc;
    double betarecyc, glund1;
    double *rc_real;    // wmean planes (ng*ny*4), weta (ny), y+/eta arrays (4*(ny+2ng)), averages (ng*ny*5)
    int *rc_maps;       // map_j_inn (ny), map_j_out (ny)
    bool wmean_set;     // the wmean planes at the head of rc_real are loaded (tags 9, 10)
    double *wrecyc;     // recycled station planes wrecyc(1:ng, 1:ny, 1:nz, 1:5) (recyc_exchange)
    // the procedure-by-procedure stage (allocated on first use)
    double *fhat;       // face fluxes, 5 ghosted variables
    double *fln;        // RK increment of the stage, 5 interior variables
    bool fln_current;   // the last stage was applied by st_update_field from fln
    // random inflow plane of the recycling inflow (rand_type != 0): the C
    // library's rand() stream, advanced on the device by jump coefficients
    int jbl;            // locateval(y(1:ny), l0): rows 2..jbl are redrawn
    int rng_n;          // draws per z sweep, 3*(jbl-1)*nz
    double *rplane;     // inflow_random_plane(1:ny, 1:nz, 1:3)
    uint32_t *rng_coef; // coefficient j of draw t at [j*rng_n + t], j = 0..30
    uint32_t *rng_win;  // two 31-word windows of the generator state
    int *rng_ctl;       // [0] window in use, [1] ticket of the draw kernel
    uint32_t rng_seeded[31];  // window right after seeding
    double *w_stat;     // running statistics w_stat(1:70, 1:nx, 1:ny), allocated on first use
    bool dt_in_stage;   // the CFL reduction of this step rides on the first stage's aux
    int visc_chunk;     // planes per tile of the z-marching viscous kernel (0: automatic)
    int march_chunk;    // points per thread of the y/z flux marches (0: automatic)
    bool fuse_xy;       // x and y Euler fluxes in one pass
    bool fuse_zv;       // viscous divergence terms folded into the z march (Laplacian form)
    int force_nblk;     // force_rhs partials written in this stage (visc_div blocks, or z march + fix-up)
    int zv_nblk;        // blocks of the last z march with the force_rhs partials (the fix-up's come after them)
    double *wallprop;   // 3 x (sx*sz), wallprop(:,:,2:4)
    int *eoc[3];        // ep_ord_change per direction, faces 0..n
    // grid
    double *x, *y, *z, *yn;
    double *dcsidx, *dcsidxs, *dcsidx2, *detady, *detadys, *detady2, *dzitdz, *dzitdzs, *dzitdz2;
    // scalars and reduction scratch
    st::DevScalars *ds;
    st::DevScalars *hs;  // pinned host mirror
    double *partials;
    int npart;           // capacity in doubles
    unsigned char *dirty;  // per-cell flag: a WENO face on either side (one sweep at a time)
    int *counters;         // [1] reduction ticket, [8 ..] segment counts of the WENO fix-up's lists
    int *fixlist;          // cells of one direction's WENO fix-up, in segments (nint entries)
    int cur;             // F[cur] is the right-hand side being built
    double rhork[3], gamrk[3];
    int last_stage;      // RK stage (0..2) of the last completed stage, -1 none
    bool grid_set;
    int64_t icyc, icyc0; // host mirror of the step counter
    // optional per-kernel timing (st_kernel_times)
    struct KTimer *timer;
    // while set, the launchers record their launches here instead of
    // enqueuing them (batched launches over several solves)
    st::Recorder *rec;
    // memory lent by the creator (st::SlotMem): [arena_lo, arena_hi) holds
    // the solver's arrays, hs is the lender's; neither is freed with the solver
    char *arena_lo, *arena_hi;
    bool borrowed_hs;
};

// HIP events around each kernel family while st_kernel_times runs
struct KTimer {
    static const int kMax = 1 << 15;
    hipEvent_t ev[kMax][2];
    int fam[kMax];
    int n;      // slots used
    int made;   // slots whose events exist
};

namespace st {

// Memory a caller lends a solver so that creating it allocates nothing:
// a device arena of at least st_solver_footprint bytes (256-byte aligned)
// carved into the solver's arrays, the stream its work is ordered on,
// a pinned scalar mirror, and pinned room for nx + ny + nz + 3 ints of
// tables that must stay untouched until the stream has passed the
// creation.  The lender keeps all of it alive until the solver is
// destroyed and its stream has drained.
struct SlotMem {
    char *arena;
    size_t bytes;
    hipStream_t stream;
    DevScalars *hs;
    int *tabs;
    double time0;   // start time and step counter of the run (st_set_time without its synchronisation)
    int64_t icyc0;
};
// st_solver_create inside lent memory: no allocation and no synchronisation
int solver_create_in(const st_params *p, const SlotMem &mem, st_solver **out);

// The host-side state that enqueuing (or recording) steps advances; a
// recording that is not launched is undone by restoring it.
struct HostMark {
    int64_t icyc;
    int cur, last_stage, force_nblk, zv_nblk;
    bool dt_in_stage, fln_current;
};
inline HostMark host_mark(const st_solver *s) {
    return HostMark{s->icyc, s->cur, s->last_stage, s->force_nblk, s->zv_nblk, s->dt_in_stage, s->fln_current};
}
inline bool same_state(const HostMark &a, const HostMark &b, bool with_step) {
    return (!with_step || a.icyc == b.icyc) && a.cur == b.cur && a.last_stage == b.last_stage &&
           a.force_nblk == b.force_nblk && a.zv_nblk == b.zv_nblk && a.dt_in_stage == b.dt_in_stage &&
           a.fln_current == b.fln_current;
}
inline void host_restore(st_solver *s, const HostMark &m) {
    s->icyc = m.icyc;
    s->cur = m.cur;
    s->last_stage = m.last_stage;
    s->force_nblk = m.force_nblk;
    s->zv_nblk = m.zv_nblk;
    s->dt_in_stage = m.dt_in_stage;
    s->fln_current = m.fln_current;
}

// parameter handling shared by the entry points (st_capi.hip)
int set_error(int code, const std::string &msg);  // message for st_last_error; returns code
int validate_params(const st_params *p);
void params_to_kpar(const st_params &p, KPar &k);
// solver: the solver's pitched layout; false: the reference layout (rows of
// nx + 2 ng values), as caller-owned device arrays hold it
Geo params_to_geo(const st_params &p, bool solver = true);

// CFL reduction over caller-given arrays: k_dt over nx x ny x kplanes
// workgroup columns (each walks k with stride kplanes) writing one block
// maximum each into partials, then the two-level maximum into *dt_out.
// ticket must point at an int the kernels may use; k_dt clears it.
hipError_t launch_dt_raw(hipStream_t stream, bool mixed, const Geo &g, const KPar &P, const double *w, void *wa,
                         const double *dx1, const double *dy1, const double *dz1, const double *dxs,
                         const double *dys, const double *dzs, double *partials, int kplanes, double *stage2,
                         int *ticket, double *dt_out);
// workgroups of one k plane of k_dt; k_dt_finish's second-level slots
int dt_blocks_per_plane(const Geo &g);
constexpr int kDtStage2 = 128;

// launchers (st_kernels.hip); all enqueue on s->stream and return a hipError_t
hipError_t launch_fill_periodic(st_solver *s, double *arr, int nv);
hipError_t launch_fill_periodic_aux(st_solver *s, int var);
hipError_t launch_extrapolate_aux(st_solver *s, int var);
hipError_t launch_aux(st_solver *s, int i0, int i1, int j0, int j1, int k0, int k1, bool image,
                      bool dt = false);
hipError_t launch_wall_ghosts(st_solver *s);
// physical ghosts of every non-periodic face in the order xmin..zmax
// (update_ghost without the periodic swap)
hipError_t launch_physical_ghosts(st_solver *s);
// characteristic boundary treatment added to fl (bc_nr)
hipError_t launch_bc_nr(st_solver *s, void *fl);
// recyc_exchange + bcrecyc of face xmin
hipError_t launch_recycling(st_solver *s);
// the random inflow plane: redraw rows 2..jbl (after a z sweep), and the
// generator window reset to its seeded state
hipError_t launch_inflow_draws(st_solver *s);
hipError_t launch_rng_reset(st_solver *s);
// procedure-by-procedure stage: face fluxes of faces f0..f1 into fhat; the
// divergence of fhat into fl over cells c0..c1; fln = -coef fl, fl = 0
// (init) or fln -= coef fl; w += fln; force_rhs and force_var on fln / w;
// the station planes into wrecyc
hipError_t launch_face_range(st_solver *s, int dir, int f0, int f1);
hipError_t launch_face_update(st_solver *s, int dir, void *fl, int c0, int c1);
hipError_t launch_flux_step(st_solver *s, void *fl, double coef, bool init);
hipError_t launch_update_field(st_solver *s);
hipError_t launch_force_rhs_piece(st_solver *s);
hipError_t launch_force_var_piece(st_solver *s);
hipError_t launch_recyc_exchange(st_solver *s);
// fix = false leaves the WENO fix-up of the sweep to launch_weno_fix
hipError_t launch_euler_x(st_solver *s, void *fl, bool init, bool fix = true);
hipError_t launch_euler_y(st_solver *s, void *fl, bool fix = true);
hipError_t launch_euler_xy(st_solver *s, void *fl, bool fix = true);
hipError_t launch_euler_z(st_solver *s, void *fl, bool fix = true);
hipError_t launch_weno_fix(st_solver *s, void *fl, int dir, int bit, int anyidx);
// z march + viscous divergence (+ force_rhs partials); fix = false leaves
// its WENO fix-up to launch_euler_zv_fix, which must follow it
hipError_t launch_euler_zv(st_solver *s, void *fl, const void *fl_prev, int stage, bool reduce, bool ord2,
                           bool fix = true);
hipError_t launch_euler_zv_fix(st_solver *s, void *fl, int stage, bool reduce);
hipError_t launch_visc(st_solver *s, void *fl, int mode, bool ghost_images);
hipError_t launch_visc_cons(st_solver *s, void *fl, int mode, bool ghost_images);
hipError_t launch_sensor(st_solver *s, bool ghost_images);
// visflx mode 1, or mode 7 (second order, conservative viscous form) with ord2
hipError_t launch_visc_div(st_solver *s, void *fl, const void *fl_prev, int stage, bool reduce, bool ord2);
hipError_t launch_force_finish(st_solver *s, int stage);
hipError_t launch_update(st_solver *s, const void *fl_prev, const void *fl, int stage);
hipError_t launch_force_var(st_solver *s);
hipError_t launch_dt(st_solver *s);
hipError_t launch_step_end(st_solver *s);
hipError_t launch_residual(st_solver *s);
hipError_t launch_count_weno(st_solver *s);
hipError_t launch_stats(st_solver *s, int itav);
hipError_t launch_sensor_refine(st_solver *s, bool direct, bool ghost_images);
hipError_t launch_zero(st_solver *s, double *a, size_t n);
hipError_t launch_init_temperature(st_solver *s);
hipError_t launch_zero_rhs(st_solver *s, void *a);

}  // namespace st
