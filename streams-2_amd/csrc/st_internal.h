// st_internal.h -- internal types and device helpers of libstreams_mi355x.
//
// Device layout: every ghosted array keeps the reference's Fortran order,
// one block per variable (x fastest, lower bound 1-ng), with its rows padded
// to a whole number of 128-byte lines of w_aux values and its first point
// placed so that i = 1 of every row starts a line: a wavefront's 64
// consecutive lanes read 64 consecutive values along x in whole lines.
// (Rows of the reference's nx + 2 ng values start anywhere in a line; a
// streaming pass over them ran 8 % slower.)  Host arrays keep the reference
// layout; the copies are pitched.  The right-hand side fl is
// double-buffered: F[cur] holds the RHS of the stage being built, F[cur^1]
// the RHS of the previous stage, which replaces the reference's fln register
// (init_flux/update_flux, kernels_cpu.F90:29-47,3400-3418).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/streams_mi355x.h"

namespace st {

// ---------------------------------------------------------------- geometry
struct Geo {
    int nx, ny, nz, ng;
    int sx, sy, sz;      // row pitch (>= nx + 2 ng), ghosted y and z extents
    int sxy;             // sx*sy
    long long nvol;      // points of one ghosted variable, padding included (offsets across variables are 64-bit)
    long long nint;      // points of one interior variable
    int rx;              // nx + 2 ng: the ghosted x extent, the row length of the reference layout
    long long nref;      // rx*sy*sz: points of one ghosted variable in the reference layout (host arrays)
};

// Row pitch of the solver layout: a multiple of one 128-byte line of w_aux
// values, 16 in FP64 and 32 in the mixed form (whose w rows then span whole
// pairs of lines).
constexpr int kRowAlign = 32;  // the larger of the two
__host__ __device__ constexpr int row_align(int precision) { return precision == ST_PREC_MIXED ? 32 : 16; }
// Bytes from an allocation's (256-byte aligned) start to the first point of
// a ghosted array of values of `bytes` each: i = 1 then falls on offset 128.
__host__ __device__ constexpr size_t lead_bytes(int ng, size_t bytes) { return 128 - (size_t)ng * bytes; }

// A pass over a box of ghosted points whose lanes walk x (k_aux) starts its
// rows at the last x position at or below i0 that opens a line (i = 1 mod
// 32) and gives each row a multiple of 64 lanes: every wavefront then reads
// and writes whole lines of one row, and the lanes outside [i0, i0 + ni)
// idle.  Boxes narrower than a wavefront (the x ghost slabs) pack their
// rows instead.
__host__ __device__ __forceinline__ int box_row_start(int i0, int ni) {
    return ni >= 64 ? i0 - (((i0 - 1) % 32) + 32) % 32 : i0;
}
__host__ __device__ __forceinline__ int box_row_lanes(int i0, int ni) {
    return ni >= 64 ? (i0 + ni - box_row_start(i0, ni) + 63) / 64 * 64 : ni;
}

// The pass over the whole ghosted box instead walks the pitched array
// linearly, lane L + aux_lead on point L, so that its wavefronts read w in
// whole lines from the line holding the first point; the lanes on the row
// padding write zeros to w_aux, so that its lines are written whole too.
__host__ __device__ __forceinline__ bool aux_whole_box(const Geo &g, int i0, int ni, int j0, int nj, int k0, int nk) {
    return i0 == 1 - g.ng && ni == g.rx && j0 == 1 - g.ng && nj == g.sy && k0 == 1 - g.ng && nk == g.sz;
}
__host__ __device__ __forceinline__ int aux_lead(const Geo &g) { return (int)(lead_bytes(g.ng, 8) / 8); }
__host__ __device__ __forceinline__ long long aux_lanes(const Geo &g) { return g.nvol + aux_lead(g); }

__host__ __device__ __forceinline__ int gidx(const Geo &g, int i, int j, int k) {
    return (i + g.ng - 1) + g.sx * ((j + g.ng - 1) + g.sy * (k + g.ng - 1));
}
__host__ __device__ __forceinline__ int iidx(const Geo &g, int i, int j, int k) {
    return (i - 1) + g.nx * ((j - 1) + g.ny * (k - 1));
}

// ------------------------------------------------------- kernel parameters
// The scalars a kernel needs, passed by value.
struct KPar {
    int lmax_base;         // ep_order/2
    int visc_lmax;         // visc_order/2
    int weno_scheme, weno_version, nkeep;
    int flux_splitting;    // WENO faces: 0 Roe characteristic, 1 Rusanov (conserved components)
    int calorically_perfect, indx_cp_l, indx_cp_r;
    int visc_model;
    int fzf[6];            // force_zero_flux
    int periodic[3];
    int wall_stag[2];      // ymin / ymax staggered isothermal wall written by k_update
    int bctag[6], nrtag[6];  // physical tag and effective characteristic tag of each face
    int eul_lo[3], eul_hi[3];
    int ndim;
    double rgas0, prandtl, t0, rho0, u0, l0, mu0;
    double p_ref;          // rho0*rgas0*t0: the mixed form stores p - p_ref
    double winf[5], winf_ps[5], xshock_imp, shock_angle, tanhfacs;  // free-stream / shock boundary data
    double aoa, ptot0, ttot0;  // total conditions of the subsonic free-stream boundary (tag 3)
    double sutherland_s, t_ref_dim, vtexp;
    double sensor_threshold, t_wall, t_bulk_target, volchan, cfl, tol_iter_nr;
    double cp[ST_MAX_CP_COEFF], cv[ST_MAX_CP_COEFF];
    double cd1[4];         // first-derivative weights of ep_order (l = 1..4)
    double vd1[4];         // first-derivative weights of visc_order
    double vd2[5];         // second-derivative weights of visc_order (l = 0..4)
    double cd1tab[4][4];   // coeff_deriv1(l, lmax) for lmax = 1..4 (boundary orders)
};

// Device-resident scalars of a run.
struct DevScalars {
    double dt, time, dpdx, rhobulk, ubulk, tbulk;
    double b1, b2;                   // bulk5g(1:2)/volchan of the last force_rhs
    double residual, rhomin, rhomax, tmin, tmax, pmin, pmax;
    double bulkt;
    unsigned long long dtinv_bits;   // running max of the CFL spectral radii
    long long icyc;
    double cnt[3];                   // WENO face counts
};

// ------------------------------------------------------------- EOS helpers
// x**n for integer n by binary exponentiation (the reference's integer power)
__device__ __forceinline__ double ipow(double x, int n) {
    bool inv = n < 0;
    unsigned un = inv ? (unsigned)(-n) : (unsigned)n;
    double r = 1.0, b = x;
    while (un) {
        if (un & 1u) r *= b;
        b *= b;
        un >>= 1;
    }
    return inv ? 1.0 / r : r;
}

__device__ __forceinline__ double cpc(const KPar &P, int l) { return P.cp[l - P.indx_cp_l]; }
__device__ __forceinline__ double cvc(const KPar &P, int l) { return P.cv[l - P.indx_cp_l]; }

// get_gamloc_dev, kernels_cpu.F90:7096-7113
__device__ __forceinline__ double gamloc_of(const KPar &P, double tt) {
    double cploc;
    if (P.calorically_perfect == 1) {
        cploc = cpc(P, 0);
    } else {
        cploc = 0.0;
        for (int l = P.indx_cp_l; l <= P.indx_cp_r; l++) cploc = cploc + cpc(P, l) * ipow(tt / P.t0, l);
    }
    return cploc / (cploc - P.rgas0);
}

__device__ __forceinline__ double cploc_of(const KPar &P, double tt) {
    if (P.calorically_perfect == 1) return cpc(P, 0);
    double cploc = 0.0;
    for (int l = P.indx_cp_l; l <= P.indx_cp_r; l++) cploc = cploc + cpc(P, l) * ipow(tt / P.t0, l);
    return cploc;
}

// get_temperature_from_e_dev, kernels_cpu.F90:7116-7152
__device__ __forceinline__ double temperature_from_e(const KPar &P, double ee, double t_start) {
    const double t0 = P.t0;
    double ebar = ee - cvc(P, P.indx_cp_r + 1) * t0;
    if (P.calorically_perfect == 1) return t0 + ebar / cvc(P, 0);
    double t_old = t_start, tt = t_start;
    for (int iter = 1; iter <= 50; iter++) {
        double den = 0.0, num = 0.0;
        for (int l = P.indx_cp_l; l <= P.indx_cp_r; l++) {
            double t_pow = ipow(t_old / t0, l);
            if (l == -1) {
                den = den + cvc(P, l) * t_pow;
                num = num + cvc(P, l) * log(t_old / t0);
            } else {
                double t_powp = (t_old / t0) * t_pow;
                den = den + cvc(P, l) * t_pow;
                num = num + cvc(P, l) * (t_powp - 1.0) / (l + 1.0);
            }
        }
        num = num * t0;
        tt = t_old + (ebar - num) / den;
        if (fabs(tt - t_old) < P.tol_iter_nr) break;
        t_old = tt;
    }
    return tt;
}

// get_e_from_temperature_dev, kernels_cpu.F90:7155-7176
__device__ __forceinline__ double e_from_temperature(const KPar &P, double tt) {
    const double t0 = P.t0;
    double ee = cvc(P, P.indx_cp_r + 1);
    if (P.calorically_perfect == 1) {
        ee = ee + cvc(P, 0) * (tt / t0 - 1.0);
    } else {
        for (int l = P.indx_cp_l; l <= P.indx_cp_r; l++) {
            if (l == -1) ee = ee + cvc(P, l) * log(tt / t0);
            else ee = ee + cvc(P, l) / (l + 1.0) * (ipow(tt / t0, l + 1) - 1.0);
        }
    }
    return ee * t0;
}

// viscosity law of eval_aux_subroutine, kernels_cpu.F90:6998-7005
__device__ __forceinline__ double viscosity_of(const KPar &P, double tt) {
    if (P.visc_model == 1) return P.mu0 * pow(tt / P.t0, P.vtexp);
    if (P.visc_model == 2)
        return P.mu0 * pow(tt / P.t0, 1.5) * (1.0 + P.sutherland_s / P.t_ref_dim) /
               (tt / P.t0 + P.sutherland_s / P.t_ref_dim);
    return 0.0;
}

}  // namespace st
