// st_kernels.hip -- gfx950 kernels of one RK stage and their launchers.
//
// Stage pipeline of the fused path (st_rk_sync), one kernel per line:
//   aux        w -> w_aux(1:7) over the ghosted box, periodic ghosts read
//              from their images (replaces bcswap + compute_aux)
//   euler_x    x faces + divergence, writes F[cur] (replaces init_flux's
//              fl=0 and the first accumulation)
//   visc       Laplacian viscous terms, w_aux(9,10) and the ghosts of
//              w_aux(10) (bcextr_var + bcswap_var)
//   euler_y    y faces, marching along y, accumulates into F[cur]
//   euler_z    z faces, marching along z
//   sensor     stage 3 only: Ducros sensor and its periodic ghosts
//   visc_div   grad(div/3) terms + block partial sums of force_rhs
//   finish     bulk sums -> dpdx, rho_b, u_b, T_b (one workgroup)
//   update     fln = -rho_k dt F[prev] - gamma_k dt F[cur] - forcing,
//              w += fln, isothermal-wall ghosts
// Reductions use per-block partials summed in a fixed order, so a run is
// bitwise reproducible; the CFL maximum uses an atomic max on the bit
// pattern of non-negative doubles, which is order independent.
#include "st_flux.h"
#include "st_solver.h"

namespace st {

constexpr int H = 3;           // stencil half width of the flux windows
constexpr int RED_THREADS = 256;

// -------------------------------------------------------------- utilities
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
    return v;
}

// Block sum of NV values in a fixed order; result valid in thread 0.
__device__ __forceinline__ int flat_block() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }
__device__ __forceinline__ int flat_thread() { return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z); }

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*lds)[NV]) {
    const int ft = flat_thread();
    const int lane = ft & 63;
    const int wid = ft >> 6;
    const int nw = (blockDim.x * blockDim.y * blockDim.z + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) lds[wid][q] = v[q];
    }
    __syncthreads();
    if (ft == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double a = lds[0][q];
            for (int w = 1; w < nw; w++) a += lds[w][q];
            v[q] = a;
        }
    }
}

// ------------------------------------------------------------ ghost fills
// Periodic copy of bcswap (base_cpu.F90:410-450, 669-745 with the local
// neighbour shortcut :1382-1417) for nv ghosted variables: faces of the
// periodic directions only, edges and corners untouched.
template <typename T>
__global__ void k_fill_periodic(Geo g, KPar P, T *__restrict__ a, int nv) {
    const int ng = g.ng;
    const int nxf = P.periodic[0] ? 2 * ng * g.ny * g.nz : 0;
    const int nyf = P.periodic[1] ? 2 * ng * g.nx * g.nz : 0;
    const int nzf = (P.periodic[2] && P.ndim == 3) ? 2 * ng * g.nx * g.ny : 0;
    const int tot = nxf + nyf + nzf;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot * nv; t += gridDim.x * blockDim.x) {
        const int m = t / tot;
        int r = t - m * tot;
        int dst, src;
        if (r < nxf) {
            const int l = r % (2 * ng);
            const int jk = r / (2 * ng);
            const int j = jk % g.ny + 1, k = jk / g.ny + 1;
            if (l < ng) { dst = gidx(g, 1 - ng + l, j, k); src = gidx(g, g.nx - ng + 1 + l, j, k); }
            else { dst = gidx(g, g.nx + 1 + (l - ng), j, k); src = gidx(g, 1 + (l - ng), j, k); }
        } else if ((r -= nxf) < nyf) {
            const int i = r % g.nx + 1;
            const int lk = r / g.nx;
            const int l = lk % (2 * ng), k = lk / (2 * ng) + 1;
            if (l < ng) { dst = gidx(g, i, 1 - ng + l, k); src = gidx(g, i, g.ny - ng + 1 + l, k); }
            else { dst = gidx(g, i, g.ny + 1 + (l - ng), k); src = gidx(g, i, 1 + (l - ng), k); }
        } else {
            r -= nyf;
            const int i = r % g.nx + 1;
            const int jl = r / g.nx;
            const int j = jl % g.ny + 1, l = jl / g.ny;
            if (l < ng) { dst = gidx(g, i, j, 1 - ng + l); src = gidx(g, i, j, g.nz - ng + 1 + l); }
            else { dst = gidx(g, i, j, g.nz + 1 + (l - ng)); src = gidx(g, i, j, 1 + (l - ng)); }
        }
        T *b = a + (size_t)m * g.nvol;
        b[dst] = b[src];
    }
}

// bcextr_var_subroutine (kernels_cpu.F90:6199-6254): zero-gradient ghosts of
// one variable on all six faces.
template <typename T>
__global__ void k_extrapolate_var(Geo g, T *__restrict__ a) {
    const int ng = g.ng;
    const int nxf = 2 * ng * g.ny * g.nz, nyf = 2 * ng * g.nx * g.nz, nzf = 2 * ng * g.nx * g.ny;
    const int tot = nxf + nyf + nzf;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x) {
        int r = t, dst, src;
        if (r < nxf) {
            const int l = r % (2 * ng), jk = r / (2 * ng);
            const int j = jk % g.ny + 1, k = jk / g.ny + 1;
            if (l < ng) { dst = gidx(g, -l, j, k); src = gidx(g, 1, j, k); }
            else { dst = gidx(g, g.nx + 1 + (l - ng), j, k); src = gidx(g, g.nx, j, k); }
        } else if ((r -= nxf) < nyf) {
            const int i = r % g.nx + 1, lk = r / g.nx;
            const int l = lk % (2 * ng), k = lk / (2 * ng) + 1;
            if (l < ng) { dst = gidx(g, i, -l, k); src = gidx(g, i, 1, k); }
            else { dst = gidx(g, i, g.ny + 1 + (l - ng), k); src = gidx(g, i, g.ny, k); }
        } else {
            r -= nyf;
            const int i = r % g.nx + 1, jl = r / g.nx;
            const int j = jl % g.ny + 1, l = jl / g.ny;
            if (l < ng) { dst = gidx(g, i, j, -l); src = gidx(g, i, j, 1); }
            else { dst = gidx(g, i, j, g.nz + 1 + (l - ng)); src = gidx(g, i, j, g.nz); }
        }
        a[dst] = a[src];
    }
}

// bcwall_staggered_subroutine (kernels_cpu.F90:6612-6678) for one ghost
// point: the mirror of interior point js, written at jg.
template <typename R>
__device__ __forceinline__ void wall_ghost(const Geo &g, const KPar &P, double *__restrict__ w,
                                           const R *__restrict__ wa, int i, int js, int jg, int k) {
    const int s = gidx(g, i, js, k), d = gidx(g, i, jg, k);
    const size_t nv = g.nvol;
    double rho = w[s];
    const double uu = w[nv + s] / rho;
    const double vv = w[2 * nv + s] / rho;
    const double ww = w[3 * nv + s] / rho;
    const double rhoe = w[4 * nv + s];
    const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
    double ee = rhoe / rho - qq;
    double tt = temperature_from_e(P, ee, double(wa[5 * nv + s]));
    const double pp = rho * tt * P.rgas0;
    tt = 2.0 * P.t_wall - tt;
    ee = e_from_temperature(P, tt);
    rho = pp / tt / P.rgas0;
    w[d] = rho;
    w[nv + d] = -rho * uu;
    w[2 * nv + d] = -rho * vv;
    w[3 * nv + d] = -rho * ww;
    w[4 * nv + d] = rho * (ee + qq);
}

template <typename R>
__global__ void k_wall_ghosts(Geo g, KPar P, double *__restrict__ w, const R *__restrict__ wa) {
    const int tot = 2 * g.ng * g.nx * g.nz;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x) {
        const int i = t % g.nx + 1;
        const int r = t / g.nx;
        const int l = r % (2 * g.ng);
        const int k = r / (2 * g.ng) + 1;
        if (l < g.ng) {
            if (P.wall_stag[0]) wall_ghost(g, P, w, wa, i, l + 1, -l, k);
        } else {
            const int ll = l - g.ng + 1;
            if (P.wall_stag[1]) wall_ghost(g, P, w, wa, i, g.ny + 1 - ll, g.ny + ll, k);
        }
    }
}

// ------------------------------------------------- physical boundary faces
// Ghost point l (1..ng) and boundary point of face ilat (1..6 = xmin..zmax)
// at tangential indices (a, b).
__device__ __forceinline__ int face_ghost(const Geo &g, int ilat, int a, int b, int l) {
    switch (ilat) {
    case 1: return gidx(g, 1 - l, a, b);
    case 2: return gidx(g, g.nx + l, a, b);
    case 3: return gidx(g, a, 1 - l, b);
    case 4: return gidx(g, a, g.ny + l, b);
    case 5: return gidx(g, a, b, 1 - l);
    default: return gidx(g, a, b, g.nz + l);
    }
}
__device__ __forceinline__ int face_boundary(const Geo &g, int ilat, int a, int b) {
    switch (ilat) {
    case 1: return gidx(g, 1, a, b);
    case 2: return gidx(g, g.nx, a, b);
    case 3: return gidx(g, a, 1, b);
    case 4: return gidx(g, a, g.ny, b);
    case 5: return gidx(g, a, b, 1);
    default: return gidx(g, a, b, g.nz);
    }
}

// One face of update_ghost (singleideal_cpu.F90:1064-1138), one thread per
// tangential point: free stream (bcfree_subroutine, kernels_cpu.F90:
// 5989-6056), extrapolation (bcextr_subroutine, :6323-6390), symmetry on
// ymin (bcsym_subroutine, :6394-6418), staggered or node isothermal wall
// on ymin (bcwall_staggered_subroutine :6612-6678, bcwall_subroutine
// :6478-6525), oblique shock on ymax (bcshock_subroutine, :6159-6196).
template <typename R>
__global__ void k_bc_face(Geo g, KPar P, double *__restrict__ w, const R *__restrict__ wa,
                          const double *__restrict__ xc, const double *__restrict__ yc, int ilat, int tag,
                          int staggered, const double *__restrict__ wmean) {
    const int na = ilat <= 2 ? g.ny : g.nx;
    const int nb = ilat <= 4 ? g.nz : g.ny;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na * nb) return;
    const int a = t % na + 1, b = t / na + 1;
    const size_t nv = g.nvol;
    const int ng = g.ng;
    if (tag == 1) {
        for (int l = 1; l <= ng; l++) {
            const int d = face_ghost(g, ilat, a, b, l);
            for (int m = 0; m < 5; m++) w[m * nv + d] = P.winf[m];
        }
    } else if (tag == 3 && (ilat == 1 || ilat == 4)) {
        // bcfree_sub_subroutine (kernels_cpu.F90:6060-6155): the boundary
        // cell's pressure with the total conditions gives M, T and |u|
        const int c = face_boundary(g, ilat, a, b);
        double rho = w[c];
        const double ri = 1.0 / rho;
        double uu = w[nv + c] * ri, vv = w[2 * nv + c] * ri, ww = w[3 * nv + c] * ri;
        const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
        double ee = w[4 * nv + c] / rho - qq;
        double tt = temperature_from_e(P, ee, double(wa[5 * nv + c]));
        const double pp = rho * tt * P.rgas0;
        const double gamloc = gamloc_of(P, tt);
        const double del = 0.5 * (gamloc - 1.0);
        const double rmfac = pow(pp / P.ptot0, -(gamloc - 1.0) / gamloc);
        const double rml = sqrt((rmfac - 1.0) / del);
        tt = P.ttot0 / rmfac;
        const double vel_mod = rml * sqrt(gamloc * P.rgas0 * tt);
        uu = vel_mod * cos(P.aoa);
        vv = vel_mod * sin(P.aoa);
        ww = 0.0;
        rho = pp / tt / P.rgas0;
        const double rhou = rho * uu, rhov = rho * vv, rhow = rho * ww;
        ee = e_from_temperature(P, tt);
        const double rhoe = rho * ee + 0.5 * (rhou * rhou + rhov * rhov + rhow * rhow) / rho;
        for (int l = 1; l <= ng; l++) {
            const int d = face_ghost(g, ilat, a, b, l);
            w[d] = rho;
            w[nv + d] = rhou;
            w[2 * nv + d] = rhov;
            w[3 * nv + d] = rhow;
            w[4 * nv + d] = rhoe;
        }
    } else if (tag == 4 && ilat != 3) {
        // bcextr_sub_subroutine (kernels_cpu.F90:5064-5164): rho and momentum
        // of the boundary cell at the reference pressure p0
        const int c = face_boundary(g, ilat, a, b);
        const double rho = w[c], rhou = w[nv + c], rhov = w[2 * nv + c], rhow = w[3 * nv + c];
        const double tt = P.p_ref / rho / P.rgas0;
        const double ee = e_from_temperature(P, tt);
        const double rhoe = rho * ee + 0.5 * (rhou * rhou + rhov * rhov + rhow * rhow) / rho;
        for (int l = 1; l <= ng; l++) {
            const int d = face_ghost(g, ilat, a, b, l);
            w[d] = rho;
            w[nv + d] = rhou;
            w[2 * nv + d] = rhov;
            w[3 * nv + d] = rhow;
            w[4 * nv + d] = rhoe;
        }
    } else if (tag == 9 && ilat == 1) {
        // bclam_subroutine (kernels_cpu.F90:5951-5985): the mean planes
        const int j = a, k = b;
        for (int l = 1; l <= ng; l++) {
            const double *wm = wmean + (ng - l) + (size_t)ng * (j - 1);
            const size_t st = (size_t)ng * g.ny;
            const double rho = wm[0], rhou = wm[st], rhov = wm[2 * st], rhow = wm[3 * st];
            const double tt = P.p_ref / rho / P.rgas0;
            const double ee = e_from_temperature(P, tt);
            const int d = gidx(g, 1 - l, j, k);
            w[d] = rho;
            w[nv + d] = rhou;
            w[2 * nv + d] = rhov;
            w[3 * nv + d] = rhow;
            w[4 * nv + d] = rho * ee + 0.5 * (rhou * rhou + rhov * rhov + rhow * rhow) / rho;
        }
    } else if (tag == 2) {
        const int s0 = face_boundary(g, ilat, a, b);
        for (int l = 1; l <= ng; l++) {
            const int d = face_ghost(g, ilat, a, b, l);
            for (int m = 0; m < 5; m++) w[m * nv + d] = w[m * nv + s0];
        }
    } else if (tag == 5 && ilat == 3) {
        for (int l = 1; l <= ng; l++) {
            const int d = gidx(g, a, 1 - l, b), s0 = gidx(g, a, 1 + l, b);
            for (int m = 0; m < 5; m++) w[m * nv + d] = w[m * nv + s0];
        }
    } else if (tag == 6 && staggered) {
        for (int l = 1; l <= ng; l++) {
            if (ilat == 3) wall_ghost(g, P, w, wa, a, l, 1 - l, b);
            else if (ilat == 4) wall_ghost(g, P, w, wa, a, g.ny + 1 - l, g.ny + l, b);
        }
    } else if (tag == 6 && ilat == 3) {
        const int c1 = gidx(g, a, 1, b);
        w[nv + c1] = 0.0;
        w[2 * nv + c1] = 0.0;
        w[3 * nv + c1] = 0.0;
        double ee = e_from_temperature(P, P.t_wall);
        w[4 * nv + c1] = w[c1] * ee;
        for (int l = 1; l <= ng; l++) {
            const int s0 = gidx(g, a, 1 + l, b), d = gidx(g, a, 1 - l, b);
            double rho = w[s0];
            const double uu = w[nv + s0] / rho, vv = w[2 * nv + s0] / rho, ww = w[3 * nv + s0] / rho;
            const double rhoe = w[4 * nv + s0];
            const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
            ee = rhoe / rho - qq;
            double tt = temperature_from_e(P, ee, double(wa[5 * nv + s0]));
            const double pp = rho * tt * P.rgas0;
            tt = 2.0 * P.t_wall - tt;
            ee = e_from_temperature(P, tt);
            rho = pp / tt / P.rgas0;
            w[d] = rho;
            w[nv + d] = -rho * uu;
            w[2 * nv + d] = -rho * vv;
            w[3 * nv + d] = -rho * ww;
            w[4 * nv + d] = rho * (ee + qq);
        }
    } else if (tag == 7 && ilat == 4) {
        const double tanhlen = 8.0 * P.tanhfacs;
        const int i = a, k = b;
        const int cn = gidx(g, i, g.ny, k);
        for (int l = 0; l <= ng; l++) {
            const double xsh = P.xshock_imp - yc[g.ny + l + ng - 1] / tan(P.shock_angle);
            const double xx = xc[i + ng - 1] - xsh;
            const int d = gidx(g, i, g.ny + l, k);
            if (fabs(xx) > tanhlen) {
                for (int m = 0; m < 5; m++) w[m * nv + d] = w[m * nv + cn];
            } else {
                const double tanhf = 0.5 * (1.0 + tanh(xx / P.tanhfacs));
                for (int m = 0; m < 5; m++) {
                    const double dwinf = P.winf_ps[m] - P.winf[m];
                    w[m * nv + d] = P.winf[m] + dwinf * tanhf;
                }
            }
        }
    }
}

// bc_nr_lat_x/y/z_kernel (kernels_cpu.F90:5169-5283, 5420-5540, 5688-5807),
// one thread per point of the boundary plane, FP64 throughout: the
// one-sided derivative of w along the normal in characteristic variables
// at the boundary state, scaled by the (clipped) eigenvalues, back through
// the right eigenvectors into fl.
template <typename R, int DIR>
__global__ void k_bc_nr(Geo g, KPar P, const double *__restrict__ w, const R *__restrict__ wa, R *__restrict__ fl,
                        const double *__restrict__ metric, int start, int nr_type) {
    const int na = DIR == 0 ? g.ny : g.nx;
    const int nb = DIR == 2 ? g.ny : g.nz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na * nb) return;
    const int a = t % na + 1, b = t / na + 1;
    const int n = DIR == 0 ? g.nx : (DIR == 1 ? g.ny : g.nz);
    const int ib = start ? 1 : n;
    const int i = DIR == 0 ? ib : a, j = DIR == 1 ? ib : (DIR == 0 ? a : b), k = DIR == 2 ? ib : b;
    const int sgn = start ? 1 : -1;
    const int st = DIR == 0 ? 1 : (DIR == 1 ? g.sx : g.sxy);
    const size_t nv = g.nvol;
    const int c = gidx(g, i, j, k);
    const double c_one[3] = {-1.5, 2.0, -0.5};
    double dw_dn[5], dw_dn_outer[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        const double *wm = w + m * nv;
        double d = 0.0;
        for (int l = 1; l <= 3; l++) d = d + sgn * c_one[l - 1] * wm[c + sgn * (l - 1) * st];
        dw_dn[m] = d;
        double w_target = 0.0;
        if (nr_type == 2) w_target = wm[c - sgn * st];
        else if (nr_type == 3) w_target = P.winf[m];
        dw_dn_outer[m] = sgn * (wm[c] - w_target);
    }
    const double rho = wa[c], uu = wa[nv + c], vv = wa[2 * nv + c], ww = wa[3 * nv + c];
    const double h = wa[4 * nv + c], tt = wa[5 * nv + c];
    const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
    const double gamloc = gamloc_of(P, tt);
    const double cc = gamloc * tt * P.rgas0;
    Roe s;
    s.c = sqrt(cc);
    s.ci = 1.0 / s.c;
    const double p_rho = tt * P.rgas0;
    const double p_e = rho * (gamloc - 1.0);
    const double etot = h - tt * P.rgas0;
    s.b3 = etot - rho * p_rho / p_e;
    s.b2 = p_e / (rho * cc);
    s.b1 = p_rho / cc - s.b2 * (etot - 2.0 * qq);
    s.h = h; s.uu = uu; s.vv = vv; s.ww = ww;
    double dwc[5], dwc_outer[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        double el[5];
        eig_left<DIR>(s, m, el);
        double x = 0.0, y = 0.0;
        for (int v = 0; v < 5; v++) {
            x = x + el[v] * dw_dn[v];
            y = y + el[v] * dw_dn_outer[v];
        }
        dwc[m] = x;
        dwc_outer[m] = y;
    }
    const double un = DIR == 0 ? uu : (DIR == 1 ? vv : ww);
    double ev[5] = {un - s.c, un, un, un, un + s.c};
    if (nr_type == 1)
        for (int l = 0; l < 5; l++) ev[l] = sgn * fmin(sgn * ev[l], 0.0);
    if (nr_type == 2 || nr_type == 3)
        for (int m = 0; m < 5; m++)
            if (sgn * ev[m] > 0.0) dwc[m] = dwc_outer[m];
    for (int m = 0; m < 5; m++) dwc[m] = ev[m] * dwc[m];
    if (nr_type == 6) {
        dwc[1] = 0.0;
        dwc[2] = 0.0;
        dwc[3] = 0.0;
        if (start) dwc[4] = dwc[0];
        else dwc[0] = dwc[4];
    }
    double df[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int mm = 0; mm < 5; mm++) {
        double er[5];
        eig_right<DIR>(s, mm, er);
        for (int m = 0; m < 5; m++) df[m] = df[m] + er[m] * dwc[mm];
    }
    const double d = metric[ib - 1];
    const int oi = iidx(g, i, j, k);
    for (int m = 0; m < 5; m++) fl[(size_t)m * g.nint + oi] = R(double(fl[(size_t)m * g.nint + oi]) + df[m] * d);
}

// ------------------------------------------------ recycling-rescaling inflow
// recyc_exchange for one block in x and z (singleideal_cpu.F90:935-999,
// recyc_exchange_subroutine_1/2/3, kernels_cpu.F90:5003-5060): the recycled
// plane i of the inflow is w(i_recyc+1-i) with the span shifted by nz/2.
__device__ __forceinline__ int recyc_k(int k, int nz) {
    const int h = (nz / 2) % nz;
    return k > h ? k - h : k + nz - h;
}

// wrecyc(i, j, k, m), i = 1..ng, of recyc_exchange (k_recyc_exchange)
__device__ __forceinline__ size_t wrc(const Geo &g, int i, int j, int k, int m) {
    return (size_t)(i - 1) + (size_t)g.ng * ((size_t)(j - 1) + (size_t)g.ny * ((size_t)(k - 1) + (size_t)g.nz * m));
}

// bcrecyc_subroutine_1 (kernels_cpu.F90:5810-5828): spanwise sums of the
// recycled planes, one thread per (i, j), k summed in order.
__global__ void k_recyc_avg(Geo g, const double *__restrict__ wrecyc, double *__restrict__ av) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ng * g.ny * 5) return;
    const int i = t % g.ng + 1, j = (t / g.ng) % g.ny + 1, m = t / (g.ng * g.ny);
    // one (i, j, m) per thread; the sum runs over k in order, its loads
    // issued eight at a time ahead of the additions
    double a = 0.0;
    int k = 1;
    for (; k + 7 <= g.nz; k += 8) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = wrecyc[wrc(g, i, j, k + q, m)];
#pragma unroll
        for (int q = 0; q < 8; q++) a = a + v[q];
    }
    for (; k <= g.nz; k++) a = a + wrecyc[wrc(g, i, j, k, m)];
    av[(i - 1) + g.ng * ((j - 1) + g.ny * m)] = a;
}

// bcrecyc_subroutine_2/3 (kernels_cpu.F90:5832-5947): one thread per
// (j, k) of the inflow; the Favre fluctuations of the recycled planes are
// formed where they are interpolated.  wmean(1-ng:0, j, m) sits at
// [(ng-i) + ng*((j-1) + ny*m)] for plane 1-i.  With u0_02 != 0 (rand_type
// != 0) each velocity fluctuation gains u0_02*(r - 0.5)*min(eta, 1) from
// the random plane r(j, k, c) (Fortran order (1:ny, 1:nz, 1:3)).
__global__ void k_recyc_inflow(Geo g, KPar P, double *__restrict__ w, const double *__restrict__ rc,
                               const int *__restrict__ maps, const double *__restrict__ wrecyc, double betarecyc,
                               const double *__restrict__ rplane, double u0_02) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ny * g.nz) return;
    const int j = t % g.ny + 1, k = t / g.ny + 1;
    const int ng = g.ng, ny = g.ny, nz = g.nz, ny2 = ny + 2 * ng;
    const size_t nv = g.nvol;
    const double *wmean = rc, *weta_a = rc + (size_t)ng * ny * 4;
    const double *ypi = weta_a + ny, *eti = ypi + ny2, *ypr = eti + ny2, *etr = ypr + ny2;
    const double *av = etr + ny2;
    const double eta = eti[j + ng - 1];
    const double weta = weta_a[j - 1];
    const double weta1 = 1.0 - weta;
    const double etamin = fmin(eta, 1.0);
    const double bdamp = 0.5 * (1.0 - tanh(4.0 * (eti[j + ng - 1] - 2.0)));
    const int j_inn = maps[j - 1], j_out = maps[ny + j - 1];
    const bool zero = j == 1 || j_inn >= ny || j_out >= ny;
    double disty_inn = 0.0, disty_out = 0.0;
    if (!zero) {
        disty_inn = (ypi[j + ng - 1] - ypr[j_inn + ng - 1]) / (ypr[j_inn + 1 + ng - 1] - ypr[j_inn + ng - 1]);
        disty_out = (eta - etr[j_out + ng - 1]) / (etr[j_out + 1 + ng - 1] - etr[j_out + ng - 1]);
    }
    const double p0 = P.rho0 * P.rgas0 * P.t0;
    for (int i = 1; i <= ng; i++) {
        const double rhomean = wmean[(ng - i) + ng * ((j - 1) + ny * 0)];
        const double uumean = wmean[(ng - i) + ng * ((j - 1) + ny * 1)] / rhomean;
        const double vvmean = wmean[(ng - i) + ng * ((j - 1) + ny * 2)] / rhomean;
        const double wwmean = wmean[(ng - i) + ng * ((j - 1) + ny * 3)] / rhomean;
        double rhofluc = 0.0, ufluc = 0.0, vfluc = 0.0, wfluc = 0.0;
        if (!zero) {
            // fluctuation of variable m (0 rho, 1..3 velocity) of plane i at row jx
            auto fluc = [&](int jx, int m) {
                const double *a = av + (i - 1) + ng * (jx - 1);
                const double rho = wrecyc[wrc(g, i, jx, k, 0)];
                if (m == 0) return rho - a[0] / nz;
                return wrecyc[wrc(g, i, jx, k, m)] / rho - a[(size_t)ng * ny * m] / a[0];
            };
            double f_in[4], f_out[4];
            for (int m = 0; m < 4; m++) {
                f_in[m] = fluc(j_inn, m) * (1.0 - disty_inn) + fluc(j_inn + 1, m) * disty_inn;
                f_out[m] = fluc(j_out, m) * (1.0 - disty_out) + fluc(j_out + 1, m) * disty_out;
            }
            rhofluc = f_in[0] * weta1 + f_out[0] * weta;
            ufluc = f_in[1] * weta1 + f_out[1] * weta;
            vfluc = f_in[2] * weta1 + f_out[2] * weta;
            wfluc = f_in[3] * weta1 + f_out[3] * weta;
            rhofluc = rhofluc * bdamp;
            ufluc = ufluc * bdamp * betarecyc;
            vfluc = vfluc * bdamp * betarecyc;
            wfluc = wfluc * bdamp * betarecyc;
            if (u0_02 != 0.0) {
                const size_t pl = (size_t)(j - 1) + (size_t)ny * (k - 1), pn = (size_t)ny * nz;
                ufluc = ufluc + u0_02 * (rplane[pl] - 0.5) * etamin;
                vfluc = vfluc + u0_02 * (rplane[pl + pn] - 0.5) * etamin;
                wfluc = wfluc + u0_02 * (rplane[pl + 2 * pn] - 0.5) * etamin;
            }
        }
        rhofluc = fmax(-0.75 * rhomean, rhofluc);
        rhofluc = fmin(4.0 * rhomean, rhofluc);
        const double rho = rhomean + rhofluc;
        const double tt = p0 / rho / P.rgas0;
        const double uu = uumean + ufluc, vv = vvmean + vfluc, ww = wwmean + wfluc;
        const double rhou = rho * uu, rhov = rho * vv, rhow = rho * ww;
        const double ee = e_from_temperature(P, tt);
        const int d = gidx(g, 1 - i, j, k);
        w[d] = rho;
        w[nv + d] = rhou;
        w[2 * nv + d] = rhov;
        w[3 * nv + d] = rhow;
        w[4 * nv + d] = rho * ee + 0.5 * (rhou * rhou + rhov * rhov + rhow * rhow) / rho;
    }
}

// ------------------------------------------------------ random inflow plane
// get_crandom_f(inflow_random_plane(2:jbl, 1:nz, 1:3)) after a z sweep
// (singleideal_cpu.F90:841): draw t of the sweep goes to row j = 2 +
// t/(3 nz), k = 1 + (t/3) % nz, component t % 3, as rand()/RAND_MAX.  The C
// library's rand() is an additive lagged-Fibonacci generator,
// r[n] = r[n-31] + r[n-3] (mod 2^32), output r[n] >> 1; draw t of the sweep
// is r[n0+t] = sum_q coef[q][t] win[q] (mod 2^32) over the 31 words
// win[q] = r[n0-31+q] before the sweep, so every draw is independent.  The
// window of the next sweep is written to the other buffer; the last
// workgroup to finish flips the buffer index and clears the ticket.
__global__ __launch_bounds__(256) void k_inflow_draws(Geo g, const uint32_t *__restrict__ coef, int n,
                                                      uint32_t *__restrict__ win2, int *__restrict__ ctl,
                                                      double *__restrict__ plane) {
    __shared__ uint32_t wl[31];
    __shared__ bool last;
    const int par = __hip_atomic_load(&ctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t *win = win2 + 31 * par;
    uint32_t *wout = win2 + 31 * (par ^ 1);
    if (threadIdx.x < 31) wl[threadIdx.x] = win[threadIdx.x];
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        uint32_t r = 0u;
        for (int q = 0; q < 31; q++) r += coef[(size_t)q * n + t] * wl[q];
        const int c = t % 3, k = (t / 3) % g.nz, j = 2 + t / (3 * g.nz);
        plane[(size_t)(j - 1) + (size_t)g.ny * ((size_t)k + (size_t)g.nz * c)] = (double)(r >> 1) / 2147483647.0;
        if (t >= n - 31) wout[t - (n - 31)] = r;
    }
    if (t < 31 && n - 31 + t < 0) wout[t] = wl[n + t];  // fewer than 31 draws: the window slides
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(&ctl[1], 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (last && threadIdx.x == 0) {
        ctl[0] = par ^ 1;
        ctl[1] = 0;
    }
}

struct Win31 {
    uint32_t v[31];
};

// init_crandom_f of run(): the window of the freshly seeded generator
__global__ void k_rng_reset(Win31 seeded, uint32_t *__restrict__ win2, int *__restrict__ ctl) {
    const int t = threadIdx.x;
    if (t < 31) win2[t] = seeded.v[t];
    if (t == 0) {
        ctl[0] = 0;
        ctl[1] = 0;
    }
}

// -------------------------------------------------------------------- aux
// eval_aux_subroutine (kernels_cpu.F90:6960-7011) on the box
// [i0,i1]x[j0,j1]x[k0,k1]; with image=true a ghost point of a periodic
// direction reads w at its periodic image, which is the value bcswap
// would have copied there.
// With DT it also forms the nine spectral radii of compute_dt_subroutine
// (kernels_cpu.F90:6814-6880) at the interior points it covers and writes
// one block maximum per workgroup: the first stage of a step that recomputes
// dt evaluates the same primitives of the same w as compute_dt's own
// compute_aux(central=1) at the start of the step.
template <typename R, bool DT = false>
__device__ __forceinline__ void k_aux_body(const dim3 blockIdx, const dim3 gridDim, Geo g, KPar P, const double *__restrict__ w, R *__restrict__ wa,
                                             int i0, int ni, int j0, int nj, int k0, int nk, int image,
                                             const double *__restrict__ dx1, const double *__restrict__ dy1,
                                             const double *__restrict__ dz1, const double *__restrict__ dxs,
                                             const double *__restrict__ dys, const double *__restrict__ dzs,
                                             double *__restrict__ partials) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double mx = 0.0;
    // lane -> point: the whole ghosted box walks the pitched array linearly
    // from the line that holds its first point (aux_lanes), padding included;
    // other boxes walk their rows (box_row_lanes)
    int i, j, k;
    bool act, pad = false;
    if (aux_whole_box(g, i0, ni, j0, nj, k0, nk)) {
        const int L = t - aux_lead(g);
        act = L >= 0 && L < g.nvol;
        const int e = act ? L % g.sx : 0, row = act ? L / g.sx : 0;
        i = e + 1 - g.ng;
        j = row % g.sy + 1 - g.ng;
        k = row / g.sy + 1 - g.ng;
        pad = act && i > g.nx + g.ng;
        act = act && !pad;
    } else {
        const int ial = box_row_start(i0, ni), rl = box_row_lanes(i0, ni);
        i = ial + t % rl;
        const int jk = t / rl;
        j = j0 + jk % nj;
        k = k0 + jk / nj;
        act = t < rl * nj * nk && i >= i0 && i < i0 + ni;
    }
    if (act) {
        int is = i, js = j, ks = k;
        if (image) {
            if (P.periodic[0]) is = i < 1 ? i + g.nx : (i > g.nx ? i - g.nx : i);
            if (P.periodic[1]) js = j < 1 ? j + g.ny : (j > g.ny ? j - g.ny : j);
            if (P.periodic[2]) ks = k < 1 ? k + g.nz : (k > g.nz ? k - g.nz : k);
        }
        const int s = gidx(g, is, js, ks), d = gidx(g, i, j, k);
        const size_t nv = g.nvol;
        const double rho = w[s], rhou = w[nv + s], rhov = w[2 * nv + s], rhow = w[3 * nv + s], rhoe = w[4 * nv + s];
        const double ri = 1.0 / rho;
        const double uu = rhou * ri, vv = rhov * ri, ww = rhow * ri;
        const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
        const double ee = rhoe / rho - qq;
        const double tt = temperature_from_e(P, ee, double(wa[5 * nv + d]));
        const double pp = rho * tt * P.rgas0;
        const double mu = viscosity_of(P, tt);
        wa[d] = R(rho);
        wa[nv + d] = R(uu);
        wa[2 * nv + d] = R(vv);
        wa[3 * nv + d] = R(ww);
        wa[4 * nv + d] = R((rhoe + pp) / rho);
        wa[5 * nv + d] = R(tt);
        wa[6 * nv + d] = R(mu);
        if constexpr (sizeof(R) != sizeof(double)) wa[10 * nv + d] = R(pp - P.p_ref);
        if constexpr (DT) {
            if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz) {
                const double nu = ri * mu;
                const double kor = nu / P.prandtl;
                const double gamloc = gamloc_of(P, tt);
                const double cs = sqrt(gamloc * P.rgas0 * tt);
                const double ax = dx1[i - 1], ay = dy1[j - 1], az = dz1[k - 1];
                const double sx = dxs[i - 1], sy = dys[j - 1], sz = dzs[k - 1];
                mx = fmax(mx, (fabs(uu) + cs) * ax);
                mx = fmax(mx, (fabs(vv) + cs) * ay);
                mx = fmax(mx, (fabs(ww) + cs) * az);
                mx = fmax(mx, nu * sx);
                mx = fmax(mx, nu * sy);
                mx = fmax(mx, nu * sz);
                mx = fmax(mx, kor * sx);
                mx = fmax(mx, kor * sy);
                mx = fmax(mx, kor * sz);
            }
        }
    } else if (pad) {
        // the row padding after x = nx + ng: zeros, so that the lines this
        // pass writes are written whole (never read)
        const int d = gidx(g, i, j, k);
        const size_t nv = g.nvol;
#pragma unroll
        for (int m = 0; m < 7; m++) wa[m * nv + d] = R(0.0);
        if constexpr (sizeof(R) != sizeof(double)) wa[10 * nv + d] = R(0.0);
    }
    if constexpr (DT) {
        __shared__ double lds[4];
        mx = wave_max(mx);
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0) partials[blockIdx.x] = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
    }
}
template <typename R, bool DT = false>
__global__ __launch_bounds__(256) void k_aux(Geo g, KPar P, const double *__restrict__ w, R *__restrict__ wa,
                                             int i0, int ni, int j0, int nj, int k0, int nk, int image,
                                             const double *__restrict__ dx1, const double *__restrict__ dy1,
                                             const double *__restrict__ dz1, const double *__restrict__ dxs,
                                             const double *__restrict__ dys, const double *__restrict__ dzs,
                                             double *__restrict__ partials) {
    k_aux_body<R, DT>(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), g, P, w, wa, i0, ni, j0, nj, k0, nk, image, dx1, dy1, dz1, dxs, dys, dzs, partials);
}

// ------------------------------------------------------- dirty-cell flags
// Faces whose sensor window exceeds the threshold take the characteristic
// WENO flux.  The main sweep kernels evaluate only central faces, so their
// register budget stays that of the KEEP split; a cell with a WENO face on
// either side gets its flag byte set and is handled by the fix-up (k_fix_compact
// lists the flagged cells, k_fix_list processes them), which
// recomputes both faces of the cell with the full hybrid logic, applies the
// same divergence as the reference and clears the flag.
// flags[nint] doubles as the any-dirty word of the sweep: one plain store per
// wave that has a dirty lane (no returned value, so the wave never waits).
__device__ __forceinline__ void mark_dirty(unsigned char *__restrict__ flags, long long nint, bool dirty, int cell) {
    const unsigned long long mask = __ballot(dirty);
    if (dirty) flags[cell] = 1;
    if (mask != 0ull && __lane_id() == __ffsll((long long)mask) - 1) flags[nint] = 1;
}

// clears both any-dirty bytes: nint (x, y and z sweeps) and nint + 1 (the y
// term of the fused x + y pass)
__device__ __forceinline__ void k_clear_any_body(const dim3 blockIdx, const dim3 gridDim, unsigned char *__restrict__ flags, int nint) {
    flags[nint] = 0;
    flags[nint + 1] = 0;
}
__global__ void k_clear_any(unsigned char *__restrict__ flags, int nint) {
    k_clear_any_body(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), flags, nint);
}

// ------------------------------------------------------------ Euler y / z
// Streaming form of the central split along DIR.  One thread per (i, row)
// marches over a chunk of points; when point b arrives it forms its three
// pairs (b-l, b), l = 1..3, once, and adds c_l(face) times each pair to the
// three faces b-3..b-1 it straddles (the pair sums of the reference's
// double loop over l and m, kernels_cpu.F90:1861-1936, regrouped by pair:
// each pair is evaluated once instead of once per face).  Face b-3 is then
// complete; its divergence goes to cell b-3 (euler_y/z_update_subroutine,
// :991-1010, 1039-1058).  Faces whose sensor window exceeds the threshold
// are left to the fix-up through the dirty-cell flags.
template <typename R>
struct Acc {
    R m, mu, mv, mw, e, pp;
};

template <typename R, int DIR, int KM>
__device__ __forceinline__ Acc<R> pair_terms(const KPar &P, const PcR<R> &a, const PcR<R> &b, R pref) {
    const R una = un_of<DIR>(a), unb = un_of<DIR>(b);
    const R rhom = a.r + b.r;
    R uv = (una + unb) * rhom;
    Acc<R> r;
    if (KM >= 0) {
        const R eem = a.e + b.e;
        R deef = R(1.0);
        if (KM > 0) {
            const R drho = R(2.0) * (b.r - a.r) / rhom;
            const R dee = R(2.0) * (b.e - a.e) / eem;
            R t_sdr = (R(0.5) * drho) * (R(0.5) * drho);
            R t_sde = (R(0.5) * dee) * (R(0.5) * dee);
            const R t2_sdr = t_sdr, t2_sde = t_sde;
            R sdr = R(1.0) + t_sdr / R(3.0), sde = R(1.0) + t_sde, sne = R(1.0) + t_sde / R(3.0);
            for (int n = 2; n <= P.nkeep; n++) {
                const int n2 = 2 * n;
                t_sdr = t2_sdr * t_sdr;
                t_sde = t2_sde * t_sde;
                sdr = sdr + t_sdr / (R(1.0) + R(n2));
                sde = sde + t_sde;
                sne = sne + t_sde / (R(1.0) + R(n2));
            }
            uv = uv * (R(1.0) / sdr);
            deef = sne / sde;
        }
        if constexpr (sizeof(R) == sizeof(double))
            r.e = uv * eem * deef + uv * (a.u * b.u + a.v * b.v + a.w * b.w) + R(4.0) * (una * b.p + unb * a.p);
        else
            r.e = uv * eem * deef + uv * (a.u * b.u + a.v * b.v + a.w * b.w) +
                  R(4.0) * (una * (b.p + pref) + unb * (a.p + pref));
    } else {
        r.e = uv * (a.e + b.e);
    }
    r.m = uv * R(2.0);
    r.mu = uv * (a.u + b.u);
    r.mv = uv * (a.v + b.v);
    r.mw = uv * (a.w + b.w);
    r.pp = R(2.0) * (a.p + b.p);
    return r;
}

template <typename R>
__device__ __forceinline__ void acc_add(Acc<R> &A, R c, const Acc<R> &p) {
    A.m += c * p.m;
    A.mu += c * p.mu;
    A.mv += c * p.mv;
    A.mw += c * p.mw;
    A.e += c * p.e;
    A.pp += c * p.pp;
}

template <typename R>
__device__ __forceinline__ Acc<R> acc_zero() {
    Acc<R> a;
    a.m = a.mu = a.mv = a.mw = a.e = a.pp = R(0.0);
    return a;
}

// coeff_deriv1(l, lmax) for l, lmax = 1..3, read once per kernel: the reads
// are uniform, so the table sits in scalar registers, and face_coeff selects
// among values.  (Selecting among table addresses by a per-lane lm instead
// turns every face into conditional vector loads from the argument block.)
template <typename R>
struct CoefTab {
    R c[3][3];
};
template <typename R>
__device__ __forceinline__ CoefTab<R> coef_tab(const KPar &P) {
    CoefTab<R> t;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) t.c[a][b] = R(P.cd1tab[a][b]);
    return t;
}

// coeff_deriv1(l, lmax) of a face with effective half width lm (1..3)
template <typename R>
__device__ __forceinline__ R face_coeff(const CoefTab<R> &T, int lm, int l) {
    const R v3 = T.c[2][l - 1], v2 = T.c[1][l - 1], v1 = T.c[0][l - 1];
    return l > lm ? R(0.0) : (lm == 3 ? v3 : (lm == 2 ? v2 : v1));
}

// Operands of the viscous divergence terms folded into the z march (VD > 0):
// visflx_div_subroutine (kernels_cpu.F90:3484-3532), or with O2 its
// second-order form visflx_div_ord2_subroutine (:3446-3480), for the cell
// the march completes, and with VD = 2 the block partials of force_rhs as
// k_visc_div forms them.
struct VDArgs {
    const void *flp;  // right-hand side of the previous stage (R)
    const double *w, *dx1, *dy1, *yn;
    const double *xc, *yc, *zc;  // node coordinates (second-order form)
    const DevScalars *ds;
    double rkr, rkg;
    double *partials;
};

template <typename R, int DIR, int KM, int VD, bool O2>
__device__ __forceinline__ void k_euler_march_body(const dim3 blockIdx, const dim3 gridDim, Geo g, KPar P, const R *__restrict__ wa,
                                                     const int *__restrict__ eoc, const double *__restrict__ metric,
                                                     R *__restrict__ fl, int chunk, int nchunk,
                                                     unsigned char *__restrict__ flags, VDArgs V) {
    static_assert(VD == 0 || DIR == 2, "the viscous divergence rides on the z march only");
    __shared__ double lds[4][5];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int no = DIR == 1 ? g.nz : g.ny;  // the other transverse extent
    const int i = t % g.nx + 1;
    const int r = t / g.nx;
    const int o = r % no + 1;
    const int c = r / no;
    const int lo = P.eul_lo[DIR], hi = P.eul_hi[DIR];
    const int c0 = lo + c * chunk;
    // threads past the grid or the last chunk do nothing, but with VD = 2
    // they stay for the block sum
    const bool active = t < g.nx * no * nchunk && c0 <= hi;
    if (VD < 2 && !active) return;
    const int c1 = min(c0 + chunk - 1, hi);
    const int n = DIR == 1 ? g.ny : g.nz;
    const int stride = DIR == 1 ? g.sx : g.sxy;
    const size_t nv = g.nvol;
    const bool keep = KM >= 0;
    const R thr = R(P.sensor_threshold);
    const R rgas0 = R(P.rgas0), pref = R(P.p_ref);
    constexpr int TS = tp_slot<R>();
    const int ws = P.weno_scheme;
    // bit s of hist = sensor flag of point b-s; a face f = b-3 uses points
    // f-ws+1 .. f+ws, i.e. bits 3-ws .. 2+ws
    const unsigned smask = ((1u << (2 * ws)) - 1u) << (3 - ws);
    const int p0 = DIR == 1 ? gidx(g, i, 0, o) : gidx(g, i, o, 0);  // offset of along-index 0
    const int lb = P.lmax_base;
    const CoefTab<R> CT = coef_tab<R>(P);
    auto lm_of = [&](int f) {
        const int ff = f < 0 ? 0 : (f > n ? n : f);
        const int v = lb + eoc[ff];
        return v < 1 ? 1 : v;
    };
    double red[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // force_rhs partial sums (VD = 2)
    if (active) {
        // The march starts three points early (b = c0-3): those iterations
        // only load the warm-up points and feed faces below c0-1, which are
        // never written.  Every point thus passes through the one loop body,
        // so a face's arithmetic is the same whatever the chunk it falls in.
        PcR<R> q0{}, q1{}, q2{};
        unsigned hist = 0;
        Acc<R> A0 = acc_zero<R>(), A1 = acc_zero<R>(), A2 = acc_zero<R>();  // faces b-3, b-2, b-1
        int lm0 = lm_of(c0 - 6), lm1 = lm_of(c0 - 5), lm2 = lm_of(c0 - 4);
        R fprev[5] = {R(0.0), R(0.0), R(0.0), R(0.0), R(0.0)};
        bool wprev = false;
        R dw[7] = {R(0.0), R(0.0), R(0.0), R(0.0), R(0.0), R(0.0), R(0.0)};  // div/3 at points b-6 .. b (VD)
        // prefetch of the next point
        int a = p0 + (c0 - 3) * stride;
        R nr = wa[a], nu = wa[nv + a], nvv = wa[2 * nv + a], nw = wa[3 * nv + a], nh = wa[4 * nv + a],
          nt = wa[TS * nv + a], ns = wa[7 * nv + a];
        R nd = VD ? wa[9 * nv + a] : R(0.0);
        for (int b = c0 - 3; b <= c1 + 3; b++) {
            // issue the read of fl at the cell completed in this iteration first,
            // so that its latency overlaps the pair arithmetic below
            R flv[5] = {R(0.0), R(0.0), R(0.0), R(0.0), R(0.0)};
            const int fc = b - 3;
            const int oic = DIR == 1 ? iidx(g, i, fc < 1 ? 1 : fc, o) : iidx(g, i, o, fc < 1 ? 1 : fc);
            // the viscous-divergence operands of cell fc: mu, the x and y
            // neighbours of div/3 (the z ones are in dw), and for VD = 2 the
            // force_rhs inputs
            R vmu = R(0.0), dxm[3], dxp[3], dym[3], dyp[3], flp1 = R(0.0), flp2 = R(0.0), tcell = R(0.0);
            double w1 = 0.0, w2 = 0.0;
            if (fc >= c0) {
#pragma unroll
                for (int m = 0; m < 5; m++) flv[m] = fl[(size_t)m * g.nint + oic];
                if (VD) {
                    const int cc = gidx(g, i, o, fc);
                    const R *D = wa + 9 * nv;
                    vmu = wa[6 * nv + cc];
#pragma unroll
                    for (int l = 1; l <= (O2 ? 1 : 3); l++) {
                        dxm[l - 1] = D[cc - l];
                        dxp[l - 1] = D[cc + l];
                        dym[l - 1] = D[cc - l * g.sx];
                        dyp[l - 1] = D[cc + l * g.sx];
                    }
                    if (VD == 2 && V.rkr != 0.0) {
                        const R *fp = (const R *)V.flp;
                        flp1 = fp[oic];
                        flp2 = fp[g.nint + oic];
                    }
                    if (VD == 2) {
                        w1 = V.w[cc];
                        w2 = V.w[nv + cc];
                        tcell = wa[5 * nv + cc];
                    }
                }
            }
            const PcR<R> qb = make_pcr<R>(nr, nu, nvv, nw, nh, nt, rgas0, pref, keep);
            hist = (hist << 1) | (ns > thr ? 1u : 0u);
            if (VD) {
#pragma unroll
                for (int s2 = 0; s2 < 6; s2++) dw[s2] = dw[s2 + 1];
                dw[6] = nd;
            }
            if (b < c1 + 3) {
                a += stride;
                nr = wa[a]; nu = wa[nv + a]; nvv = wa[2 * nv + a]; nw = wa[3 * nv + a];
                nh = wa[4 * nv + a]; nt = wa[TS * nv + a]; ns = wa[7 * nv + a];
                if (VD) nd = wa[9 * nv + a];
            }
            {
                const Acc<R> p3 = pair_terms<R, DIR, KM>(P, q0, qb, pref);
                acc_add(A0, face_coeff<R>(CT, lm0, 3), p3);
                acc_add(A1, face_coeff<R>(CT, lm1, 3), p3);
                acc_add(A2, face_coeff<R>(CT, lm2, 3), p3);
            }
            {
                const Acc<R> p2 = pair_terms<R, DIR, KM>(P, q1, qb, pref);
                acc_add(A1, face_coeff<R>(CT, lm1, 2), p2);
                acc_add(A2, face_coeff<R>(CT, lm2, 2), p2);
            }
            {
                const Acc<R> p1 = pair_terms<R, DIR, KM>(P, q2, qb, pref);
                acc_add(A2, face_coeff<R>(CT, lm2, 1), p1);
            }
            const int f = b - 3;  // face complete
            if (f >= c0 - 1) {
                // (the fl values of cell f were loaded at the top of the iteration)
                const bool wcur = (hist & smask) != 0u;
                R fh[5];
                fh[0] = R(0.25) * A0.m;
                fh[1] = R(0.25) * A0.mu;
                fh[2] = R(0.25) * A0.mv;
                fh[3] = R(0.25) * A0.mw;
                fh[4] = R(0.25) * A0.e;
                if ((f == 0 && P.fzf[2 * DIR] == 1) || (f == n && P.fzf[2 * DIR + 1] == 1)) {
#pragma unroll
                    for (int m = 0; m < 5; m++) fh[m] = R(0.0);
                }
                fh[1 + DIR] = fh[1 + DIR] + R(0.5) * A0.pp;
                bool dirty = false;
                if (f >= c0) {
                    dirty = wcur || wprev;
                    if (VD == 0) {
                        if (!dirty) {
                            const R d = R(metric[f - 1]);
#pragma unroll
                            for (int m = 0; m < 5; m++) fl[(size_t)m * g.nint + oic] = flv[m] + (fh[m] - fprev[m]) * d;
                        }
                    } else {
                        // z divergence (clean cells; a WENO cell gets it from
                        // the fix-up), then the viscous divergence terms in
                        // the order of k_visc_div
                        R tv[5];
                        const R d = R(metric[f - 1]);
#pragma unroll
                        for (int m = 0; m < 5; m++) tv[m] = dirty ? flv[m] : flv[m] + (fh[m] - fprev[m]) * d;
                        R sigx, sigy, sigz;
                        if (O2) {
                            const int ng = g.ng;
                            const R dxl = vmu / R(V.xc[i + 1 + ng - 1] - V.xc[i - 1 + ng - 1]);
                            const R dyl = vmu / R(V.yc[o + 1 + ng - 1] - V.yc[o - 1 + ng - 1]);
                            const R dzl = vmu / R(V.zc[f + 1 + ng - 1] - V.zc[f - 1 + ng - 1]);
                            sigx = dxl * (dxp[0] - dxm[0]);
                            sigy = dyl * (dyp[0] - dym[0]);
                            sigz = dzl * (dw[4] - dw[2]);
                        } else {
                            R dxv = R(0.0), dyv = R(0.0), dzv = R(0.0);
#pragma unroll
                            for (int l = 1; l <= 3; l++) {
                                if (l <= P.visc_lmax) {
                                    const R ccl = R(P.vd1[l - 1]);
                                    dxv = dxv + ccl * (dxp[l - 1] - dxm[l - 1]);
                                    dyv = dyv + ccl * (dyp[l - 1] - dym[l - 1]);
                                    dzv = dzv + ccl * (dw[3 + l] - dw[3 - l]);
                                }
                            }
                            dxv = dxv * R(V.dx1[i - 1]);
                            dyv = dyv * R(V.dy1[o - 1]);
                            dzv = dzv * R(metric[f - 1]);
                            sigx = vmu * dxv;
                            sigy = vmu * dyv;
                            sigz = vmu * dzv;
                        }
                        const R sigq = sigx * q0.u + sigy * q0.v + sigz * q0.w;
                        tv[1] = tv[1] - sigx;
                        tv[2] = tv[2] - sigy;
                        tv[3] = tv[3] - sigz;
                        tv[4] = tv[4] - sigq;
#pragma unroll
                        for (int m = 0; m < 5; m++)
                            if (m > 0 || !dirty) fl[(size_t)m * g.nint + oic] = tv[m];
                        if (VD == 2) {
                            const double dt = V.ds->dt;
                            const double rhodt = V.rkr * dt, gamdt = V.rkg * dt;
                            const double dy = V.yn[o] - V.yn[o - 1];
                            const double fln1 = -rhodt * double(flp1) - gamdt * double(tv[0]);
                            const double fln2 = -rhodt * double(flp2) - gamdt * double(tv[1]);
                            red[0] += fln1 * dy;
                            red[1] += fln2 * dy;
                            red[2] += w1 * dy;
                            red[3] += w2 * dy;
                            red[4] += w2 * dy * double(tcell);
                        }
                    }
                }
                mark_dirty(flags, g.nint, dirty, oic);
#pragma unroll
                for (int m = 0; m < 5; m++) fprev[m] = fh[m];
                wprev = wcur;
            }
            A0 = A1;
            A1 = A2;
            A2 = acc_zero<R>();
            lm0 = lm1;
            lm1 = lm2;
            lm2 = lm_of(b);
            q0 = q1;
            q1 = q2;
            q2 = qb;
        }
    }
    if (VD == 2) {
        block_sum<5>(red, lds);
        if (flat_thread() == 0) {
            double *pp = V.partials + (size_t)(int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)) * 5;
#pragma unroll
            for (int q = 0; q < 5; q++) pp[q] = red[q];
        }
    }
}
template <typename R, int DIR, int KM, int VD, bool O2>
__global__ __launch_bounds__(256) void k_euler_march(Geo g, KPar P, const R *__restrict__ wa,
                                                     const int *__restrict__ eoc, const double *__restrict__ metric,
                                                     R *__restrict__ fl, int chunk, int nchunk,
                                                     unsigned char *__restrict__ flags, VDArgs V) {
    k_euler_march_body<R, DIR, KM, VD, O2>(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), g, P, wa, eoc, metric, fl, chunk, nchunk, flags, V);
}

// --------------------------------------------------------------- Euler x
// Lane-parallel streaming form along x.  A wave covers 64 consecutive
// points i0-3 .. i0+60 of one (j,k) row: each lane loads one point, forms
// the three pairs (b-l, b) ending at its point and their coefficient-weighted
// sums for the faces b-1, b-2, b-3; face f is then assembled from three
// lanes with cross-lane shuffles, and the divergence of cell f takes face
// f-1 from the neighbouring lane (euler_x_fluxes_hybrid_kernel +
// euler_x_update_subroutine, kernels_cpu.F90:425-689, 944-963, with the
// pair sums regrouped by pair).  Faces 59 per wave, cells 58.
constexpr int XW = 58;

// Lane moves of the x faces by DPP wave shifts (wave_shr:1 and wave_shl:1
// of the GFX9 family): lane i receives the value of lane i-1 (i+1); a lane
// whose source lies outside the wave receives 0 (bound_ctrl), which spares
// the copy that would stage its old value.  A move by d chains d shifts, so
// lanes below d (above 63-d) hold values the x faces never read.  These are
// VALU moves: no LDS round trip on the face chain.
constexpr int kDppUp = 0x138;    // wave_shr:1
constexpr int kDppDown = 0x130;  // wave_shl:1

template <int CTRL>
__device__ __forceinline__ int dpp_move(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = dpp_move<CTRL>(__double2loint(v));
    const int hi = dpp_move<CTRL>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(dpp_move<CTRL>(__float_as_int(v)));
}

template <typename R>
__device__ __forceinline__ R lane_up(R v, int d) {
    for (int s = 0; s < d; s++) v = dpp_move<kDppUp>(v);
    return v;
}
template <typename R>
__device__ __forceinline__ R lane_down(R v, int d) {
    for (int s = 0; s < d; s++) v = dpp_move<kDppDown>(v);
    return v;
}

template <typename R>
__device__ __forceinline__ PcR<R> shfl_up_pc(const PcR<R> &q, int d) {
    PcR<R> r;
    r.r = lane_up(q.r, d);
    r.u = lane_up(q.u, d);
    r.v = lane_up(q.v, d);
    r.w = lane_up(q.w, d);
    r.p = lane_up(q.p, d);
    r.e = lane_up(q.e, d);
    return r;
}

template <typename R>
__device__ __forceinline__ Acc<R> shfl_down_acc(const Acc<R> &q, int d) {
    Acc<R> r;
    r.m = lane_down(q.m, d);
    r.mu = lane_down(q.mu, d);
    r.mv = lane_down(q.mv, d);
    r.mw = lane_down(q.mw, d);
    r.e = lane_down(q.e, d);
    r.pp = lane_down(q.pp, d);
    return r;
}

template <typename R, bool INIT, int KM>
__global__ __launch_bounds__(256) void k_euler_x(Geo g, KPar P, const R *__restrict__ wa,
                                                 const int *__restrict__ eoc, const double *__restrict__ dcsidx,
                                                 R *__restrict__ fl, unsigned char *__restrict__ flags) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    const int nseg = (g.nx + XW - 1) / XW;
    const int nwork = nseg * g.ny * g.nz;
    const int lo = P.eul_lo[0], hi = P.eul_hi[0];
    const bool keep = KM >= 0;
    const size_t nv = g.nvol;
    const int ws = P.weno_scheme;
    const unsigned long long wmask = (1ull << (2 * ws)) - 1ull;
    const R pref = R(P.p_ref);
    const int lb = P.lmax_base;
    const CoefTab<R> CT = coef_tab<R>(P);
    auto lm_of = [&](int f) {
        const int ff = f < 0 ? 0 : (f > g.nx ? g.nx : f);
        const int v = lb + eoc[ff];
        return v < 1 ? 1 : v;
    };
    // point of this lane in work item wk (clamped into the ghosted row)
    auto point_of = [&](int wk_) {
        const int seg = wk_ % nseg;
        const int row = wk_ / nseg;
        const int pt_ = 1 + XW * seg - 3 + lane;
        return gidx(g, pt_ > g.nx + 3 ? g.nx + 3 : pt_, row % g.ny + 1, row / g.ny + 1);
    };
    // the loads of the next work item are issued before the work on the
    // current one, so their latency overlaps the pair arithmetic
    R nr = R(0), nu = R(0), nvv = R(0), nw = R(0), nh = R(0), nt = R(0), ns = R(0);
    if (wave < nwork) {
        const int a = point_of(wave);
        nr = wa[a]; nu = wa[nv + a]; nvv = wa[2 * nv + a]; nw = wa[3 * nv + a];
        nh = wa[4 * nv + a]; nt = wa[tp_slot<R>() * nv + a]; ns = wa[7 * nv + a];
    }
    for (int wk = wave; wk < nwork; wk += nwaves) {
        const int seg = wk % nseg;
        const int row = wk / nseg;
        const int j = row % g.ny + 1, k = row / g.ny + 1;
        const int i0 = 1 + XW * seg;
        const int pt = i0 - 3 + lane;                     // point held by this lane
        const PcR<R> q = make_pcr<R>(nr, nu, nvv, nw, nh, nt, R(P.rgas0), pref, keep);
        const bool flag = ns > R(P.sensor_threshold);
        if (wk + nwaves < nwork) {
            const int a = point_of(wk + nwaves);
            nr = wa[a]; nu = wa[nv + a]; nvv = wa[2 * nv + a]; nw = wa[3 * nv + a];
            nh = wa[4 * nv + a]; nt = wa[tp_slot<R>() * nv + a]; ns = wa[7 * nv + a];
        }
        const PcR<R> q1 = shfl_up_pc(q, 1), q2 = shfl_up_pc(q, 2), q3 = shfl_up_pc(q, 3);
        const int lmA = lm_of(pt - 1), lmB = lm_of(pt - 2), lmC = lm_of(pt - 3);
        Acc<R> Q1 = acc_zero<R>(), Q2 = acc_zero<R>(), Q3 = acc_zero<R>();
        {
            const Acc<R> p3 = pair_terms<R, 0, KM>(P, q3, q, pref);
            acc_add(Q1, face_coeff<R>(CT, lmA, 3), p3);
            acc_add(Q2, face_coeff<R>(CT, lmB, 3), p3);
            acc_add(Q3, face_coeff<R>(CT, lmC, 3), p3);
        }
        {
            const Acc<R> p2 = pair_terms<R, 0, KM>(P, q2, q, pref);
            acc_add(Q1, face_coeff<R>(CT, lmA, 2), p2);
            acc_add(Q2, face_coeff<R>(CT, lmB, 2), p2);
        }
        {
            const Acc<R> p1 = pair_terms<R, 0, KM>(P, q1, q, pref);
            acc_add(Q1, face_coeff<R>(CT, lmA, 1), p1);
        }
        // face f = pt: pairs from the lanes of points f+1, f+2, f+3
        Acc<R> F = shfl_down_acc(Q1, 1);
        {
            const Acc<R> t2 = shfl_down_acc(Q2, 2);
            F.m += t2.m; F.mu += t2.mu; F.mv += t2.mv; F.mw += t2.mw; F.e += t2.e; F.pp += t2.pp;
            const Acc<R> t3 = shfl_down_acc(Q3, 3);
            F.m += t3.m; F.mu += t3.mu; F.mv += t3.mv; F.mw += t3.mw; F.e += t3.e; F.pp += t3.pp;
        }
        const int f = pt;
        R fh[5];
        fh[0] = R(0.25) * F.m;
        fh[1] = R(0.25) * F.mu;
        fh[2] = R(0.25) * F.mv;
        fh[3] = R(0.25) * F.mw;
        fh[4] = R(0.25) * F.e;
        if ((f == 0 && P.fzf[0] == 1) || (f == g.nx && P.fzf[1] == 1)) {
#pragma unroll
            for (int m = 0; m < 5; m++) fh[m] = R(0.0);
        }
        fh[1] = fh[1] + R(0.5) * F.pp;
        // WENO face: a sensor flag among points f-ws+1 .. f+ws
        const unsigned long long fm = __ballot(flag);
        const int sh = lane - ws + 1;
        const bool wf = lane >= 2 && lane <= 60 && ((fm >> (sh < 0 ? 0 : sh)) & wmask) != 0ull;
        // cell c = pt, needs face c-1 from the lane below
        R fl_[5];
#pragma unroll
        for (int m = 0; m < 5; m++) fl_[m] = lane_up(fh[m], 1);
        const bool wl = dpp_move<kDppUp>(wf ? 1 : 0) != 0;
        const int c = pt;
        const bool cell = lane >= 3 && lane <= 60 && c <= g.nx;
        const bool in = cell && c >= lo && c <= hi;
        const bool dirty = in && (wf || wl);
        if (cell) {
            const int o = iidx(g, c, j, k);
            const R d = R(dcsidx[c - 1]);
#pragma unroll
            for (int m = 0; m < 5; m++) {
                if (INIT) fl[(size_t)m * g.nint + o] = (in && !dirty) ? (fh[m] - fl_[m]) * d : R(0.0);
                else if (in && !dirty) fl[(size_t)m * g.nint + o] += (fh[m] - fl_[m]) * d;
            }
        }
        mark_dirty(flags, g.nint, dirty, cell ? iidx(g, c, j, k) : 0);
    }
}

// ------------------------------------------------------------- Euler x + y
// The x sweep and the y march in one pass.  A wave holds 64 consecutive
// points along x (the x sweep's segment: 58 cells and 3 halo points on each
// side) and marches them along y over a chunk of rows, with the register
// march of k_euler_march<1> in every lane.  When the march completes row r
// (the oldest point of each lane's window), the lanes hold row r across the
// segment, so the x faces of row r are formed right there with the
// shuffles of k_euler_x.  Cell (i, r) then receives x-div + y-div in one
// store of fl (INIT: the stage's right-hand side starts here); a direction
// whose faces need WENO leaves its term to the fix-up through bit 0 (x) or
// bit 1 (y) of the cell's flag byte and the any-dirty bytes nint / nint+1.
// Both terms and their order of addition are those of the separate sweeps.
// Rows outside the y range of the Euler fluxes (cut by a characteristic
// boundary) receive the x term only, as from the two sweeps.
template <typename R, int KM>
__device__ __forceinline__ void k_euler_xy_body(const dim3 blockIdx, const dim3 gridDim, Geo g, KPar P, const R *__restrict__ wa,
                                                  const int *__restrict__ eocx, const int *__restrict__ eocy,
                                                  const double *__restrict__ dcsidx, const double *__restrict__ detady,
                                                  R *__restrict__ fl, int chunk, int nchunk,
                                                  unsigned char *__restrict__ flags) {
    const int lane = threadIdx.x & 63;
    const int gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nseg = (g.nx + XW - 1) / XW;
    if (gw >= nseg * g.nz * nchunk) return;  // whole waves; no workgroup barrier below
    const int seg = gw % nseg;
    const int rr = gw / nseg;
    const int k = rr % g.nz + 1;
    const int cidx = rr / g.nz;
    const int i0 = 1 + XW * seg;
    const int pt = i0 - 3 + lane;
    const int pl = pt > g.nx + 3 ? g.nx + 3 : pt;
    const int lox = P.eul_lo[0], hix = P.eul_hi[0];
    const int loy = P.eul_lo[1], hiy = P.eul_hi[1];
    // the march covers every row (the x faces); the y divergence goes to
    // the rows of the y range only (a characteristic boundary cuts it)
    const int c0 = 1 + cidx * chunk;
    if (c0 > g.ny) return;
    const int c1 = min(c0 + chunk - 1, g.ny);
    const int n = g.ny;
    const int stride = g.sx;
    const size_t nv = g.nvol;
    const bool keep = KM >= 0;
    const R thr = R(P.sensor_threshold);
    const R rgas0 = R(P.rgas0), pref = R(P.p_ref);
    constexpr int TS = tp_slot<R>();
    const int ws = P.weno_scheme;
    const unsigned smask = ((1u << (2 * ws)) - 1u) << (3 - ws);
    const unsigned long long wmask = (1ull << (2 * ws)) - 1ull;
    const int lb = P.lmax_base;
    const CoefTab<R> CT = coef_tab<R>(P);
    // every lane of a wave marches the same row index, so the y-face order
    // is uniform across the wave: scalar table reads below
    auto lmy_of = [&](int f) {
        const int ff = f < 0 ? 0 : (f > n ? n : f);
        const int v = lb + eocy[ff];
        return __builtin_amdgcn_readfirstlane(v < 1 ? 1 : v);
    };
    auto lmx_of = [&](int f) {
        const int ff = f < 0 ? 0 : (f > g.nx ? g.nx : f);
        const int v = lb + eocx[ff];
        return v < 1 ? 1 : v;
    };
    // x-face coefficients of this lane's faces (the same for every row)
    const int lmA = lmx_of(pt - 1), lmB = lmx_of(pt - 2), lmC = lmx_of(pt - 3);
    const R cA1 = face_coeff<R>(CT, lmA, 1), cA2 = face_coeff<R>(CT, lmA, 2), cA3 = face_coeff<R>(CT, lmA, 3);
    const R cB2 = face_coeff<R>(CT, lmB, 2), cB3 = face_coeff<R>(CT, lmB, 3), cC3 = face_coeff<R>(CT, lmC, 3);
    const bool cell = lane >= 3 && lane <= 60 && pt <= g.nx;
    const bool inx = cell && pt >= lox && pt <= hix;
    const double dxc = dcsidx[(pt < 1 ? 1 : (pt > g.nx ? g.nx : pt)) - 1];
    const int p0 = gidx(g, pl, 0, k);  // offset of row index 0 of this lane's column
    PcR<R> q0{}, q1{}, q2{};
    unsigned hist = 0;
    Acc<R> A0 = acc_zero<R>(), A1 = acc_zero<R>(), A2 = acc_zero<R>();  // y faces b-3, b-2, b-1
    int lm0 = lmy_of(c0 - 6), lm1 = lmy_of(c0 - 5), lm2 = lmy_of(c0 - 4);
    R fprev[5] = {R(0.0), R(0.0), R(0.0), R(0.0), R(0.0)};
    bool wprev = false;
    unsigned char anyx = 0, anyy = 0;
    int a = p0 + (c0 - 3) * stride;
    R nr = wa[a], nu = wa[nv + a], nvv = wa[2 * nv + a], nw = wa[3 * nv + a], nh = wa[4 * nv + a],
      nt = wa[TS * nv + a], ns = wa[7 * nv + a];
    for (int b = c0 - 3; b <= c1 + 3; b++) {
        const PcR<R> qb = make_pcr<R>(nr, nu, nvv, nw, nh, nt, rgas0, pref, keep);
        hist = (hist << 1) | (ns > thr ? 1u : 0u);
        if (b < c1 + 3) {
            a += stride;
            nr = wa[a]; nu = wa[nv + a]; nvv = wa[2 * nv + a]; nw = wa[3 * nv + a];
            nh = wa[4 * nv + a]; nt = wa[TS * nv + a]; ns = wa[7 * nv + a];
        }
        // y pairs ending at point b (k_euler_march)
        {
            const Acc<R> p3 = pair_terms<R, 1, KM>(P, q0, qb, pref);
            acc_add(A0, face_coeff<R>(CT, lm0, 3), p3);
            acc_add(A1, face_coeff<R>(CT, lm1, 3), p3);
            acc_add(A2, face_coeff<R>(CT, lm2, 3), p3);
        }
        {
            const Acc<R> p2 = pair_terms<R, 1, KM>(P, q1, qb, pref);
            acc_add(A1, face_coeff<R>(CT, lm1, 2), p2);
            acc_add(A2, face_coeff<R>(CT, lm2, 2), p2);
        }
        {
            const Acc<R> p1 = pair_terms<R, 1, KM>(P, q2, qb, pref);
            acc_add(A2, face_coeff<R>(CT, lm2, 1), p1);
        }
        const int f = b - 3;  // y face f complete; row f = the window's oldest point q0
        if (f >= c0 - 1) {
            const bool wcur = (hist & smask) != 0u;
            R fh[5];
            fh[0] = R(0.25) * A0.m;
            fh[1] = R(0.25) * A0.mu;
            fh[2] = R(0.25) * A0.mv;
            fh[3] = R(0.25) * A0.mw;
            fh[4] = R(0.25) * A0.e;
            if ((f == 0 && P.fzf[2] == 1) || (f == n && P.fzf[3] == 1)) {
#pragma unroll
                for (int m = 0; m < 5; m++) fh[m] = R(0.0);
            }
            fh[2] = fh[2] + R(0.5) * A0.pp;
            if (f >= c0) {
                // x faces of row f (k_euler_x), from the lanes' points of that row
                // pairs (pt-l, pt), l = 3, 2, 1, from one chain of left
                // neighbours; the face of lane L gathers Q1(L+1) + Q2(L+2) +
                // Q3(L+3), shifted down in Horner form ((Q3 + Q2) + Q1): three
                // lane moves per value instead of six
                const PcR<R> x1 = shfl_up_pc(q0, 1);
                const PcR<R> x2 = shfl_up_pc(x1, 1);
                const PcR<R> x3 = shfl_up_pc(x2, 1);
                Acc<R> Q1 = acc_zero<R>(), Q2 = acc_zero<R>(), F = acc_zero<R>();
                {
                    const Acc<R> p3 = pair_terms<R, 0, KM>(P, x3, q0, pref);
                    acc_add(Q1, cA3, p3);
                    acc_add(Q2, cB3, p3);
                    acc_add(F, cC3, p3);  // Q3
                }
                {
                    const Acc<R> p2 = pair_terms<R, 0, KM>(P, x2, q0, pref);
                    acc_add(Q1, cA2, p2);
                    acc_add(Q2, cB2, p2);
                }
                F = shfl_down_acc(F, 1);
                F.m += Q2.m; F.mu += Q2.mu; F.mv += Q2.mv; F.mw += Q2.mw; F.e += Q2.e; F.pp += Q2.pp;
                {
                    const Acc<R> p1 = pair_terms<R, 0, KM>(P, x1, q0, pref);
                    acc_add(Q1, cA1, p1);
                }
                F = shfl_down_acc(F, 1);
                F.m += Q1.m; F.mu += Q1.mu; F.mv += Q1.mv; F.mw += Q1.mw; F.e += Q1.e; F.pp += Q1.pp;
                F = shfl_down_acc(F, 1);
                R gx[5];
                gx[0] = R(0.25) * F.m;
                gx[1] = R(0.25) * F.mu;
                gx[2] = R(0.25) * F.mv;
                gx[3] = R(0.25) * F.mw;
                gx[4] = R(0.25) * F.e;
                if ((pt == 0 && P.fzf[0] == 1) || (pt == g.nx && P.fzf[1] == 1)) {
#pragma unroll
                    for (int m = 0; m < 5; m++) gx[m] = R(0.0);
                }
                gx[1] = gx[1] + R(0.5) * F.pp;
                // x WENO face: a sensor flag among points pt-ws+1 .. pt+ws of row f
                const unsigned long long fm = __ballot(((hist >> 3) & 1u) != 0u);
                const int sh = lane - ws + 1;
                const bool wf = lane >= 2 && lane <= 60 && ((fm >> (sh < 0 ? 0 : sh)) & wmask) != 0ull;
                R gl[5];
#pragma unroll
                for (int m = 0; m < 5; m++) gl[m] = lane_up(gx[m], 1);
                const bool wl = dpp_move<kDppUp>(wf ? 1 : 0) != 0;
                const bool xdirty = inx && (wf || wl);
                const bool yin = f >= loy && f <= hiy;
                const bool ydirty = cell && yin && (wcur || wprev);
                const unsigned long long mx = __ballot(xdirty), my = __ballot(ydirty);
                if (mx != 0ull) anyx = 1;
                if (my != 0ull) anyy = 1;
                if (cell) {
                    const int oi = iidx(g, pt, f, k);
                    const R d = R(detady[f - 1]);
                    const R dx = R(dxc);
#pragma unroll
                    for (int m = 0; m < 5; m++) {
                        const R xv = (inx && !xdirty) ? (gx[m] - gl[m]) * dx : R(0.0);
                        fl[(size_t)m * g.nint + oi] = (ydirty || !yin) ? xv : xv + (fh[m] - fprev[m]) * d;
                    }
                    if (xdirty || ydirty) flags[oi] = (unsigned char)((xdirty ? 1 : 0) | (ydirty ? 2 : 0));
                }
            }
#pragma unroll
            for (int m = 0; m < 5; m++) fprev[m] = fh[m];
            wprev = wcur;
        }
        A0 = A1;
        A1 = A2;
        A2 = acc_zero<R>();
        lm0 = lm1;
        lm1 = lm2;
        lm2 = lmy_of(b);
        q0 = q1;
        q1 = q2;
        q2 = qb;
    }
    if (lane == 0) {
        if (anyx) flags[g.nint] = 1;
        if (anyy) flags[g.nint + 1] = 1;
    }
}
template <typename R, int KM>
__global__ __launch_bounds__(256) void k_euler_xy(Geo g, KPar P, const R *__restrict__ wa,
                                                  const int *__restrict__ eocx, const int *__restrict__ eocy,
                                                  const double *__restrict__ dcsidx, const double *__restrict__ detady,
                                                  R *__restrict__ fl, int chunk, int nchunk,
                                                  unsigned char *__restrict__ flags) {
    k_euler_xy_body<R, KM>(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), g, P, wa, eocx, eocy, dcsidx, detady, fl, chunk, nchunk, flags);
}

// ------------------------------------------------------------ WENO fix-up
// Hybrid flux of the face between window points W-1 and W of q (offsets
// -W+1..+W): central when no sensor value of the reference's switch window
// exceeds the threshold, characteristic WENO otherwise (the branch of the
// reference kernels).
template <int DIR, int W = H>
__device__ __forceinline__ void hybrid_face(const KPar &P, const Pt (&q)[2 * W], const double (&sen)[2 * W], int f,
                                            int e, int n, double fh[5]) {
    bool shk = false;
#pragma unroll
    for (int s = 0; s < 2 * W; s++)
        if (s >= W - P.weno_scheme && s < W + P.weno_scheme) shk |= sen[s] > P.sensor_threshold;
    if (!shk) {
        Pc c[2 * W];
#pragma unroll
        for (int s = 0; s < 2 * W; s++) c[s] = pc_of(q[s]);
        int lmax = P.lmax_base + e;
        lmax = lmax < 1 ? 1 : lmax;
        const bool zf = (f == 0 && P.fzf[2 * DIR] == 1) || (f == n && P.fzf[2 * DIR + 1] == 1);
        if (P.nkeep == 0) central_flux<DIR, W, Pc, 0>(P, c, lmax, zf, fh);
        else if (P.nkeep > 0) central_flux<DIR, W, Pc, 1>(P, c, lmax, zf, fh);
        else central_flux<DIR, W, Pc, -1>(P, c, lmax, zf, fh);
    } else if (P.flux_splitting == 1) {
        weno_flux_rusanov<DIR, W>(P, q, e, fh);
    } else {
        weno_flux<DIR, W>(P, q, e, fh);
    }
}

// Flagged cells: both faces recomputed with the hybrid logic, then
// fl += (fhat(c) - fhat(c-1)) * metric(c), and the flag cleared.  One
// 64-lane workgroup per grid-stride lane set; flagged cells are compacted
// into an LDS queue with a ballot so the WENO work runs on full waves.
// With RED the change the fix-up makes to fl(1) and fl(2) enters the
// force_rhs sums as k_visc_div forms them: r0, r1 += -gamdt*dfl*dy.
template <typename R, int DIR, bool RED = false>
__device__ __forceinline__ void fix_cell(const Geo &g, const KPar &P, const R *__restrict__ wa,
                                         const double *__restrict__ w, const int *__restrict__ eoc,
                                         const double *__restrict__ metric, R *__restrict__ fl, int oi,
                                         double *r0 = nullptr, double *r1 = nullptr, double gamdt = 0.0,
                                         const double *__restrict__ yn = nullptr) {
    const size_t nv = g.nvol;
    const bool keep = P.nkeep >= 0;
    const int n = DIR == 0 ? g.nx : (DIR == 1 ? g.ny : g.nz);
    const int stride = DIR == 0 ? 1 : (DIR == 1 ? g.sx : g.sxy);
    const int i = oi % g.nx + 1;
    const int j = (oi / g.nx) % g.ny + 1;
    const int k = oi / (g.nx * g.ny) + 1;
    const int c = DIR == 0 ? i : (DIR == 1 ? j : k);
    const int base = gidx(g, i, j, k);
    // Every load of the stencil is issued before any arithmetic: the seven
    // points of a y or z stencil lie in seven different pages, and waiting
    // on them one after another would serialize seven memory latencies.
    double raw[7][6], sen[7];
#pragma unroll
    for (int s = 0; s < 7; s++) {
        const int a = base + (s - 3) * stride;
        if constexpr (sizeof(R) == sizeof(double)) {
#pragma unroll
            for (int v = 0; v < 6; v++) raw[s][v] = double(wa[v * nv + a]);
        } else {
            // the mixed form stores w_aux in FP32: the WENO stencil takes the
            // FP64 primitives of w instead, formed as eval_aux forms them
            // (periodic neighbours from their images); raw[s][5] is the
            // stored temperature, the start of the cp(T) inversion
            int pi = i, pj = j, pk = k;
            if (DIR == 0) pi += s - 3;
            else if (DIR == 1) pj += s - 3;
            else pk += s - 3;
            if (P.periodic[0]) pi = pi < 1 ? pi + g.nx : (pi > g.nx ? pi - g.nx : pi);
            if (P.periodic[1]) pj = pj < 1 ? pj + g.ny : (pj > g.ny ? pj - g.ny : pj);
            if (P.periodic[2]) pk = pk < 1 ? pk + g.nz : (pk > g.nz ? pk - g.nz : pk);
            const int src = gidx(g, pi, pj, pk);
#pragma unroll
            for (int v = 0; v < 5; v++) raw[s][v] = w[v * nv + src];
            raw[s][5] = double(wa[5 * nv + a]);
        }
        sen[s] = double(wa[7 * nv + a]);
    }
    const int elo = eoc[c - 1], ehi = eoc[c];
    const double d = metric[c - 1];
    double flv[5];
#pragma unroll
    for (int m = 0; m < 5; m++) flv[m] = double(fl[(size_t)m * g.nint + oi]);
    Pt q[7];
#pragma unroll
    for (int s = 0; s < 7; s++) {
        if constexpr (sizeof(R) == sizeof(double)) {
            q[s] = make_pt(raw[s][0], raw[s][1], raw[s][2], raw[s][3], raw[s][4], raw[s][5], P.rgas0, keep);
        } else {
            const double rho = raw[s][0], rhoe = raw[s][4];
            const double ri = 1.0 / rho;
            const double uu = raw[s][1] * ri, vv = raw[s][2] * ri, ww = raw[s][3] * ri;
            const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
            const double ee = rhoe / rho - qq;
            const double tt = temperature_from_e(P, ee, raw[s][5]);
            const double pp = rho * tt * P.rgas0;
            q[s] = make_pt(rho, uu, vv, ww, (rhoe + pp) / rho, tt, P.rgas0, keep);
        }
    }
    // Face A is a WENO face of the cell where it has one (the lower one when
    // its window is shocked, else the upper one), face B the other.  Cells
    // of a wave mostly have one WENO face, at either end; taking that face
    // first keeps the WENO branch of a wave on all its lanes at once instead
    // of running it twice on about half of them.  Each face is evaluated
    // exactly as before, so the results do not depend on the order.
    bool shlo = false, shhi = false;
#pragma unroll
    for (int s = 0; s < 6; s++)
        if (s >= H - P.weno_scheme && s < H + P.weno_scheme) {  // the windows of hybrid_face
            shlo |= sen[s] > P.sensor_threshold;
            shhi |= sen[s + 1] > P.sensor_threshold;
        }
    const bool alo = shlo || !shhi;  // face A is the lower face
    double flo[5], fhi[5];
    // one loop body for both faces (its code is large; two inlined copies
    // would double the instruction footprint of the kernel)
#pragma unroll 1
    for (int t = 0; t < 2; t++) {
        const bool lo = (t == 0) == alo;
        Pt qw[6];
        double sw[6];
#pragma unroll
        for (int s = 0; s < 6; s++) {
            qw[s] = lo ? q[s] : q[s + 1];
            sw[s] = lo ? sen[s] : sen[s + 1];
        }
        double fh[5];
        hybrid_face<DIR>(P, qw, sw, lo ? c - 1 : c, lo ? elo : ehi, n, fh);
#pragma unroll
        for (int m = 0; m < 5; m++) {
            if (lo) flo[m] = fh[m];
            else fhi[m] = fh[m];
        }
    }
    R fnew[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        fnew[m] = R(flv[m] + (fhi[m] - flo[m]) * d);
        fl[(size_t)m * g.nint + oi] = fnew[m];
    }
    if (RED) {
        const double dy = yn[j] - yn[j - 1];
        *r0 += -gamdt * (double(fnew[0]) - flv[0]) * dy;
        *r1 += -gamdt * (double(fnew[1]) - flv[1]) * dy;
    }
}

// The x fix-up of a wave of listed cells.  Lanes 0..62 own the cells of
// 63 consecutive list entries; lane 63 takes the next entry as a helper.
// Every lane evaluates the lower face of its cell; when the next lane holds
// the next cell along x (a run of flagged cells), that face is the upper
// face of this lane's cell and comes with a lane shift, so inside a run
// each face is evaluated once.  Only the last cell of a run evaluates its
// upper face itself, which is then mostly a central face.  Each face is
// evaluated exactly as fix_cell evaluates it (the same window and branch),
// so the results are those of fix_cell.
template <typename R>
__device__ __forceinline__ void fix_cell_x(const Geo &g, const KPar &P, const R *__restrict__ wa,
                                           const double *__restrict__ w, const int *__restrict__ eoc,
                                           const double *__restrict__ metric, R *__restrict__ fl, int oi, bool act,
                                           bool owner) {
    const size_t nv = g.nvol;
    const bool keep = P.nkeep >= 0;
    const int n = g.nx;
    const int i = act ? oi % g.nx + 1 : 1;
    const int j = act ? (oi / g.nx) % g.ny + 1 : 1;
    const int k = act ? oi / (g.nx * g.ny) + 1 : 1;
    const int base = gidx(g, i, j, k);
    Pt q[7];
    double sen[7];
    double flo[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (act) {
        double raw[7][6];
#pragma unroll
        for (int s = 0; s < 7; s++) {
            const int a = base + (s - 3);
            if constexpr (sizeof(R) == sizeof(double)) {
#pragma unroll
                for (int v = 0; v < 6; v++) raw[s][v] = double(wa[v * nv + a]);
            } else {
                int pi = i + s - 3;
                if (P.periodic[0]) pi = pi < 1 ? pi + g.nx : (pi > g.nx ? pi - g.nx : pi);
                const int src = gidx(g, pi, j, k);
#pragma unroll
                for (int v = 0; v < 5; v++) raw[s][v] = w[v * nv + src];
                raw[s][5] = double(wa[5 * nv + a]);
            }
            sen[s] = double(wa[7 * nv + a]);
        }
#pragma unroll
        for (int s = 0; s < 7; s++) {
            if constexpr (sizeof(R) == sizeof(double)) {
                q[s] = make_pt(raw[s][0], raw[s][1], raw[s][2], raw[s][3], raw[s][4], raw[s][5], P.rgas0, keep);
            } else {
                const double rho = raw[s][0], rhoe = raw[s][4];
                const double ri = 1.0 / rho;
                const double uu = raw[s][1] * ri, vv = raw[s][2] * ri, ww = raw[s][3] * ri;
                const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
                const double ee = rhoe / rho - qq;
                const double tt = temperature_from_e(P, ee, raw[s][5]);
                const double pp = rho * tt * P.rgas0;
                q[s] = make_pt(rho, uu, vv, ww, (rhoe + pp) / rho, tt, P.rgas0, keep);
            }
        }
        Pt qw[6];
        double sw[6];
#pragma unroll
        for (int s = 0; s < 6; s++) {
            qw[s] = q[s];
            sw[s] = sen[s];
        }
        hybrid_face<0>(P, qw, sw, i - 1, eoc[i - 1], n, flo);
    }
    // the next lane's lower face is this cell's upper face when the next
    // lane holds the next cell of the same x row
    const int oin = __shfl_down(oi, 1, 64);
    double nb[5];
#pragma unroll
    for (int m = 0; m < 5; m++) nb[m] = __shfl_down(flo[m], 1, 64);
    if (!act || !owner) return;
    const bool shared = __lane_id() < 63 && oin == oi + 1 && i < n;
    double fhi[5];
    if (shared) {
#pragma unroll
        for (int m = 0; m < 5; m++) fhi[m] = nb[m];
    } else {
        Pt qw[6];
        double sw[6];
#pragma unroll
        for (int s = 0; s < 6; s++) {
            qw[s] = q[s + 1];
            sw[s] = sen[s + 1];
        }
        hybrid_face<0>(P, qw, sw, i, eoc[i], n, fhi);
    }
    const double d = metric[i - 1];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        const double flv = double(fl[(size_t)m * g.nint + oi]);
        fl[(size_t)m * g.nint + oi] = R(flv + (fhi[m] - flo[m]) * d);
    }
}

// The WENO fix-up of one direction in two kernels.  k_fix_compact: block b
// takes segment b of the flag array (wpb 64-bit words = 8 * wpb cells, each
// thread a run of consecutive words), lists the cells whose flag byte holds
// `bit` into list[b * 8 * wpb ...] in increasing order, stores their number
// in counts[b] and clears the bit.  k_fix_list: every block forms the prefix
// sums of counts in LDS and its waves take batches of 64 entries of the
// concatenated lists in turn, so the work is spread over the chip however
// the shocked cells are distributed.  No atomics: the lists and their order
// are a function of the flags alone.
constexpr int kFixMaxSeg = 4096;
constexpr int kFixBlocks = 256;  // workgroups of the fix-up (one per CU at its occupancy)

__device__ __forceinline__ unsigned dirty_bytes(unsigned long long wd, int bit, int cell0, int ncell) {
    unsigned mine = 0;  // bit b: byte b of the word is a flagged cell
#pragma unroll
    for (int b = 0; b < 8; b++)
        if (((wd >> (8 * b)) & (unsigned long long)bit) != 0ull && cell0 + b < ncell) mine |= 1u << b;
    return mine;
}

// exclusive prefix over the block's 256 threads of v; *total receives the sum
__device__ __forceinline__ int block_excl_scan256(int v, int *wsum, int *total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int before = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (q < wv) before += wsum[q];
        tot += wsum[q];
    }
    __syncthreads();
    *total = tot;
    return before + x - v;
}

__device__ __forceinline__ void k_fix_compact_body(const dim3 blockIdx, const dim3 gridDim, unsigned char *__restrict__ flags, int ncell, int bit,
                                                     int anyidx, int wpb, int *__restrict__ list,
                                                     int *__restrict__ counts) {
    __shared__ int wsum[4];
    if (flags[ncell + anyidx] == 0) return;  // no dirty cell of this direction in this sweep
    const int nword = (ncell + 7) >> 3;
    const int wpt = wpb >> 8;                // words per thread
    const int w0 = blockIdx.x * wpb + threadIdx.x * wpt;
    const int w1 = min(w0 + wpt, nword);
    unsigned long long *words = (unsigned long long *)flags;
    int c = 0;
    for (int wi = w0; wi < w1; wi++) c += __popc(dirty_bytes(words[wi], bit, 8 * wi, ncell));
    int total;
    int off = block_excl_scan256(c, wsum, &total);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
    if (c == 0) return;
    int *seg = list + (size_t)blockIdx.x * 8 * wpb;
    for (int wi = w0; wi < w1; wi++) {
        const unsigned long long wd = words[wi];
        const unsigned mine = dirty_bytes(wd, bit, 8 * wi, ncell);
        if (mine == 0u) continue;
        // clear the bit in the listed bytes only: the last word may also
        // hold the any-dirty bytes, which are not cells
        unsigned long long clr = 0ull;
#pragma unroll
        for (int b = 0; b < 8; b++)
            if ((mine >> b) & 1u) {
                seg[off++] = 8 * wi + b;
                clr |= (unsigned long long)bit << (8 * b);
            }
        words[wi] = wd & ~clr;
    }
}
__global__ __launch_bounds__(256) void k_fix_compact(unsigned char *__restrict__ flags, int ncell, int bit,
                                                     int anyidx, int wpb, int *__restrict__ list,
                                                     int *__restrict__ counts) {
    k_fix_compact_body(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), flags, ncell, bit, anyidx, wpb, list, counts);
}

// RED: the block's partial force_rhs sums of the fix-up's changes to fl(1)
// and fl(2) go to partials[blockIdx.x] (5 values, the last three zero),
// written also when there is nothing to fix.
template <typename R, int DIR, bool RED>
__device__ __forceinline__ void k_fix_list_body(const dim3 blockIdx, const dim3 gridDim, Geo g, KPar P, const R *__restrict__ wa,
                                                  const double *__restrict__ w, const int *__restrict__ eoc,
                                                  const double *__restrict__ metric, R *__restrict__ fl,
                                                  const unsigned char *__restrict__ flags, int anyidx,
                                                  const int *__restrict__ list, const int *__restrict__ counts,
                                                  int nseg, int seglen, const DevScalars *__restrict__ ds,
                                                  double rkg, const double *__restrict__ yn,
                                                  double *__restrict__ partials) {
    __shared__ int pre[kFixMaxSeg + 1];
    __shared__ int wsum[4];
    __shared__ double lds[4][5];
    double r0 = 0.0, r1 = 0.0;
    const bool any = flags[g.nint + anyidx] != 0;  // uniform
    if (!RED && !any) return;
    const int t = threadIdx.x;
    if (any) {
        // prefix sums of the segment counts: thread t sums segments 16t .. 16t+15
        int cs[16], sum = 0;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int sgi = 16 * t + q;
            cs[q] = sgi < nseg ? counts[sgi] : 0;
            sum += cs[q];
        }
        int total;
        int run = block_excl_scan256(sum, wsum, &total);
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int sgi = 16 * t + q;
            if (sgi < nseg) pre[sgi] = run;
            run += cs[q];
        }
        if (t == 0) pre[nseg] = total;
        __syncthreads();
        const double gamdt = RED ? rkg * ds->dt : 0.0;
        const int lane = t & 63;
        const int wave = (blockIdx.x * blockDim.x + t) >> 6;
        const int nwaves = (gridDim.x * blockDim.x) >> 6;
        auto entry = [&](int e) {
            int lo = 0, hi = nseg;  // pre[lo] <= e < pre[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (pre[mid] <= e) lo = mid;
                else hi = mid;
            }
            return list[(size_t)lo * seglen + (e - pre[lo])];
        };
        if constexpr (DIR == 0 && !RED) {
            // batches of 63 cells, lane 63 the helper of the next entry (fix_cell_x)
            for (int base = wave * 63; base < total; base += nwaves * 63) {
                const int e = base + lane;
                const bool act = e < total;
                fix_cell_x<R>(g, P, wa, w, eoc, metric, fl, act ? entry(e) : -2, act, lane < 63);
            }
        } else {
            for (int base = wave * 64; base < total; base += nwaves * 64) {
                const int e = base + lane;
                if (e < total) fix_cell<R, DIR, RED>(g, P, wa, w, eoc, metric, fl, entry(e), &r0, &r1, gamdt, yn);
            }
        }
    }
    if (RED) {
        double red[5] = {r0, r1, 0.0, 0.0, 0.0};
        block_sum<5>(red, lds);
        if (t == 0) {
            double *pp = partials + (size_t)blockIdx.x * 5;
#pragma unroll
            for (int q = 0; q < 5; q++) pp[q] = red[q];
        }
    }
}
template <typename R, int DIR, bool RED>
__global__ __launch_bounds__(256) void k_fix_list(Geo g, KPar P, const R *__restrict__ wa,
                                                  const double *__restrict__ w, const int *__restrict__ eoc,
                                                  const double *__restrict__ metric, R *__restrict__ fl,
                                                  const unsigned char *__restrict__ flags, int anyidx,
                                                  const int *__restrict__ list, const int *__restrict__ counts,
                                                  int nseg, int seglen, const DevScalars *__restrict__ ds,
                                                  double rkg, const double *__restrict__ yn,
                                                  double *__restrict__ partials) {
    k_fix_list_body<R, DIR, RED>(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), g, P, wa, w, eoc, metric, fl, flags, anyidx, list, counts, nseg, seglen, ds, rkg, yn, partials);
}

// ---------------------------------------------------------------- viscous
// Ghost images of an interior value: periodic copy for periodic directions
// (bcswap_var), zero-gradient copy otherwise (bcextr_var).
template <typename R>
__device__ __forceinline__ void write_images(const Geo &g, const KPar &P, R *__restrict__ a, int i, int j, int k,
                                             R v, bool extrapolate) {
    const int ng = g.ng;
    if (P.periodic[0]) {
        if (i <= ng) a[gidx(g, i + g.nx, j, k)] = v;
        if (i > g.nx - ng) a[gidx(g, i - g.nx, j, k)] = v;
    } else if (extrapolate) {
        if (i == 1) for (int l = 1; l <= ng; l++) a[gidx(g, 1 - l, j, k)] = v;
        if (i == g.nx) for (int l = 1; l <= ng; l++) a[gidx(g, g.nx + l, j, k)] = v;
    }
    if (P.periodic[1]) {
        if (j <= ng) a[gidx(g, i, j + g.ny, k)] = v;
        if (j > g.ny - ng) a[gidx(g, i, j - g.ny, k)] = v;
    } else if (extrapolate) {
        if (j == 1) for (int l = 1; l <= ng; l++) a[gidx(g, i, 1 - l, k)] = v;
        if (j == g.ny) for (int l = 1; l <= ng; l++) a[gidx(g, i, g.ny + l, k)] = v;
    }
    if (P.periodic[2] && P.ndim == 3) {
        if (k <= ng) a[gidx(g, i, j, k + g.nz)] = v;
        if (k > g.nz - ng) a[gidx(g, i, j, k - g.nz)] = v;
    } else if (extrapolate) {
        if (k == 1) for (int l = 1; l <= ng; l++) a[gidx(g, i, j, 1 - l)] = v;
        if (k == g.nz) for (int l = 1; l <= ng; l++) a[gidx(g, i, j, g.nz + l)] = v;
    }
}

// Point values of the viscous stencils: u, v, w, T, mu (w_aux 2, 3, 4, 6, 7).
template <typename R>
struct VP {
    R u, v, w, t, m;
};
template <typename R>
__device__ __forceinline__ VP<R> load_vp(const R *__restrict__ wa, size_t nv, int c) {
    VP<R> p;
    p.u = wa[nv + c];
    p.v = wa[2 * nv + c];
    p.w = wa[3 * nv + c];
    p.t = wa[5 * nv + c];
    p.m = wa[6 * nv + c];
    return p;
}

// visflx_subroutine (mode 0, kernels_cpu.F90:3598-3777, SENSOR) and
// visflx_nosensor_subroutine (mode 4, :3996-4180), z-marching: a VBX x VBY
// column tile walks a chunk of planes.  The z stencil of u, v, w, T, mu
// lives in a 7-plane register window (each plane is read from HBM once per
// chunk); the x and y stencils come from an LDS copy of the current plane
// with a 3-point halo.  Loads of plane k+1 (centre, halo, and the fl of
// plane k) are issued before the work on plane k, and the LDS tile is
// double-buffered, so one barrier per plane suffices.  ``images`` also
// writes the ghost images of w_aux(10) (or of the sensor with SENSOR).
constexpr int VBX = 32, VBY = 8;
// each lane fills at most one halo point of the plane tile
static_assert(VBX * VBY >= 2 * H * (VBX + VBY), "fewer lanes than halo points");

// The FP32 form fits 128 VGPRs without spills under a 4-waves bound (137
// unbounded, 3 waves/SIMD): 0.630 -> 0.604 ms on 512x256x256.  The FP64 form
// stays unbounded (209 VGPRs); a 3-waves bound spills 136 B/lane there.
template <typename R, bool SENSOR>
__device__ __forceinline__ void k_visc_zm_body(const dim3 blockIdx, const dim3 gridDim, Geo g, KPar P, R *__restrict__ wa, R *__restrict__ fl,
                                                      const double *__restrict__ dx1, const double *__restrict__ dxs,
                                                      const double *__restrict__ dx2, const double *__restrict__ dy1,
                                                      const double *__restrict__ dys, const double *__restrict__ dy2,
                                                      const double *__restrict__ dz1, const double *__restrict__ dzs,
                                                      const double *__restrict__ dz2, double *__restrict__ wallprop,
                                                      int images, int chunk) {
    // two plane buffers: plane k is read while plane k+1 is written
    __shared__ R tiles[2][5][VBY + 2 * H][VBX + 2 * H];
    const int tx = threadIdx.x, ty = threadIdx.y;
    // Workgroups are dealt round-robin to the 8 XCDs in linear order.  When
    // the tile rows split evenly, renumber them so that each XCD takes a band
    // of whole tile rows (every x tile of gridDim.y/8 consecutive y tiles) of
    // each z chunk: most halo points a tile reads then belong to tiles of the
    // same XCD and come from its L2 (1.13 -> 1.09 ms on 512x256x256).
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (gridDim.y % 8 == 0) {
        const int band = gridDim.y / 8;
        const int per = gridDim.x * band;  // tiles of one XCD in one z chunk
        const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        const int r = (b >> 3) % per;
        bz = (b >> 3) / per;
        bx = r % gridDim.x;
        by = (b & 7) * band + r / gridDim.x;
    }
    const int i = bx * VBX + tx + 1;
    const int j = by * VBY + ty + 1;
    const int k0 = bz * chunk + 1;
    const int k1 = min(k0 + chunk - 1, g.nz);
    const bool act = i <= g.nx && j <= g.ny;
    // loads stay inside the ghosted box; only inactive lanes see clamped values
    const int ii = min(i, g.nx + H), jj = min(j, g.ny + H);
    const size_t nv = g.nvol;
    const int sz = g.sxy;
    const int lmax = P.visc_lmax;
    const int col = gidx(g, ii, jj, 0);  // + k*sz
    // halo point of this lane (x halo for the first/last H lanes of a row,
    // y halo for the first/last H rows), as a column offset + LDS slot
    // halo point of this lane, as an LDS slot (hly, hlx) and a column
    // offset: lanes 0 .. 2H*VBX-1 take the y halo rows (H above the tile, H
    // below), the next 2H*VBY lanes the x halo columns, the rest none
    const int tid = ty * VBX + tx;
    const bool hl = tid < 2 * H * (VBX + VBY);
    int hlx = 0, hly = 0;
    if (tid < 2 * H * VBX) {
        const int r = tid / VBX;
        hly = r < H ? r : r + VBY;
        hlx = tid % VBX + H;
    } else {
        const int q = tid - 2 * H * VBX, c6 = q % (2 * H);
        hly = q / (2 * H) + H;
        hlx = c6 < H ? c6 : c6 + VBX;
    }
    const int colh = gidx(g, min(bx * VBX + 1 - H + hlx, g.nx + H), min(by * VBY + 1 - H + hly, g.ny + H), 0);
    VP<R> win[2 * H + 1];
#pragma unroll
    for (int q = 0; q < 2 * H; q++) win[q + 1] = load_vp(wa, nv, col + (k0 - H + q) * sz);
    VP<R> nxt = load_vp(wa, nv, col + (k0 + H) * sz);
    // The halo points of plane k are the centre points of the neighbouring
    // tiles, which those load H planes ahead, into their windows.  Loaded
    // one plane ahead, the halo lines were fetched by the neighbours three
    // planes earlier and have left L2 (the halos were re-fetched from HBM,
    // 1.4x the kernel's algorithmic reads).  The FP64 form therefore loads
    // them KH = H planes ahead, in step with the neighbours, and keeps KH
    // planes of its halo point in registers; the FP32 form, at its 128-VGPR
    // bound, loads them one plane ahead (KH = 1).
    constexpr int KH = sizeof(R) == 8 ? H : 1;
    VP<R> hv[KH];  // halo point of planes k .. k+KH-1 at the top of iteration k
#pragma unroll
    for (int q = 0; q < KH; q++) hv[q] = load_vp(wa, nv, (hl ? colh : col) + min(k0 + q, g.nz + H) * sz);
    const R ax = R(dx1[min(i, g.nx) - 1]), ay = R(dy1[min(j, g.ny) - 1]);
    const R sx_ = R(dxs[min(i, g.nx) - 1]), sy_ = R(dys[min(j, g.ny) - 1]);
    const R qx = R(dx2[min(i, g.nx) - 1]), qy = R(dy2[min(j, g.ny) - 1]);
    int buf = 0;
    for (int k = k0; k <= k1; k++, buf ^= 1) {
#pragma unroll
        for (int q = 0; q < 2 * H; q++) win[q] = win[q + 1];
        win[2 * H] = nxt;
        const VP<R> &c0 = win[H];
        R(*tile)[VBY + 2 * H][VBX + 2 * H] = tiles[buf];
        tile[0][ty + H][tx + H] = c0.u;
        tile[1][ty + H][tx + H] = c0.v;
        tile[2][ty + H][tx + H] = c0.w;
        tile[3][ty + H][tx + H] = c0.t;
        tile[4][ty + H][tx + H] = c0.m;
        if (hl) {
            tile[0][hly][hlx] = hv[0].u; tile[1][hly][hlx] = hv[0].v; tile[2][hly][hlx] = hv[0].w;
            tile[3][hly][hlx] = hv[0].t; tile[4][hly][hlx] = hv[0].m;
        }
        // right-hand side of this plane, read ahead of the stencil work
        const int oi = iidx(g, min(i, g.nx), min(j, g.ny), k);
        const size_t ni = g.nint;
        const R fo2 = fl[ni + oi], fo3 = fl[2 * ni + oi], fo4 = fl[3 * ni + oi], fo5 = fl[4 * ni + oi];
        // loads of plane k+1 (centre) and k+KH (halo) overlap the work on plane k
#pragma unroll
        for (int q = 0; q + 1 < KH; q++) hv[q] = hv[q + 1];
        if (k < k1) nxt = load_vp(wa, nv, col + (k + 1 + H) * sz);
        if (hl && k + KH <= k1) hv[KH - 1] = load_vp(wa, nv, colh + (k + KH) * sz);
        __syncthreads();
        if (act) {
            const R uu = c0.u, vv = c0.v, ww = c0.w, tt = c0.t, mu = c0.m;
            R ux = R(0), vx = R(0), wx = R(0), tx_ = R(0), mux = R(0);
            R uy = R(0), vy = R(0), wy = R(0), ty_ = R(0), muy = R(0);
            R uz = R(0), vz = R(0), wz = R(0), tz = R(0), muz = R(0);
            const R c20 = R(P.vd2[0]);
            R ulapx = c20 * uu, ulapy = ulapx, ulapz = ulapx;
            R vlapx = c20 * vv, vlapy = vlapx, vlapz = vlapx;
            R wlapx = c20 * ww, wlapy = wlapx, wlapz = wlapx;
            R tlapx = c20 * tt, tlapy = tlapx, tlapz = tlapx;
#pragma unroll
            for (int l = 1; l <= H; l++) {
                if (l <= lmax) {
                    const R ccl = R(P.vd1[l - 1]), clapl = R(P.vd2[l]);
                    const int r = ty + H, cp = tx + H + l, cm = tx + H - l;
                    ux = ux + ccl * (tile[0][r][cp] - tile[0][r][cm]);
                    ulapx = ulapx + clapl * (tile[0][r][cp] + tile[0][r][cm]);
                    vx = vx + ccl * (tile[1][r][cp] - tile[1][r][cm]);
                    vlapx = vlapx + clapl * (tile[1][r][cp] + tile[1][r][cm]);
                    wx = wx + ccl * (tile[2][r][cp] - tile[2][r][cm]);
                    wlapx = wlapx + clapl * (tile[2][r][cp] + tile[2][r][cm]);
                    tx_ = tx_ + ccl * (tile[3][r][cp] - tile[3][r][cm]);
                    tlapx = tlapx + clapl * (tile[3][r][cp] + tile[3][r][cm]);
                    mux = mux + ccl * (tile[4][r][cp] - tile[4][r][cm]);
                }
            }
#pragma unroll
            for (int l = 1; l <= H; l++) {
                if (l <= lmax) {
                    const R ccl = R(P.vd1[l - 1]), clapl = R(P.vd2[l]);
                    const int c = tx + H, rp = ty + H + l, rm = ty + H - l;
                    uy = uy + ccl * (tile[0][rp][c] - tile[0][rm][c]);
                    ulapy = ulapy + clapl * (tile[0][rp][c] + tile[0][rm][c]);
                    vy = vy + ccl * (tile[1][rp][c] - tile[1][rm][c]);
                    vlapy = vlapy + clapl * (tile[1][rp][c] + tile[1][rm][c]);
                    wy = wy + ccl * (tile[2][rp][c] - tile[2][rm][c]);
                    wlapy = wlapy + clapl * (tile[2][rp][c] + tile[2][rm][c]);
                    ty_ = ty_ + ccl * (tile[3][rp][c] - tile[3][rm][c]);
                    tlapy = tlapy + clapl * (tile[3][rp][c] + tile[3][rm][c]);
                    muy = muy + ccl * (tile[4][rp][c] - tile[4][rm][c]);
                }
            }
#pragma unroll
            for (int l = 1; l <= H; l++) {
                if (l <= lmax) {
                    const R ccl = R(P.vd1[l - 1]), clapl = R(P.vd2[l]);
                    const VP<R> &pp = win[H + l], &pm = win[H - l];
                    uz = uz + ccl * (pp.u - pm.u);
                    ulapz = ulapz + clapl * (pp.u + pm.u);
                    vz = vz + ccl * (pp.v - pm.v);
                    vlapz = vlapz + clapl * (pp.v + pm.v);
                    wz = wz + ccl * (pp.w - pm.w);
                    wlapz = wlapz + clapl * (pp.w + pm.w);
                    tz = tz + ccl * (pp.t - pm.t);
                    tlapz = tlapz + clapl * (pp.t + pm.t);
                    muz = muz + ccl * (pp.m - pm.m);
                }
            }
            const R az = R(dz1[k - 1]);
            ux = ux * ax; vx = vx * ax; wx = wx * ax; tx_ = tx_ * ax; mux = mux * ax;
            uy = uy * ay; vy = vy * ay; wy = wy * ay; ty_ = ty_ * ay; muy = muy * ay;
            uz = uz * az; vz = vz * az; wz = wz * az; tz = tz * az; muz = muz * az;
            const R sz_ = R(dzs[k - 1]), qz = R(dz2[k - 1]);
            ulapx = ulapx * sx_ + ux * qx; vlapx = vlapx * sx_ + vx * qx;
            wlapx = wlapx * sx_ + wx * qx; tlapx = tlapx * sx_ + tx_ * qx;
            ulapy = ulapy * sy_ + uy * qy; vlapy = vlapy * sy_ + vy * qy;
            wlapy = wlapy * sy_ + wy * qy; tlapy = tlapy * sy_ + ty_ * qy;
            ulapz = ulapz * sz_ + uz * qz; vlapz = vlapz * sz_ + vz * qz;
            wlapz = wlapz * sz_ + wz * qz; tlapz = tlapz * sz_ + tz * qz;
            const R ulap = ulapx + ulapy + ulapz;
            const R vlap = vlapx + vlapy + vlapz;
            const R wlap = wlapx + wlapy + wlapz;
            const R tlap = tlapx + tlapy + tlapz;
            const R div = ux + vy + wz;
            const R div3l = div / R(3.0);
            const int c = gidx(g, i, j, k);
            wa[9 * nv + c] = div3l;
            const R omegax = wy - vz, omegay = uz - wx, omegaz = vx - uy;
            const R omod2 = omegax * omegax + omegay * omegay + omegaz * omegaz;
            wa[8 * nv + c] = sqrt(omod2);
            if (SENSOR) {
                const R u0l0 = R(P.u0 / P.l0);
                R sv = -div / sqrt(omod2 + div * div + u0l0 * u0l0);
                sv = fmax(sv, R(0.0));
                wa[7 * nv + c] = sv * sv;
                if (images) write_images(g, P, wa + 7 * nv, i, j, k, sv * sv, false);
            }
            if (images && !SENSOR) write_images(g, P, wa + 9 * nv, i, j, k, div3l, true);
            const R cploc = R(cploc_of(P, double(tt)));
            const R sig11 = R(2.0) * (ux - div3l), sig12 = uy + vx, sig13 = uz + wx;
            const R sig22 = R(2.0) * (vy - div3l), sig23 = vz + wy, sig33 = R(2.0) * (wz - div3l);
            const R sigx = mux * sig11 + muy * sig12 + muz * sig13 + mu * ulap;
            const R sigy = mux * sig12 + muy * sig22 + muz * sig23 + mu * vlap;
            const R sigz = mux * sig13 + muy * sig23 + muz * sig33 + mu * wlap;
            const R sigqt = (mux * tx_ + muy * ty_ + muz * tz + mu * tlap) * cploc / R(P.prandtl);
            const R sigah = (sig11 * ux + sig12 * uy + sig13 * uz + sig12 * vx + sig22 * vy + sig23 * vz +
                                  sig13 * wx + sig23 * wy + sig33 * wz) * mu;
            const R sigq = sigx * uu + sigy * vv + sigz * ww + sigah + sigqt;
            fl[ni + oi] = fo2 - sigx;
            fl[2 * ni + oi] = fo3 - sigy;
            fl[3 * ni + oi] = fo4 - sigz;
            fl[4 * ni + oi] = fo5 - sigq;
            if (j == 1) {
                const size_t wsz = (size_t)g.sx * g.sz;
                const size_t wp = (size_t)(i + g.ng - 1) + (size_t)g.sx * (k + g.ng - 1);
                wallprop[wp] = double(mu * uy);
                wallprop[wsz + wp] = double(mu * wy);
                wallprop[2 * wsz + wp] = double(mu * ty_ * cploc / R(P.prandtl));
            }
        }
    }
}
template <typename R, bool SENSOR>
__global__ __launch_bounds__(VBX *VBY) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 1))) void k_visc_zm(Geo g, KPar P, R *__restrict__ wa, R *__restrict__ fl,
                                                      const double *__restrict__ dx1, const double *__restrict__ dxs,
                                                      const double *__restrict__ dx2, const double *__restrict__ dy1,
                                                      const double *__restrict__ dys, const double *__restrict__ dy2,
                                                      const double *__restrict__ dz1, const double *__restrict__ dzs,
                                                      const double *__restrict__ dz2, double *__restrict__ wallprop,
                                                      int images, int chunk) {
    k_visc_zm_body<R, SENSOR>(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), g, P, wa, fl, dx1, dxs, dx2, dy1, dys, dy2, dz1, dzs, dz2, wallprop, images, chunk);
}

// visflx (mode 0, with SENSOR) and visflx_nosensor (mode 4) for any viscous
// order up to 8 (lmax <= 4): one thread per interior cell, the stencil read
// from w_aux directly.  It serves four ghost layers (visc_order 8); the
// z-marching tile kernel above serves lmax <= 3.  Same arithmetic, in the
// same order, as k_visc_zm.
template <typename R, bool SENSOR>
__global__ __launch_bounds__(256) void k_visc_gen(Geo g, KPar P, R *__restrict__ wa, R *__restrict__ fl,
                                                  const double *__restrict__ dx1, const double *__restrict__ dxs,
                                                  const double *__restrict__ dx2, const double *__restrict__ dy1,
                                                  const double *__restrict__ dys, const double *__restrict__ dy2,
                                                  const double *__restrict__ dz1, const double *__restrict__ dzs,
                                                  const double *__restrict__ dz2, double *__restrict__ wallprop,
                                                  int images) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const size_t nv = g.nvol;
    const int c = gidx(g, i, j, k);
    const int lmax = P.visc_lmax;
    const int stride[3] = {1, g.sx, g.sxy};
    const R uu = wa[nv + c], vv = wa[2 * nv + c], ww = wa[3 * nv + c], tt = wa[5 * nv + c], mu = wa[6 * nv + c];
    const R c20 = R(P.vd2[0]);
    // per direction: first derivatives of u, v, w, T, mu and the
    // second-derivative sums of u, v, w, T
    R d[3][5], lap[3][4];
    for (int dir = 0; dir < 3; dir++) {
        for (int q = 0; q < 5; q++) d[dir][q] = R(0);
        lap[dir][0] = c20 * uu; lap[dir][1] = c20 * vv; lap[dir][2] = c20 * ww; lap[dir][3] = c20 * tt;
        const int st = stride[dir];
        for (int l = 1; l <= 4; l++) {
            if (l <= lmax) {
                const R ccl = R(P.vd1[l - 1]), clapl = R(P.vd2[l]);
                const int cp = c + l * st, cm = c - l * st;
                const R up = wa[nv + cp], um = wa[nv + cm];
                const R vp = wa[2 * nv + cp], vm = wa[2 * nv + cm];
                const R wp = wa[3 * nv + cp], wm = wa[3 * nv + cm];
                const R tp = wa[5 * nv + cp], tm = wa[5 * nv + cm];
                const R mp = wa[6 * nv + cp], mm = wa[6 * nv + cm];
                d[dir][0] = d[dir][0] + ccl * (up - um);
                lap[dir][0] = lap[dir][0] + clapl * (up + um);
                d[dir][1] = d[dir][1] + ccl * (vp - vm);
                lap[dir][1] = lap[dir][1] + clapl * (vp + vm);
                d[dir][2] = d[dir][2] + ccl * (wp - wm);
                lap[dir][2] = lap[dir][2] + clapl * (wp + wm);
                d[dir][3] = d[dir][3] + ccl * (tp - tm);
                lap[dir][3] = lap[dir][3] + clapl * (tp + tm);
                d[dir][4] = d[dir][4] + ccl * (mp - mm);
            }
        }
    }
    const R ax = R(dx1[i - 1]), ay = R(dy1[j - 1]), az = R(dz1[k - 1]);
    const R sa[3] = {R(dxs[i - 1]), R(dys[j - 1]), R(dzs[k - 1])};
    const R qa[3] = {R(dx2[i - 1]), R(dy2[j - 1]), R(dz2[k - 1])};
    const R a1[3] = {ax, ay, az};
    for (int dir = 0; dir < 3; dir++) {
        for (int q = 0; q < 5; q++) d[dir][q] = d[dir][q] * a1[dir];
        for (int q = 0; q < 4; q++) lap[dir][q] = lap[dir][q] * sa[dir] + d[dir][q] * qa[dir];
    }
    const R ux = d[0][0], vx = d[0][1], wx = d[0][2], tx_ = d[0][3], mux = d[0][4];
    const R uy = d[1][0], vy = d[1][1], wy = d[1][2], ty_ = d[1][3], muy = d[1][4];
    const R uz = d[2][0], vz = d[2][1], wz = d[2][2], tz = d[2][3], muz = d[2][4];
    const R ulap = lap[0][0] + lap[1][0] + lap[2][0];
    const R vlap = lap[0][1] + lap[1][1] + lap[2][1];
    const R wlap = lap[0][2] + lap[1][2] + lap[2][2];
    const R tlap = lap[0][3] + lap[1][3] + lap[2][3];
    const R div = ux + vy + wz;
    const R div3l = div / R(3.0);
    wa[9 * nv + c] = div3l;
    const R omegax = wy - vz, omegay = uz - wx, omegaz = vx - uy;
    const R omod2 = omegax * omegax + omegay * omegay + omegaz * omegaz;
    wa[8 * nv + c] = sqrt(omod2);
    if (SENSOR) {
        const R u0l0 = R(P.u0 / P.l0);
        R sv = -div / sqrt(omod2 + div * div + u0l0 * u0l0);
        sv = fmax(sv, R(0.0));
        wa[7 * nv + c] = sv * sv;
        if (images) write_images(g, P, wa + 7 * nv, i, j, k, sv * sv, false);
    }
    if (images && !SENSOR) write_images(g, P, wa + 9 * nv, i, j, k, div3l, true);
    const R cploc = R(cploc_of(P, double(tt)));
    const R sig11 = R(2.0) * (ux - div3l), sig12 = uy + vx, sig13 = uz + wx;
    const R sig22 = R(2.0) * (vy - div3l), sig23 = vz + wy, sig33 = R(2.0) * (wz - div3l);
    const R sigx = mux * sig11 + muy * sig12 + muz * sig13 + mu * ulap;
    const R sigy = mux * sig12 + muy * sig22 + muz * sig23 + mu * vlap;
    const R sigz = mux * sig13 + muy * sig23 + muz * sig33 + mu * wlap;
    const R sigqt = (mux * tx_ + muy * ty_ + muz * tz + mu * tlap) * cploc / R(P.prandtl);
    const R sigah = (sig11 * ux + sig12 * uy + sig13 * uz + sig12 * vx + sig22 * vy + sig23 * vz + sig13 * wx +
                     sig23 * wy + sig33 * wz) * mu;
    const R sigq = sigx * uu + sigy * vv + sigz * ww + sigah + sigqt;
    const int oi = iidx(g, i, j, k);
    const size_t ni = g.nint;
    fl[ni + oi] = fl[ni + oi] - sigx;
    fl[2 * ni + oi] = fl[2 * ni + oi] - sigy;
    fl[3 * ni + oi] = fl[3 * ni + oi] - sigz;
    fl[4 * ni + oi] = fl[4 * ni + oi] - sigq;
    if (j == 1) {
        const size_t wsz = (size_t)g.sx * g.sz;
        const size_t wp = (size_t)(i + g.ng - 1) + (size_t)g.sx * (k + g.ng - 1);
        wallprop[wp] = double(mu * uy);
        wallprop[wsz + wp] = double(mu * wy);
        wallprop[2 * wsz + wp] = double(mu * ty_ * cploc / R(P.prandtl));
    }
}

// ------------------------------------------------ conservative viscous form
// Face flux between points a (left) and b (right) at node distance h
// (kernels_cpu.F90:4805-4836): mu_f (du, dv, dw) and
// mu_f (cp/Pr dT + d(|u|^2/2)), mu_f = (mu_a + mu_b)/2/h.
template <typename R>
__device__ __forceinline__ void cons_face(const KPar &P, const VP<R> &a, const VP<R> &b, R h, R s[4]) {
    const R qq = R(0.5) * (a.u * a.u + a.v * a.v + a.w * a.w);
    const R qqp = R(0.5) * (b.u * b.u + b.v * b.v + b.w * b.w);
    const R cploc = R(P.calorically_perfect == 1 ? P.cp[0 - P.indx_cp_l] : cploc_of(P, 0.5 * double(a.t + b.t)));
    R muf = a.m + b.m;
    muf = R(0.5) * muf / h;
    s[0] = (b.u - a.u) * muf;
    s[1] = (b.v - a.v) * muf;
    s[2] = (b.w - a.w) * muf;
    s[3] = ((b.t - a.t) * cploc / R(P.prandtl) + (qqp - qq)) * muf;
}

// Conservative viscous terms of one cell.
//   CELL:  visflx_reduced_ord2_subroutine (kernels_cpu.F90:4623-4721,
//          visflx modes 2/6): grad(mu).sigma + dissipation, div/3 and |omega|
//          into w_aux(10)/(9), wall shear into wallprop, Ducros sensor when
//          SENSOR (mode 2);
//   FACES: visflx_x/y/z_subroutine (:4779-4999, mode 5), the two x faces,
//          then the y faces, then the z faces of the cell, with the
//          reference's per-cell summation order.
// ``images`` also writes the ghost images of w_aux(10) (bcextr_var mode 1
// on walls, bcswap_var on periodic sides).
template <typename R, bool CELL, bool FACES, bool SENSOR>
__global__ __launch_bounds__(256) void k_visc_cons(Geo g, KPar P, R *__restrict__ wa, R *__restrict__ fl,
                                                   const double *__restrict__ xc, const double *__restrict__ yc,
                                                   const double *__restrict__ zc, double *__restrict__ wallprop,
                                                   int images) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const size_t nv = g.nvol;
    const int c = gidx(g, i, j, k);
    const int sy = g.sx, sz = g.sxy;
    const VP<R> p0 = load_vp(wa, nv, c);
    const VP<R> xm = load_vp(wa, nv, c - 1), xp = load_vp(wa, nv, c + 1);
    const VP<R> ym = load_vp(wa, nv, c - sy), yp = load_vp(wa, nv, c + sy);
    const VP<R> zm = load_vp(wa, nv, c - sz), zp = load_vp(wa, nv, c + sz);
    const int ng = g.ng;
    const double xi_m = xc[i - 1 + ng - 1], xi = xc[i + ng - 1], xi_p = xc[i + 1 + ng - 1];
    const double yj_m = yc[j - 1 + ng - 1], yj = yc[j + ng - 1], yj_p = yc[j + 1 + ng - 1];
    const double zk_m = zc[k - 1 + ng - 1], zk = zc[k + ng - 1], zk_p = zc[k + 1 + ng - 1];
    const int oi = iidx(g, i, j, k);
    const size_t ni = g.nint;
    R f2 = fl[ni + oi], f3 = fl[2 * ni + oi], f4 = fl[3 * ni + oi], f5 = fl[4 * ni + oi];
    if (CELL) {
        const R cploc = R(cploc_of(P, double(p0.t)));
        const R dxl = R(1.0) / R(xi_p - xi_m);
        const R dyl = R(1.0) / R(yj_p - yj_m);
        const R dzl = R(1.0) / R(zk_p - zk_m);
        const R ux = dxl * (xp.u - xm.u), vx = dxl * (xp.v - xm.v), wx = dxl * (xp.w - xm.w);
        const R mux = dxl * (xp.m - xm.m);
        const R uy = dyl * (yp.u - ym.u), vy = dyl * (yp.v - ym.v), wy = dyl * (yp.w - ym.w);
        const R ty = dyl * (yp.t - ym.t), muy = dyl * (yp.m - ym.m);
        const R uz = dzl * (zp.u - zm.u), vz = dzl * (zp.v - zm.v), wz = dzl * (zp.w - zm.w);
        const R muz = dzl * (zp.m - zm.m);
        const R mu = p0.m;
        if (j == 1) {
            const size_t wsz = (size_t)g.sx * g.sz;
            const size_t wp = (size_t)(i + ng - 1) + (size_t)g.sx * (k + ng - 1);
            wallprop[wp] = double(mu * uy);
            wallprop[wsz + wp] = double(mu * wy);
            wallprop[2 * wsz + wp] = double(mu * ty * cploc / R(P.prandtl));
        }
        const R div = ux + vy + wz;
        const R div3l = div / R(3.0);
        wa[9 * nv + c] = div3l;
        if (images) write_images(g, P, wa + 9 * nv, i, j, k, div3l, true);
        const R omegax = wy - vz, omegay = uz - wx, omegaz = vx - uy;
        const R omod2 = omegax * omegax + omegay * omegay + omegaz * omegaz;
        wa[8 * nv + c] = sqrt(omod2);
        if (SENSOR) {
            const R u0l0 = R(P.u0 / P.l0);
            R sv = -div / sqrt(omod2 + div * div + u0l0 * u0l0);
            sv = fmax(sv, R(0.0));
            wa[7 * nv + c] = sv * sv;
        }
        const R sig11 = ux - R(2.0) * div3l, sig12 = vx, sig13 = wx;
        const R sig21 = uy, sig22 = vy - R(2.0) * div3l, sig23 = wy;
        const R sig31 = uz, sig32 = vz, sig33 = wz - R(2.0) * div3l;
        const R sigx = mux * sig11 + muy * sig12 + muz * sig13;
        const R sigy = mux * sig21 + muy * sig22 + muz * sig23;
        const R sigz = mux * sig31 + muy * sig32 + muz * sig33;
        const R sigah = (sig11 * ux + sig12 * uy + sig13 * uz + sig21 * vx + sig22 * vy + sig23 * vz +
                              sig31 * wx + sig32 * wy + sig33 * wz) * mu;
        const R sigq = sigx * p0.u + sigy * p0.v + sigz * p0.w + sigah;
        f2 = f2 - sigx;
        f3 = f3 - sigy;
        f4 = f4 - sigz;
        f5 = f5 - sigq;
    }
    if (FACES) {
        R sl[4], sr[4];
        // x: fhat = -flux, fl += dxhl*(fhat(i) - fhat(i-1))
        cons_face(P, xm, p0, R(xi - xi_m), sl);
        cons_face(P, p0, xp, R(xi_p - xi), sr);
        const R dxhl = R(2.0) / R(xi_p - xi_m);
        f2 = f2 + dxhl * (-sr[0] - -sl[0]);
        f3 = f3 + dxhl * (-sr[1] - -sl[1]);
        f4 = f4 + dxhl * (-sr[2] - -sl[2]);
        f5 = f5 + dxhl * (-sr[3] - -sl[3]);
        // y: fl = fl + flux(j-1/2)*dyhl - flux(j+1/2)*dyhl
        cons_face(P, ym, p0, R(yj - yj_m), sl);
        cons_face(P, p0, yp, R(yj_p - yj), sr);
        const R dyhl = R(2.0) / R(yj_p - yj_m);
        f2 = f2 + sl[0] * dyhl - sr[0] * dyhl;
        f3 = f3 + sl[1] * dyhl - sr[1] * dyhl;
        f4 = f4 + sl[2] * dyhl - sr[2] * dyhl;
        f5 = f5 + sl[3] * dyhl - sr[3] * dyhl;
        cons_face(P, zm, p0, R(zk - zk_m), sl);
        cons_face(P, p0, zp, R(zk_p - zk), sr);
        const R dzhl = R(2.0) / R(zk_p - zk_m);
        f2 = f2 + sl[0] * dzhl - sr[0] * dzhl;
        f3 = f3 + sl[1] * dzhl - sr[1] * dzhl;
        f4 = f4 + sl[2] * dzhl - sr[2] * dzhl;
        f5 = f5 + sl[3] * dzhl - sr[3] * dzhl;
    }
    fl[ni + oi] = f2;
    fl[2 * ni + oi] = f3;
    fl[3 * ni + oi] = f4;
    fl[4 * ni + oi] = f5;
}

// sensor_subroutine (kernels_cpu.F90:4726-4744) + bcswap_var of w_aux(8)
template <typename R>
__device__ __forceinline__ void k_sensor_body(const dim3 blockIdx, const dim3 gridDim, Geo g, KPar P, R *__restrict__ wa, int images) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const size_t nv = g.nvol;
    const int c = gidx(g, i, j, k);
    const R om = wa[8 * nv + c];
    const R omod2 = om * om;
    const R div = R(3.0) * wa[9 * nv + c];
    const R u0l0 = R(P.u0 / P.l0);
    R sv = -div / sqrt(omod2 + div * div + u0l0 * u0l0);
    sv = fmax(sv, R(0.0));
    wa[7 * nv + c] = sv * sv;
    if (images) write_images(g, P, wa + 7 * nv, i, j, k, sv * sv, false);
}
template <typename R>
__global__ __launch_bounds__(256) void k_sensor(Geo g, KPar P, R *__restrict__ wa, int images) {
    k_sensor_body<R>(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), g, P, wa, images);
}

// Ducros sensor in FP64 from the conserved state w at interior cell
// (i, j, k), as the FP64 form computes it: velocities u = rho*u * (1/rho) as
// eval_aux forms them (periodic neighbours from their images), the first
// derivatives of the viscous kernel in the same order, then the sensor
// either directly (DIRECT: as visflx mode 0 does at the start of a run,
// kernels_cpu.F90:3598-3777) or through |omega| and div/3 as they are
// stored (sensor_subroutine, :4726-4744, at the third stage).
template <bool DIRECT>
__device__ double sensor_fp64(const Geo &g, const KPar &P, const double *__restrict__ w,
                              const double *__restrict__ dx1, const double *__restrict__ dy1,
                              const double *__restrict__ dz1, int i, int j, int k) {
    const size_t nv = g.nvol;
    auto at = [&](int ii, int jj, int kk) {
        if (P.periodic[0]) ii = ii < 1 ? ii + g.nx : (ii > g.nx ? ii - g.nx : ii);
        if (P.periodic[1]) jj = jj < 1 ? jj + g.ny : (jj > g.ny ? jj - g.ny : jj);
        if (P.periodic[2]) kk = kk < 1 ? kk + g.nz : (kk > g.nz ? kk - g.nz : kk);
        return gidx(g, ii, jj, kk);
    };
    auto vel = [&](int c, int m) {
        const double ri = 1.0 / w[c];
        return w[m * nv + c] * ri;
    };
    double ux = 0, vx = 0, wx = 0, uy = 0, vy = 0, wy = 0, uz = 0, vz = 0, wz = 0;
    const int lmax = P.visc_lmax;
    for (int l = 1; l <= lmax; l++) {
        const double ccl = P.vd1[l - 1];
        const int xp = at(i + l, j, k), xm = at(i - l, j, k);
        ux = ux + ccl * (vel(xp, 1) - vel(xm, 1));
        vx = vx + ccl * (vel(xp, 2) - vel(xm, 2));
        wx = wx + ccl * (vel(xp, 3) - vel(xm, 3));
    }
    for (int l = 1; l <= lmax; l++) {
        const double ccl = P.vd1[l - 1];
        const int yp = at(i, j + l, k), ym = at(i, j - l, k);
        uy = uy + ccl * (vel(yp, 1) - vel(ym, 1));
        vy = vy + ccl * (vel(yp, 2) - vel(ym, 2));
        wy = wy + ccl * (vel(yp, 3) - vel(ym, 3));
    }
    for (int l = 1; l <= lmax; l++) {
        const double ccl = P.vd1[l - 1];
        const int zp = at(i, j, k + l), zm = at(i, j, k - l);
        uz = uz + ccl * (vel(zp, 1) - vel(zm, 1));
        vz = vz + ccl * (vel(zp, 2) - vel(zm, 2));
        wz = wz + ccl * (vel(zp, 3) - vel(zm, 3));
    }
    const double ax = dx1[i - 1], ay = dy1[j - 1], az = dz1[k - 1];
    ux = ux * ax; vx = vx * ax; wx = wx * ax;
    uy = uy * ay; vy = vy * ay; wy = wy * ay;
    uz = uz * az; vz = vz * az; wz = wz * az;
    double div = ux + vy + wz;
    const double omegax = wy - vz, omegay = uz - wx, omegaz = vx - uy;
    double omod2 = omegax * omegax + omegay * omegay + omegaz * omegaz;
    if (!DIRECT) {
        const double om = sqrt(omod2);
        const double div3l = div / 3.0;
        omod2 = om * om;
        div = 3.0 * div3l;
    }
    const double u0l0 = P.u0 / P.l0;
    double sv = -div / sqrt(omod2 + div * div + u0l0 * u0l0);
    sv = fmax(sv, 0.0);
    return sv * sv;
}

// Shock-sensor refinement of the mixed form.  The FP32 sensor decides the
// central/WENO switch like the FP64 one except within its own round-off of
// the threshold, so the cells whose FP32 value lies within kSensorBand of
// the threshold take the FP64 value of sensor_fp64 and store it snapped
// to the side of the threshold that value is on: the sweeps compare the
// stored value with the FP32 threshold, the WENO fix-up with the FP64 one,
// and both then give the FP64 form's decision.  Far from the threshold the
// FP32 error (well below 1e-5) cannot change a decision.
constexpr double kSensorBand = 1e-3;

template <bool DIRECT>
__device__ __forceinline__ void k_sensor_refine_body(const dim3 blockIdx, const dim3 gridDim, Geo g, KPar P, const double *__restrict__ w,
                                                       float *__restrict__ wa, const double *__restrict__ dx1,
                                                       const double *__restrict__ dy1,
                                                       const double *__restrict__ dz1, int images) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const size_t nv = g.nvol;
    const int c = gidx(g, i, j, k);
    const double thr = P.sensor_threshold;
    if (fabs(double(wa[7 * nv + c]) - thr) >= kSensorBand) return;
    const double s64 = sensor_fp64<DIRECT>(g, P, w, dx1, dy1, dz1, i, j, k);
    const float rt = float(thr);
    float st = float(s64);
    if (s64 > thr) {
        st = fmaxf(st, nextafterf(rt, INFINITY));          // > thr in both precisions
    } else {
        const float lo = double(rt) > thr ? nextafterf(rt, -INFINITY) : rt;
        st = fminf(st, lo);                                   // <= thr in both precisions
    }
    wa[7 * nv + c] = st;
    if (images) write_images(g, P, wa + 7 * nv, i, j, k, st, false);
}
template <bool DIRECT>
__global__ __launch_bounds__(256) void k_sensor_refine(Geo g, KPar P, const double *__restrict__ w,
                                                       float *__restrict__ wa, const double *__restrict__ dx1,
                                                       const double *__restrict__ dy1,
                                                       const double *__restrict__ dz1, int images) {
    k_sensor_refine_body<DIRECT>(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), g, P, w, wa, dx1, dy1, dz1, images);
}

// visflx_div_subroutine (kernels_cpu.F90:3484-3532).  With REDUCE the
// kernel also forms fln = -rho_k dt F[prev] - gamma_k dt F[cur] of
// update_flux (:3400-3418) for rho and rho*u and writes the block partial
// sums of force_rhs_1_subroutine (:3157-3193).
template <typename R, bool REDUCE>
__global__ __launch_bounds__(256) void k_visc_div(Geo g, KPar P, const R *__restrict__ wa, R *__restrict__ fl,
                                                  const R *__restrict__ flp, const double *__restrict__ w,
                                                  const double *__restrict__ dx1, const double *__restrict__ dy1,
                                                  const double *__restrict__ dz1, const double *__restrict__ yn,
                                                  const DevScalars *__restrict__ ds, double rkr, double rkg,
                                                  double *__restrict__ partials) {
    __shared__ double lds[4][5];
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (i <= g.nx && j <= g.ny) {
        const size_t nv = g.nvol;
        const int c = gidx(g, i, j, k);
        const R *D = wa + 9 * nv;
        const R uu = wa[nv + c], vv = wa[2 * nv + c], ww = wa[3 * nv + c], mu = wa[6 * nv + c];
        R dxv = R(0.0), dyv = R(0.0), dzv = R(0.0);
        const int sy = g.sx, sz = g.sxy;
#pragma unroll
        for (int l = 1; l <= 4; l++) {
            if (l <= P.visc_lmax) {
                const R ccl = R(P.vd1[l - 1]);
                dxv = dxv + ccl * (D[c + l] - D[c - l]);
                dyv = dyv + ccl * (D[c + l * sy] - D[c - l * sy]);
                dzv = dzv + ccl * (D[c + l * sz] - D[c - l * sz]);
            }
        }
        dxv = dxv * R(dx1[i - 1]);
        dyv = dyv * R(dy1[j - 1]);
        dzv = dzv * R(dz1[k - 1]);
        const R sigx = mu * dxv, sigy = mu * dyv, sigz = mu * dzv;
        const R sigq = sigx * uu + sigy * vv + sigz * ww;
        const int oi = iidx(g, i, j, k);
        const size_t ni = g.nint;
        const R f2 = fl[ni + oi] - sigx;
        fl[ni + oi] = f2;
        fl[2 * ni + oi] -= sigy;
        fl[3 * ni + oi] -= sigz;
        fl[4 * ni + oi] -= sigq;
        if (REDUCE) {
            const double dt = ds->dt;
            const double rhodt = rkr * dt, gamdt = rkg * dt;
            const double dy = yn[j] - yn[j - 1];
            const double fln1 = -rhodt * double(flp[oi]) - gamdt * double(fl[oi]);
            const double fln2 = -rhodt * double(flp[ni + oi]) - gamdt * double(f2);
            const double w1 = w[c], w2 = w[nv + c];
            acc[0] = fln1 * dy;
            acc[1] = fln2 * dy;
            acc[2] = w1 * dy;
            acc[3] = w2 * dy;
            acc[4] = w2 * dy * double(wa[5 * nv + c]);
        }
    }
    if (REDUCE) {
        block_sum<5>(acc, lds);
        if (flat_thread() == 0) {
            double *pp = partials + (size_t)flat_block() * 5;
#pragma unroll
            for (int q = 0; q < 5; q++) pp[q] = acc[q];
        }
    }
}

// visflx_div_ord2_subroutine (kernels_cpu.F90:3446-3480, visflx mode 7);
// REDUCE as in k_visc_div.
template <typename R, bool REDUCE>
__global__ __launch_bounds__(256) void k_visc_div2(Geo g, KPar P, const R *__restrict__ wa, R *__restrict__ fl,
                                                   const R *__restrict__ flp, const double *__restrict__ w,
                                                   const double *__restrict__ xc, const double *__restrict__ yc,
                                                   const double *__restrict__ zc, const double *__restrict__ yn,
                                                   const DevScalars *__restrict__ ds, double rkr, double rkg,
                                                   double *__restrict__ partials) {
    __shared__ double lds[4][5];
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (i <= g.nx && j <= g.ny) {
        const size_t nv = g.nvol;
        const int c = gidx(g, i, j, k);
        const R *D = wa + 9 * nv;
        const R uu = wa[nv + c], vv = wa[2 * nv + c], ww = wa[3 * nv + c], mu = wa[6 * nv + c];
        const int sy = g.sx, sz = g.sxy, ng = g.ng;
        const R dxl = mu / R(xc[i + 1 + ng - 1] - xc[i - 1 + ng - 1]);
        const R dyl = mu / R(yc[j + 1 + ng - 1] - yc[j - 1 + ng - 1]);
        const R dzl = mu / R(zc[k + 1 + ng - 1] - zc[k - 1 + ng - 1]);
        const R sigx = dxl * (D[c + 1] - D[c - 1]);
        const R sigy = dyl * (D[c + sy] - D[c - sy]);
        const R sigz = dzl * (D[c + sz] - D[c - sz]);
        const R sigq = sigx * uu + sigy * vv + sigz * ww;
        const int oi = iidx(g, i, j, k);
        const size_t ni = g.nint;
        const R f2 = fl[ni + oi] - sigx;
        fl[ni + oi] = f2;
        fl[2 * ni + oi] -= sigy;
        fl[3 * ni + oi] -= sigz;
        fl[4 * ni + oi] -= sigq;
        if (REDUCE) {
            const double dt = ds->dt;
            const double rhodt = rkr * dt, gamdt = rkg * dt;
            const double dy = yn[j] - yn[j - 1];
            const double fln1 = -rhodt * double(flp[oi]) - gamdt * double(fl[oi]);
            const double fln2 = -rhodt * double(flp[ni + oi]) - gamdt * double(f2);
            const double w1 = w[c], w2 = w[nv + c];
            acc[0] = fln1 * dy;
            acc[1] = fln2 * dy;
            acc[2] = w1 * dy;
            acc[3] = w2 * dy;
            acc[4] = w2 * dy * double(wa[5 * nv + c]);
        }
    }
    if (REDUCE) {
        block_sum<5>(acc, lds);
        if (flat_thread() == 0) {
            double *pp = partials + (size_t)flat_block() * 5;
#pragma unroll
            for (int q = 0; q < 5; q++) pp[q] = acc[q];
        }
    }
}

// Two-level sum of per-block partials: RB workgroups each add a fixed range
// of partials, the last workgroup to finish (atomic ticket) adds the RB
// results in index order, so the sum does not depend on scheduling.
constexpr int RB = 128;
static_assert(RB <= RED_THREADS, "the last block reads one second-level value per thread");

template <int NV>
__device__ __forceinline__ bool two_level_sum(const dim3 blockIdx, const dim3 gridDim, const double *__restrict__ partials,
                                              int nblk, double *__restrict__ stage2,
                                              int *__restrict__ ticket, double (&out)[NV], double (*lds)[NV]) {
    __shared__ bool last;
    const int per = (nblk + gridDim.x - 1) / gridDim.x;
    const int b0 = blockIdx.x * per, b1 = min(b0 + per, nblk);
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) acc[q] = 0.0;
    for (int b = b0 + (int)threadIdx.x; b < b1; b += blockDim.x) {
#pragma unroll
        for (int q = 0; q < NV; q++) acc[q] += partials[(size_t)b * NV + q];
    }
    block_sum<NV>(acc, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) stage2[blockIdx.x * NV + q] = acc[q];
        __threadfence();
        last = atomicAdd(ticket, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return false;
    // the last block: one second-level value per thread (gridDim.x <=
    // blockDim.x), read past the non-coherent caches, then the same
    // fixed-order tree as the first level
    __threadfence();
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; q++)
        v[q] = threadIdx.x < gridDim.x
                   ? __hip_atomic_load(&stage2[threadIdx.x * NV + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                   : 0.0;
    block_sum<NV>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) out[q] = v[q];
        *ticket = 0;
    }
    return threadIdx.x == 0;
}

// force_rhs (singleideal_cpu.F90:845-889): bulk sums -> dpdx, rho_b, u_b, T_b
__device__ __forceinline__ void k_force_finish_body(const dim3 blockIdx, const dim3 gridDim, KPar P, const double *__restrict__ partials, int nblk,
                                                              double *__restrict__ stage2, int *__restrict__ ticket,
                                                              DevScalars *ds, int first_stage) {
    __shared__ double lds[RED_THREADS / 64][5];
    double acc[5];
    if (two_level_sum<5>(blockIdx, gridDim, partials, nblk, stage2, ticket, acc, lds)) {
        const double v = P.volchan;
        const double b1 = acc[0] / v, b2 = acc[1] / v, b3 = acc[2] / v, b4 = acc[3] / v, b5 = acc[4] / v;
        ds->dpdx = (first_stage ? 0.0 : ds->dpdx) + b2;
        ds->rhobulk = b3;
        ds->ubulk = b4 / ds->rhobulk;
        ds->tbulk = b5 / ds->rhobulk / ds->ubulk;
        ds->b1 = b1;
        ds->b2 = b2;
    }
}
__global__ __launch_bounds__(RED_THREADS) void k_force_finish(KPar P, const double *__restrict__ partials, int nblk,
                                                              double *__restrict__ stage2, int *__restrict__ ticket,
                                                              DevScalars *ds, int first_stage) {
    k_force_finish_body(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), P, partials, nblk, stage2, ticket, ds, first_stage);
}

// update_flux + force_rhs_2 + update_field (kernels_cpu.F90:3400-3442,
// 3092-3116) fused, then the isothermal-wall ghosts of update_ghost
// (bcwall_staggered) from the new state when force_var is off.  With
// force_var on, also recovers T from the new E (force_var_1_subroutine,
// :3241-3278) and writes the block partial sums of rho*u*T*dy.
template <typename R, bool FORCE_VAR>
__device__ __forceinline__ void k_update_body(const dim3 blockIdx, const dim3 gridDim, Geo g, KPar P, double *__restrict__ w, R *__restrict__ wa,
                                                const R *__restrict__ flp, const R *__restrict__ fl,
                                                const DevScalars *__restrict__ ds, double rkr, double rkg, int channel,
                                                const double *__restrict__ yn, double *__restrict__ partials) {
    __shared__ double lds[4][1];
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    double acc[1] = {0.0};
    if (i <= g.nx && j <= g.ny) {
        const size_t nv = g.nvol, ni = g.nint;
        const int c = gidx(g, i, j, k);
        const int oi = iidx(g, i, j, k);
        const double dt = ds->dt;
        const double rhodt = rkr * dt, gamdt = rkg * dt;
        double fln[5];
        // the first Wray stage has rho = 0: the previous right-hand side is
        // not read there (0 * flp adds nothing to -gamdt * fl).  Both
        // right-hand sides are read with the non-temporal hint: neither is
        // read again before a full stage of traffic has passed.
        if (rkr != 0.0) {
#pragma unroll
            for (int m = 0; m < 5; m++)
                fln[m] = -rhodt * double(__builtin_nontemporal_load(&flp[m * ni + oi])) -
                         gamdt * double(__builtin_nontemporal_load(&fl[m * ni + oi]));
        } else {
#pragma unroll
            for (int m = 0; m < 5; m++) fln[m] = -gamdt * double(__builtin_nontemporal_load(&fl[m * ni + oi]));
        }
        double wo[5];
#pragma unroll
        for (int m = 0; m < 5; m++) wo[m] = w[m * nv + c];
        if (channel) {
            // u of the stage's state, formed as k_aux forms w_aux(2) from the
            // same w (so not read back from w_aux)
            const double b1 = ds->b1, b2 = ds->b2;
            const double uu = wo[1] * (1.0 / wo[0]);
            fln[0] = fln[0] - b1;
            fln[1] = fln[1] - b2;
            fln[4] = fln[4] - uu * b2;
        }
        double wn[5];
#pragma unroll
        for (int m = 0; m < 5; m++) {
            wn[m] = wo[m] + fln[m];
            w[m * nv + c] = wn[m];
        }
        if (FORCE_VAR) {
            const double dy = yn[j] - yn[j - 1];
            const double ri = 1.0 / wn[0];
            const double uu = wn[1] * ri, vv = wn[2] * ri, ww = wn[3] * ri;
            const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
            const double ee = wn[4] / wn[0] - qq;
            const double tt = temperature_from_e(P, ee, double(wa[5 * nv + c]));
            wa[5 * nv + c] = R(tt);
            acc[0] = wn[1] * tt * dy;
        } else {
            if (P.wall_stag[0] && j <= g.ng) wall_ghost(g, P, w, wa, i, j, 1 - j, k);
            if (P.wall_stag[1] && j > g.ny - g.ng) wall_ghost(g, P, w, wa, i, j, 2 * g.ny + 1 - j, k);
        }
    }
    if (FORCE_VAR) {
        block_sum<1>(acc, lds);
        if (flat_thread() == 0) partials[(int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z))] = acc[0];
    }
}
template <typename R, bool FORCE_VAR>
__global__ __launch_bounds__(256) void k_update(Geo g, KPar P, double *__restrict__ w, R *__restrict__ wa,
                                                const R *__restrict__ flp, const R *__restrict__ fl,
                                                const DevScalars *__restrict__ ds, double rkr, double rkg, int channel,
                                                const double *__restrict__ yn, double *__restrict__ partials) {
    k_update_body<R, FORCE_VAR>(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), g, P, w, wa, flp, fl, ds, rkr, rkg, channel, yn, partials);
}

// force_var (singleideal_cpu.F90:890-933, grid_dim 1) bulk temperature
__global__ __launch_bounds__(RED_THREADS) void k_force_var_finish(KPar P, const double *__restrict__ partials, int nblk,
                                                                  double *__restrict__ stage2, int *__restrict__ ticket,
                                                                  DevScalars *ds) {
    __shared__ double lds[RED_THREADS / 64][1];
    double acc[1];
    if (two_level_sum<1>(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), partials,
                         nblk, stage2, ticket, acc, lds)) {
        double bulktg = acc[0] / P.volchan;
        bulktg = bulktg / ds->rhobulk / ds->ubulk;
        ds->bulkt = P.t_bulk_target - bulktg;
    }
}

// force_var_2_subroutine (kernels_cpu.F90:3282-3314) + wall ghosts
template <typename R>
__global__ __launch_bounds__(256) void k_force_var2(Geo g, KPar P, double *__restrict__ w, const R *__restrict__ wa,
                                                    const DevScalars *__restrict__ ds) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const size_t nv = g.nvol;
    const int c = gidx(g, i, j, k);
    const double rho = w[c], rhou = w[nv + c], rhov = w[2 * nv + c], rhow = w[3 * nv + c];
    const double ttnew = double(wa[5 * nv + c]) + ds->bulkt;
    const double ee = e_from_temperature(P, ttnew);
    w[4 * nv + c] = rho * ee + 0.5 * (rhou * rhou + rhov * rhov + rhow * rhow) / rho;
}

// ------------------------------------------------------------ statistics
// compute_stats (singleideal.F90:2003-2179, grid_dim 1): one thread per
// (i, j) walks k = 1..nz in order, so each spanwise sum is accumulated in
// the reference's order; the 70 running sums live in LDS, one column per
// thread.  Velocity gradients from w (with ghosts), T and mu from
// w_aux(6), w_aux(7).  Then w_stat = (w_stat*itav + sum/nz)/(itav + 1),
// w_stat(1:70, 1:nx, 1:ny) in Fortran order.
constexpr int NV_STAT = 70;
constexpr int STAT_BX = 64;

template <typename R>
__global__ __launch_bounds__(STAT_BX) void k_stats(Geo g, KPar P, const double *__restrict__ w,
                                                   const R *__restrict__ wa, const double *__restrict__ dx1,
                                                   const double *__restrict__ dy1, const double *__restrict__ dz1,
                                                   double *__restrict__ w_stat, int itav) {
    __shared__ double zs[NV_STAT][STAT_BX];
    const int t = threadIdx.x;
    const int i = blockIdx.x * STAT_BX + t + 1;
    const int j = blockIdx.y + 1;
    if (i > g.nx) return;  // no barrier below: each thread owns its column
#pragma unroll
    for (int q = 0; q < NV_STAT; q++) zs[q][t] = 0.0;
    const size_t nv = g.nvol;
    const int lmax = P.visc_lmax;
    const double dx = dx1[i - 1], dy = dy1[j - 1];
    const double rgas0 = P.rgas0;
    auto vel = [&](int c, int m) { return w[m * nv + c] / w[c]; };
    for (int k = 1; k <= g.nz; k++) {
        const int c = gidx(g, i, j, k);
        double ux = 0, vx = 0, wx = 0, uy = 0, vy = 0, wy = 0, uz = 0, vz = 0, wz = 0;
        for (int l = 1; l <= lmax; l++) {
            const double ccl = P.vd1[l - 1];
            const int xp = c + l, xm = c - l, yp = c + l * g.sx, ym = c - l * g.sx, zp = c + l * g.sxy,
                      zm = c - l * g.sxy;
            ux = ux + ccl * (vel(xp, 1) - vel(xm, 1));
            vx = vx + ccl * (vel(xp, 2) - vel(xm, 2));
            wx = wx + ccl * (vel(xp, 3) - vel(xm, 3));
            uy = uy + ccl * (vel(yp, 1) - vel(ym, 1));
            vy = vy + ccl * (vel(yp, 2) - vel(ym, 2));
            wy = wy + ccl * (vel(yp, 3) - vel(ym, 3));
            uz = uz + ccl * (vel(zp, 1) - vel(zm, 1));
            vz = vz + ccl * (vel(zp, 2) - vel(zm, 2));
            wz = wz + ccl * (vel(zp, 3) - vel(zm, 3));
        }
        const double dz = dz1[k - 1];
        ux = ux * dx; vx = vx * dx; wx = wx * dx;
        uy = uy * dy; vy = vy * dy; wy = wy * dy;
        uz = uz * dz; vz = vz * dz; wz = wz * dz;
        const double omx = wy - vz, omy = uz - wx, omz = vx - uy;
        const double rho = w[c], rhou = w[nv + c], rhov = w[2 * nv + c], rhow = w[3 * nv + c];
        const double ri = 1.0 / rho;
        const double uu = rhou * ri, vv = rhov * ri, ww = rhow * ri;
        const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
        const double tt = double(wa[5 * nv + c]);
        const double pp = rho * rgas0 * tt;
        const double mu = double(wa[6 * nv + c]);
        const double nu = mu / rho;
        const double gamloc = gamloc_of(P, tt);
        const double cploc = gamloc / (gamloc - 1.0) * rgas0;
        const double cs = sqrt(gamloc * rgas0 * tt);
        const double machlocal = sqrt(2.0 * qq) / cs;
        const double t_tot = tt + qq * (gamloc - 1.0) / gamloc;
        const double rho2 = rho * rho, uu2 = uu * uu, vv2 = vv * vv, ww2 = ww * ww;
        const double pp2 = pp * pp, tt2 = tt * tt;
        const double divl = ux + vy + wz;
        const double div3l = divl / 3.0;
        double s11 = 2.0 * (ux - div3l), s12 = uy + vx, s13 = uz + wx;
        double s22 = 2.0 * (vy - div3l), s23 = vz + wy, s33 = 2.0 * (wz - div3l);
        s11 = s11 * mu; s12 = s12 * mu; s13 = s13 * mu;
        s22 = s22 * mu; s23 = s23 * mu; s33 = s33 * mu;
        const double s21 = s12, s31 = s13, s32 = s23;
        int q = 0;
        auto add = [&](double v) { zs[q][t] = zs[q][t] + v; q++; };
        add(rho); add(uu); add(vv); add(ww); add(pp); add(tt); add(rho2); add(uu2); add(vv2); add(ww2);
        add(pp2); add(tt2); add(rhou); add(rhov); add(rhow); add(rhou * uu); add(rhov * vv); add(rhow * ww);
        add(rhou * vv); add(mu);
        add(nu); add(omx * omx); add(omy * omy); add(omz * omz); add(rho * tt); add(rho * tt2); add(t_tot);
        add(rho * t_tot); add(t_tot * t_tot); add(rhou * tt);
        add(rhov * tt); add(rhow * tt); add(machlocal); add(machlocal * machlocal); add(rhou * uu2);
        add(rhov * uu2); add(rhou * vv2); add(rhov * vv2); add(rhou * ww2); add(rhov * ww2);
        add(pp * uu); add(pp * vv); add(s11); add(s12); add(s13); add(s22); add(s23); add(s33); add(s11 * uu);
        add(s12 * uu);
        add(s21 * vv); add(s22 * vv); add(s31 * ww); add(s32 * ww); add(s11 * vv + s21 * uu);
        add(s12 * vv + s22 * uu); add(s11 * ux + s12 * uy + s13 * uz); add(s21 * vx + s22 * vy + s23 * vz);
        add(s31 * wx + s32 * wy + s33 * wz); add(s11 * vx + s12 * (ux + vy) + s22 * uy + s13 * vz + s23 * uz);
        add(pp * ux); add(pp * vy); add(pp * wz); add(pp * (uy + vx)); add(divl * divl); add(rho * tt2 * tt);
        add(rho * tt2 * tt2); add(rhou * uu2 * uu); add(cploc); add(gamloc);
    }
    double *ws = w_stat + (size_t)NV_STAT * ((size_t)(i - 1) + (size_t)g.nx * (size_t)(j - 1));
    const double nzd = (double)g.nz, ta = (double)itav, tb = (double)(itav + 1);
#pragma unroll
    for (int q = 0; q < NV_STAT; q++) ws[q] = (ws[q] * ta + zs[q][t] / nzd) / tb;
}

// ---------------------------------------------------------------- CFL dt
// compute_aux(central=1, ghost=0) + compute_dt_subroutine (kernels_cpu.F90:
// 6960-7011, 6814-6880): primitives recomputed from w in registers (w_aux
// written only for T when cp(T) needs it as the Newton start), block max of
// the nine spectral radii, atomic max of the block maxima.
// A workgroup covers 64 x 4 points of (i, j) and walks k from blockIdx.z+1
// with stride gridDim.z (one plane when gridDim.z = nz).  Workgroup 0
// clears the ticket of k_dt_finish, which runs after this kernel.
template <typename R>
__global__ __launch_bounds__(256) void k_dt(Geo g, KPar P, const double *__restrict__ w, R *__restrict__ wa,
                                            const double *__restrict__ dx1, const double *__restrict__ dy1,
                                            const double *__restrict__ dz1, const double *__restrict__ dxs,
                                            const double *__restrict__ dys, const double *__restrict__ dzs,
                                            double *__restrict__ partials, int *__restrict__ ticket) {
    __shared__ double lds[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (flat_block() == 0 && flat_thread() == 0) *ticket = 0;
    double mx = 0.0;
    for (int k = blockIdx.z + 1; k <= g.nz; k += gridDim.z) {
        if (i > g.nx || j > g.ny) break;
        const size_t nv = g.nvol;
        const int c = gidx(g, i, j, k);
        const double rho = w[c];
        const double ri = 1.0 / rho;
        const double uu = w[nv + c] * ri, vv = w[2 * nv + c] * ri, ww = w[3 * nv + c] * ri;
        const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
        const double ee = w[4 * nv + c] / rho - qq;
        const double tt = temperature_from_e(P, ee, double(wa[5 * nv + c]));
        if (P.calorically_perfect != 1) wa[5 * nv + c] = R(tt);
        const double mu = viscosity_of(P, tt);
        const double nu = ri * mu;
        const double kor = nu / P.prandtl;
        const double gamloc = gamloc_of(P, tt);
        const double cs = sqrt(gamloc * P.rgas0 * tt);
        const double ax = dx1[i - 1], ay = dy1[j - 1], az = dz1[k - 1];
        const double sx = dxs[i - 1], sy = dys[j - 1], sz = dzs[k - 1];
        mx = fmax(mx, (fabs(uu) + cs) * ax);
        mx = fmax(mx, (fabs(vv) + cs) * ay);
        mx = fmax(mx, (fabs(ww) + cs) * az);
        mx = fmax(mx, nu * sx);
        mx = fmax(mx, nu * sy);
        mx = fmax(mx, nu * sz);
        mx = fmax(mx, kor * sx);
        mx = fmax(mx, kor * sy);
        mx = fmax(mx, kor * sz);
    }
    mx = wave_max(mx);
    const int lane = flat_thread() & 63, wid = flat_thread() >> 6;
    if (lane == 0) lds[wid] = mx;
    __syncthreads();
    if (flat_thread() == 0) {
        const int nw = (blockDim.x * blockDim.y * blockDim.z + 63) >> 6;
        double m = lds[0];
        for (int q = 1; q < nw; q++) m = fmax(m, lds[q]);
        partials[flat_block()] = m;
    }
}

// Maximum of this block's threads, in thread 0 (order-independent).
__device__ __forceinline__ double block_max(double m, double *lds) {
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) m = fmax(m, lds[q]);
    return m;
}

// dt = cfl / max (singleideal_cpu.F90:1507-1514) over the block maxima:
// gridDim.x blocks take a slice each, the last one to finish combines the
// slice maxima (max is exact in any order).
__device__ __forceinline__ void k_dt_finish_body(const dim3 blockIdx, const dim3 gridDim, KPar P, const double *__restrict__ partials, int nblk,
                                                           double *__restrict__ stage2, int *__restrict__ ticket,
                                                           double *__restrict__ dt_out) {
    __shared__ double lds[RED_THREADS / 64];
    __shared__ bool last;
    const int per = (nblk + gridDim.x - 1) / gridDim.x;
    const int b0 = blockIdx.x * per, b1 = min(b0 + per, nblk);
    double m = 0.0;
    for (int b = b0 + (int)threadIdx.x; b < b1; b += blockDim.x) m = fmax(m, partials[b]);
    m = block_max(m, lds);
    if (threadIdx.x == 0) {
        stage2[blockIdx.x] = m;
        __threadfence();
        last = atomicAdd(ticket, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    m = threadIdx.x < gridDim.x ? __hip_atomic_load(&stage2[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                : 0.0;
    m = block_max(m, lds);
    if (threadIdx.x == 0) {
        const double dt_min = 1.0 / m;
        *dt_out = P.cfl * dt_min;
        *ticket = 0;
    }
}
__global__ __launch_bounds__(RED_THREADS) void k_dt_finish(KPar P, const double *__restrict__ partials, int nblk,
                                                           double *__restrict__ stage2, int *__restrict__ ticket,
                                                           double *__restrict__ dt_out) {
    k_dt_finish_body(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), P, partials, nblk, stage2, ticket, dt_out);
}

// end of a step of run(): icyc += 1, time += dt (singleideal_cpu.F90:1604,1642)
__device__ __forceinline__ void k_step_end_body(const dim3 blockIdx, const dim3 gridDim, DevScalars *ds) {
    ds->icyc = ds->icyc + 1;
    ds->time = ds->time + ds->dt;
}
__global__ void k_step_end(DevScalars *ds) {
    k_step_end_body(dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z), ds);
}

// ------------------------------------------------------------- residual
// compute_residual_subroutine + compute_rho_t_p_minmax_subroutine
// (kernels_cpu.F90:6682-6701, 6765-6810): fln(2) of the last stage is
// rebuilt from the two stored right-hand sides and the forcing.
template <typename R>
__global__ __launch_bounds__(256) void k_residual(Geo g, KPar P, const R *__restrict__ wa,
                                                  const R *__restrict__ flp, const R *__restrict__ fl,
                                                  const DevScalars *__restrict__ ds, double rkr, double rkg, int channel,
                                                  double *__restrict__ partials, const double *__restrict__ fln) {
    __shared__ double lds[4][7];
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    double r = 0.0;
    double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
    double mxv[3] = {-100.0, -100.0, -100.0};
    if (i <= g.nx && j <= g.ny) {
        const size_t nv = g.nvol, ni = g.nint;
        const int c = gidx(g, i, j, k), oi = iidx(g, i, j, k);
        const double dt = ds->dt;
        double fln2;
        if (fln) {  // the stored increment of the procedure-by-procedure stage
            fln2 = fln[ni + oi];
        } else {
            fln2 = -(rkr * dt) * double(flp[ni + oi]) - (rkg * dt) * double(fl[ni + oi]);
            if (channel) fln2 = fln2 - ds->b2;
        }
        const double q = fln2 / dt;
        r = q * q;
        const double rho = double(wa[c]), tt = double(wa[5 * nv + c]);
        const double pp = rho * tt * P.rgas0;
        mn[0] = rho; mn[1] = tt; mn[2] = pp;
        mxv[0] = rho; mxv[1] = tt; mxv[2] = pp;
    }
    r = wave_sum(r);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        mn[q] = wave_min(mn[q]);
        mxv[q] = wave_max(mxv[q]);
    }
    const int lane = flat_thread() & 63, wid = flat_thread() >> 6;
    if (lane == 0) {
        lds[wid][0] = r;
        for (int q = 0; q < 3; q++) { lds[wid][1 + q] = mn[q]; lds[wid][4 + q] = mxv[q]; }
    }
    __syncthreads();
    if (flat_thread() == 0) {
        const int nw = (blockDim.x * blockDim.y * blockDim.z + 63) >> 6;
        double o[7];
        for (int q = 0; q < 7; q++) o[q] = lds[0][q];
        for (int wv = 1; wv < nw; wv++) {
            o[0] += lds[wv][0];
            for (int q = 1; q < 4; q++) o[q] = fmin(o[q], lds[wv][q]);
            for (int q = 4; q < 7; q++) o[q] = fmax(o[q], lds[wv][q]);
        }
        double *pp = partials + (size_t)flat_block() * 7;
        for (int q = 0; q < 7; q++) pp[q] = o[q];
    }
}

__global__ __launch_bounds__(RED_THREADS) void k_residual_finish(Geo g, const double *__restrict__ partials, int nblk,
                                                                 DevScalars *ds) {
    __shared__ double lds[RED_THREADS][7];
    double o[7] = {0.0, 1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308, -100.0, -100.0, -100.0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        const double *pp = partials + (size_t)b * 7;
        o[0] += pp[0];
        for (int q = 1; q < 4; q++) o[q] = fmin(o[q], pp[q]);
        for (int q = 4; q < 7; q++) o[q] = fmax(o[q], pp[q]);
    }
    for (int q = 0; q < 7; q++) lds[threadIdx.x][q] = o[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int t = 1; t < (int)blockDim.x; t++) {
            o[0] += lds[t][0];
            for (int q = 1; q < 4; q++) o[q] = fmin(o[q], lds[t][q]);
            for (int q = 4; q < 7; q++) o[q] = fmax(o[q], lds[t][q]);
        }
        const double restemp = (double)g.nx * (double)g.ny * (double)g.nz;
        ds->residual = sqrt(o[0] / restemp);
        ds->rhomin = o[1]; ds->tmin = o[2]; ds->pmin = o[3];
        ds->rhomax = o[4]; ds->tmax = o[5]; ds->pmax = o[6];
    }
}

// count_weno_subroutine (kernels_cpu.F90:52-113)
template <typename R>
__global__ __launch_bounds__(256) void k_count_weno(Geo g, KPar P, const R *__restrict__ wa, DevScalars *ds) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    double cnt[3] = {0.0, 0.0, 0.0};
    if (i <= g.nx && j <= g.ny) {
        const R *S = wa + 7 * (size_t)g.nvol;
        const int c = gidx(g, i, j, k);
        const int ws = P.weno_scheme;
        const int st[3] = {1, g.sx, g.sxy};
        const int idx[3] = {i, j, k};
        for (int d = 0; d < 3; d++) {
            if (idx[d] < P.eul_lo[d] || idx[d] > P.eul_hi[d]) continue;
            bool shk = false;
            for (int s = -ws + 1; s <= ws; s++) shk |= S[c + s * st[d]] > R(P.sensor_threshold);
            cnt[d] = shk ? 1.0 : 0.0;
        }
    }
    for (int d = 0; d < 3; d++) {
        double v = wave_sum(cnt[d]);
        if ((flat_thread() & 63) == 0 && v > 0.0) atomicAdd(&ds->cnt[d], v);
    }
}

template <typename T>
__global__ void k_zero(T *__restrict__ a, size_t n) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) a[t] = T(0);
}

template <typename T>
__global__ void k_fill(T *__restrict__ a, size_t n, T v) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) a[t] = v;
}

// ======================================== piecewise stage (procedure by procedure)
// The reference's rk_sync calls its procedures one by one on fl/fln/fhat
// (singleideal_cpu.F90:332-409).  These kernels carry that form for callers
// that split the stage (ranges of euler_x as rk_async does, or their own
// ordering); st_rk_sync runs the fused stage instead.

// Hybrid face fluxes of faces f0..f1 along DIR into fhat (ghosted layout,
// the face index in the DIR coordinate), every interior line of the other
// two directions (euler_*_fluxes_hybrid_kernel, kernels_cpu.F90:425-689 and
// its y/z twins): one thread per face, the window of cells f-W+1..f+W
// (W = 3; W = 4 with four ghost layers, for WENO7 and ep_order 8).
template <typename R, int DIR, int W>
__global__ __launch_bounds__(256) void k_face_range(Geo g, KPar P, const R *__restrict__ wa,
                                                    const double *__restrict__ w, const int *__restrict__ eoc,
                                                    double *__restrict__ fhat, int f0, int f1) {
    const int nf = f1 - f0 + 1;
    const int n1 = DIR == 0 ? nf : g.nx, n2 = DIR == 1 ? nf : g.ny, n3 = DIR == 2 ? nf : g.nz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n1 * n2 * n3) return;
    const int a1 = (int)(t % n1), a2 = (int)((t / n1) % n2), a3 = (int)(t / ((long long)n1 * n2));
    const int i = DIR == 0 ? f0 + a1 : 1 + a1, j = DIR == 1 ? f0 + a2 : 1 + a2, k = DIR == 2 ? f0 + a3 : 1 + a3;
    const int f = DIR == 0 ? i : (DIR == 1 ? j : k);
    const int n = DIR == 0 ? g.nx : (DIR == 1 ? g.ny : g.nz);
    const int stride = DIR == 0 ? 1 : (DIR == 1 ? g.sx : g.sxy);
    const size_t nv = g.nvol;
    const bool keep = P.nkeep >= 0;
    const int base = gidx(g, i, j, k);
    Pt q[2 * W];
    double sen[2 * W];
#pragma unroll
    for (int s = 0; s < 2 * W; s++) {
        const int a = base + (s - (W - 1)) * stride;
        sen[s] = double(wa[7 * nv + a]);
        if constexpr (sizeof(R) == sizeof(double)) {
            q[s] = make_pt(double(wa[a]), double(wa[nv + a]), double(wa[2 * nv + a]), double(wa[3 * nv + a]),
                           double(wa[4 * nv + a]), double(wa[5 * nv + a]), P.rgas0, keep);
        } else {
            // FP32 w_aux: the primitives from w in FP64 as eval_aux forms them
            int pi = i, pj = j, pk = k;
            if (DIR == 0) pi += s - (W - 1);
            else if (DIR == 1) pj += s - (W - 1);
            else pk += s - (W - 1);
            if (P.periodic[0]) pi = pi < 1 ? pi + g.nx : (pi > g.nx ? pi - g.nx : pi);
            if (P.periodic[1]) pj = pj < 1 ? pj + g.ny : (pj > g.ny ? pj - g.ny : pj);
            if (P.periodic[2]) pk = pk < 1 ? pk + g.nz : (pk > g.nz ? pk - g.nz : pk);
            const int src = gidx(g, pi, pj, pk);
            const double rho = w[src], rhoe = w[4 * nv + src];
            const double ri = 1.0 / rho;
            const double uu = w[nv + src] * ri, vv = w[2 * nv + src] * ri, ww = w[3 * nv + src] * ri;
            const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
            const double tt = temperature_from_e(P, rhoe / rho - qq, double(wa[5 * nv + a]));
            const double pp = rho * tt * P.rgas0;
            q[s] = make_pt(rho, uu, vv, ww, (rhoe + pp) / rho, tt, P.rgas0, keep);
        }
    }
    double fh[5];
    hybrid_face<DIR, W>(P, q, sen, f, eoc[f], n, fh);
#pragma unroll
    for (int m = 0; m < 5; m++) fhat[m * nv + base] = fh[m];
}

// euler_*_update_subroutine (kernels_cpu.F90:944-1058): cells c0..c1 along
// DIR take fl += (fhat(c) - fhat(c-1)) * metric(c).
template <typename R, int DIR>
__global__ __launch_bounds__(256) void k_face_update(Geo g, const double *__restrict__ fhat,
                                                     const double *__restrict__ metric, R *__restrict__ fl, int c0,
                                                     int c1) {
    const int nc = c1 - c0 + 1;
    const int n1 = DIR == 0 ? nc : g.nx, n2 = DIR == 1 ? nc : g.ny, n3 = DIR == 2 ? nc : g.nz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n1 * n2 * n3) return;
    const int a1 = (int)(t % n1), a2 = (int)((t / n1) % n2), a3 = (int)(t / ((long long)n1 * n2));
    const int i = DIR == 0 ? c0 + a1 : 1 + a1, j = DIR == 1 ? c0 + a2 : 1 + a2, k = DIR == 2 ? c0 + a3 : 1 + a3;
    const int c = DIR == 0 ? i : (DIR == 1 ? j : k);
    const int stride = DIR == 0 ? 1 : (DIR == 1 ? g.sx : g.sxy);
    const size_t nv = g.nvol;
    const int gc = gidx(g, i, j, k), oi = iidx(g, i, j, k);
    const double d = metric[c - 1];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        const size_t q = (size_t)m * g.nint + oi;
        fl[q] = R(double(fl[q]) + (fhat[m * nv + gc] - fhat[m * nv + gc - stride]) * d);
    }
}

// init_flux_subroutine (kernels_cpu.F90:29-47): fln = -rhodt fl, fl = 0;
// update_flux_subroutine (:3400-3418): fln = fln - gamdt fl.
template <typename R>
__global__ void k_init_flux(size_t n, R *__restrict__ fl, double *__restrict__ fln, double rhodt, int init) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        if (init) {
            fln[t] = -rhodt * double(fl[t]);
            fl[t] = R(0);
        } else {
            fln[t] = fln[t] - rhodt * double(fl[t]);
        }
    }
}

// update_field_subroutine (:3422-3442): w = w + fln on the interior.
__global__ void k_update_field(Geo g, double *__restrict__ w, const double *__restrict__ fln) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const int c = gidx(g, i, j, k), oi = iidx(g, i, j, k);
#pragma unroll
    for (int m = 0; m < 5; m++) w[m * g.nvol + c] = w[m * g.nvol + c] + fln[(size_t)m * g.nint + oi];
}

// force_rhs_1_subroutine (:3157-3193): block partials of fln(1) dy,
// fln(2) dy, w(1) dy, w(2) dy and w(2) dy w_aux(6).
template <typename R>
__global__ __launch_bounds__(256) void k_force_rhs_sums(Geo g, const double *__restrict__ w, const R *__restrict__ wa,
                                                        const double *__restrict__ fln, const double *__restrict__ yn,
                                                        double *__restrict__ partials) {
    __shared__ double lds[4][5];
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (i <= g.nx && j <= g.ny) {
        const int c = gidx(g, i, j, k), oi = iidx(g, i, j, k);
        const double dy = yn[j] - yn[j - 1];
        acc[0] = fln[oi] * dy;
        acc[1] = fln[(size_t)g.nint + oi] * dy;
        acc[2] = w[c] * dy;
        acc[3] = w[g.nvol + c] * dy;
        acc[4] = w[g.nvol + c] * dy * double(wa[5 * (size_t)g.nvol + c]);
    }
    block_sum<5>(acc, lds);
    if (flat_thread() == 0)
        for (int q = 0; q < 5; q++) partials[(size_t)flat_block() * 5 + q] = acc[q];
}

// force_rhs_2_subroutine (:3092-3116)
template <typename R>
__global__ void k_force_rhs_apply(Geo g, const R *__restrict__ wa, double *__restrict__ fln,
                                  const DevScalars *__restrict__ ds) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const int c = gidx(g, i, j, k), oi = iidx(g, i, j, k);
    const double b1 = ds->b1, b2 = ds->b2;
    const double uu = double(wa[g.nvol + c]);
    fln[oi] = fln[oi] - b1;
    fln[(size_t)g.nint + oi] = fln[(size_t)g.nint + oi] - b2;
    fln[4 * (size_t)g.nint + oi] = fln[4 * (size_t)g.nint + oi] - uu * b2;
}

// force_var_1_subroutine (:3241-3278): T from w (Newton from w_aux(6)),
// stored in w_aux(6), and the block partials of rho*u*T*dy.
template <typename R>
__global__ __launch_bounds__(256) void k_force_var_sums(Geo g, KPar P, const double *__restrict__ w,
                                                        R *__restrict__ wa, const double *__restrict__ yn,
                                                        double *__restrict__ partials) {
    __shared__ double lds[4][1];
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    double acc[1] = {0.0};
    if (i <= g.nx && j <= g.ny) {
        const size_t nv = g.nvol;
        const int c = gidx(g, i, j, k);
        const double dy = yn[j] - yn[j - 1];
        const double rho = w[c], rhou = w[nv + c];
        const double ri = 1.0 / rho;
        const double uu = rhou * ri, vv = w[2 * nv + c] * ri, ww = w[3 * nv + c] * ri;
        const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
        const double ee = w[4 * nv + c] / rho - qq;
        const double tt = temperature_from_e(P, ee, double(wa[5 * nv + c]));
        wa[5 * nv + c] = R(tt);
        acc[0] = rhou * tt * dy;
    }
    block_sum<1>(acc, lds);
    if (flat_thread() == 0) partials[flat_block()] = acc[0];
}

// recyc_exchange (singleideal_cpu.F90:935-999) for one block in x and z:
// wrecyc(i, j, k, m) = w(i_recyc+1-i, j, k shifted by nz/2, m)
__global__ void k_recyc_exchange(Geo g, const double *__restrict__ w, int i_recyc, double *__restrict__ wrecyc) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)g.ng * g.ny * g.nz;
    if (t >= n) return;
    const int i = (int)(t % g.ng) + 1, j = (int)((t / g.ng) % g.ny) + 1, k = (int)(t / ((long long)g.ng * g.ny)) + 1;
    const int src = gidx(g, i_recyc + 1 - i, j, recyc_k(k, g.nz));
    for (int m = 0; m < 5; m++) wrecyc[(size_t)m * n + t] = w[(size_t)m * g.nvol + src];
}

// ================================================================ launchers
static inline dim3 cell_grid(const Geo &g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz); }
static inline int cell_blocks(const Geo &g) { return ((g.nx + 63) / 64) * ((g.ny + 3) / 4) * g.nz; }
static const dim3 kCellBlock(64, 4, 1);

// Workgroups of a two-level reduction of n partials: 64 partials or more
// each, at most RB (its second-level slots).  A function of the grid
// alone, so the summation order of a solve does not depend on how it is
// launched.
static inline int red_blocks(long long n) {
    const long long b = (n + 63) / 64;
    return (int)(b < 1 ? 1 : (b > RB ? RB : b));
}
// Workgroups of the WENO fix-up: one per 16384 cells, at most one per CU
static inline int fix_blocks(const Geo &g) {
    const long long b = (g.nint + 16383) / 16384;
    return (int)(b < 1 ? 1 : (b > kFixBlocks ? kFixBlocks : b));
}

static inline int grid1d(size_t n, int bs) {
    size_t b = (n + bs - 1) / bs;
    return (int)(b > 65535 * 16 ? 65535 * 16 : (b < 1 ? 1 : b));
}

// The element type of w_aux and of the right-hand sides follows the
// precision of the solve: double (FP64 form) or float (mixed form).  Each
// public launcher dispatches to its typed instance.
#define ST_DISPATCH(call)                         \
    do {                                          \
        if (s->mixed) {                           \
            using R = float;                      \
            call;                                 \
        } else {                                  \
            using R = double;                     \
            call;                                 \
        }                                         \
    } while (0)

template <typename T>
static hipError_t fill_periodic_t(st_solver *s, T *arr, int nv) {
    ST_NOREC();
    const Geo &g = s->g;
    size_t tot = (size_t)2 * g.ng * ((size_t)g.ny * g.nz + (size_t)g.nx * g.nz + (size_t)g.nx * g.ny) * nv;
    hipLaunchKernelGGL(k_fill_periodic<T>, dim3(grid1d(tot, 256)), dim3(256), 0, s->stream, g, s->kp, arr, nv);
    return hipGetLastError();
}

hipError_t launch_fill_periodic(st_solver *s, double *arr, int nv) { return fill_periodic_t(s, arr, nv); }

hipError_t launch_fill_periodic_aux(st_solver *s, int var) {
    ST_DISPATCH(return fill_periodic_t(s, (R *)s->w_aux + (size_t)(var - 1) * s->g.nvol, 1));
}

hipError_t launch_extrapolate_aux(st_solver *s, int var) {
    ST_NOREC();
    const Geo &g = s->g;
    size_t tot = (size_t)2 * g.ng * ((size_t)g.ny * g.nz + (size_t)g.nx * g.nz + (size_t)g.nx * g.ny);
    ST_DISPATCH(hipLaunchKernelGGL(k_extrapolate_var<R>, dim3(grid1d(tot, 256)), dim3(256), 0, s->stream, g,
                                   (R *)s->w_aux + (size_t)(var - 1) * g.nvol));
    return hipGetLastError();
}

hipError_t launch_aux(st_solver *s, int i0, int i1, int j0, int j1, int k0, int k1, bool image, bool dt) {
    const int ni = i1 - i0 + 1, nj = j1 - j0 + 1, nk = k1 - k0 + 1;
    if (ni <= 0 || nj <= 0 || nk <= 0) return hipSuccess;
    const size_t tot = aux_whole_box(s->g, i0, ni, j0, nj, k0, nk) ? aux_lanes(s->g)
                                                                    : (size_t)box_row_lanes(i0, ni) * nj * nk;
    const unsigned nblk = (unsigned)((tot + 255) / 256);
#define ST_LAUNCH_AUX(D_)                                                                                          \
    ST_DISPATCH(ST_GO(k_aux, ST_T(R, D_), 256, 0, dim3(nblk), dim3(256), s->g, s->kp, s->w, (R *)s->w_aux, i0, ni,  \
                      j0, nj, k0, nk, image ? 1 : 0, s->dcsidx, s->detady, s->dzitdz, s->dcsidxs, s->detadys,        \
                      s->dzitdzs, s->partials))
    if (dt && s->p.cfl > 0) ST_LAUNCH_AUX(true);
    else ST_LAUNCH_AUX(false);
#undef ST_LAUNCH_AUX
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !(dt && s->p.cfl > 0)) return e;
    ST_GO(k_dt_finish, , RED_THREADS, 0, dim3(red_blocks(nblk)), dim3(RED_THREADS), s->kp, s->partials, (int)nblk,
          s->partials + s->npart - RB * 8, s->counters + 1, &s->ds->dt);
    return hipGetLastError();
}

hipError_t launch_wall_ghosts(st_solver *s) {
    ST_NOREC();
    const Geo &g = s->g;
    if (!s->kp.wall_stag[0] && !s->kp.wall_stag[1]) return hipSuccess;
    size_t tot = (size_t)2 * g.ng * g.nx * g.nz;
    ST_DISPATCH(hipLaunchKernelGGL(k_wall_ghosts<R>, dim3(grid1d(tot, 256)), dim3(256), 0, s->stream, g, s->kp, s->w,
                                   (const R *)s->w_aux));
    return hipGetLastError();
}

template <typename R>
static hipError_t launch_fix(st_solver *s, R *fl, int dir, int bit = 1, int anyidx = 0, double rkg = 0.0,
                             double *partials = nullptr) {
    const Geo &g = s->g;
    // segments of 2048 words (16384 cells) or longer, at most kFixMaxSeg
    const int nword = (int)((g.nint + 7) / 8);
    int nseg = (nword + 2047) / 2048;
    if (nseg > kFixMaxSeg) nseg = kFixMaxSeg;
    if (nseg < 1) nseg = 1;
    int wpb = (nword + nseg - 1) / nseg;
    wpb = (wpb + 255) / 256 * 256;
    nseg = (nword + wpb - 1) / wpb;
    int *counts = s->counters + 8;
    ST_GO(k_fix_compact, , 256, 0, dim3((unsigned)nseg), dim3(256), s->dirty, g.nint, bit, anyidx, wpb, s->fixlist,
          counts);
    // the fix-up holds one wave per SIMD (its register budget): at most one
    // 4-wave block per CU
    const int nb = fix_blocks(g);
    const int seglen = 8 * wpb;
    const R *wa = (const R *)s->w_aux;
#define ST_LAUNCH_F(D_, RED_, E_, M_)                                                                              \
    ST_GO(k_fix_list, ST_T(R, D_, RED_), 256, 0, dim3(nb), dim3(256), g, s->kp, wa, s->w, E_, M_, fl, s->dirty,     \
          anyidx, s->fixlist, counts, nseg, seglen, s->ds, rkg, s->yn, partials)
    if (dir == 0) ST_LAUNCH_F(0, false, s->eoc[0], s->dcsidx);
    else if (dir == 1) ST_LAUNCH_F(1, false, s->eoc[1], s->detady);
    else if (partials) ST_LAUNCH_F(2, true, s->eoc[2], s->dzitdz);
    else ST_LAUNCH_F(2, false, s->eoc[2], s->dzitdz);
#undef ST_LAUNCH_F
    return hipGetLastError();
}

template <typename R>
static hipError_t euler_x_t(st_solver *s, R *fl, bool init, bool fix) {
    ST_NOREC();
    const Geo &g = s->g;
    ST_GO(k_clear_any, , 1024, 0, dim3(1), dim3(1), s->dirty, g.nint);
    if (s->kp.eul_hi[0] < s->kp.eul_lo[0]) {
        if (init) hipLaunchKernelGGL(k_zero<R>, dim3(grid1d((size_t)5 * g.nint, 256)), dim3(256), 0, s->stream, fl,
                                     (size_t)5 * g.nint);
        return hipGetLastError();
    }
    const long nwork = (long)((g.nx + XW - 1) / XW) * g.ny * g.nz;
    long nblk = (nwork + 3) / 4;
    if (nblk > 256L * 4) nblk = 256L * 4;  // resident workgroups at 4 waves per SIMD
    dim3 grid((unsigned)nblk);
    const int km = s->kp.nkeep == 0 ? 0 : (s->kp.nkeep > 0 ? 1 : -1);
#define ST_LAUNCH_X(I, K)                                                                                          \
    hipLaunchKernelGGL((k_euler_x<R, I, K>), grid, dim3(256), 0, s->stream, g, s->kp, (const R *)s->w_aux, s->eoc[0], \
                       s->dcsidx, fl, s->dirty)
    if (init) {
        if (km == 0) ST_LAUNCH_X(true, 0); else if (km > 0) ST_LAUNCH_X(true, 1); else ST_LAUNCH_X(true, -1);
    } else {
        if (km == 0) ST_LAUNCH_X(false, 0); else if (km > 0) ST_LAUNCH_X(false, 1); else ST_LAUNCH_X(false, -1);
    }
#undef ST_LAUNCH_X
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !fix) return e;
    return launch_fix(s, fl, 0);
}

hipError_t launch_euler_x(st_solver *s, void *fl, bool init, bool fix) {
    ST_DISPATCH(return euler_x_t(s, (R *)fl, init, fix));
}

template <typename R>
static hipError_t launch_march(st_solver *s, R *fl, int dir, bool fix, int vd = 0, const R *fl_prev = nullptr,
                               int stage = 0, bool ord2 = false) {
    const Geo &g = s->g;
    const int lo = s->kp.eul_lo[dir], hi = s->kp.eul_hi[dir];
    ST_GO(k_clear_any, , 1024, 0, dim3(1), dim3(1), s->dirty, g.nint);
    if (hi < lo) return hipGetLastError();
    const int ncell = hi - lo + 1;
    const int no = dir == 1 ? g.nz : g.ny;
    // chunks of ~32 cells (3 warm-up points each) so that the grid holds
    // several times the resident waves of the chip
    int chunk = ncell < 32 ? ncell : 32;
    const long minthreads = 256L * 16 * 64 * 4;
    while (chunk > 8 && (long)g.nx * no * ((ncell + chunk - 1) / chunk) < minthreads) chunk /= 2;
    // the march that carries the force_rhs partials sums them per chunk, so
    // it keeps the automatic chunk (a function of the grid alone): a tuned
    // chunk must not change results
    if (s->march_chunk > 0 && vd == 0) chunk = s->march_chunk < ncell ? s->march_chunk : ncell;
    const int nchunk = (ncell + chunk - 1) / chunk;
    const size_t tot = (size_t)g.nx * no * nchunk;
    const unsigned nblk = (unsigned)((tot + 255) / 256);
    dim3 grid(nblk);
    const int km = s->kp.nkeep == 0 ? 0 : (s->kp.nkeep > 0 ? 1 : -1);
    VDArgs V{};
    if (vd) {
        V.flp = fl_prev;
        V.w = s->w;
        V.dx1 = s->dcsidx;
        V.dy1 = s->detady;
        V.yn = s->yn;
        V.ds = s->ds;
        V.rkr = s->rhork[stage];
        V.rkg = s->gamrk[stage];
        V.partials = s->partials;
        V.xc = s->x;
        V.yc = s->y;
        V.zc = s->z;
    }
#define ST_LAUNCH_M(D, K, VD_, O2_, E, M)                                                                          \
    ST_GO(k_euler_march, ST_T(R, D, K, VD_, O2_), 256, 0, grid, dim3(256), g, s->kp, (const R *)s->w_aux, E, M, fl, \
          chunk, nchunk, s->dirty, V)
#define ST_LAUNCH_MK(D, VD_, O2_, E, M)                      \
    do {                                                     \
        if (km == 0) ST_LAUNCH_M(D, 0, VD_, O2_, E, M);      \
        else if (km > 0) ST_LAUNCH_M(D, 1, VD_, O2_, E, M);  \
        else ST_LAUNCH_M(D, -1, VD_, O2_, E, M);             \
    } while (0)
    if (dir == 1) ST_LAUNCH_MK(1, 0, false, s->eoc[1], s->detady);
    else if (vd == 2 && ord2) ST_LAUNCH_MK(2, 2, true, s->eoc[2], s->dzitdz);
    else if (vd == 2) ST_LAUNCH_MK(2, 2, false, s->eoc[2], s->dzitdz);
    else if (vd == 1 && ord2) ST_LAUNCH_MK(2, 1, true, s->eoc[2], s->dzitdz);
    else if (vd == 1) ST_LAUNCH_MK(2, 1, false, s->eoc[2], s->dzitdz);
    else ST_LAUNCH_MK(2, 0, false, s->eoc[2], s->dzitdz);
#undef ST_LAUNCH_MK
#undef ST_LAUNCH_M
    hipError_t e = hipGetLastError();
    // the force_rhs partials of this stage: the march's blocks, then the
    // z fix-up's (it always writes them, zero when nothing is flagged)
    if (vd == 2) {
        s->force_nblk = (int)nblk + fix_blocks(g);
        s->zv_nblk = (int)nblk;
    }
    if (e != hipSuccess || !fix) return e;
    if (vd == 2) return launch_fix(s, fl, dir, 1, 0, s->gamrk[stage], s->partials + (size_t)nblk * 5);
    return launch_fix(s, fl, dir);
}

// The fused x + y pass (INIT: the stage's right-hand side starts here).
template <typename R>
static hipError_t launch_xy(st_solver *s, R *fl, bool fix) {
    const Geo &g = s->g;
    const int ncell = g.ny;
    const long nseg = (g.nx + XW - 1) / XW;
    int chunk = ncell < 32 ? ncell : 32;
    const long minwaves = 256L * 16 * 4;
    while (chunk > 8 && nseg * g.nz * ((ncell + chunk - 1) / chunk) < minwaves) chunk /= 2;
    if (s->march_chunk > 0) chunk = s->march_chunk < ncell ? s->march_chunk : ncell;
    const int nchunk = (ncell + chunk - 1) / chunk;
    const long nw = nseg * g.nz * nchunk;
    ST_GO(k_clear_any, , 1024, 0, dim3(1), dim3(1), s->dirty, g.nint);
    const dim3 grid((unsigned)((nw + 3) / 4));
    const int km = s->kp.nkeep == 0 ? 0 : (s->kp.nkeep > 0 ? 1 : -1);
#define ST_LAUNCH_XY(K)                                                                                          \
    ST_GO(k_euler_xy, ST_T(R, K), 256, 0, grid, dim3(256), g, s->kp, (const R *)s->w_aux, s->eoc[0], s->eoc[1],    \
          s->dcsidx, s->detady, fl, chunk, nchunk, s->dirty)
    if (km == 0) ST_LAUNCH_XY(0); else if (km > 0) ST_LAUNCH_XY(1); else ST_LAUNCH_XY(-1);
#undef ST_LAUNCH_XY
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !fix) return e;
    e = launch_fix(s, fl, 0, 1, 0);
    if (e != hipSuccess) return e;
    return launch_fix(s, fl, 1, 2, 1);
}

hipError_t launch_euler_xy(st_solver *s, void *fl, bool fix) { ST_DISPATCH(return launch_xy(s, (R *)fl, fix)); }
hipError_t launch_euler_y(st_solver *s, void *fl, bool fix) { ST_DISPATCH(return launch_march(s, (R *)fl, 1, fix)); }
hipError_t launch_euler_z(st_solver *s, void *fl, bool fix) { ST_DISPATCH(return launch_march(s, (R *)fl, 2, fix)); }

// z march with the viscous divergence terms (visflx mode 1, or mode 7 with
// ord2) of the same cells and, in the channel case, the force_rhs partials;
// the z fix-up is part of it (its changes to fl(1:2) enter the partials).
// Requires a z range spanning all planes, and visc_lmax <= 3 for mode 1.
hipError_t launch_euler_zv(st_solver *s, void *fl, const void *fl_prev, int stage, bool reduce, bool ord2,
                           bool fix) {
    ST_DISPATCH(return launch_march(s, (R *)fl, 2, fix, reduce ? 2 : 1, (const R *)fl_prev, stage, ord2));
}

// The z fix-up of launch_euler_zv(fix = false): with the force_rhs partials,
// its blocks write theirs after the march's
hipError_t launch_euler_zv_fix(st_solver *s, void *fl, int stage, bool reduce) {
    if (reduce)
        ST_DISPATCH(return launch_fix(s, (R *)fl, 2, 1, 0, s->gamrk[stage], s->partials + (size_t)s->zv_nblk * 5));
    ST_DISPATCH(return launch_fix(s, (R *)fl, 2));
}

// The fix-up of the last sweep of direction dir, for callers that time it
// apart from the sweep: bit 1 / any-dirty byte 0 after a single-direction
// sweep; the fused x + y pass leaves x under (1, 0) and y under (2, 1).
hipError_t launch_weno_fix(st_solver *s, void *fl, int dir, int bit, int anyidx) {
    ST_DISPATCH(return launch_fix(s, (R *)fl, dir, bit, anyidx));
}

hipError_t launch_visc(st_solver *s, void *fl, int mode, bool ghost_images) {
    const Geo &g = s->g;
    if (s->kp.visc_lmax > H) {
        // viscous order 8 (four ghost layers): the generic per-cell kernel
        ST_NOREC();
#define ST_LAUNCH_VG(S_)                                                                                           \
    hipLaunchKernelGGL((k_visc_gen<R, S_>), cell_grid(g), kCellBlock, 0, s->stream, g, s->kp, (R *)s->w_aux,       \
                       (R *)fl, s->dcsidx, s->dcsidxs, s->dcsidx2, s->detady, s->detadys, s->detady2, s->dzitdz,   \
                       s->dzitdzs, s->dzitdz2, s->wallprop, ghost_images ? 1 : 0)
        if (mode == 0) ST_DISPATCH(ST_LAUNCH_VG(true));
        else ST_DISPATCH(ST_LAUNCH_VG(false));
#undef ST_LAUNCH_VG
        return hipGetLastError();
    }
    const int bxy = ((g.nx + VBX - 1) / VBX) * ((g.ny + VBY - 1) / VBY);
    // chunks of up to 32 planes (6 warm-up planes each), enough of them to
    // give every CU several tiles
    int chunk = 32;
    while (chunk > 8 && (long)bxy * ((g.nz + chunk - 1) / chunk) < 2048) chunk /= 2;
    if (s->visc_chunk > 0) chunk = s->visc_chunk;
    const dim3 grid((g.nx + VBX - 1) / VBX, (g.ny + VBY - 1) / VBY, (g.nz + chunk - 1) / chunk);
    const dim3 block(VBX, VBY);
#define ST_LAUNCH_V(S_)                                                                                            \
    ST_GO(k_visc_zm, ST_T(R, S_), VBX * VBY, (sizeof(R) == 4 ? 4 : 1), grid, block, g, s->kp, (R *)s->w_aux,       \
          (R *)fl, s->dcsidx, s->dcsidxs, s->dcsidx2, s->detady, s->detadys, s->detady2, s->dzitdz, s->dzitdzs,     \
          s->dzitdz2, s->wallprop, ghost_images ? 1 : 0, chunk)
    if (mode == 0) ST_DISPATCH(ST_LAUNCH_V(true));
    else ST_DISPATCH(ST_LAUNCH_V(false));
#undef ST_LAUNCH_V
    return hipGetLastError();
}

// Conservative viscous form: mode 2 (cell terms + sensor), 6 (cell terms),
// 5 (face terms) and 56 (6 then 5 in one pass, the rk_sync order).
hipError_t launch_visc_cons(st_solver *s, void *fl, int mode, bool ghost_images) {
    ST_NOREC();
    const Geo &g = s->g;
#define ST_LAUNCH_VC(C_, F_, S_)                                                                                      \
    ST_DISPATCH(hipLaunchKernelGGL((k_visc_cons<R, C_, F_, S_>), cell_grid(g), kCellBlock, 0, s->stream, g, s->kp,   \
                                   (R *)s->w_aux, (R *)fl, s->x, s->y, s->z, s->wallprop, ghost_images ? 1 : 0))
    switch (mode) {
    case 2: ST_LAUNCH_VC(true, false, true); break;
    case 5: ST_LAUNCH_VC(false, true, false); break;
    case 6: ST_LAUNCH_VC(true, false, false); break;
    case 56: ST_LAUNCH_VC(true, true, false); break;
    default: return hipErrorInvalidValue;
    }
#undef ST_LAUNCH_VC
    return hipGetLastError();
}

hipError_t launch_sensor(st_solver *s, bool ghost_images) {
    ST_DISPATCH(ST_GO(k_sensor, ST_T(R), 256, 0, cell_grid(s->g), kCellBlock, s->g, s->kp, (R *)s->w_aux,
                      ghost_images ? 1 : 0));
    return hipGetLastError();
}

hipError_t launch_sensor_refine(st_solver *s, bool direct, bool ghost_images) {
    if (!s->mixed) return hipSuccess;
    if (direct)
        ST_GO(k_sensor_refine, ST_T(true), 256, 0, cell_grid(s->g), kCellBlock, s->g, s->kp, s->w, (float *)s->w_aux,
              s->dcsidx, s->detady, s->dzitdz, ghost_images ? 1 : 0);
    else
        ST_GO(k_sensor_refine, ST_T(false), 256, 0, cell_grid(s->g), kCellBlock, s->g, s->kp, s->w, (float *)s->w_aux,
              s->dcsidx, s->detady, s->dzitdz, ghost_images ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_visc_div(st_solver *s, void *fl, const void *fl_prev, int stage, bool reduce, bool ord2) {
    ST_NOREC();
    const Geo &g = s->g;
    s->force_nblk = cell_blocks(g);
    const double rkr = stage >= 0 ? s->rhork[stage] : 0.0, rkg = stage >= 0 ? s->gamrk[stage] : 0.0;
#define ST_LAUNCH_D(K_, RED_, A_, B_, C_)                                                                           \
    ST_DISPATCH(hipLaunchKernelGGL((K_<R, RED_>), cell_grid(g), kCellBlock, 0, s->stream, g, s->kp,                  \
                                   (const R *)s->w_aux, (R *)fl, (const R *)fl_prev, s->w, A_, B_, C_, s->yn, s->ds,  \
                                   rkr, rkg, s->partials))
    if (ord2) {
        if (reduce) ST_LAUNCH_D(k_visc_div2, true, s->x, s->y, s->z);
        else ST_LAUNCH_D(k_visc_div2, false, s->x, s->y, s->z);
    } else {
        if (reduce) ST_LAUNCH_D(k_visc_div, true, s->dcsidx, s->detady, s->dzitdz);
        else ST_LAUNCH_D(k_visc_div, false, s->dcsidx, s->detady, s->dzitdz);
    }
#undef ST_LAUNCH_D
    return hipGetLastError();
}

hipError_t launch_force_finish(st_solver *s, int stage) {
    ST_GO(k_force_finish, , RED_THREADS, 0, dim3(red_blocks(s->force_nblk)), dim3(RED_THREADS), s->kp, s->partials,
          s->force_nblk,
          s->partials + s->npart - RB * 8, s->counters + 1, s->ds, stage == 0 ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_update(st_solver *s, const void *fl_prev, const void *fl, int stage) {
    const Geo &g = s->g;
    const bool fv = s->p.channel_case && s->p.theta_wall >= -1.0;
#define ST_LAUNCH_U(FV_)                                                                                             \
    ST_DISPATCH(ST_GO(k_update, ST_T(R, FV_), 256, 0, cell_grid(g), kCellBlock, g, s->kp, s->w, (R *)s->w_aux,      \
                      (const R *)fl_prev, (const R *)fl, s->ds, s->rhork[stage], s->gamrk[stage], s->p.channel_case, \
                      s->yn, s->partials))
    if (fv) ST_LAUNCH_U(true);
    else ST_LAUNCH_U(false);
#undef ST_LAUNCH_U
    return hipGetLastError();
}

hipError_t launch_recycling(st_solver *s) {
    ST_NOREC();
    const Geo &g = s->g;
    double *av = s->rc_real + (size_t)g.ng * g.ny * 4 + g.ny + 4 * (size_t)(g.ny + 2 * g.ng);
    const int n1 = g.ng * g.ny, n2 = g.ny * g.nz;
    hipError_t e = launch_recyc_exchange(s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_recyc_avg, dim3((5 * n1 + 127) / 128), dim3(128), 0, s->stream, g, s->wrecyc, av);
    const double u0_02 = s->p.rand_type != 0 && s->rplane ? 0.02 * s->p.u0 : 0.0;
    hipLaunchKernelGGL(k_recyc_inflow, dim3((n2 + 127) / 128), dim3(128), 0, s->stream, g, s->kp, s->w, s->rc_real,
                       s->rc_maps, s->wrecyc, s->betarecyc, s->rplane, u0_02);
    return hipGetLastError();
}

hipError_t launch_inflow_draws(st_solver *s) {
    ST_NOREC();
    const int n = s->rng_n;
    if (!s->rplane || n <= 0) return hipSuccess;
    const int nt = n > 31 ? n : 31;
    hipLaunchKernelGGL(k_inflow_draws, dim3((nt + 255) / 256), dim3(256), 0, s->stream, s->g, s->rng_coef, n,
                       s->rng_win, s->rng_ctl, s->rplane);
    return hipGetLastError();
}

hipError_t launch_rng_reset(st_solver *s) {
    ST_NOREC();
    if (!s->rplane) return hipSuccess;
    Win31 seeded;
    for (int q = 0; q < 31; q++) seeded.v[q] = s->rng_seeded[q];
    hipLaunchKernelGGL(k_rng_reset, dim3(1), dim3(64), 0, s->stream, seeded, s->rng_win, s->rng_ctl);
    return hipGetLastError();
}

hipError_t launch_physical_ghosts(st_solver *s) {
    ST_NOREC();
    const Geo &g = s->g;
    if (s->recyc) {
        hipError_t e = launch_recycling(s);  // recyc_exchange, then face xmin
        if (e != hipSuccess) return e;
    }
    for (int ilat = 1; ilat <= 2 * s->kp.ndim; ilat++) {
        const int tag = s->kp.bctag[ilat - 1];
        if (tag == 0 || tag == 8 || tag == 10) continue;
        if (tag == 9 && !s->wmean_set) continue;
        const int n = (ilat <= 2 ? g.ny : g.nx) * (ilat <= 4 ? g.nz : g.ny);
        ST_DISPATCH(hipLaunchKernelGGL(k_bc_face<R>, dim3((n + 127) / 128), dim3(128), 0, s->stream, g, s->kp, s->w,
                                       (const R *)s->w_aux, s->x, s->y, ilat, tag, s->p.is_y_staggered,
                                       (const double *)s->rc_real));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_bc_nr(st_solver *s, void *fl) {
    ST_NOREC();
    const Geo &g = s->g;
    for (int ilat = 1; ilat <= 2 * s->kp.ndim; ilat++) {
        const int nr = s->kp.nrtag[ilat - 1];
        if (nr == 0) continue;
        const int dir = (ilat - 1) / 2, start = (ilat - 1) % 2 == 0 ? 1 : 0;
        const int n = (dir == 0 ? g.ny : g.nx) * (dir == 2 ? g.ny : g.nz);
        const dim3 grid((n + 127) / 128), block(128);
        if (dir == 0)
            ST_DISPATCH(hipLaunchKernelGGL((k_bc_nr<R, 0>), grid, block, 0, s->stream, g, s->kp, s->w,
                                           (const R *)s->w_aux, (R *)fl, s->dcsidx, start, nr));
        else if (dir == 1)
            ST_DISPATCH(hipLaunchKernelGGL((k_bc_nr<R, 1>), grid, block, 0, s->stream, g, s->kp, s->w,
                                           (const R *)s->w_aux, (R *)fl, s->detady, start, nr));
        else
            ST_DISPATCH(hipLaunchKernelGGL((k_bc_nr<R, 2>), grid, block, 0, s->stream, g, s->kp, s->w,
                                           (const R *)s->w_aux, (R *)fl, s->dzitdz, start, nr));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_force_var(st_solver *s) {
    ST_NOREC();
    const Geo &g = s->g;
    hipLaunchKernelGGL(k_force_var_finish, dim3(red_blocks(cell_blocks(g))), dim3(RED_THREADS), 0, s->stream, s->kp,
                       s->partials, cell_blocks(g), s->partials + s->npart - RB * 8, s->counters + 1, s->ds);
    ST_DISPATCH(hipLaunchKernelGGL(k_force_var2<R>, cell_grid(g), kCellBlock, 0, s->stream, g, s->kp, s->w,
                                   (const R *)s->w_aux, s->ds));
    return launch_wall_ghosts(s);
}

int dt_blocks_per_plane(const Geo &g) { return ((g.nx + 63) / 64) * ((g.ny + 3) / 4); }

hipError_t launch_dt_raw(hipStream_t stream, bool mixed, const Geo &g, const KPar &P, const double *w, void *wa,
                         const double *dx1, const double *dy1, const double *dz1, const double *dxs,
                         const double *dys, const double *dzs, double *partials, int kplanes, double *stage2,
                         int *ticket, double *dt_out) {
    static_assert(kDtStage2 == RB, "second-level slots of k_dt_finish");
    const dim3 grid((g.nx + 63) / 64, (g.ny + 3) / 4, kplanes);
    if (mixed)
        hipLaunchKernelGGL(k_dt<float>, grid, kCellBlock, 0, stream, g, P, w, (float *)wa, dx1, dy1, dz1, dxs, dys,
                           dzs, partials, ticket);
    else
        hipLaunchKernelGGL(k_dt<double>, grid, kCellBlock, 0, stream, g, P, w, (double *)wa, dx1, dy1, dz1, dxs, dys,
                           dzs, partials, ticket);
    hipLaunchKernelGGL(k_dt_finish, dim3(red_blocks((long long)dt_blocks_per_plane(g) * kplanes)), dim3(RED_THREADS),
                       0, stream, P, partials,
                       dt_blocks_per_plane(g) * kplanes, stage2, ticket, dt_out);
    return hipGetLastError();
}

hipError_t launch_dt(st_solver *s) {
    ST_NOREC();
    return launch_dt_raw(s->stream, s->mixed, s->g, s->kp, s->w, s->w_aux, s->dcsidx, s->detady, s->dzitdz,
                         s->dcsidxs, s->detadys, s->dzitdzs, s->partials, s->g.nz, s->partials + s->npart - RB * 8,
                         s->counters + 1, &s->ds->dt);
}

hipError_t launch_step_end(st_solver *s) {
    ST_GO(k_step_end, , 1024, 0, dim3(1), dim3(1), s->ds);
    return hipGetLastError();
}

hipError_t launch_residual(st_solver *s) {
    ST_NOREC();
    const Geo &g = s->g;
    const int st = s->last_stage < 0 ? 0 : s->last_stage;
    const double rkr = s->last_stage < 0 ? 0.0 : s->rhork[st];
    const double rkg = s->last_stage < 0 ? 0.0 : s->gamrk[st];
    ST_DISPATCH(hipLaunchKernelGGL(k_residual<R>, cell_grid(g), kCellBlock, 0, s->stream, g, s->kp,
                                   (const R *)s->w_aux, (const R *)s->F[s->cur ^ 1], (const R *)s->F[s->cur], s->ds,
                                   rkr, rkg, s->p.channel_case, s->partials,
                                   s->fln_current ? (const double *)s->fln : nullptr));
    hipLaunchKernelGGL(k_residual_finish, dim3(1), dim3(RED_THREADS), 0, s->stream, g, s->partials, cell_blocks(g), s->ds);
    return hipGetLastError();
}

hipError_t launch_count_weno(st_solver *s) {
    ST_NOREC();
    ST_DISPATCH(hipLaunchKernelGGL(k_count_weno<R>, cell_grid(s->g), kCellBlock, 0, s->stream, s->g, s->kp,
                                   (const R *)s->w_aux, s->ds));
    return hipGetLastError();
}

hipError_t launch_stats(st_solver *s, int itav) {
    ST_NOREC();
    const Geo &g = s->g;
    const dim3 grid((g.nx + STAT_BX - 1) / STAT_BX, g.ny);
    ST_DISPATCH(hipLaunchKernelGGL(k_stats<R>, grid, dim3(STAT_BX), 0, s->stream, g, s->kp, s->w,
                                   (const R *)s->w_aux, s->dcsidx, s->detady, s->dzitdz, s->w_stat, itav));
    return hipGetLastError();
}

// initialize (singleideal.F90:3032): w_aux(:,:,:,6) = t0, the first Newton
// start of T(e) for a thermally perfect gas
hipError_t launch_init_temperature(st_solver *s) {
    ST_NOREC();
    const size_t n = s->g.nvol;
    ST_DISPATCH(hipLaunchKernelGGL(k_fill<R>, dim3(grid1d(n, 256)), dim3(256), 0, s->stream,
                                   (R *)s->w_aux + 5 * n, n, R(s->p.t0)));
    return hipGetLastError();
}

hipError_t launch_zero(st_solver *s, double *a, size_t n) {
    ST_NOREC();
    hipLaunchKernelGGL(k_zero<double>, dim3(grid1d(n, 256)), dim3(256), 0, s->stream, a, n);
    return hipGetLastError();
}

hipError_t launch_zero_rhs(st_solver *s, void *a) {
    ST_NOREC();
    const size_t n = (size_t)5 * s->g.nint;
    ST_DISPATCH(hipLaunchKernelGGL(k_zero<R>, dim3(grid1d(n, 256)), dim3(256), 0, s->stream, (R *)a, n));
    return hipGetLastError();
}

// ----------------------------------------------------- piecewise stage
template <typename R, int DIR>
static hipError_t face_range_t(st_solver *s, int f0, int f1) {
    ST_NOREC();
    const Geo &g = s->g;
    const long long n1 = DIR == 0 ? f1 - f0 + 1 : g.nx, n2 = DIR == 1 ? f1 - f0 + 1 : g.ny,
                    n3 = DIR == 2 ? f1 - f0 + 1 : g.nz;
    const long long tot = n1 * n2 * n3;
    if (tot <= 0) return hipSuccess;
    const dim3 grid((unsigned)((tot + 255) / 256));
    if (g.ng == 4)
        hipLaunchKernelGGL((k_face_range<R, DIR, 4>), grid, dim3(256), 0, s->stream, g, s->kp, (const R *)s->w_aux,
                           s->w, s->eoc[DIR], s->fhat, f0, f1);
    else
        hipLaunchKernelGGL((k_face_range<R, DIR, 3>), grid, dim3(256), 0, s->stream, g, s->kp, (const R *)s->w_aux,
                           s->w, s->eoc[DIR], s->fhat, f0, f1);
    return hipGetLastError();
}

template <typename R, int DIR>
static hipError_t face_update_t(st_solver *s, void *fl, int c0, int c1) {
    ST_NOREC();
    const Geo &g = s->g;
    const long long n1 = DIR == 0 ? c1 - c0 + 1 : g.nx, n2 = DIR == 1 ? c1 - c0 + 1 : g.ny,
                    n3 = DIR == 2 ? c1 - c0 + 1 : g.nz;
    const long long tot = n1 * n2 * n3;
    if (tot <= 0) return hipSuccess;
    const double *metric = DIR == 0 ? s->dcsidx : (DIR == 1 ? s->detady : s->dzitdz);
    hipLaunchKernelGGL((k_face_update<R, DIR>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s->stream, g,
                       s->fhat, metric, (R *)fl, c0, c1);
    return hipGetLastError();
}

hipError_t launch_face_range(st_solver *s, int dir, int f0, int f1) {
    if (dir == 0) ST_DISPATCH(return (face_range_t<R, 0>(s, f0, f1)));
    if (dir == 1) ST_DISPATCH(return (face_range_t<R, 1>(s, f0, f1)));
    ST_DISPATCH(return (face_range_t<R, 2>(s, f0, f1)));
}

hipError_t launch_face_update(st_solver *s, int dir, void *fl, int c0, int c1) {
    if (dir == 0) ST_DISPATCH(return (face_update_t<R, 0>(s, fl, c0, c1)));
    if (dir == 1) ST_DISPATCH(return (face_update_t<R, 1>(s, fl, c0, c1)));
    ST_DISPATCH(return (face_update_t<R, 2>(s, fl, c0, c1)));
}

hipError_t launch_flux_step(st_solver *s, void *fl, double coef, bool init) {
    ST_NOREC();
    const size_t n = 5 * (size_t)s->g.nint;
    ST_DISPATCH(hipLaunchKernelGGL(k_init_flux<R>, dim3(grid1d(n, 256)), dim3(256), 0, s->stream, n, (R *)fl,
                                   s->fln, coef, init ? 1 : 0));
    return hipGetLastError();
}

hipError_t launch_update_field(st_solver *s) {
    ST_NOREC();
    hipLaunchKernelGGL(k_update_field, cell_grid(s->g), kCellBlock, 0, s->stream, s->g, s->w, (const double *)s->fln);
    return hipGetLastError();
}

hipError_t launch_force_rhs_piece(st_solver *s) {
    ST_NOREC();
    const Geo &g = s->g;
    ST_DISPATCH(hipLaunchKernelGGL(k_force_rhs_sums<R>, cell_grid(g), kCellBlock, 0, s->stream, g, s->w,
                                   (const R *)s->w_aux, (const double *)s->fln, s->yn, s->partials));
    hipLaunchKernelGGL(k_force_finish, dim3(red_blocks(cell_blocks(g))), dim3(RED_THREADS), 0, s->stream, s->kp,
                       s->partials, cell_blocks(g),
                       s->partials + s->npart - RB * 8, s->counters + 1, s->ds, 0);
    ST_DISPATCH(hipLaunchKernelGGL(k_force_rhs_apply<R>, cell_grid(g), kCellBlock, 0, s->stream, g,
                                   (const R *)s->w_aux, s->fln, s->ds));
    return hipGetLastError();
}

hipError_t launch_force_var_piece(st_solver *s) {
    ST_NOREC();
    const Geo &g = s->g;
    ST_DISPATCH(hipLaunchKernelGGL(k_force_var_sums<R>, cell_grid(g), kCellBlock, 0, s->stream, g, s->kp, s->w,
                                   (R *)s->w_aux, s->yn, s->partials));
    hipLaunchKernelGGL(k_force_var_finish, dim3(red_blocks(cell_blocks(g))), dim3(RED_THREADS), 0, s->stream, s->kp,
                       s->partials, cell_blocks(g), s->partials + s->npart - RB * 8, s->counters + 1, s->ds);
    ST_DISPATCH(hipLaunchKernelGGL(k_force_var2<R>, cell_grid(g), kCellBlock, 0, s->stream, g, s->kp, s->w,
                                   (const R *)s->w_aux, s->ds));
    return hipGetLastError();
}

hipError_t launch_recyc_exchange(st_solver *s) {
    ST_NOREC();
    const Geo &g = s->g;
    const long long n = (long long)g.ng * g.ny * g.nz;
    hipLaunchKernelGGL(k_recyc_exchange, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->stream, g, s->w,
                       s->i_recyc, s->wrecyc);
    return hipGetLastError();
}

}  // namespace st
