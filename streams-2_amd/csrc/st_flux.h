// st_flux.h -- convective face flux of the hybrid scheme (device code).
//
// One face between stencil points s=H-1 and s=H of a 2H-point window
// (H = 3: offsets -2..+3 around the face).  The window holds the primitive
// variables rho,u,v,w,H,T of w_aux(1:6) plus the pressure p = T*rho*rgas0
// and the internal energy e = H - p/rho - |u|^2/2 of each point, computed
// once per point instead of once per point pair as in the reference (the
// value is the same).  DIR selects the normal direction; the arithmetic of
// each branch follows the reference kernel of that direction:
//   central KEEP / non-KEEP split : kernels_cpu.F90:466-618 (x), 1854-2004 (y), 1109-1257 (z)
//   Roe + WENO characteristic flux: kernels_cpu.F90:619-682 (x), 2005-2064 (y), 1258-1318 (z)
//   compute_roe_average           : kernels_cpu.F90:3013-3089
//   eigenvectors_x/y/z            : kernels_cpu.F90:2728-2782, 2842-2896, 2956-3010
//   wenorec_1d                    : kernels_cpu.F90:2526-2724 (orders 1-3, WENO-JS and WENO-Z)
#pragma once

#include "st_internal.h"

namespace st {

struct Pt {
    double r, u, v, w, h, t;  // w_aux(1:6)
    double p, e;              // pressure and internal energy
};

// e carries the internal energy for the KEEP split (nkeep >= 0) and the
// total enthalpy H for the non-KEEP split (nkeep < 0), the only energy each
// split uses.
__device__ __forceinline__ Pt make_pt(double r, double u, double v, double w, double h, double t, double rgas0,
                                      bool keep) {
    Pt q;
    q.r = r; q.u = u; q.v = v; q.w = w; q.h = h; q.t = t;
    q.p = t * r * rgas0;
    q.e = keep ? h - q.p / r - 0.5 * (u * u + v * v + w * w) : h;
    return q;
}

template <int DIR>
__device__ __forceinline__ double un_of(const Pt &q) {
    return DIR == 0 ? q.u : (DIR == 1 ? q.v : q.w);
}

// The six values the central split needs at a point: rho, u, v, w, p, e
// (H and T enter only through p and e).  R is the arithmetic type of the
// central sweeps: double in the FP64 form, float in the mixed form.
template <typename R>
struct PcR {
    R r, u, v, w, p, e;
};
using Pc = PcR<double>;

// tp is T in the FP64 form (p = T*rho*rgas0) and the stored deviation
// p - p_ref in the mixed form (FP32), where q.p then holds that deviation:
// a constant part of p cancels from the momentum-flux difference, so only
// the energy flux adds p_ref back (pair_terms).
template <typename R>
__device__ __forceinline__ PcR<R> make_pcr(R r, R u, R v, R w, R h, R tp, R rgas0, R pref, bool keep) {
    PcR<R> q;
    q.r = r; q.u = u; q.v = v; q.w = w;
    if constexpr (sizeof(R) == sizeof(double)) {
        q.p = tp * r * rgas0;
        q.e = keep ? h - q.p / r - R(0.5) * (u * u + v * v + w * w) : h;
    } else {
        q.p = tp;
        q.e = keep ? h - (tp + pref) / r - R(0.5) * (u * u + v * v + w * w) : h;
    }
    return q;
}

// w_aux slot read as tp by make_pcr: T (6th variable) in FP64, p - p_ref
// (11th) in the mixed form
template <typename R>
constexpr int tp_slot() { return sizeof(R) == sizeof(double) ? 5 : 10; }

__device__ __forceinline__ Pc make_pc(double r, double u, double v, double w, double h, double t, double rgas0,
                                      bool keep) {
    return make_pcr<double>(r, u, v, w, h, t, rgas0, 0.0, keep);
}

__device__ __forceinline__ Pc pc_of(const Pt &a) {
    Pc q;
    q.r = a.r; q.u = a.u; q.v = a.v; q.w = a.w; q.p = a.p; q.e = a.e;
    return q;
}

template <int DIR, typename R>
__device__ __forceinline__ R un_of(const PcR<R> &q) {
    return DIR == 0 ? q.u : (DIR == 1 ? q.v : q.w);
}

__device__ __forceinline__ double sqr(double a) { return a * a; }

// ------------------------------------------------------------ central flux
// KM selects the split at compile time: 0 KEEP (nkeep = 0, the default),
// 1 KEEP(n) with nkeep > 0, -1 the non-KEEP split (nkeep < 0); the runtime
// nkeep of the parameters must agree with it.
template <int DIR, int H, class PT, int KM>
__device__ __forceinline__ void central_flux(const KPar &P, const PT (&q)[2 * H], int lmax, bool zero_flux,
                                             double fh[5]) {
    double ft1 = 0.0, ft2 = 0.0, ft3 = 0.0, ft4 = 0.0, ft5 = 0.0, ft6 = 0.0;
    // coefficients selected by value: a per-lane row address into the
    // argument block would make every read a vector load
    auto cd_of = [&](int l) {
        const double v3 = P.cd1tab[2][l - 1], v2 = P.cd1tab[1][l - 1], v1 = P.cd1tab[0][l - 1];
        if constexpr (H >= 4) {
            if (lmax >= 4) return P.cd1tab[3][l - 1];  // ep_order 8
        }
        return lmax >= 3 ? v3 : (lmax == 2 ? v2 : v1);
    };
    const int nkeep = KM == 0 ? 0 : (KM < 0 ? -1 : P.nkeep);
#pragma unroll
    for (int l = 1; l <= H; l++) {
        if (l <= lmax) {
            double uvs1 = 0.0, uvs2 = 0.0, uvs3 = 0.0, uvs4 = 0.0, uvs6 = 0.0;
            double uvs5_i = 0.0, uvs5_k = 0.0, uvs5_p = 0.0, uvs5 = 0.0;
#pragma unroll
            for (int m = 0; m < l; m++) {
                const PT &a = q[H - 1 - m];
                const PT &b = q[H - 1 - m + l];
                const double una = un_of<DIR>(a), unb = un_of<DIR>(b);
                const double rhom = a.r + b.r;
                if (nkeep >= 0) {
                    const double eem = a.e + b.e;
                    double uv_part = (una + unb) * rhom;
                    double deef = 1.0;
                    if (KM > 0) {
                        // KEEP(n) correction factors (kernels_cpu.F90:505-531)
                        double drho = 2.0 * (b.r - a.r) / rhom;
                        double dee = 2.0 * (b.e - a.e) / eem;
                        double t_sdr = (0.5 * drho) * (0.5 * drho);
                        double t_sde = (0.5 * dee) * (0.5 * dee);
                        const double t2_sdr = t_sdr, t2_sde = t_sde;
                        double sdr = 1.0 + t_sdr / 3.0;
                        double sde = 1.0 + t_sde;
                        double sne = 1.0 + t_sde / 3.0;
                        for (int n = 2; n <= nkeep; n++) {
                            int n2 = 2 * n;
                            t_sdr = t2_sdr * t_sdr;
                            t_sde = t2_sde * t_sde;
                            sdr = sdr + t_sdr / (1.0 + n2);
                            sde = sde + t_sde;
                            sne = sne + t_sde / (1.0 + n2);
                        }
                        uv_part = uv_part * (1.0 / sdr);
                        deef = sne / sde;
                    }
                    uvs1 = uvs1 + uv_part * 2.0;
                    uvs2 = uvs2 + uv_part * (a.u + b.u);
                    uvs3 = uvs3 + uv_part * (a.v + b.v);
                    uvs4 = uvs4 + uv_part * (a.w + b.w);
                    uvs5_i = uvs5_i + uv_part * eem * deef;
                    uvs5_k = uvs5_k + uv_part * (a.u * b.u + a.v * b.v + a.w * b.w);
                    uvs5_p = uvs5_p + 4.0 * (una * b.p + unb * a.p);
                    uvs6 = uvs6 + 2.0 * (a.p + b.p);
                } else {
                    const double uv_part = (una + unb) * rhom;
                    uvs1 = uvs1 + uv_part * 2.0;
                    uvs2 = uvs2 + uv_part * (a.u + b.u);
                    uvs3 = uvs3 + uv_part * (a.v + b.v);
                    uvs4 = uvs4 + uv_part * (a.w + b.w);
                    uvs5 = uvs5 + uv_part * (a.e + b.e);
                    uvs6 = uvs6 + 2.0 * (a.p + b.p);
                }
            }
            const double c = cd_of(l);
            ft1 = ft1 + c * uvs1;
            ft2 = ft2 + c * uvs2;
            ft3 = ft3 + c * uvs3;
            ft4 = ft4 + c * uvs4;
            ft5 = ft5 + c * (nkeep >= 0 ? (uvs5_i + uvs5_k + uvs5_p) : uvs5);
            ft6 = ft6 + c * uvs6;
        }
    }
    fh[0] = 0.25 * ft1;
    fh[1] = 0.25 * ft2;
    fh[2] = 0.25 * ft3;
    fh[3] = 0.25 * ft4;
    fh[4] = 0.25 * ft5;
    if (zero_flux) {
        fh[0] = 0.0; fh[1] = 0.0; fh[2] = 0.0; fh[3] = 0.0; fh[4] = 0.0;
    }
    fh[1 + DIR] = fh[1 + DIR] + 0.5 * ft6;
}

// ------------------------------------------------------------ Roe average
struct Roe {
    double b1, b2, b3, c, ci, h, uu, vv, ww;
};

// compute_roe_average between points A (left) and B (right)
__device__ __forceinline__ Roe roe_average(const KPar &P, const Pt &A, const Pt &B) {
    Roe s;
    double r = B.r / A.r;
    r = sqrt(r);
    const double rho = r * A.r;
    const double rp1 = 1.0 / (r + 1.0);
    const double uu = (r * B.u + A.u) * rp1;
    const double vv = (r * B.v + A.v) * rp1;
    const double ww = (r * B.w + A.w) * rp1;
    const double h = (r * B.h + A.h) * rp1;
    const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
    const double hbar = h - qq - cpc(P, P.indx_cp_r + 1) * P.t0;
    double tt, gamloc;
    if (P.calorically_perfect == 1) {
        tt = P.t0 + hbar / cpc(P, 0);
        gamloc = cpc(P, 0) / (cpc(P, 0) - P.rgas0);
    } else {
        const double t0 = P.t0;
        tt = (r * B.t + A.t) * rp1;
        double told = tt;
        for (int iter = 1; iter <= 50; iter++) {
            double num = 0.0, den = 0.0;
            for (int ll = P.indx_cp_l; ll <= P.indx_cp_r; ll++) {
                double tpow = ipow(told / t0, ll);
                if (ll == -1) {
                    den = den + cpc(P, ll) * tpow;
                    num = num + cpc(P, ll) * log(told / t0);
                } else {
                    double tpowp = (told / t0) * tpow;
                    den = den + cpc(P, ll) * tpow;
                    num = num + cpc(P, ll) * (tpowp - 1.0) / (ll + 1.0);
                }
            }
            num = num * t0;
            tt = told + (hbar - num) / den;
            if (fabs(tt - told) < P.tol_iter_nr) break;
            told = tt;
        }
        double cploc = 0.0;
        for (int ll = P.indx_cp_l; ll <= P.indx_cp_r; ll++) cploc = cploc + cpc(P, ll) * ipow(tt / t0, ll);
        gamloc = cploc / (cploc - P.rgas0);
    }
    const double cc = gamloc * tt * P.rgas0;
    const double gm1loc = gamloc - 1.0;
    s.c = sqrt(cc);
    s.ci = 1.0 / s.c;
    const double p_rho = tt * P.rgas0;
    const double p_e = rho * gm1loc;
    const double etot = h - tt * P.rgas0;
    s.b3 = etot - rho * p_rho / p_e;
    s.b2 = p_e / (rho * cc);
    s.b1 = p_rho / cc - s.b2 * (etot - 2.0 * qq);
    s.h = h;
    s.uu = uu;
    s.vv = vv;
    s.ww = ww;
    return s;
}

// Left eigenvector m (component v = 0..4) and right eigenvector m of the
// eigenvectors_x/y/z routines.  The y and z sets are the x set with the
// velocity components permuted (normal, tangential 1, tangential 2) =
// (v,u,w) for y and (w,u,v) for z; the component order stays (rho, rho*u,
// rho*v, rho*w, rho*E) so the sums below run in the reference order.
template <int DIR>
__device__ __forceinline__ void eig_left(const Roe &s, int m, double el[5]) {
    const double un = DIR == 0 ? s.uu : (DIR == 1 ? s.vv : s.ww);
    const double t1 = DIR == 0 ? s.vv : s.uu;
    const double t2 = DIR == 2 ? s.vv : s.ww;
    const int in = 1 + DIR;                    // component of the normal momentum
    const int it1 = DIR == 0 ? 2 : 1;          // component of tangential momentum 1
    const int it2 = DIR == 2 ? 2 : 3;          // component of tangential momentum 2
    double col[5];
    if (m == 0 || m == 4) {
        const double sg = m == 0 ? 1.0 : -1.0;
        col[0] = 0.5 * (s.b1 + sg * un * s.ci);
        col[in] = -0.5 * (s.b2 * un + sg * s.ci);
        col[it1] = -0.5 * (s.b2 * t1);
        col[it2] = -0.5 * (s.b2 * t2);
        col[4] = 0.5 * s.b2;
    } else if (m == 1) {
        col[0] = 1.0 - s.b1;
        col[1] = s.b2 * s.uu;
        col[2] = s.b2 * s.vv;
        col[3] = s.b2 * s.ww;
        col[4] = -s.b2;
    } else if (m == 2) {
        col[0] = -t1; col[in] = 0.0; col[it1] = 1.0; col[it2] = 0.0; col[4] = 0.0;
    } else {
        col[0] = -t2; col[in] = 0.0; col[it1] = 0.0; col[it2] = 1.0; col[4] = 0.0;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) el[v] = col[v];
}

template <int DIR>
__device__ __forceinline__ void eig_right(const Roe &s, int m, double er[5]) {
    const double un = DIR == 0 ? s.uu : (DIR == 1 ? s.vv : s.ww);
    const double t1 = DIR == 0 ? s.vv : s.uu;
    const double t2 = DIR == 2 ? s.vv : s.ww;
    const int in = 1 + DIR;
    const int it1 = DIR == 0 ? 2 : 1;
    const int it2 = DIR == 2 ? 2 : 3;
    double row[5];
    if (m == 0 || m == 4) {
        const double sg = m == 0 ? -1.0 : 1.0;
        row[0] = 1.0;
        row[in] = m == 0 ? un - s.c : un + s.c;
        row[it1] = t1;
        row[it2] = t2;
        row[4] = m == 0 ? s.h - un * s.c : s.h + un * s.c;
        (void)sg;
    } else if (m == 1) {
        row[0] = 1.0; row[1] = s.uu; row[2] = s.vv; row[3] = s.ww; row[4] = s.b3;
    } else if (m == 2) {
        row[0] = 0.0; row[in] = 0.0; row[it1] = 1.0; row[it2] = 0.0; row[4] = t1;
    } else {
        row[0] = 0.0; row[in] = 0.0; row[it1] = 0.0; row[it2] = 1.0; row[4] = t2;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) er[v] = row[v];
}

// wenorec_1d for one characteristic field, iweno = 3 (window of 6 values:
// vp/vm index 0..5 <-> l = 1..6 of the reference, i = 3 <-> index 2).
__device__ __forceinline__ double wenorec3(const double vp[6], const double vm[6], int ord, int version,
                                           double betascale) {
#define VP(o) vp[2 + (o)]
#define VM(o) vm[2 + (o)]
    if (ord <= 1) return VP(0) + VM(1);
    if (ord == 2) {
        const double d0 = 1.0 / 3.0, d1 = 2.0 / 3.0;
        double bp0 = sqr(VP(0) - VP(-1)), bp1 = sqr(VP(1) - VP(0));
        bp1 = betascale * bp1;
        bp0 = betascale * bp0;
        double bm0 = sqr(VM(2) - VM(1)), bm1 = sqr(VM(1) - VM(0));
        bm1 = betascale * bm1;
        bm0 = betascale * bm0;
        bp0 = d0 / sqr(0.000001 + bp0);
        bm0 = d0 / sqr(0.000001 + bm0);
        double sump = bp0, summ = bm0;
        bp1 = d1 / sqr(0.000001 + bp1);
        bm1 = d1 / sqr(0.000001 + bm1);
        sump = sump + bp1;
        summ = summ + bm1;
        bp0 = bp0 / sump; bp1 = bp1 / sump;
        bm0 = bm0 / summ; bm1 = bm1 / summ;
        const double vminus = bp0 * (-VP(-1) + 3.0 * VP(0)) + bp1 * (VP(0) + VP(1));
        const double vplus = bm0 * (-VM(2) + 3.0 * VM(1)) + bm1 * (VM(0) + VM(1));
        return 0.5 * (vminus + vplus);
    }
    const double dw0 = 1.0 / 10.0, dw1 = 6.0 / 10.0, dw2 = 3.0 / 10.0;
    const double d0 = 13.0 / 12.0, d1 = 1.0 / 4.0;
    const double c0 = 1.0 / 3.0, c1 = 5.0 / 6.0, c2 = -1.0 / 6.0, c3 = -7.0 / 6.0, c4 = 11.0 / 6.0;
    double bp2 = d0 * sqr(VP(0) - 2.0 * VP(1) + VP(2)) + d1 * sqr(3.0 * VP(0) - 4.0 * VP(1) + VP(2));
    double bp1 = d0 * sqr(VP(-1) - 2.0 * VP(0) + VP(1)) + d1 * sqr(VP(-1) - VP(1));
    double bp0 = d0 * sqr(VP(0) - 2.0 * VP(-1) + VP(-2)) + d1 * sqr(3.0 * VP(0) - 4.0 * VP(-1) + VP(-2));
    bp2 = betascale * bp2;
    bp1 = betascale * bp1;
    bp0 = betascale * bp0;
    double bm2 = d0 * sqr(VM(1) - 2.0 * VM(0) + VM(-1)) + d1 * sqr(3.0 * VM(1) - 4.0 * VM(0) + VM(-1));
    double bm1 = d0 * sqr(VM(2) - 2.0 * VM(1) + VM(0)) + d1 * sqr(VM(2) - VM(0));
    double bm0 = d0 * sqr(VM(1) - 2.0 * VM(2) + VM(3)) + d1 * sqr(3.0 * VM(1) - 4.0 * VM(2) + VM(3));
    bm2 = betascale * bm2;
    bm1 = betascale * bm1;
    bm0 = betascale * bm0;
    if (version == 0) {
        bp0 = dw0 / sqr(0.000001 + bp0);
        bm0 = dw0 / sqr(0.000001 + bm0);
        bp1 = dw1 / sqr(0.000001 + bp1);
        bm1 = dw1 / sqr(0.000001 + bm1);
        bp2 = dw2 / sqr(0.000001 + bp2);
        bm2 = dw2 / sqr(0.000001 + bm2);
    } else {
        // WENO-Z; eps40 is the single-precision literal 1.e-35 of the reference
        const double eps40 = (double)1.e-35f;
        const double tau5p = fabs(bp0 - bp2) + eps40;
        const double tau5m = fabs(bm0 - bm2) + eps40;
        bp0 = (bp0 + eps40) / (bp0 + tau5p);
        bm0 = (bm0 + eps40) / (bm0 + tau5m);
        bp1 = (bp1 + eps40) / (bp1 + tau5p);
        bm1 = (bm1 + eps40) / (bm1 + tau5m);
        bp2 = (bp2 + eps40) / (bp2 + tau5p);
        bm2 = (bm2 + eps40) / (bm2 + tau5m);
        bp0 = dw0 / bp0; bm0 = dw0 / bm0;
        bp1 = dw1 / bp1; bm1 = dw1 / bm1;
        bp2 = dw2 / bp2; bm2 = dw2 / bm2;
    }
    const double sump = bp0 + bp1 + bp2;
    const double summ = bm0 + bm1 + bm2;
    bp0 = bp0 / sump; bp1 = bp1 / sump; bp2 = bp2 / sump;
    bm0 = bm0 / summ; bm1 = bm1 / summ; bm2 = bm2 / summ;
    const double vminus = bp2 * (c0 * VP(0) + c1 * VP(1) + c2 * VP(2)) +
                          bp1 * (c2 * VP(-1) + c1 * VP(0) + c0 * VP(1)) +
                          bp0 * (c0 * VP(-2) + c3 * VP(-1) + c4 * VP(0));
    const double vplus = bm2 * (c0 * VM(1) + c1 * VM(0) + c2 * VM(-1)) +
                         bm1 * (c2 * VM(2) + c1 * VM(1) + c0 * VM(0)) +
                         bm0 * (c0 * VM(3) + c3 * VM(2) + c4 * VM(1));
    return vminus + vplus;
#undef VP
#undef VM
}

// wenorec_1d for one characteristic field with iweno = 4 (window of 8
// values, l = 1..8 of the reference, i = 4 <-> index 3): orders 1-3 as
// wenorec3 on the middle six values; order 4 is WENO7 (Jiang-Shu weights
// (1, 12, 18, 4)/35 and the Balsara-Shu smoothness indicators,
// kernels_cpu.F90:2668-2724; the reference takes WENO-Z for order 3 only).
__device__ __forceinline__ double wenorec4(const double vp[8], const double vm[8], int ord, int version,
                                           double betascale) {
    if (ord <= 3) return wenorec3(vp + 1, vm + 1, ord, version, betascale);
#define A(o) vp[3 + (o)]
#define B(o) vm[3 + (o)]
    const double d1 = 1.0 / 36.0, d2 = 13.0 / 12.0, d3 = 781.0 / 720.0;
    const double dw0 = 1.0 / 35.0, dw1 = 12.0 / 35.0, dw2 = 18.0 / 35.0, dw3 = 4.0 / 35.0;
    double bp3 = d1 * sqr(-11.0 * A(0) + 18.0 * A(1) - 9.0 * A(2) + 2.0 * A(3)) +
                 d2 * sqr(2.0 * A(0) - 5.0 * A(1) + 4.0 * A(2) - A(3)) + d3 * sqr(-A(0) + 3.0 * A(1) - 3.0 * A(2) + A(3));
    double bp2 = d1 * sqr(-2.0 * A(-1) - 3.0 * A(0) + 6.0 * A(1) - A(2)) + d2 * sqr(A(-1) - 2.0 * A(0) + A(1)) +
                 d3 * sqr(-A(-1) + 3.0 * A(0) - 3.0 * A(1) + A(2));
    double bp1 = d1 * sqr(A(-2) - 6.0 * A(-1) + 3.0 * A(0) + 2.0 * A(1)) + d2 * sqr(A(-1) - 2.0 * A(0) + A(1)) +
                 d3 * sqr(-A(-2) + 3.0 * A(-1) - 3.0 * A(0) + A(1));
    double bp0 = d1 * sqr(-2.0 * A(-3) + 9.0 * A(-2) - 18.0 * A(-1) + 11.0 * A(0)) +
                 d2 * sqr(-A(-3) + 4.0 * A(-2) - 5.0 * A(-1) + 2.0 * A(0)) +
                 d3 * sqr(-A(-3) + 3.0 * A(-2) - 3.0 * A(-1) + A(0));
    double bm3 = d1 * sqr(-11.0 * B(1) + 18.0 * B(0) - 9.0 * B(-1) + 2.0 * B(-2)) +
                 d2 * sqr(2.0 * B(1) - 5.0 * B(0) + 4.0 * B(-1) - B(-2)) + d3 * sqr(-B(1) + 3.0 * B(0) - 3.0 * B(-1) + B(-2));
    double bm2 = d1 * sqr(-2.0 * B(2) - 3.0 * B(1) + 6.0 * B(0) - B(-1)) + d2 * sqr(B(2) - 2.0 * B(1) + B(0)) +
                 d3 * sqr(-B(2) + 3.0 * B(1) - 3.0 * B(0) + B(-1));
    double bm1 = d1 * sqr(B(3) - 6.0 * B(2) + 3.0 * B(1) + 2.0 * B(0)) + d2 * sqr(B(2) - 2.0 * B(1) + B(0)) +
                 d3 * sqr(-B(3) + 3.0 * B(2) - 3.0 * B(1) + B(0));
    double bm0 = d1 * sqr(-2.0 * B(4) + 9.0 * B(3) - 18.0 * B(2) + 11.0 * B(1)) +
                 d2 * sqr(-B(4) + 4.0 * B(3) - 5.0 * B(2) + 2.0 * B(1)) + d3 * sqr(-B(4) + 3.0 * B(3) - 3.0 * B(2) + B(1));
    bp3 = betascale * bp3; bp2 = betascale * bp2; bp1 = betascale * bp1; bp0 = betascale * bp0;
    bm3 = betascale * bm3; bm2 = betascale * bm2; bm1 = betascale * bm1; bm0 = betascale * bm0;
    bp0 = dw0 / sqr(0.000001 + bp0); bm0 = dw0 / sqr(0.000001 + bm0);
    bp1 = dw1 / sqr(0.000001 + bp1); bm1 = dw1 / sqr(0.000001 + bm1);
    bp2 = dw2 / sqr(0.000001 + bp2); bm2 = dw2 / sqr(0.000001 + bm2);
    bp3 = dw3 / sqr(0.000001 + bp3); bm3 = dw3 / sqr(0.000001 + bm3);
    const double sump = bp0 + bp1 + bp2 + bp3, summ = bm0 + bm1 + bm2 + bm3;
    bp0 = bp0 / sump; bp1 = bp1 / sump; bp2 = bp2 / sump; bp3 = bp3 / sump;
    bm0 = bm0 / summ; bm1 = bm1 / summ; bm2 = bm2 / summ; bm3 = bm3 / summ;
    const double vminus = bp3 * (6.0 * A(0) + 26.0 * A(1) - 10.0 * A(2) + 2.0 * A(3)) +
                          bp2 * (-2.0 * A(-1) + 14.0 * A(0) + 14.0 * A(1) - 2.0 * A(2)) +
                          bp1 * (2.0 * A(-2) - 10.0 * A(-1) + 26.0 * A(0) + 6.0 * A(1)) +
                          bp0 * (-6.0 * A(-3) + 26.0 * A(-2) - 46.0 * A(-1) + 50.0 * A(0));
    const double vplus = bm3 * (6.0 * B(1) + 26.0 * B(0) - 10.0 * B(-1) + 2.0 * B(-2)) +
                         bm2 * (-2.0 * B(2) + 14.0 * B(1) + 14.0 * B(0) - 2.0 * B(-1)) +
                         bm1 * (2.0 * B(3) - 10.0 * B(2) + 26.0 * B(1) + 6.0 * B(0)) +
                         bm0 * (-6.0 * B(4) + 26.0 * B(3) - 46.0 * B(2) + 50.0 * B(1));
    return (vminus + vplus) / 24.0;
#undef A
#undef B
}

// reconstruction of a window of 2W values (W = 3: WENO up to order 3; W = 4: up to order 4)
template <int W>
__device__ __forceinline__ double wenorec_w(const double (&vp)[2 * W], const double (&vm)[2 * W], int ord,
                                            int version, double betascale) {
    if constexpr (W == 3) return wenorec3(vp, vm, ord, version, betascale);
    else return wenorec4(vp, vm, ord, version, betascale);
}

// Characteristic WENO flux at the face of a 2W-point window (the face lies
// between window points W-1 and W).  The eigenvalue bounds run over the
// weno_size = 2*weno_scheme points of the reference's stencil, window
// points W-ws .. W-1+ws.
template <int DIR, int W = 3>
__device__ __forceinline__ void weno_flux(const KPar &P, const Pt (&q)[2 * W], int eoc, double fh[5]) {
    const Roe s = roe_average(P, q[W - 1], q[W]);
    double evmax[3] = {-1.0, -1.0, -1.0};  // |un-c|, |un|, |un+c|
    const int ws = P.weno_scheme;
#pragma unroll
    for (int l = 0; l < 2 * W; l++) {
        if (l < W - ws || l > W - 1 + ws) continue;
        const double un = un_of<DIR>(q[l]);
        const double tt = q[l].t;
        const double gamloc = gamloc_of(P, tt);
        const double c = sqrt(gamloc * P.rgas0 * tt);
        evmax[0] = fmax(fabs(un - c), evmax[0]);
        evmax[1] = fmax(fabs(un), evmax[1]);
        evmax[2] = fmax(fabs(un + c), evmax[2]);
    }
    int ord = P.weno_scheme + eoc;
    if (ord < 1) ord = 1;
    const double u0_2 = P.u0 * P.u0;
    const double r2u2 = P.rho0 * P.rho0 * u0_2;
    const double r2u4 = r2u2 * u0_2;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // one characteristic field per iteration, not unrolled: the unrolled
    // form of the five fields (twice per fixed-up cell) outgrows the
    // instruction cache and the register file
#pragma unroll 1
    for (int m = 0; m < 5; m++) {
        double el[5];
        eig_left<DIR>(s, m, el);
        const double ev = m == 0 ? evmax[0] : (m == 4 ? evmax[2] : evmax[1]);
        double gp[2 * W], gm[2 * W];
#pragma unroll
        for (int l = 0; l < 2 * W; l++) {
            const Pt &a = q[l];
            const double rhoun = a.r * un_of<DIR>(a);
            const double pp = a.r * a.t * P.rgas0;
            double f1 = a.u * rhoun, f2 = a.v * rhoun, f3 = a.w * rhoun;
            if (DIR == 0) f1 = f1 + pp;
            if (DIR == 1) f2 = f2 + pp;
            if (DIR == 2) f3 = f3 + pp;
            double wc = 0.0, gc = 0.0;
            wc = wc + el[0] * a.r;
            gc = gc + el[0] * rhoun;
            wc = wc + el[1] * a.r * a.u;
            gc = gc + el[1] * f1;
            wc = wc + el[2] * a.r * a.v;
            gc = gc + el[2] * f2;
            wc = wc + el[3] * a.r * a.w;
            gc = gc + el[3] * f3;
            wc = wc + el[4] * (a.r * a.h - pp);
            gc = gc + el[4] * (a.h * rhoun);
            const double c = 0.5 * (gc + ev * wc);
            gp[l] = c;
            gm[l] = gc - c;
        }
        const double bs = (m == 2 || m == 3) ? 1.0 / r2u4 : 1.0 / r2u2;
        const double fi = wenorec_w<W>(gp, gm, ord, P.weno_version, bs);
        double er[5];
        eig_right<DIR>(s, m, er);
#pragma unroll
        for (int v = 0; v < 5; v++) acc[v] = acc[v] + er[v] * fi;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) fh[v] = acc[v];
}

// Rusanov WENO flux (the WENO branch of the *_hybrid_rusanov_kernel
// routines, kernels_cpu.F90:887-934, 1521-1568, 2268-2315): one bound
// max(|un-c|, |un+c|) over the stencil splits each conserved component's
// flux as 1/2 (f +- evmax u), and wenorec_1d_rusanov (:2325-2523)
// reconstructs it with the component scales 1/(rho0^2 u0^2), 1/(rho0^2
// u0^4) for the momenta and 1/(rho0^2 u0^6) for the energy.
template <int DIR, int W = 3>
__device__ __forceinline__ void weno_flux_rusanov(const KPar &P, const Pt (&q)[2 * W], int eoc, double fh[5]) {
    const int ws = P.weno_scheme;
    double evmax = -1.0;
#pragma unroll
    for (int l = 0; l < 2 * W; l++) {
        if (l < W - ws || l > W - 1 + ws) continue;
        const double un = un_of<DIR>(q[l]);
        const double tt = q[l].t;
        const double gamloc = gamloc_of(P, tt);
        const double c = sqrt(gamloc * P.rgas0 * tt);
        const double evm = fmax(fabs(un - c), fabs(un + c));
        evmax = fmax(evm, evmax);
    }
    int ord = P.weno_scheme + eoc;
    if (ord < 1) ord = 1;
    const double u0_2 = P.u0 * P.u0;
    const double r2u2 = P.rho0 * P.rho0 * u0_2;
    const double r2u4 = r2u2 * u0_2;
#pragma unroll 1
    for (int m = 0; m < 5; m++) {
        double gp[2 * W], gm[2 * W];
#pragma unroll
        for (int l = 0; l < 2 * W; l++) {
            const Pt &a = q[l];
            const double un = un_of<DIR>(a);
            const double rhoun = a.r * un;
            const double pp = a.r * a.t * P.rgas0;
            const double rhoevm = a.r * evmax;
            double evm, c;
            if (m == 0) {
                evm = rhoun;
                c = 0.5 * (evm + rhoevm);
            } else if (m == 4) {
                evm = a.h * rhoun;
                c = 0.5 * (evm + evmax * (a.r * a.h - pp));
            } else {
                const double vel = m == 1 ? a.u : (m == 2 ? a.v : a.w);
                evm = vel * rhoun;
                if (m == 1 + DIR) evm = evm + pp;
                c = 0.5 * (evm + rhoevm * vel);
            }
            gp[l] = c;
            gm[l] = evm - c;
        }
        const double bs = m == 0 ? 1.0 / r2u2 : (m == 4 ? 1.0 / r2u4 / u0_2 : 1.0 / r2u4);
        fh[m] = wenorec_w<W>(gp, gm, ord, P.weno_version, bs);
    }
}

}  // namespace st
