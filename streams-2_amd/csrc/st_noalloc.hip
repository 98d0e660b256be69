// st_noalloc.hip -- the allocation-free variant of st_compute_dt for callers
// that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops, harnesses that forbid device allocation).
//
// The CFL reduction's block maxima, second-level slots and ticket live in
// the file-scope array `scratch` below, one copy per GPU, held until the
// process ends.  Its device address is looked up once per device with
// hipGetSymbolAddress (the name is a host-side stand-in that must never be
// passed to a kernel) and cached.  Every call on one GPU shares that GPU's
// array: calls that overlap on one GPU (two streams, two threads) race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "st_solver.h"

namespace {

// block maxima of k_dt: at most this many workgroups per launch
constexpr int kPartials = 32768;
constexpr int kStage2 = st::kDtStage2;
// partials, second-level maxima, then the ticket (an int in its own 256-byte slot)
constexpr int kScratchDoubles = kPartials + kStage2 + 32;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double *g_scratch[kMaxDevices];  // device address of `scratch` per device, set once

// Device address of this device's scratch array (looked up on the first
// call made with the device current).
int scratch_of_current_device(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return st::set_error(ST_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return st::set_error(ST_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return st::set_error(ST_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = (double *)ptr;
    }
    *out = g_scratch[dev];
    return ST_OK;
}

}  // namespace

extern "C" int st_compute_dt_noalloc(const st_params *p, void *hip_stream, const double *w, void *w_aux,
                                     const double *dcsidx, const double *detady, const double *dzitdz,
                                     const double *dcsidxs, const double *detadys, const double *dzitdzs,
                                     double *dt_out) {
    int rc = st::validate_params(p);
    if (rc) return rc;
    if (!w || !w_aux || !dcsidx || !detady || !dzitdz || !dcsidxs || !detadys || !dzitdzs || !dt_out)
        return st::set_error(ST_ERR_ARG, "null device array");
    const st::Geo g = st::params_to_geo(*p, false);  // the caller's arrays: the reference layout
    const int per_plane = st::dt_blocks_per_plane(g);
    if (per_plane > kPartials)
        return st::set_error(ST_ERR_ARG, "grid too large for st_compute_dt_noalloc: ceil(nx/64)*ceil(ny/4) = " +
                                             std::to_string(per_plane) + " exceeds " + std::to_string(kPartials));
    if (p->cfl < 0) return ST_OK;  // fixed time step: dt is not recomputed (as st_compute_dt)
    double *base = nullptr;
    rc = scratch_of_current_device(&base);
    if (rc) return rc;
    st::KPar k{};
    st::params_to_kpar(*p, k);
    const int kplanes = per_plane * g.nz <= kPartials ? g.nz : kPartials / per_plane;
    hipError_t e = st::launch_dt_raw((hipStream_t)hip_stream, p->precision == ST_PREC_MIXED, g, k, w, w_aux, dcsidx,
                                     detady, dzitdz, dcsidxs, detadys, dzitdzs, base, kplanes, base + kPartials,
                                     (int *)(base + kPartials + kStage2), dt_out);
    if (e != hipSuccess) return st::set_error(ST_ERR_HIP, std::string("st_compute_dt_noalloc: ") + hipGetErrorString(e));
    return ST_OK;
}
