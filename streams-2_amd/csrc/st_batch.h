// st_batch.h -- one launch over a table of solves.
//
// The stage kernels of the channel step are written as a __device__ body
// that takes its workgroup index and grid as arguments
// (K_body(blockIdx, gridDim, args...)) and a __global__ wrapper that passes
// the hardware ones.  A solver whose `rec` is set does not launch: each
// launch is recorded (body, grid, workgroup shape, arguments packed into a
// flat struct).  The server records one chunk of steps of every solve of a
// group whose launch sequences agree, then issues, for each launch of the
// sequence, one k_batched<body> launch whose grid is the concatenation of
// the solves' grids.  A workgroup finds its solve through an owner table
// (one entry per 8 workgroups; each solve's range is padded to a multiple
// of 8, so the round-robin placement of workgroups on the 8 XCDs is the
// same as in the solve's own launch) and runs the body with its solve's
// arguments and local workgroup index.  Results are those of the solve's
// own launches, bit for bit: the body is the same code, and every
// reduction indexes its partials by the local workgroup index.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <type_traits>
#include <vector>

namespace st {

// ------------------------------------------------------ packed arguments
// Pack<P...> holds one value of each parameter type, in order, as a plain
// aggregate: the same layout in host and device code, copied as bytes.
template <typename... P> struct Pack;
template <> struct Pack<> {};
template <typename H, typename... T> struct Pack<H, T...> {
    H h;
    Pack<T...> t;
};

inline Pack<> make_pack(Pack<> *) { return {}; }
template <typename H, typename... T, typename A0, typename... A>
Pack<H, T...> make_pack(Pack<H, T...> *, A0 &&a0, A &&...a) {
    return Pack<H, T...>{static_cast<H>(a0), make_pack((Pack<T...> *)nullptr, static_cast<A &&>(a)...)};
}

// parameter types of a body void(dim3 blockIdx, dim3 gridDim, P...)
template <typename F, F f> struct BodySig;
template <typename... P, void (*f)(dim3, dim3, P...)> struct BodySig<void (*)(dim3, dim3, P...), f> {
    using PackT = Pack<std::decay_t<P>...>;
};

// The body's arguments are read in place from the table through a
// constant-address-space pointer, as a kernel reads its own arguments: a
// struct argument indexed at run time (KPar's coefficient tables) is then
// not copied to scratch.  (The qualifier belongs to the device pass; the
// host pass only type-checks these bodies.)
#if defined(__HIP_DEVICE_COMPILE__)
#define ST_AS4 __attribute__((address_space(4)))
#else
#define ST_AS4
#endif
template <auto Body, typename... D>
__device__ __forceinline__ void call_body(const dim3 &b, const dim3 &g, const ST_AS4 Pack<> *, const D &...d) {
    Body(b, g, d...);
}
template <auto Body, typename H, typename... T, typename... D>
__device__ __forceinline__ void call_body(const dim3 &b, const dim3 &g, const ST_AS4 Pack<H, T...> *p,
                                          const D &...d) {
    call_body<Body>(b, g, &p->t, d..., p->h);
}

// ------------------------------------------------------ device table
// One solve of a batched launch: its first workgroup in the batched grid,
// its workgroup count and its own grid.
struct BatchEntry {
    int b0, nb, gx, gy, gz, pad[3];
};

constexpr int kBatchGroup = 8;   // workgroups per owner entry (= XCDs)
constexpr int kBatchMax = 64;    // solves in one batched launch

template <auto Body, typename PackT>
__device__ __forceinline__ void batched_entry(const BatchEntry *__restrict__ ent, const PackT *__restrict__ packs,
                                              const int *__restrict__ owner) {
    const int b = (int)blockIdx.x;
    const int r = ((const ST_AS4 int *)owner)[b / kBatchGroup];
    const ST_AS4 BatchEntry *e = (const ST_AS4 BatchEntry *)ent + r;
    const int l = b - e->b0;
    if (l >= e->nb) return;  // padding of the solve's range
    const int gx = e->gx, gy = e->gy, gxy = gx * gy;
    // the local index is uniform: keep it in scalar registers
    const int bz = __builtin_amdgcn_readfirstlane(l / gxy);
    const int rem = l - bz * gxy;
    const int by = __builtin_amdgcn_readfirstlane(rem / gx);
    const int bx = __builtin_amdgcn_readfirstlane(rem - by * gx);
    const dim3 g((unsigned)gx, (unsigned)gy, (unsigned)e->gz);
    call_body<Body>(dim3((unsigned)bx, (unsigned)by, (unsigned)bz), g, (const ST_AS4 PackT *)(packs + r));
}

template <auto Body, typename PackT, int LB>
__global__ __launch_bounds__(LB) void k_batched(const BatchEntry *__restrict__ ent, const PackT *__restrict__ packs,
                                                const int *__restrict__ owner) {
    batched_entry<Body, PackT>(ent, packs, owner);
}

// the same, for bodies whose own kernel sets a waves-per-SIMD bound
template <auto Body, typename PackT, int LB, int WPE>
__global__ __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(WPE))) void k_batched_w(
    const BatchEntry *__restrict__ ent, const PackT *__restrict__ packs, const int *__restrict__ owner) {
    batched_entry<Body, PackT>(ent, packs, owner);
}

// ------------------------------------------------------ host records
// Launches k_batched for one recorded launch: `total` workgroups of shape
// `block`, tables at device addresses ent/packs/owner.
using BatchLaunchFn = hipError_t (*)(unsigned total, dim3 block, hipStream_t stream, const void *ent,
                                     const void *packs, const int *owner);

struct BatchRec {
    BatchLaunchFn fn;  // also identifies the body instance
    dim3 grid, block;
    size_t align;
    std::vector<unsigned char> pack;
};

struct Recorder {
    std::vector<BatchRec> recs;
    bool bad = false;  // a launch that has no body form was reached
};

template <auto Body, typename PackT, int LB, int WPE>
hipError_t batched_launch(unsigned total, dim3 block, hipStream_t stream, const void *ent, const void *packs,
                          const int *owner) {
    if constexpr (WPE > 0)
        hipLaunchKernelGGL((k_batched_w<Body, PackT, LB, WPE>), dim3(total), block, 0, stream,
                           (const BatchEntry *)ent, (const PackT *)packs, owner);
    else
        hipLaunchKernelGGL((k_batched<Body, PackT, LB>), dim3(total), block, 0, stream, (const BatchEntry *)ent,
                           (const PackT *)packs, owner);
    return hipGetLastError();
}

template <auto Body, int LB, int WPE, typename... A>
void record_launch(Recorder *rec, dim3 grid, dim3 block, A &&...a) {
    using PackT = typename BodySig<decltype(Body), Body>::PackT;
    static_assert(std::is_trivially_copyable<PackT>::value, "batched arguments are copied as bytes");
    const PackT p = make_pack((PackT *)nullptr, static_cast<A &&>(a)...);
    BatchRec r;
    r.fn = &batched_launch<Body, PackT, LB, WPE>;
    r.grid = grid;
    r.block = block;
    r.align = alignof(PackT);
    r.pack.resize(sizeof(PackT));
    std::memcpy(r.pack.data(), &p, sizeof(PackT));
    rec->recs.push_back(std::move(r));
}

}  // namespace st

// Launch NAME TARGS on the solver's stream, or record it while the solver
// records.  NAME##_body is its device body; LB is the kernel's workgroup
// bound and WPE its waves-per-SIMD bound (0: none).
#define ST_T(...) <__VA_ARGS__>
#define ST_GO(NAME, TARGS, LB, WPE, GRID, BLOCK, ...)                                                   \
    do {                                                                                               \
        if (s->rec) st::record_launch<NAME##_body TARGS, LB, WPE>(s->rec, GRID, BLOCK, __VA_ARGS__);  \
        else hipLaunchKernelGGL((NAME TARGS), GRID, BLOCK, 0, s->stream, __VA_ARGS__);                \
    } while (0)
// a launch with no body form: a recording solver marks its sequence unbatchable
#define ST_NOREC()                          \
    do {                                    \
        if (s->rec) {                       \
            s->rec->bad = true;             \
            return hipSuccess;              \
        }                                   \
    } while (0)
