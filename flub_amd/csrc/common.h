// Shared helpers of the flub HIP library: error reporting, launch geometry,
// index decoding and the neighbour stencil used by every tagged kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/flub.h"

#define FLUB_API extern "C" __attribute__((visibility("default")))

// Error codes of the C-ABI: 0 ok, < 0 argument errors, > 0 a hipError_t.
enum {
    FLUB_OK = 0,
    FLUB_EINVAL = -1,     // bad argument (null pointer, size, id)
    FLUB_ESHAPE = -2,     // grid shape not supported (odd/too small/too large)
    FLUB_ESTATE = -3,     // call out of order (e.g. solve before set_tags)
    FLUB_ENOTCONV = -4,   // solver hit its iteration cap (result still valid)
    FLUB_ENONFINITE = -5, // NaN/Inf met in an input or a reduction
};

void flub_set_error(const char *fmt, ...);

#define FLUB_CHECK_HIP(expr)                                                          \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            flub_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),     \
                           __FILE__, __LINE__);                                       \
            return (int)e_;                                                           \
        }                                                                             \
    } while (0)

#define FLUB_REQUIRE(cond, code, ...)                                                 \
    do {                                                                              \
        if (!(cond)) {                                                                \
            flub_set_error(__VA_ARGS__);                                              \
            return (code);                                                            \
        }                                                                             \
    } while (0)

// After a launch: report configuration errors immediately.
#define FLUB_CHECK_LAUNCH() FLUB_CHECK_HIP(hipGetLastError())

static inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

static inline unsigned blocks_for(size_t n, unsigned bs)
{
    return (unsigned)((n + bs - 1) / bs);
}

// Tags in HBM are one byte per cell: 0 AIR, 1 FLUID, 2 SOLID.
enum : uint8_t { TAG_AIR = 0, TAG_FLUID = 1, TAG_SOLID = 2 };

// Cell index ((i*ny)+j)*nz+k, k fastest (the reference's wp.array3d layout).
__device__ __forceinline__ uint32_t cidx(int i, int j, int k, int ny, int nz)
{
    return ((uint32_t)i * (uint32_t)ny + (uint32_t)j) * (uint32_t)nz + (uint32_t)k;
}

// The 7-point tagged neighbour sum of multigrid/mgpcg.py:118-142
// (neighbor_dot_tag), evaluated in the reference's exact operation order:
// per axis a = (x'_+ - x_I)*f_+, b = (x'_- - x_I)*f_-, ret += a + b, with
// clamped neighbours (a domain wall reads the cell itself), f = [tag < 2] and
// x' = 0 on AIR.  divd (mgpcg.py:51-58,144-160) is the count of in-domain
// non-solid neighbours; it is produced from the same tag loads.
template <typename LoadX, typename LoadT>
__device__ __forceinline__ float ndt_stencil(LoadX X, LoadT T, int i, int j, int k, int nx, int ny,
                                             int nz, float xI, int tI, int *divd_out)
{
    float ret = 0.0f;
    int divd = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int pi = i, pj = j, pk = k, mi = i, mj = j, mk = k;
        bool pin, min_;
        if (a == 0) { pin = i + 1 < nx; min_ = i > 0; pi = pin ? i + 1 : i; mi = min_ ? i - 1 : i; }
        else if (a == 1) { pin = j + 1 < ny; min_ = j > 0; pj = pin ? j + 1 : j; mj = min_ ? j - 1 : j; }
        else { pin = k + 1 < nz; min_ = k > 0; pk = pin ? k + 1 : k; mk = min_ ? k - 1 : k; }
        const int tp = pin ? T(pi, pj, pk) : tI;
        const int tm = min_ ? T(mi, mj, mk) : tI;
        const float xp_raw = pin ? X(pi, pj, pk) : xI;
        const float xm_raw = min_ ? X(mi, mj, mk) : xI;
        const float fp = tp < 2 ? 1.0f : 0.0f;
        const float fm = tm < 2 ? 1.0f : 0.0f;
        const float xp = tp == 0 ? 0.0f : xp_raw;
        const float xm = tm == 0 ? 0.0f : xm_raw;
        const float sa = __fmul_rn(xp - xI, fp);
        const float sb = __fmul_rn(xm - xI, fm);
        ret = ret + (sa + sb);
        divd += (pin && tp < 2) + (min_ && tm < 2);
    }
    if (tI != 1) ret = 0.0f;
    if (divd_out) *divd_out = tI == 1 ? divd : 0;
    return ret;
}

// mgpcg.py:15-28 neighbor_sum (tag-free): per axis ret += x_+ + x_- - 2 x_I.
template <typename LoadX>
__device__ __forceinline__ float nsum_stencil(LoadX X, int i, int j, int k, int nx, int ny, int nz,
                                              float xI)
{
    float ret = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float xp, xm;
        if (a == 0) { xp = i + 1 < nx ? X(i + 1, j, k) : xI; xm = i > 0 ? X(i - 1, j, k) : xI; }
        else if (a == 1) { xp = j + 1 < ny ? X(i, j + 1, k) : xI; xm = j > 0 ? X(i, j - 1, k) : xI; }
        else { xp = k + 1 < nz ? X(i, j, k + 1) : xI; xm = k > 0 ? X(i, j, k - 1) : xI; }
        ret = ret + ((xp + xm) - __fmul_rn(xI, 2.0f));
    }
    return ret;
}

// XCD-aware work split.  The GPU hands workgroup b to XCD b % 8 (each XCD
// has its own L2); with a grid that is a multiple of 8 workgroups, XCD x owns
// the x-th contiguous eighth of the items and its workgroups stride through
// that slab together (so HBM still sees streaming order).  Consecutive sweeps
// over the same level then find their data in the same XCD's L2.
struct Chunk {
    uint32_t begin, end, first, stride;
};
__device__ __forceinline__ Chunk xcd_chunk(uint32_t total)
{
    const uint32_t nb8 = gridDim.x >> 3, x = blockIdx.x & 7u, lb = blockIdx.x >> 3;
    const uint32_t slab = (total + 7u) >> 3;
    const uint32_t begin = x * slab;
    const uint32_t end = begin + slab < total ? begin + slab : total;
    return {begin, end, begin + lb * 256u + threadIdx.x, nb8 * 256u};
}

// One group of a single-block pass over nblk partials, thread t summing
// part[t], part[t + BS], ... in that order: v[q] = part[g + q * BS + t] (0
// past the end, which leaves a sum or a max of non-negative values unchanged).
// The Q loads are issued together: the partials come from other XCDs' L2, so a
// loop that waits for each one in turn is latency-bound.
template <int BS, int Q>
__device__ __forceinline__ void load_partials(const double *__restrict__ part, int g, int nblk, double (&v)[Q])
{
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int b = g + q * BS + (int)threadIdx.x;
        v[q] = b < nblk ? part[b] : 0.0;
    }
}

// Deterministic block reduction of a double (fixed shuffle/LDS tree).
template <int BS>
__device__ __forceinline__ double block_sum(double v, double *lds)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < BS / 64; ++w) s += lds[w];
    }
    __syncthreads();
    return s;   // valid in thread 0
}
