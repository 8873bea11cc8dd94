// Device helpers shared by the tagged multigrid kernels (smooth.hip,
// blocked.hip): storage-type conversion, vector loads/stores of 2 or 4 cells
// along k, and the neighbour sum evaluated from a 16-bit stencil code.
//
// Stencil code of a cell (built once per set_tags by k_make_codes): for the
// neighbours i+, i-, j+, j-, k+, k- two bits each (0 = SOLID or outside the
// domain, 1 = AIR, 2 = FLUID), bits 12-14 the count divd of in-domain
// non-solid neighbours (mgpcg.py:51-58,144-160), bit 15 set for a FLUID cell.
#pragma once
#include "common.h"

template <typename T> __device__ __forceinline__ float ldv(T v);
template <> __device__ __forceinline__ float ldv<float>(float v) { return v; }
template <> __device__ __forceinline__ float ldv<__half>(__half v) { return __half2float(v); }
template <> __device__ __forceinline__ float ldv<__hip_bfloat16>(__hip_bfloat16 v) { return __bfloat162float(v); }
template <typename T> __device__ __forceinline__ T stv(float v);
template <> __device__ __forceinline__ float stv<float>(float v) { return v; }
template <> __device__ __forceinline__ __half stv<__half>(float v) { return __float2half_rn(v); }
template <> __device__ __forceinline__ __hip_bfloat16 stv<__hip_bfloat16>(float v) { return __float2bfloat16(v); }

template <typename T>
__device__ __forceinline__ void ld4(const T *p, float v[4])
{
    if constexpr (sizeof(T) == 4) {
        const float4 a = *reinterpret_cast<const float4 *>(p);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    } else {
        const uint2 w = *reinterpret_cast<const uint2 *>(p);
        const T *e = reinterpret_cast<const T *>(&w);
        v[0] = ldv<T>(e[0]); v[1] = ldv<T>(e[1]); v[2] = ldv<T>(e[2]); v[3] = ldv<T>(e[3]);
    }
}
template <typename T>
__device__ __forceinline__ void st4(T *p, const float v[4])
{
    if constexpr (sizeof(T) == 4) {
        *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
        uint2 w;
        T *e = reinterpret_cast<T *>(&w);
        e[0] = stv<T>(v[0]); e[1] = stv<T>(v[1]); e[2] = stv<T>(v[2]); e[3] = stv<T>(v[3]);
        *reinterpret_cast<uint2 *>(p) = w;
    }
}
template <typename T>
__device__ __forceinline__ void ld2(const T *p, float v[2])
{
    if constexpr (sizeof(T) == 4) {
        const float2 a = *reinterpret_cast<const float2 *>(p);
        v[0] = a.x; v[1] = a.y;
    } else {
        const uint32_t w = *reinterpret_cast<const uint32_t *>(p);
        const T *e = reinterpret_cast<const T *>(&w);
        v[0] = ldv<T>(e[0]); v[1] = ldv<T>(e[1]);
    }
}

// N consecutive values of type T (N * sizeof(T) = 4, 8, 16 or 32 bytes,
// aligned to that size up to 16) as floats, and back.
template <typename T, int N>
__device__ __forceinline__ void ldn(const T *p, float v[N])
{
    constexpr int B = N * (int)sizeof(T), PER = 16 / (int)sizeof(T);
    if constexpr (B >= 16) {
#pragma unroll
        for (int c = 0; c < B / 16; ++c) {
            const uint4 w = reinterpret_cast<const uint4 *>(p)[c];
            const T *e = reinterpret_cast<const T *>(&w);
#pragma unroll
            for (int q = 0; q < PER; ++q) v[c * PER + q] = ldv<T>(e[q]);
        }
    } else if constexpr (B == 8) {
        const uint2 w = *reinterpret_cast<const uint2 *>(p);
        const T *e = reinterpret_cast<const T *>(&w);
#pragma unroll
        for (int q = 0; q < N; ++q) v[q] = ldv<T>(e[q]);
    } else {
        const uint32_t w = *reinterpret_cast<const uint32_t *>(p);
        const T *e = reinterpret_cast<const T *>(&w);
#pragma unroll
        for (int q = 0; q < N; ++q) v[q] = ldv<T>(e[q]);
    }
}
template <typename T, int N>
__device__ __forceinline__ void stn(T *p, const float v[N])
{
    constexpr int B = N * (int)sizeof(T), PER = 16 / (int)sizeof(T);
    if constexpr (B >= 16) {
#pragma unroll
        for (int c = 0; c < B / 16; ++c) {
            uint4 w;
            T *e = reinterpret_cast<T *>(&w);
#pragma unroll
            for (int q = 0; q < PER; ++q) e[q] = stv<T>(v[c * PER + q]);
            reinterpret_cast<uint4 *>(p)[c] = w;
        }
    } else if constexpr (B == 8) {
        uint2 w;
        T *e = reinterpret_cast<T *>(&w);
#pragma unroll
        for (int q = 0; q < N; ++q) e[q] = stv<T>(v[q]);
        *reinterpret_cast<uint2 *>(p) = w;
    } else {
        uint32_t w;
        T *e = reinterpret_cast<T *>(&w);
#pragma unroll
        for (int q = 0; q < N; ++q) e[q] = stv<T>(v[q]);
        *reinterpret_cast<uint32_t *>(p) = w;
    }
}
// N consecutive 16-bit stencil codes (N = 4 or 8).
template <int N>
__device__ __forceinline__ void codesn(const uint16_t *code, uint32_t base, uint32_t c[N])
{
    if constexpr (N == 8) {
        const uint4 w = *reinterpret_cast<const uint4 *>(code + base);
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) { c[2 * q] = ws[q] & 0xffffu; c[2 * q + 1] = ws[q] >> 16; }
    } else {
        const uint2 w = *reinterpret_cast<const uint2 *>(code + base);
        c[0] = w.x & 0xffffu; c[1] = w.x >> 16; c[2] = w.y & 0xffffu; c[3] = w.y >> 16;
    }
}

__device__ __forceinline__ void codes4(const uint16_t *code, uint32_t base, uint32_t c[4])
{
    const uint2 w = *reinterpret_cast<const uint2 *>(code + base);
    c[0] = w.x & 0xffffu; c[1] = w.x >> 16; c[2] = w.y & 0xffffu; c[3] = w.y >> 16;
}

// neighbor_dot_tag (mgpcg.py:118-142) from the stencil code (the self-FLUID
// test is the caller's): per axis ret += (a + b) with a = (x'_+ - x_I) for an
// AIR/FLUID neighbour (x' = 0 on AIR) and a = 0 for a SOLID/outside one.  The
// reference multiplies the latter difference by 0.0 instead; the signed zero
// that gives never changes ret (ret starts at +0 and is never -0), so results
// are bitwise those of the reference kernels.
//
// The selections are written as bit masks (v & 0 is +0.0, the value the
// reference's selects give), so the compiler emits no branches for them.
__device__ __forceinline__ float keep_if(float v, uint32_t mask) { return __uint_as_float(__float_as_uint(v) & mask); }

__device__ __forceinline__ float nsum_code(uint32_t cd, float xI, float xip, float xim, float xjp, float xjm,
                                           float xkp, float xkm)
{
    float s = 0.0f;
    const float xs[6] = {xip, xim, xjp, xjm, xkp, xkm};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t cp = (cd >> (4 * a)) & 3u, cm = (cd >> (4 * a + 2)) & 3u;
        // class 2 keeps x, class 1 reads 0 (AIR), class 0 contributes 0
        const float ta = keep_if(keep_if(xs[2 * a], 0u - (cp >> 1)) - xI, 0u - ((cp | (cp >> 1)) & 1u));
        const float tb = keep_if(keep_if(xs[2 * a + 1], 0u - (cm >> 1)) - xI, 0u - ((cm | (cm >> 1)) & 1u));
        s = s + (ta + tb);
    }
    return s;
}

// nsum_code for neighbour values that hold +0 on every non-FLUID cell (the
// state every red-black sweep leaves: the reference's smoother writes 0 there,
// mgpcg.py:72-86): an AIR neighbour then reads 0 by itself, so only SOLID /
// outside neighbours are masked -- one mask per neighbour instead of two.
// Bitwise equal to nsum_code under that condition.
__device__ __forceinline__ float nsum_code_air0(uint32_t cd, float xI, float xip, float xim, float xjp, float xjm,
                                                float xkp, float xkm)
{
    const uint32_t nz = (cd | (cd >> 1)) & 0x555u;   // bit 2n: neighbour n is AIR or FLUID
    const float xs[6] = {xip, xim, xjp, xjm, xkp, xkm};
    float s = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t mp = (uint32_t)((int32_t)(nz << (31 - 4 * a)) >> 31);
        const uint32_t mm = (uint32_t)((int32_t)(nz << (29 - 4 * a)) >> 31);
        s = s + (keep_if(xs[2 * a] - xI, mp) + keep_if(xs[2 * a + 1] - xI, mm));
    }
    return s;
}

__device__ __forceinline__ int code_divd(uint32_t cd) { return (int)((cd >> 12) & 7u); }
__device__ __forceinline__ bool code_fluid(uint32_t cd) { return (cd >> 15) != 0; }

// The red-black update in direct form, for neighbour values that hold +0 on
// every non-FLUID cell and outside the domain: z' = (r + sum of the six
// neighbours) * rcp(divd) on FLUID cells, +0 elsewhere.  divd counts the
// neighbours that are neither SOLID nor outside (mgpcg.py:145-160) and AIR
// neighbours hold 0, so this is algebraically the reference's
// z + (r + ndt(z)) / divd (mgpcg.py:72-86, 119-142) without the cell's own old
// value and without a mask per neighbour.  Only the reduced-precision sweeps
// use it: they have no reference result to match bitwise, and the masks were
// most of their instructions.
__device__ __forceinline__ float gs_direct(uint32_t cd, float r, float xip, float xim, float xjp, float xjm, float xkp,
                                           float xkm)
{
    const float s = ((xip + xim) + (xjp + xjm)) + (xkp + xkm);
    const uint32_t divd = (cd >> 12) & 7u;
    return keep_if((r + s) * __builtin_amdgcn_rcpf((float)divd), 0u - (uint32_t)(divd != 0u));
}

// One red-black Gauss-Seidel update of a cell (mgpcg.py:72-86):
// z + (r + ndt(z)) / divd on FLUID cells, +0 elsewhere (divd is 0 there; the
// quotient computed for them is discarded).  RCP: the quotient as a multiply
// by the hardware reciprocal of divd (within an ulp of the IEEE quotient; the
// reduced-precision sweeps, whose results are rounded to 16 bits or feed a
// V-cycle with 16-bit storage, use it: it has no reference counterpart to
// match bitwise, and the IEEE division is most of their instruction count).
template <bool AIR0 = false, bool RCP = false>
__device__ __forceinline__ float gs_update(uint32_t cd, float z, float r, float xip, float xim, float xjp, float xjm,
                                           float xkp, float xkm)
{
    const float s = AIR0 ? nsum_code_air0(cd, z, xip, xim, xjp, xjm, xkp, xkm)
                         : nsum_code(cd, z, xip, xim, xjp, xjm, xkp, xkm);
    const uint32_t divd = (cd >> 12) & 7u;
    const float q = RCP ? (r + s) * __builtin_amdgcn_rcpf((float)divd) : (r + s) / (float)divd;
    return keep_if(z + q, 0u - (uint32_t)(divd != 0u));
}
