// Clamped array sampling and tricubic interpolation of the reference
// (utils/util.py:20-28 sample_3d, :121-159 cubic / _tricubic), shared by the
// advection kernels and the cached-level-set resampling of the geometry
// kernels.  Every expression keeps the reference's order; the including
// sources build with -ffp-contract=off, so each operation rounds as written.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// A sampled array with its clamped-index byte offsets: sample (u, v, w) reads
// q at oi(u) + oj(v) + ok(w) bytes (utils/util.py:20-28 sample_3d clamps each
// index to the array).  The offsets of each distinct index are formed once per
// sample and the loads add them, so a load costs two 32-bit adds from the
// array's (uniform) base; the launchers keep every array below 4 GB.
struct Arr {
    const float *q;
    int gx, gy, gz;
    __device__ __forceinline__ uint32_t oi(int u) const
    {
        return (uint32_t)clampi(u, 0, gx - 1) * ((uint32_t)gy * (uint32_t)gz * 4u);
    }
    __device__ __forceinline__ uint32_t oj(int v) const { return (uint32_t)clampi(v, 0, gy - 1) * ((uint32_t)gz * 4u); }
    __device__ __forceinline__ uint32_t ok(int w) const { return (uint32_t)clampi(w, 0, gz - 1) * 4u; }
    __device__ __forceinline__ float at(uint32_t off) const
    {
        return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(q) + off);
    }
};

// utils/util.py:121-128 cubic
__device__ __forceinline__ float cubic(float a, float b, float c, float d, float s)
{
    const float t0 = -1.0f / 3.0f * s + 0.5f * s * s - 1.0f / 6.0f * s * s * s;
    const float t1 = 1.0f - s * s + 0.5f * (s * s * s - s);
    const float t2 = s + 0.5f * (s * s - s * s * s);
    const float t3 = 1.0f / 6.0f * (s * s * s - s);
    return a * t0 + b * t1 + c * t2 + d * t3;
}

// utils/util.py:130-159 _tricubic at (s, t, r), index units
__device__ float tricubic(const Arr &q, float s, float t, float r)
{
    const int iu = (int)floorf(s), iv = (int)floorf(t), iw = (int)floorf(r);
    const float fu = s - (float)iu, fv = t - (float)iv, fw = r - (float)iw;
    uint32_t oi[4], oj[4], ok[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) { oi[a] = q.oi(iu + a - 1); oj[a] = q.oj(iv + a - 1); ok[a] = q.ok(iw + a - 1); }
    float zs[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float ys[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t jk = oj[b] + ok[c];
            ys[b] = cubic(q.at(oi[0] + jk), q.at(oi[1] + jk), q.at(oi[2] + jk), q.at(oi[3] + jk), fu);
        }
        zs[c] = cubic(ys[0], ys[1], ys[2], ys[3], fv);
    }
    return cubic(zs[0], zs[1], zs[2], zs[3], fw);
}

}  // namespace
