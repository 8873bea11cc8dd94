// Semi-Lagrangian / MacCormack advection, the step before the projection in
// the reference's frame (advect.py:11-137), with its interpolation
// (utils/util.py:20-159): RK3 backtrace through the MAC velocity sampled
// trilinearly, tricubic (4-point Lagrange) sampling of the advected field.
//
// advect_mc is three launches in the reference: q_out = q(back(p, dt)),
// q_tmp = q(back(p, -dt)), then q_out += (q_in - q_tmp)/2 with a min/max
// limiter that falls back to q(back(p, dt)).  One kernel does all three per
// point: the first and third launches trace the same point the same way, and
// the limiter's test min < q_out < max cannot pass unless min < max, so the
// backward trace and its sample are computed only where min < max.  (The
// limiter's corners come from p / h with p already in grid units, so they
// clamp to the far corner of the array -- min == max -- everywhere except
// within one cell of the low faces; reproduced as the reference has it.)
// Every expression is evaluated in the reference's order with
// -ffp-contract=off, so the results are bitwise those of the three launches.
#include "common.h"
#include "interp.h"

namespace {

// wp.lerp
__device__ __forceinline__ float lerp_(float a, float b, float t) { return a * (1.0f - t) + b * t; }

__device__ __forceinline__ float tri8(const Arr &q, int i, int j, int k, float fi, float fj, float fk)
{
    const uint32_t i0 = q.oi(i), i1 = q.oi(i + 1), j0 = q.oj(j), j1 = q.oj(j + 1), k0 = q.ok(k), k1 = q.ok(k + 1);
    const uint32_t r00 = i0 + j0, r10 = i1 + j0, r01 = i0 + j1, r11 = i1 + j1;
    const float a = q.at(r00 + k0), b = q.at(r10 + k0), c = q.at(r01 + k0), d = q.at(r11 + k0);
    const float e = q.at(r00 + k1), f = q.at(r10 + k1), g = q.at(r01 + k1), h = q.at(r11 + k1);
    return lerp_(lerp_(lerp_(a, b, fi), lerp_(c, d, fi), fj), lerp_(lerp_(e, f, fi), lerp_(g, h, fi), fj), fk);
}

// utils/util.py:33-86 sample_velocity_3d at p (grid units)
__device__ void sample_velocity(const Arr &U, const Arr &V, const Arr &W, const float p[3], float out[3])
{
    const int iu = (int)floorf(p[0]), iv = (int)floorf(p[1]), iw = (int)floorf(p[2]);
    const float fu = p[0] - (float)iu, fv = p[1] - (float)iv, fw = p[2] - (float)iw;
    const float s = p[0] - 0.5f, t = p[1] - 0.5f, r = p[2] - 0.5f;
    const int is = (int)floorf(s), it = (int)floorf(t), ir = (int)floorf(r);
    const float fs = s - (float)is, ft = t - (float)it, fr = r - (float)ir;
    out[0] = tri8(U, iu, it, ir, fu, ft, fr);
    out[1] = tri8(V, is, iv, ir, fs, fv, fr);
    out[2] = tri8(W, is, it, iw, fs, ft, fw);
}

// advect.py:11-22 backtrace (RK3), p in grid units
__device__ void backtrace(const Arr &U, const Arr &V, const Arr &W, float p[3], float dt, float h)
{
    float v1[3], v2[3], v3[3], p1[3], p2[3];
    sample_velocity(U, V, W, p, v1);
    const float c1 = 0.5f * dt / h, c2 = 0.75f * dt / h, c3 = dt / h;
#pragma unroll
    for (int a = 0; a < 3; ++a) p1[a] = p[a] - c1 * v1[a];
    sample_velocity(U, V, W, p1, v2);
#pragma unroll
    for (int a = 0; a < 3; ++a) p2[a] = p[a] - c2 * v2[a];
    sample_velocity(U, V, W, p2, v3);
    const float k1 = 2.0f / 9.0f, k2 = 1.0f / 3.0f, k3 = 4.0f / 9.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) p[a] = p[a] - c3 * (k1 * v1[a] + k2 * v2[a] + k3 * v3[a]);
}

__device__ __forceinline__ void decode3(uint32_t t, int sy, int sz, int &i, int &j, int &k)
{
    const uint32_t row = t / (uint32_t)sz;
    k = (int)(t - row * (uint32_t)sz);
    i = (int)(row / (uint32_t)sy);
    j = (int)(row - (uint32_t)i * sy);
}

struct Vel {
    Arr U, V, W;
};

__device__ __forceinline__ Vel make_vel(const float *u, const float *v, const float *w, int nx, int ny, int nz)
{
    return Vel{Arr{u, nx + 1, ny, nz}, Arr{v, nx, ny + 1, nz}, Arr{w, nx, ny, nz + 1}};
}

// advect.py:58-112 the three MacCormack launches in one (see the top of this file)
__global__ void __launch_bounds__(256) k_advect_mc(const float *__restrict__ u, const float *__restrict__ v,
                                                   const float *__restrict__ w, int nx, int ny, int nz,
                                                   const float *__restrict__ q_in, float *__restrict__ q_out, int gx,
                                                   int gy, int gz, float stx, float sty, float stz, float dt, float h0)
{
    const uint32_t n = (uint32_t)gx * gy * gz;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    decode3(t, gy, gz, i, j, k);
    const Vel vel = make_vel(u, v, w, nx, ny, nz);
    const Arr Q{q_in, gx, gy, gz};
    const float st[3] = {stx, sty, stz};
    float p1[3] = {(float)i + stx, (float)j + sty, (float)k + stz};
    backtrace(vel.U, vel.V, vel.W, p1, dt, h0);
    const float q_sl = tricubic(Q, p1[0] - st[0], p1[1] - st[1], p1[2] - st[2]);
    // advect.py:114-131 sample_minmax(q_in, p, h0, stagger)
    const float rs = p1[0] / h0 - st[0], rt = p1[1] / h0 - st[1], rr = p1[2] / h0 - st[2];
    const int ci = (int)floorf(rs), cj = (int)floorf(rt), ck = (int)floorf(rr);
    const uint32_t i0 = Q.oi(ci), i1 = Q.oi(ci + 1), j0 = Q.oj(cj), j1 = Q.oj(cj + 1), k0 = Q.ok(ck), k1 = Q.ok(ck + 1);
    const float a = Q.at(i0 + j0 + k0), b = Q.at(i1 + j0 + k0), c = Q.at(i0 + j1 + k0), d = Q.at(i1 + j1 + k0),
                e = Q.at(i0 + j0 + k1), f = Q.at(i1 + j0 + k1), g = Q.at(i0 + j1 + k1), hh = Q.at(i1 + j1 + k1);
    const float mn = fminf(fminf(fminf(fminf(a, b), c), d), fminf(fminf(fminf(e, f), g), hh));
    const float mx = fmaxf(fmaxf(fmaxf(fmaxf(a, b), c), d), fmaxf(fmaxf(fmaxf(e, f), g), hh));
    float out = q_sl;
    if (mn < mx) {
        float p2[3] = {(float)i + stx, (float)j + sty, (float)k + stz};
        backtrace(vel.U, vel.V, vel.W, p2, -dt, h0);
        const float q_tmp = tricubic(Q, p2[0] - st[0], p2[1] - st[1], p2[2] - st[2]);
        const float qo = q_sl + 0.5f * (q_in[t] - q_tmp);
        if (mn < qo && qo < mx) out = qo;
    }
    q_out[t] = out;
}

// advect.py:38-56 advect_phi
__global__ void __launch_bounds__(256) k_advect_phi(const float *__restrict__ u, const float *__restrict__ v,
                                                    const float *__restrict__ w, int nx, int ny, int nz,
                                                    const float *__restrict__ phi, float *__restrict__ new_phi,
                                                    float stx, float sty, float stz, float dt, float h0)
{
    const uint32_t n = (uint32_t)nx * ny * nz;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    decode3(t, ny, nz, i, j, k);
    const Vel vel = make_vel(u, v, w, nx, ny, nz);
    float p[3] = {(float)i + stx, (float)j + sty, (float)k + stz};
    backtrace(vel.U, vel.V, vel.W, p, dt, h0);
    const bool inb = 0.0f <= p[0] && p[0] < (float)nx && 0.0f <= p[1] && p[1] < (float)ny && 0.0f <= p[2] &&
                     p[2] < (float)nz;
    new_phi[t] = inb ? tricubic(Arr{phi, nx, ny, nz}, p[0] - stx, p[1] - sty, p[2] - stz) : h0;
}

// advect.py:25-35 advect_simple_3d (trilinear sampling, utils/util.py:90-112)
__global__ void __launch_bounds__(256) k_advect_simple(const float *__restrict__ u, const float *__restrict__ v,
                                                       const float *__restrict__ w, const float *__restrict__ q,
                                                       float *__restrict__ new_q, int nx, int ny, int nz, float dt,
                                                       float h)
{
    const uint32_t n = (uint32_t)nx * ny * nz;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    decode3(t, ny, nz, i, j, k);
    const Vel vel = make_vel(u, v, w, nx, ny, nz);
    float p[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
    backtrace(vel.U, vel.V, vel.W, p, dt, h);
    const float s = p[0] - 0.5f, tt = p[1] - 0.5f, r = p[2] - 0.5f;
    const int iu = (int)floorf(s), iv = (int)floorf(tt), iw = (int)floorf(r);
    new_q[t] = tri8(Arr{q, nx, ny, nz}, iu, iv, iw, s - (float)iu, tt - (float)iv, r - (float)iw);
}

// positive sizes, and the array's bytes addressable by 32-bit offsets
bool dims_ok(int a, int b, int c) { return a > 0 && b > 0 && c > 0 && (uint64_t)a * b * c * 4u < (1ull << 32); }

}  // namespace

FLUB_API int flub_advect_mc(const float *u, const float *v, const float *w, int nx, int ny, int nz, const float *q_in,
                            float *q_out, int gx, int gy, int gz, float stx, float sty, float stz, float dt, float h0,
                            void *stream)
{
    FLUB_REQUIRE(u && v && w && q_in && q_out && q_in != q_out, FLUB_EINVAL, "flub_advect_mc: bad pointers");
    FLUB_REQUIRE(dims_ok(nx + 1, ny + 1, nz + 1) && dims_ok(gx, gy, gz), FLUB_EINVAL, "flub_advect_mc: bad sizes");
    const uint32_t n = (uint32_t)gx * gy * gz;
    k_advect_mc<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(u, v, w, nx, ny, nz, q_in, q_out, gx, gy, gz, stx,
                                                                  sty, stz, dt, h0);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_advect_phi(const float *u, const float *v, const float *w, int nx, int ny, int nz, const float *phi,
                             float *new_phi, float stx, float sty, float stz, float dt, float h0, void *stream)
{
    FLUB_REQUIRE(u && v && w && phi && new_phi && phi != new_phi, FLUB_EINVAL, "flub_advect_phi: bad pointers");
    FLUB_REQUIRE(dims_ok(nx + 1, ny + 1, nz + 1), FLUB_EINVAL, "flub_advect_phi: bad sizes");
    const uint32_t n = (uint32_t)nx * ny * nz;
    k_advect_phi<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(u, v, w, nx, ny, nz, phi, new_phi, stx, sty, stz,
                                                                   dt, h0);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_advect_simple_3d(const float *u, const float *v, const float *w, const float *q, float *new_q,
                                   int nx, int ny, int nz, float dt, float h, void *stream)
{
    FLUB_REQUIRE(u && v && w && q && new_q && q != new_q, FLUB_EINVAL, "flub_advect_simple_3d: bad pointers");
    FLUB_REQUIRE(dims_ok(nx + 1, ny + 1, nz + 1), FLUB_EINVAL, "flub_advect_simple_3d: bad sizes");
    const uint32_t n = (uint32_t)nx * ny * nz;
    k_advect_simple<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(u, v, w, q, new_q, nx, ny, nz, dt, h);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}
