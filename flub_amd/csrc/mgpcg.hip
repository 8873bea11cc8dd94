// MGPCG pressure solver (multigrid/mgpcg.py:234-490) for one MI355X.
//
// Handle = level pyramid (r, z in FP32, one-byte tags), x, p, Ap, FP64 device
// scalars and a partial-sum buffer for the deterministic FP64 reductions.
// Per-cell arithmetic follows the reference kernels operation by operation,
// so every kernel here is bitwise equal to the FP32 oracle on equal inputs;
// the dot products accumulate in FP64 (fixed two-stage tree) instead of
// Warp's FP32 array_inner.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>

#include "common.h"
#include "grid_ops.h"
#include "mgpcg_internal.h"
#include "smooth.h"

// ---------------------------------------------------------------- tags
__global__ void k_pack_tags(const int32_t *__restrict__ in, uint8_t *__restrict__ out, uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = (uint8_t)in[t];
}
// 4 tags per thread (n % 4 == 0, 16-byte aligned input)
__global__ void k_pack_tags4(const int32_t *__restrict__ in, uint8_t *__restrict__ out, uint32_t n4)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n4) return;
    const int4 a = reinterpret_cast<const int4 *>(in)[t];
    reinterpret_cast<uchar4 *>(out)[t] = make_uchar4((uint8_t)a.x, (uint8_t)a.y, (uint8_t)a.z, (uint8_t)a.w);
}

// mgpcg.py:162-174 restrict_tag, literally: t = 2; a 0 child -> 0; else a 1
// child -> 1 (equal to min over the children for tags in {0,1,2}).
__global__ void k_restrict_tag(const uint8_t *__restrict__ f, uint8_t *__restrict__ c, int cx, int cy,
                               int cz)
{
    const uint32_t n = (uint32_t)cx * cy * cz;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int K = t % cz, J = (t / cz) % cy, I = t / ((uint32_t)cz * cy);
    const int ny = cy * 2, nz = cz * 2;
    int tt = 2;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                const int tg = f[cidx(2 * I + ii, 2 * J + jj, 2 * K + kk, ny, nz)];
                if (tg == 0) tt = 0;
                else if (tt > 0 && tg == 1) tt = 1;
            }
    c[t] = (uint8_t)tt;
}

// mgpcg.py:51-58 _compute_divd (attribute view; the solver derives divd from
// the tags inside each stencil and never stores it).
__global__ void k_divd(const uint8_t *__restrict__ tag, float *__restrict__ divd, int nx, int ny, int nz)
{
    const uint32_t n = (uint32_t)nx * ny * nz;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = t % nz, j = (t / nz) % ny, i = t / ((uint32_t)nz * ny);
    int d = 0;
    if (tag[t] == 1) {
        d += (i + 1 < nx && tag[cidx(i + 1, j, k, ny, nz)] < 2);
        d += (i > 0 && tag[cidx(i - 1, j, k, ny, nz)] < 2);
        d += (j + 1 < ny && tag[cidx(i, j + 1, k, ny, nz)] < 2);
        d += (j > 0 && tag[cidx(i, j - 1, k, ny, nz)] < 2);
        d += (k + 1 < nz && tag[cidx(i, j, k + 1, ny, nz)] < 2);
        d += (k > 0 && tag[cidx(i, j, k - 1, ny, nz)] < 2);
    }
    divd[t] = (float)d;
}

// ---------------------------------------------------------------- elementwise
__global__ void k_fill_f32(float *__restrict__ p, float v, uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void k_scale_copy(float *__restrict__ dst, const float *__restrict__ src, float k, uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t] * k;
}

__global__ void k_mask_interior(const uint8_t *__restrict__ tag, float *__restrict__ r, uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && tag[t] != 1) r[t] = 0.0f;
}

// p = ((x*h)*(1/h))*h : the stored pressure of example_fluid.py:134,145.
__global__ void k_store_pressure(float *__restrict__ p, const float *__restrict__ x, float h, float inv_h,
                                 uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = ((x[t] * h) * inv_h) * h;
}

// ---------------------------------------------------------------- stencils
struct GX {
    const float *x;
    int ny, nz;
    __device__ float operator()(int i, int j, int k) const { return x[cidx(i, j, k, ny, nz)]; }
};
struct GT {
    const uint8_t *t;
    int ny, nz;
    __device__ int operator()(int i, int j, int k) const { return t[cidx(i, j, k, ny, nz)]; }
};

// mgpcg.py:72-86 _smooth_mcadams over (nx, ny, nz/2): phase p updates the
// cells with (i+j+k)&1 == p^1.  divd from the tags (equal to _compute_divd).
__global__ void k_smooth_mc(float *__restrict__ z, const float *__restrict__ r,
                            const uint8_t *__restrict__ tag, int nx, int ny, int nz, int phase)
{
    const int nzh = nz >> 1;
    const uint32_t n = (uint32_t)nx * ny * nzh;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int kk = t % nzh, j = (t / nzh) % ny, i = t / ((uint32_t)nzh * ny);
    const int k = (((i + j) & 1) == phase) ? kk * 2 + 1 : kk * 2;
    const uint32_t I = cidx(i, j, k, ny, nz);
    const int tI = tag[I];
    const float zI = z[I];
    int divd;
    const float nd = ndt_stencil(GX{z, ny, nz}, GT{tag, ny, nz}, i, j, k, nx, ny, nz, zI, tI, &divd);
    z[I] = divd != 0 ? zI + (r[I] + nd) / (float)divd : 0.0f;
}

// mgpcg.py:41-49 _smooth (tag-free): updates (i+j+k)&1 == phase, with the
// reference's res = shape[0] on every axis.
__global__ void k_smooth_def(float *__restrict__ z, const float *__restrict__ r, int nx, int ny, int nz,
                             int phase)
{
    const int nzh = nz >> 1;
    const uint32_t n = (uint32_t)nx * ny * nzh;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int kk = t % nzh, j = (t / nzh) % ny, i = t / ((uint32_t)nzh * ny);
    const int k = kk * 2 + ((i + j + phase) & 1);
    const int res = nx;
    const float dx = (i == 0 || i == res - 1) ? 1.0f : 2.0f;
    const float dy = (j == 0 || j == res - 1) ? 1.0f : 2.0f;
    const float dz = (k == 0 || k == res - 1) ? 1.0f : 2.0f;
    const uint32_t I = cidx(i, j, k, ny, nz);
    const float zI = z[I];
    const float ns = nsum_stencil(GX{z, ny, nz}, i, j, k, nx, ny, nz, zI);
    z[I] = zI + (r[I] + ns) / (dx + dy + dz);
}

// mgpcg.py:94-101 _restrict_mcadams as a gather: a FLUID parent sums
// r - (-ndt(z)) of its 8 children in the serial launch order (kk fastest)
// starting from 0; any other parent gets 0.  (nx,ny,nz) are fine dims.
template <bool TAGGED>
__global__ void k_restrict(const float *__restrict__ r, const float *__restrict__ z,
                           const uint8_t *__restrict__ tag, float *__restrict__ rc,
                           const uint8_t *__restrict__ tagc, int nx, int ny, int nz, float cs)
{
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
    const uint32_t n = (uint32_t)cx * cy * cz;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (TAGGED && tagc[t] != 1) { rc[t] = 0.0f; return; }
    const int K = t % cz, J = (t / cz) % cy, I = t / ((uint32_t)cz * cy);
    float s = 0.0f;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                const int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                const uint32_t c = cidx(i, j, k, ny, nz);
                const float zc = z[c];
                float a;
                if (TAGGED) a = ndt_stencil(GX{z, ny, nz}, GT{tag, ny, nz}, i, j, k, nx, ny, nz, zc, tag[c], nullptr);
                else a = nsum_stencil(GX{z, ny, nz}, i, j, k, nx, ny, nz, zc);
                s += r[c] - (-a);
            }
    rc[t] = s * cs;
}

// mgpcg.py:113-116 _prolongate.
__global__ void k_prolong(float *__restrict__ z, const float *__restrict__ zc, int nx, int ny, int nz)
{
    const uint32_t n = (uint32_t)nx * ny * nz;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = t % nz, j = (t / nz) % ny, i = t / ((uint32_t)nz * ny);
    z[t] = z[t] + zc[cidx(i >> 1, j >> 1, k >> 1, ny >> 1, nz >> 1)];
}

// Bottom level (mgpcg.py:319-336) in one workgroup: nb x (phase 0, 1) then
// nb x (phase 1, 0), a workgroup barrier between half-sweeps.
template <bool TAGGED>
__global__ void __launch_bounds__(1024) k_bottom(float *__restrict__ z, const float *__restrict__ r,
                                                 const uint8_t *__restrict__ tag, int nx, int ny, int nz,
                                                 int nb)
{
    const int nzh = nz >> 1;
    const uint32_t n = (uint32_t)nx * ny * nzh;
    for (int s = 0; s < 4 * nb; ++s) {
        const int pair = s >> 1, second = s & 1;
        const int phase = pair < nb ? second : 1 - second;
        for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
            const int kk = t % nzh, j = (t / nzh) % ny, i = t / ((uint32_t)nzh * ny);
            if (TAGGED) {
                const int k = (((i + j) & 1) == phase) ? kk * 2 + 1 : kk * 2;
                const uint32_t I = cidx(i, j, k, ny, nz);
                const int tI = tag[I];
                const float zI = z[I];
                int divd;
                const float nd = ndt_stencil(GX{z, ny, nz}, GT{tag, ny, nz}, i, j, k, nx, ny, nz, zI, tI, &divd);
                z[I] = divd != 0 ? zI + (r[I] + nd) / (float)divd : 0.0f;
            } else {
                const int k = kk * 2 + ((i + j + phase) & 1);
                const int res = nx;
                const float dx = (i == 0 || i == res - 1) ? 1.0f : 2.0f;
                const float dy = (j == 0 || j == res - 1) ? 1.0f : 2.0f;
                const float dz = (k == 0 || k == res - 1) ? 1.0f : 2.0f;
                const uint32_t I = cidx(i, j, k, ny, nz);
                const float zI = z[I];
                const float ns = nsum_stencil(GX{z, ny, nz}, i, j, k, nx, ny, nz, zI);
                z[I] = zI + (r[I] + ns) / (dx + dy + dz);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- reductions
// Every reduction: RBS-thread blocks, grid-stride over the cells with a grid
// of fixed size, one FP64 partial per block, then k_finish sums the partials
// in a fixed order.  Results are identical run to run.

// Ap = ndt(p) (compute_Ap_mcadams) or nsum(p) (compute_Ap); partial p.Ap.
template <bool TAGGED>
__global__ void __launch_bounds__(RBS) k_Ap_dot(float *__restrict__ Ap, const float *__restrict__ p,
                                                const uint8_t *__restrict__ tag, int nx, int ny, int nz,
                                                double *__restrict__ part)
{
    __shared__ double lds[RBS / 64];
    const uint32_t n = (uint32_t)nx * ny * nz;
    double acc = 0.0;
    for (uint32_t t = blockIdx.x * RBS + threadIdx.x; t < n; t += gridDim.x * RBS) {
        const int k = t % nz, j = (t / nz) % ny, i = t / ((uint32_t)nz * ny);
        const float pI = p[t];
        float a;
        if (TAGGED) a = ndt_stencil(GX{p, ny, nz}, GT{tag, ny, nz}, i, j, k, nx, ny, nz, pI, tag[t], nullptr);
        else a = nsum_stencil(GX{p, ny, nz}, i, j, k, nx, ny, nz, pI);
        Ap[t] = a;
        acc += (double)pI * (double)a;
    }
    const double s = block_sum<RBS>(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(RBS) k_dot(const float *__restrict__ a, const float *__restrict__ b,
                                             uint32_t n, double *__restrict__ part)
{
    __shared__ double lds[RBS / 64];
    double acc = 0.0;
    for (uint32_t t = blockIdx.x * RBS + threadIdx.x; t < n; t += gridDim.x * RBS)
        acc += (double)a[t] * (double)b[t];
    const double s = block_sum<RBS>(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Sum over FLUID cells of q and their count (mgpcg.py:191-196 _mean).
__global__ void __launch_bounds__(RBS) k_masked_sum(const float *__restrict__ q, const uint8_t *__restrict__ tag,
                                                    uint32_t n, double *__restrict__ part)
{
    __shared__ double lds[RBS / 64];
    double acc = 0.0, cnt = 0.0;
    for (uint32_t t = blockIdx.x * RBS + threadIdx.x; t < n; t += gridDim.x * RBS)
        if (tag[t] == 1) { acc += (double)q[t]; cnt += 1.0; }
    const double s = block_sum<RBS>(acc, lds);
    const double c = block_sum<RBS>(cnt, lds);
    if (threadIdx.x == 0) { part[blockIdx.x] = s; part[gridDim.x + blockIdx.x] = c; }
}

// mgpcg.py:198-205 _sub: q -= mean on FLUID, 0 elsewhere; mean = sum/count in FP32.
__global__ void k_sub_mean(float *__restrict__ q, const uint8_t *__restrict__ tag, const double *__restrict__ sc,
                           uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float mean = (float)sc[SC_MEAN] / (float)sc[SC_COUNT];
    q[t] = tag[t] == 1 ? q[t] - mean : 0.0f;
}

// r -= alpha*Ap and, with DOX, x += alpha*p (the two are independent of the
// preconditioner between them in mgpcg.py:421-457), partial r.r.
template <bool DOX>
__global__ void __launch_bounds__(RBS) k_update_rx_dot(float *__restrict__ r, float *__restrict__ x,
                                                       const float *__restrict__ Ap, const float *__restrict__ p,
                                                       const double *__restrict__ sc, uint32_t n,
                                                       double *__restrict__ part)
{
    __shared__ double lds[RBS / 64];
    const float alpha = (float)sc[SC_ALPHA];
    double acc = 0.0;
    for (uint32_t t = blockIdx.x * RBS + threadIdx.x; t < n; t += gridDim.x * RBS) {
        const float rn = fmaf(-alpha, Ap[t], r[t]);
        r[t] = rn;
        if (DOX) x[t] = fmaf(alpha, p[t], x[t]);
        acc += (double)rn * (double)rn;
    }
    const double s = block_sum<RBS>(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// mgpcg.py:224-227 _update_r, :207-210 _update_x, :229-232 _update_p.
__global__ void k_update_r(float *__restrict__ r, const float *__restrict__ Ap, const double *__restrict__ sc,
                           uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) r[t] = fmaf(-(float)sc[SC_ALPHA], Ap[t], r[t]);
}
__global__ void k_update_x(float *__restrict__ x, const float *__restrict__ p, const double *__restrict__ sc,
                           uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] = fmaf((float)sc[SC_ALPHA], p[t], x[t]);
}
__global__ void k_update_p(float *__restrict__ p, const float *__restrict__ z, const double *__restrict__ sc,
                           uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = fmaf((float)sc[SC_BETA], p[t], z[t]);
}

// Sum the partials (fixed order) and apply the scalar update `op`.
// With host set, the sum also goes to host[0] (pinned host memory mapped for
// the device): the host reads it after an event instead of a copy.
__global__ void __launch_bounds__(RBS) k_finish(const double *__restrict__ part, int nblk, double *__restrict__ sc,
                                                int op, double eps, double *__restrict__ host = nullptr)
{
    __shared__ double lds[RBS / 64];
    double acc = 0.0, acc2 = 0.0;
    for (int g = 0; g < nblk; g += 8 * RBS) {
        double v[8], w[8] = {};
        load_partials<RBS>(part, g, nblk, v);
        if (op == FIN_MEAN) load_partials<RBS>(part + nblk, g, nblk, w);
#pragma unroll
        for (int q = 0; q < 8; ++q) { acc += v[q]; acc2 += w[q]; }
    }
    const double s = block_sum<RBS>(acc, lds);
    const double s2 = op == FIN_MEAN ? block_sum<RBS>(acc2, lds) : 0.0;
    if (threadIdx.x != 0) return;
    if (host) host[0] = s;
    switch (op) {
    case FIN_SUM: sc[SC_SUM] = s; break;
    case FIN_ALPHA: sc[SC_SUM] = s; sc[SC_ALPHA] = sc[SC_OLDZTR] / (s + eps); break;
    case FIN_BETA: sc[SC_SUM] = s; sc[SC_BETA] = s / (sc[SC_OLDZTR] + eps); sc[SC_OLDZTR] = s; break;
    case FIN_OLDZTR: sc[SC_OLDZTR] = s; break;
    case FIN_RTR: sc[SC_RTR] = s; break;
    case FIN_MEAN: sc[SC_MEAN] = s; sc[SC_COUNT] = s2; break;
    case FIN_OUT: sc[0] = s; break;
    }
}

// mgpcg.py:212-222 _update_alpha / _update_beta on the handle scalars.
__global__ void k_alpha(double *sc, double eps) { sc[SC_ALPHA] = sc[SC_OLDZTR] / (sc[SC_SUM] + eps); }
__global__ void k_beta(double *sc, double eps)
{
    const double nz = sc[SC_SUM];
    sc[SC_BETA] = nz / (sc[SC_OLDZTR] + eps);
    sc[SC_OLDZTR] = nz;
}

// ================================================================ host side
static inline unsigned nb256(uint32_t n) { return blocks_for(n, 256); }

static int red_blocks(uint32_t n)
{
    int b = (int)blocks_for(n, RBS);
    return b < MAXBLK ? (b < 1 ? 1 : b) : MAXBLK;
}

static int launch_finish(flub_mgpcg *h, int nblk, int op, double eps, double *dst = nullptr)
{
    k_finish<<<1, RBS, 0, h->s>>>(h->part, nblk, dst ? dst : h->sc, op, eps);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

// sc[SC_RTR] = sum of the r.r partials, on its way to hsc[SC_RTR] for the
// host (h->ev marks its arrival).
static int finish_rtr_to_host(flub_mgpcg *h, int nblk)
{
    k_finish<<<1, RBS, 0, h->s>>>(h->part, nblk, h->sc, FIN_RTR, 0.0, h->hsc_dev ? h->hsc_dev + SC_RTR : nullptr);
    FLUB_CHECK_LAUNCH();
    return mg_rtr_sent(h);
}

double mg_host_partial_sum(const double *part, int nblk)
{
    double acc[RBS];
    for (int t = 0; t < RBS; ++t) {
        double a = 0.0;
        for (int b = t; b < nblk; b += RBS) a += part[b];
        acc[t] = a;
    }
    double s = 0.0;
    for (int w = 0; w < RBS / 64; ++w) {
        double v[64], nv[64];
        for (int l = 0; l < 64; ++l) v[l] = acc[64 * w + l];
        for (int off = 32; off > 0; off >>= 1) {   // __shfl_down: lanes past the end keep their value
            for (int l = 0; l < 64; ++l) nv[l] = v[l] + (l + off < 64 ? v[l + off] : v[l]);
            for (int l = 0; l < 64; ++l) v[l] = nv[l];
        }
        s += v[0];
    }
    return s;
}

int mg_rtr_to_device(flub_mgpcg *h, double rTr)
{
    h->hsc[SC_RTR] = rTr;
    FLUB_CHECK_HIP(hipMemcpyAsync(h->sc + SC_RTR, h->hsc + SC_RTR, sizeof(double), hipMemcpyHostToDevice, h->s));
    return FLUB_OK;
}

int mg_rtr_sent(flub_mgpcg *h)
{
    if (!h->hsc_dev)
        FLUB_CHECK_HIP(hipMemcpyAsync(h->hsc + SC_RTR, h->sc + SC_RTR, sizeof(double), hipMemcpyDeviceToHost, h->s));
    FLUB_CHECK_HIP(hipEventRecord(h->ev, h->s));
    return FLUB_OK;
}

int mg_finish_alpha(flub_mgpcg *h, int nblk, double eps) { return launch_finish(h, nblk, FIN_ALPHA, eps); }

int mg_dot(flub_mgpcg *h, const float *a, const float *b, int op, double eps)
{
    const uint32_t n = h->lv[0].n;
    const int nb = red_blocks(n);
    k_dot<<<nb, RBS, 0, h->s>>>(a, b, n, h->part);
    FLUB_CHECK_LAUNCH();
    return launch_finish(h, nb, op, eps);
}

int mg_zero_centered(flub_mgpcg *h)
{
    Level &L0 = h->lv[0];
    const int nb = red_blocks(L0.n);
    k_masked_sum<<<nb, RBS, 0, h->s>>>(L0.r, L0.tag, L0.n, h->part);
    FLUB_CHECK_LAUNCH();
    int rc = launch_finish(h, nb, FIN_MEAN, 0.0);
    if (rc) return rc;
    k_sub_mean<<<nb256(L0.n), 256, 0, h->s>>>(L0.r, L0.tag, h->sc, L0.n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

static int half_sweep(flub_mgpcg *h, int l, int tagged, int phase)
{
    Level &L = h->lv[l];
    const uint32_t nh = L.n / 2;
    if (tagged) k_smooth_mc<<<nb256(nh), 256, 0, h->s>>>(L.z, L.r, L.tag, L.nx, L.ny, L.nz, phase);
    else k_smooth_def<<<nb256(nh), 256, 0, h->s>>>(L.z, L.r, L.nx, L.ny, L.nz, phase);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

#define TRY(x)                    \
    do {                          \
        int rc_ = (x);            \
        if (rc_) return rc_;      \
    } while (0)

// Bottom level (mgpcg.py:319-336) from z = 0 (zeroed with the restriction at
// mgpcg.py:312): one workgroup when it is small, else one launch per sweep.
static void bottom_solve(flub_mgpcg *h, int tagged)
{
    const int lb = h->L - 1;
    Level &B = h->lv[lb];
    const int nb = h->bottom / 2;
    if (h->L > 1) k_fill_f32<<<nb256(B.n), 256, 0, h->s>>>(B.z, 0.0f, B.n);
    if (tagged && h->fast && B.nz % 4 == 0 && B.n / 2 > BOTTOM_ONE_WG_MAX) {
        // quad sweeps over many workgroups (a single workgroup is latency-bound here)
        for (int it = 0; it < nb; ++it) {
            fast_sweep(h->s, B, B.z, B.z, nullptr, 0, SWEEP_PLAIN, nullptr, nullptr);
            fast_sweep(h->s, B, B.z, B.z, nullptr, 1, SWEEP_PLAIN, nullptr, nullptr);
        }
        for (int it = 0; it < nb; ++it) {
            fast_sweep(h->s, B, B.z, B.z, nullptr, 1, SWEEP_PLAIN, nullptr, nullptr);
            fast_sweep(h->s, B, B.z, B.z, nullptr, 0, SWEEP_PLAIN, nullptr, nullptr);
        }
        return;
    }
    if (B.n / 2 <= BOTTOM_ONE_WG_MAX) {
        if (nb > 0) {
            if (tagged) k_bottom<true><<<1, 1024, 0, h->s>>>(B.z, B.r, B.tag, B.nx, B.ny, B.nz, nb);
            else k_bottom<false><<<1, 1024, 0, h->s>>>(B.z, B.r, B.tag, B.nx, B.ny, B.nz, nb);
        }
    } else {
        for (int it = 0; it < nb; ++it) { half_sweep(h, lb, tagged, 0); half_sweep(h, lb, tagged, 1); }
        for (int it = 0; it < nb; ++it) { half_sweep(h, lb, tagged, 1); half_sweep(h, lb, tagged, 0); }
    }
}

// mgpcg.py:296-349, launched kernel by kernel (captured into a graph by the caller).
static int vcycle_fp32(flub_mgpcg *h, int smoother)
{
    const int tagged = smoother != FLUB_SMOOTHER_DEFAULT;
    Level &L0 = h->lv[0];
    k_fill_f32<<<nb256(L0.n), 256, 0, h->s>>>(L0.z, 0.0f, L0.n);
    FLUB_CHECK_LAUNCH();
    for (int l = 0; l + 1 < h->L; ++l) {
        for (int it = 0; it < mg_sweeps(h, l); ++it) {
            TRY(half_sweep(h, l, tagged, 0));
            TRY(half_sweep(h, l, tagged, 1));
        }
        Level &C = h->lv[l + 1];
        k_fill_f32<<<nb256(C.n), 256, 0, h->s>>>(C.z, 0.0f, C.n);
        FLUB_CHECK_LAUNCH();
        Level &F = h->lv[l];
        if (tagged) k_restrict<true><<<nb256(C.n), 256, 0, h->s>>>(F.r, F.z, F.tag, C.r, C.tag, F.nx, F.ny, F.nz, C.cs);
        else k_restrict<false><<<nb256(C.n), 256, 0, h->s>>>(F.r, F.z, F.tag, C.r, C.tag, F.nx, F.ny, F.nz, C.cs);
        FLUB_CHECK_LAUNCH();
    }
    bottom_solve(h, tagged);
    for (int l = h->L - 2; l >= 0; --l) {
        Level &F = h->lv[l];
        k_prolong<<<nb256(F.n), 256, 0, h->s>>>(F.z, h->lv[l + 1].z, F.nx, F.ny, F.nz);
        FLUB_CHECK_LAUNCH();
        for (int it = 0; it < mg_sweeps(h, l); ++it) {
            TRY(half_sweep(h, l, tagged, 1));
            TRY(half_sweep(h, l, tagged, 0));
        }
    }
    return FLUB_OK;
}

static void bottom_solve(flub_mgpcg *h, int tagged);

// The same V-cycle on the quad kernels of smooth.hip, with the fusions listed
// there.  Pre-smoothing of level l runs in place on zpre (its first sweep
// starts from z = 0), the restriction reads zpre, the prolongating first
// post-sweep reads zpre + z[l+1] and writes z[l], the rest runs in place on
// z[l]; the last level-0 sweep also yields the partials of z.r.
static int vcycle_fast(flub_mgpcg *h, bool split, int zslot = 0)
{
    // Levels whose shape fits a tile run their smoothing temporally blocked
    // (blocked.hip); the pre-smoothed z then lands in the buffer that the
    // first post-smoothing launch does not write.  With `split`, level 0 is
    // the colour-split copy (split.hip): r in sp.r, pre-smoothing in place on
    // sp.zpre, post-smoothing into sp.z[zslot]; levels 1 .. nsplit-1 are
    // colour-split in their own r, zpre and z (the restriction writes their r
    // split); the last post-sweep of such a level writes its whole z in the
    // reference layout into its zpre, which the finer level's prolongating
    // sweep reads.
    const void *pre_z[MAXL];
    uint8_t ph[MG_MAX_PHASES];
    const SplitBuf &S = h->sp;
    // the coarsest levels, from level T on, in one workgroup (tail.hip)
    const int T = tail_first_level(h, split ? std::max(h->nsplit, h->mx_l1 ? 2 : 1) : 1);
    const int lend = T > 0 ? T : h->L - 1;
    for (int l = 0; l < lend; ++l) {
        Level &F = h->lv[l];
        const int s = mg_sweeps(h, l);
        if (l == 0 && split && h->prec != FLUB_PREC_FP32) {
            // FP16/BF16 pre-smoothing (z and r), the first two sweeps converting r
            SplitMixedSweep m{};
            // the 16-bit r of this V-cycle in its slot's half of sp.rh: the final
            // post-sweep of the next cycle reads it as r_old
            void *RH = static_cast<uint16_t *>(S.rh) + (size_t)zslot * F.n;
            m.code = S.code; m.rh = RH; m.r32 = S.r; m.sc = h->sc; m.zin = S.zh; m.zout = S.zh;
            const bool have_rh = h->rh_ready == zslot;   // the r update stored this cycle's T r
            // ... and ran its first half-sweep (colour 1 of zh from z = 0)
            const bool have_zh0 = have_rh && h->zh0_ready == zslot;
            for (int it = 0; it < s; ++it)
                for (int ph = 0; ph < 2; ++ph) {
                    if (it == 0 && ph == 0 && have_zh0) continue;
                    m.phase = ph;
                    m.kind = it > 0 ? SMX_PRE
                                    : ph == 0 ? (have_rh ? SMX_PRE_ZERO_RH : SMX_PRE_ZERO_SRC)
                                              : (have_rh ? SMX_PRE_OWN0_RH : SMX_PRE_SRC);
                    TRY(split_sweep_mixed(h->s, F, h->prec, m));
                }
            TRY(split_restrict_mixed(h->s, F, h->prec, RH, S.zh, S.code, h->lv[1], h->nsplit > 1,
                                     h->mx_l1 ? S.r1h : nullptr));
            continue;
        }
        if (l == 1 && split && h->mx_l1) {
            // BF16 pre-smoothing of level 1: the pairs alternate between its two
            // BF16 z buffers, the last into zp1h; then the restriction into
            // level 2's FP32 r
            const void *in = nullptr;
            for (int it = 0; it < s; ++it) {
                void *out = ((s - 1 - it) & 1) ? S.zt1h : S.zp1h;
                TRY(split_sweep_pair(h->s, F, in, out, nullptr, S.r1h, S.lcode[1], 0, it == 0 ? SWEEP_ZERO : SWEEP_PLAIN,
                                     false, PAIR_BF16));
                in = out;
            }
            TRY(split_restrict_mixed(h->s, F, h->prec, S.r1h, S.zp1h, S.lcode[1], h->lv[2], 2 < h->nsplit));
            continue;
        }
        if (split && l < h->nsplit) {   // level 0 in sp.*, the next split levels in their own buffers
            float *R = l == 0 ? S.r : F.r, *ZP = l == 0 ? S.zpre : F.zpre;
            const uint16_t *CD = S.lcode[l];
            if (l == 0 && h->pair_l0) {
                // the first two half-sweeps in one launch (the second colour
                // reads the first from LDS instead of HBM), the others one
                // launch each: on level 0 a pair of plain sweeps streams
                // slower than two single launches
                TRY(split_sweep_pair(h->s, F, nullptr, ZP, nullptr, R, CD, 0, SWEEP_ZERO));
                for (int it = 1; it < s; ++it) {
                    TRY(probe_begin(h));
                    TRY(split_sweep(h->s, F, ZP, ZP, nullptr, R, CD, 0, SWEEP_PLAIN, nullptr, nullptr));
                    TRY(probe_end(h));
                    TRY(probe_begin(h));
                    TRY(split_sweep(h->s, F, ZP, ZP, nullptr, R, CD, 1, SWEEP_PLAIN, nullptr, nullptr));
                    TRY(probe_end(h));
                }
                TRY(split_restrict(h->s, F, R, ZP, CD, h->lv[l + 1], l + 1 < h->nsplit));
                continue;
            }
            if (l > 0 && S.zt[l] && s >= 1) {   // the sweeps in pairs, alternating with zt, the last pair into ZP
                const float *in = nullptr;
                for (int it = 0; it < s; ++it) {
                    float *out = ((s - 1 - it) & 1) ? S.zt[l] : ZP;
                    TRY(split_sweep_pair(h->s, F, in, out, nullptr, R, CD, 0, it == 0 ? SWEEP_ZERO : SWEEP_PLAIN));
                    in = out;
                }
                TRY(split_restrict(h->s, F, R, ZP, CD, h->lv[l + 1], l + 1 < h->nsplit));
                continue;
            }
            TRY(split_sweep(h->s, F, nullptr, ZP, nullptr, R, CD, 0, SWEEP_ZERO, nullptr, nullptr));
            TRY(split_sweep(h->s, F, ZP, ZP, nullptr, R, CD, 1, SWEEP_PLAIN_OWN0, nullptr, nullptr));
            for (int it = 1; it < s; ++it) {
                TRY(split_sweep(h->s, F, ZP, ZP, nullptr, R, CD, 0, SWEEP_PLAIN, nullptr, nullptr));
                TRY(split_sweep(h->s, F, ZP, ZP, nullptr, R, CD, 1, SWEEP_PLAIN, nullptr, nullptr));
            }
            TRY(split_restrict(h->s, F, R, ZP, CD, h->lv[l + 1], l + 1 < h->nsplit));
            continue;
        }
        if (mg_use_tb(h, l, 2 * s, true)) {
            const int n = mg_phase_pairs(ph, s, 0);
            float *X = (tb_launches(F, n, h->tb_g8) & 1) ? F.zpre : F.z;
            TRY(tb_chain(h->s, h->tb_order, F, FLUB_PREC_FP32, SWEEP_ZERO, nullptr, nullptr, 1.0f, F.r, ph, n, X,
                         X == F.z ? F.zpre : F.z, nullptr, nullptr, h->tb_g8));
            pre_z[l] = X;
        } else {
            TRY(fast_sweep(h->s, F, F.zpre, F.zpre, nullptr, 0, SWEEP_ZERO, nullptr, nullptr));
            TRY(fast_sweep(h->s, F, F.zpre, F.zpre, nullptr, 1, SWEEP_PLAIN, nullptr, nullptr));
            for (int it = 1; it < s; ++it) {
                TRY(fast_sweep(h->s, F, F.zpre, F.zpre, nullptr, 0, SWEEP_PLAIN, nullptr, nullptr));
                TRY(fast_sweep(h->s, F, F.zpre, F.zpre, nullptr, 1, SWEEP_PLAIN, nullptr, nullptr));
            }
            pre_z[l] = F.zpre;
        }
        TRY(fast_restrict(h->s, F, (const float *)pre_z[l], h->lv[l + 1]));
    }
    const int lb = h->L - 1;
    if (T > 0) {
        TRY(tail_cycle(h->s, h, T));
    } else if (h->L > 1 && mg_use_tb(h, lb, 2 * (h->bottom / 2) * 2)) {
        Level &B = h->lv[lb];
        const int n = mg_bottom_phases(ph, h->bottom / 2);
        TRY(tb_chain(h->s, h->tb_order, B, FLUB_PREC_FP32, SWEEP_ZERO, nullptr, nullptr, 1.0f, B.r, ph, n, B.z, B.zpre, nullptr, nullptr, h->tb_g8));
    } else {
        bottom_solve(h, 1);
    }
    for (int l = lend - 1; l >= 0; --l) {
        Level &F = h->lv[l];
        const int s = mg_sweeps(h, l);
        // the coarse z in the reference layout: a split level leaves it in its zpre
        const float *zc = split && l + 1 < h->nsplit ? h->lv[l + 1].zpre : h->lv[l + 1].z;
        if (l == 0 && split && h->prec != FLUB_PREC_FP32) {
            // FP32 post-smoothing from the FP16/BF16 pre-smoothed z into the
            // slot's z buffer, which keeps z times the storage scale; the last
            // sweep also stores 1/scale and the partials of (z/scale).r, .q
            SplitMixedSweep m{};
            float *Z = S.z[zslot];
            m.code = S.code; m.rh = static_cast<uint16_t *>(S.rh) + (size_t)zslot * F.n; m.r32 = S.r; m.sc = h->sc;
            m.rold = static_cast<const uint16_t *>(S.rh) + (size_t)(zslot ^ 1) * F.n;
            m.zinv_old = h->sc + SC_ZINV0 + (zslot ^ 1);
            for (int it = 0; it < s; ++it)
                for (int ph = 1; ph >= 0; --ph) {
                    const bool last = it + 1 == s && ph == 0;
                    m.phase = ph;
                    m.zin = it == 0 && ph == 1 ? S.zh : (const void *)Z;
                    m.zout = Z;
                    m.zc = h->nsplit > 1 ? h->lv[1].zpre : h->lv[1].z;   // reference layout
                    m.kind = it == 0 && ph == 1 ? SMX_POST_PROLONG : (last ? SMX_FINAL : SMX_POST);
                    if (last) { m.zinv = h->sc + SC_ZINV0 + zslot; m.part = h->part; m.nblk = &h->vc_dot_nblk; }
                    TRY(split_sweep_mixed(h->s, F, h->prec, m));
                }
            continue;
        }
        if (split && l < h->nsplit) {
            float *Z = l == 0 ? S.z[zslot] : F.z, *R = l == 0 ? S.r : F.r, *ZP = l == 0 ? S.zpre : F.zpre;
            const uint16_t *CD = S.lcode[l];
            if (l > 0 && S.zt[l] && s >= 2) {
                // the sweeps in pairs from ZP through zt and z (alternating),
                // the last pair writing the whole z, reference layout, into ZP;
                // a BF16 level 1 starts from its BF16 pre-smoothed z and reads
                // its BF16 r
                const bool h1 = l == 1 && h->mx_l1;
                const void *in = h1 ? S.zp1h : (const void *)ZP;
                const void *RR = h1 ? S.r1h : (const void *)R;
                for (int it = 0; it < s; ++it) {
                    const bool last = it + 1 == s;
                    float *out = last ? ZP : ((it & 1) ? Z : S.zt[l]);
                    const int ty = !h1 ? PAIR_F32 : (it == 0 ? PAIR_BF16_IN : PAIR_BF16_R);
                    TRY(split_sweep_pair(h->s, F, in, out, it == 0 ? zc : nullptr, RR, CD, 1,
                                         it == 0 ? SWEEP_PROLONG : SWEEP_PLAIN, last, ty));
                    in = out;
                }
                continue;
            }
            if (l == 0 && h->pair_l0 && s >= 2) {
                // the prolongating pair from ZP into Z, then one launch per
                // half-sweep in place, the final one with the partials of z.r
                TRY(split_sweep_pair(h->s, F, ZP, Z, zc, R, CD, 1, SWEEP_PROLONG));
                for (int it = 1; it < s; ++it) {
                    const bool last = it + 1 == s;
                    TRY(probe_begin(h));
                    TRY(split_sweep(h->s, F, Z, Z, nullptr, R, CD, 1, SWEEP_PLAIN, nullptr, nullptr));
                    TRY(probe_end(h));
                    TRY(split_sweep(h->s, F, Z, Z, nullptr, R, CD, 0, SWEEP_PLAIN, last ? h->part : nullptr,
                                    last ? &h->vc_dot_nblk : nullptr));
                }
                continue;
            }
            for (int it = 0; it < s; ++it) {
                const bool last = it + 1 == s;
                if (it == 0) TRY(split_sweep(h->s, F, ZP, Z, zc, R, CD, 1, SWEEP_PROLONG, nullptr, nullptr));
                else TRY(split_sweep(h->s, F, Z, Z, nullptr, R, CD, 1, SWEEP_PLAIN, nullptr, nullptr));
                if (last && l > 0)   // the whole z, reference layout, into zpre (read by the finer level)
                    TRY(split_sweep(h->s, F, Z, ZP, nullptr, R, CD, 0, SWEEP_PLAIN, nullptr, nullptr, true));
                else   // level 0's last sweep also yields z.r
                    TRY(split_sweep(h->s, F, Z, Z, nullptr, R, CD, 0, SWEEP_PLAIN, last ? h->part : nullptr,
                                    last ? &h->vc_dot_nblk : nullptr));
            }
            continue;
        }
        if (mg_use_tb(h, l, 2 * s, true)) {
            const int n = mg_phase_pairs(ph, s, 1);
            TRY(tb_chain(h->s, h->tb_order, F, FLUB_PREC_FP32, SWEEP_PROLONG, pre_z[l], zc, 1.0f, F.r, ph, n, F.z, F.zpre,
                         l == 0 ? h->part : nullptr, l == 0 ? &h->vc_dot_nblk : nullptr, h->tb_g8));
            continue;
        }
        for (int it = 0; it < s; ++it) {
            const bool last = l == 0 && it + 1 == s;
            if (it == 0) TRY(fast_sweep(h->s, F, F.zpre, F.z, zc, 1, SWEEP_PROLONG, nullptr, nullptr));
            else TRY(fast_sweep(h->s, F, F.z, F.z, nullptr, 1, SWEEP_PLAIN, nullptr, nullptr));
            TRY(fast_sweep(h->s, F, F.z, F.z, nullptr, 0, SWEEP_PLAIN, last ? h->part : nullptr,
                           last ? &h->vc_dot_nblk : nullptr));
        }
    }
    return FLUB_OK;
}

static bool use_fast(const flub_mgpcg *h, int smoother)
{
    return h->fast && smoother != FLUB_SMOOTHER_DEFAULT && h->pre_post >= 1;
}

int mg_precondition(flub_mgpcg *h, int smoother, bool use_graph)
{
    if (h->prec != FLUB_PREC_FP32) return mg_vcycle_mixed(h, smoother);
    const bool fast = use_fast(h, smoother);
    if (!use_graph || h->no_graph) return fast ? vcycle_fast(h, false) : vcycle_fp32(h, smoother);
    if (!h->graph_ok[smoother]) {
        // Capture on the handle's private stream (the work stream may be the
        // null stream, which cannot be captured); replay on the work stream.
        hipGraph_t g;
        hipStream_t work = h->s;
        FLUB_CHECK_HIP(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
        h->s = h->cap;
        int rc = fast ? vcycle_fast(h, false) : vcycle_fp32(h, smoother);
        h->s = work;
        hipError_t e = hipStreamEndCapture(h->cap, &g);
        if (rc) { if (e == hipSuccess) hipGraphDestroy(g); return rc; }
        FLUB_CHECK_HIP(e);
        FLUB_CHECK_HIP(hipGraphInstantiate(&h->graph[smoother], g, nullptr, nullptr, 0));
        FLUB_CHECK_HIP(hipGraphDestroy(g));
        h->graph_ok[smoother] = true;
    }
    FLUB_CHECK_HIP(hipGraphLaunch(h->graph[smoother], h->s));
    return FLUB_OK;
}

// The V-cycle on the colour-split level 0 (sp.r -> sp.z[zslot]), as a graph
// in slot 3 + zslot.
// Which of its first sweeps the mixed V-cycle into slot zslot runs (they
// depend on what the r update stored): a captured cycle is replayed only in
// the state it was captured in.
static int mx_state(const flub_mgpcg *h, int zslot)
{
    return h->prec == FLUB_PREC_FP32 ? 0 : (h->rh_ready == zslot ? 1 : 0) + (h->zh0_ready == zslot ? 2 : 0);
}

int mg_precondition_split(flub_mgpcg *h, int zslot)
{
    if (h->no_graph || !h->split_graph) return vcycle_fast(h, true, zslot);
    const int gs = 3 + zslot;
    if (h->graph_ok[gs] && h->graph_rh[gs] != mx_state(h, zslot)) return vcycle_fast(h, true, zslot);
    if (!h->graph_ok[gs]) {
        hipGraph_t g;
        hipStream_t work = h->s;
        FLUB_CHECK_HIP(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
        h->s = h->cap;
        int rc = vcycle_fast(h, true, zslot);
        h->s = work;
        hipError_t e = hipStreamEndCapture(h->cap, &g);
        if (rc) { if (e == hipSuccess) hipGraphDestroy(g); return rc; }
        FLUB_CHECK_HIP(e);
        FLUB_CHECK_HIP(hipGraphInstantiate(&h->graph[gs], g, nullptr, nullptr, 0));
        FLUB_CHECK_HIP(hipGraphDestroy(g));
        h->graph_ok[gs] = true;
        h->graph_rh[gs] = mx_state(h, zslot);
    }
    FLUB_CHECK_HIP(hipGraphLaunch(h->graph[gs], h->s));
    return FLUB_OK;
}

static void drop_graphs(flub_mgpcg *h)
{
    for (int s = 0; s < 5; ++s)
        if (h->graph_ok[s]) { hipGraphExecDestroy(h->graph[s]); h->graph_ok[s] = false; }
}

FLUB_API int flub_mgpcg_create(flub_mgpcg **out, int nx, int ny, int nz, int levels, int precision,
                               void *stream)
{
    FLUB_REQUIRE(out, FLUB_EINVAL, "flub_mgpcg_create: null output");
    *out = nullptr;
    FLUB_REQUIRE(levels >= 1 && levels <= MAXL, FLUB_EINVAL, "n_mg_levels must be in [1, %d]", MAXL);
    FLUB_REQUIRE(precision >= 0 && precision <= 2, FLUB_EINVAL, "precision must be 0, 1 or 2");
    FLUB_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && (double)(nx + 1) * (ny + 1) * (nz + 1) < 2147483647.0,
                 FLUB_ESHAPE, "grid %dx%dx%d out of range", nx, ny, nz);
    for (int l = 0; l < levels; ++l) {
        const int a = nx >> l, b = ny >> l, c = nz >> l;
        FLUB_REQUIRE(a >= 1 && b >= 1 && c >= 2 && (c % 2) == 0, FLUB_ESHAPE,
                     "level %d shape %dx%dx%d: the red-black smoother needs an even z extent", l, a, b, c);
        if (l + 1 < levels)
            FLUB_REQUIRE(a % 2 == 0 && b % 2 == 0, FLUB_ESHAPE,
                         "level %d shape %dx%dx%d must be even to restrict", l, a, b, c);
    }
    flub_mgpcg *h = (flub_mgpcg *)calloc(1, sizeof(flub_mgpcg));
    FLUB_REQUIRE(h, FLUB_EINVAL, "out of host memory");
    h->L = levels;
    h->prec = precision;
    h->rh_ready = h->zh0_ready = -1;
    h->pre_post = 6;
    h->bottom = 6;
    h->no_graph = getenv("FLUB_NO_GRAPH") != nullptr;
    h->split_graph = getenv("FLUB_SPLIT_GRAPH") != nullptr;
    h->no_tb = getenv("FLUB_NO_TB") != nullptr;
    h->no_tb_l0 = getenv("FLUB_NO_TB_L0") != nullptr;
    h->tb_g8 = getenv("FLUB_TB_G8") != nullptr;
    h->no_tail = getenv("FLUB_NO_TAIL") != nullptr;
    {
        const char *e = getenv("FLUB_PAP_IB");
        h->pap_ib = getenv("FLUB_NO_PAP") ? -1 : (e ? atoi(e) : 0);
    }
    int rc = FLUB_OK;
#define ALLOC(ptr, bytes)                                              \
    do {                                                               \
        hipError_t e_ = hipMalloc((void **)&(ptr), (bytes));           \
        if (e_ != hipSuccess) {                                        \
            flub_set_error("hipMalloc(%zu): %s", (size_t)(bytes), hipGetErrorString(e_)); \
            rc = (int)e_;                                              \
            goto fail;                                                 \
        }                                                              \
    } while (0)
    {
        // All work is ordered on the caller's stream (NULL = the null stream);
        // a private stream only records the V-cycle graphs.
        h->s = as_stream(stream);
        {
            hipError_t e = hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking);
            if (e != hipSuccess) { flub_set_error("hipStreamCreate: %s", hipGetErrorString(e)); rc = (int)e; goto fail; }
        }
        for (int l = 0; l < levels; ++l) {
            Level &L = h->lv[l];
            L.nx = nx >> l; L.ny = ny >> l; L.nz = nz >> l;
            L.n = (uint32_t)L.nx * L.ny * L.nz;
            L.cs = 1.0f;
            ALLOC(L.tag, L.n);
            ALLOC(L.code, (size_t)L.n * 2);
            ALLOC(L.r, (size_t)L.n * 4);
            ALLOC(L.z, (size_t)L.n * 4);
            ALLOC(L.zpre, (size_t)L.n * 4);
            if (precision != FLUB_PREC_FP32) {
                ALLOC(L.rh, (size_t)L.n * 2);
                ALLOC(L.zh, (size_t)L.n * 2);
            }
        }
        const size_t n0 = h->lv[0].n;
        ALLOC(h->x, n0 * 4);
        ALLOC(h->p, n0 * 4);
        ALLOC(h->Ap, n0 * 4);
        ALLOC(h->p_alt, n0 * 4);
        h->fast = levels >= 2 && (nz % 4) == 0;
        ALLOC(h->sc, SC_NUM * sizeof(double));
        ALLOC(h->part, 2 * MAXPART * sizeof(double));
        // The host mirror is mapped for the device and coherent, so that the
        // residual norm's finish kernel can store into it directly; without
        // the mapping the solves copy the value instead.
        hipError_t e = hipHostMalloc((void **)&h->hsc, SC_NUM * sizeof(double),
                                     hipHostMallocMapped | hipHostMallocCoherent);
        if (e == hipSuccess) {
            if (hipHostGetDevicePointer((void **)&h->hsc_dev, h->hsc, 0) != hipSuccess) h->hsc_dev = nullptr;
        } else {
            (void)hipGetLastError();
            h->hsc_dev = nullptr;
            e = hipHostMalloc((void **)&h->hsc, SC_NUM * sizeof(double), hipHostMallocDefault);
        }
        if (e != hipSuccess) { flub_set_error("hipHostMalloc: %s", hipGetErrorString(e)); rc = (int)e; goto fail; }
        if (getenv("FLUB_RTR_COPY")) h->hsc_dev = nullptr;
        if (h->hsc_dev) {   // the partials' host buffer, mapped the same way (else: finish kernels)
            if (hipHostMalloc((void **)&h->hpart, 2 * MAXBLK * sizeof(double),
                              hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess) {
                if (hipHostGetDevicePointer((void **)&h->hpart_dev, h->hpart, 0) != hipSuccess) h->hpart_dev = nullptr;
            } else {
                (void)hipGetLastError();
                h->hpart = nullptr;
            }
        }
        e = hipMemsetAsync(h->sc, 0, SC_NUM * sizeof(double), h->s);
        if (e == hipSuccess) e = hipMemsetAsync(h->x, 0, n0 * 4, h->s);
        if (e == hipSuccess) e = hipStreamSynchronize(h->s);
        if (e != hipSuccess) { flub_set_error("init: %s", hipGetErrorString(e)); rc = (int)e; goto fail; }
        rc = mg_mixed_setup(h);
        if (rc) goto fail;
        if (!h->no_tb) {
            rc = tb_tables_create(&h->tb_order);
            if (rc) goto fail;
        }
        e = hipEventCreateWithFlags(&h->ev, hipEventDisableTiming);
        if (e != hipSuccess) { flub_set_error("hipEventCreate: %s", hipGetErrorString(e)); rc = (int)e; goto fail; }
        h->split = h->fast && split_supported(h->lv[0]) && getenv("FLUB_NO_SPLIT") == nullptr;
        if (h->split) {
            ALLOC(h->sp.r, n0 * 4);
            ALLOC(h->sp.z[0], n0 * 4);
            ALLOC(h->sp.z[1], n0 * 4);
            ALLOC(h->sp.zpre, n0 * 4);
            ALLOC(h->sp.x, n0 * 8);
            ALLOC(h->x64n, n0 * 8);
            {   // the p ring: p, p_alt and up to 30 more, within 8 GiB
                const size_t budget = (size_t)8 << 30;
                size_t np = budget / (n0 * 4);
                np = np < 3 ? 3 : (np > MAXPH ? MAXPH : np);
                const char *ev = getenv("FLUB_PH");   // ring size override (3..32)
                if (ev) np = (size_t)(atoi(ev) < 3 ? 3 : (atoi(ev) > MAXPH ? MAXPH : atoi(ev)));
                h->sp.nph = (int)np;
                for (int i = 2; i < h->sp.nph; ++i) ALLOC(h->sp.ph[i], n0 * 4);
                ALLOC(h->sp.ahist, MAXPH * sizeof(double));
            }
            ALLOC(h->sp.b, n0 * 4);
            ALLOC(h->sp.code, n0 * 2);
            h->sp.lcode[0] = h->sp.code;
            // (the 16-bit sweeps of the mixed V-cycle stream better one launch
            // each: as pairs they measured slower)
            h->pair_l0 = precision == FLUB_PREC_FP32 && split_pair_supported(h->lv[0]) &&
                         getenv("FLUB_NO_PAIR") == nullptr && getenv("FLUB_NO_PAIR_L0") == nullptr;
            h->mx_unit = precision == FLUB_PREC_BF16 && getenv("FLUB_MX_SRC32") == nullptr;
            h->mx_fuse0 = h->mx_unit && getenv("FLUB_MX_NO_FUSE0") == nullptr;
            // The next levels are colour-split too while they hold at least
            // min_cells (default 2^20): there one launch per half-sweep on the
            // split layout streams faster than the blocked natural-layout
            // smoother (128^3: 4.5 us per half-sweep against 8.4).  The
            // bottom level stays in the reference layout.
            const char *mc = getenv("FLUB_SPLIT_MIN_CELLS");
            const uint32_t min_cells = mc ? (uint32_t)strtoul(mc, nullptr, 10) : (1u << 20);
            h->nsplit = 1;
            while (h->nsplit + 1 < levels && split_supported(h->lv[h->nsplit]) && h->lv[h->nsplit].n >= min_cells) {
                ALLOC(h->sp.lcode[h->nsplit], (size_t)h->lv[h->nsplit].n * 2);
                if (split_pair_supported(h->lv[h->nsplit]) && getenv("FLUB_NO_PAIR") == nullptr)
                    ALLOC(h->sp.zt[h->nsplit], (size_t)h->lv[h->nsplit].n * 4);
                ++h->nsplit;
            }
            if (precision != FLUB_PREC_FP32) {
                ALLOC(h->sp.rh, n0 * 2 * 2);   // one half per z slot
                ALLOC(h->sp.zh, n0 * 2);
                ALLOC(h->sp.zw, n0 * 4);
                // r_old of a solve's first final post-sweep is read but not used
                // (its beta is 0): zeros keep those partials finite
                hipError_t e_ = hipMemsetAsync(h->sp.rh, 0, n0 * 2 * 2, h->s);
                if (e_ == hipSuccess) e_ = hipStreamSynchronize(h->s);
                if (e_ != hipSuccess) { flub_set_error("init: %s", hipGetErrorString(e_)); rc = (int)e_; goto fail; }
            }
            // BF16: a split level 1 that sweeps in pairs pre-smooths in BF16 too
            // (its r and pre-smoothed z; the post-smoothing stays FP32)
            h->mx_l1 = h->mx_unit && h->nsplit > 1 && h->sp.zt[1] && getenv("FLUB_MX_NO_L1") == nullptr;
            if (h->mx_l1) {
                const size_t n1 = h->lv[1].n;
                ALLOC(h->sp.r1h, n1 * 2);
                ALLOC(h->sp.zp1h, n1 * 2);
                ALLOC(h->sp.zt1h, n1 * 2);
            }
        }
    }
#undef ALLOC
    *out = h;
    return FLUB_OK;
fail:
    flub_mgpcg_destroy(h);
    return rc;
}

FLUB_API int flub_mgpcg_destroy(flub_mgpcg *h)
{
    if (!h) return FLUB_OK;
    hipStreamSynchronize(h->s);
    drop_graphs(h);
    mg_mixed_teardown(h);
    for (int l = 0; l < h->L; ++l) {
        Level &L = h->lv[l];
        hipFree(L.tag); hipFree(L.code); hipFree(L.r); hipFree(L.z); hipFree(L.zpre); hipFree(L.rh); hipFree(L.zh);
    }
    hipFree(h->x); hipFree(h->p); hipFree(h->p_alt); hipFree(h->Ap); hipFree(h->sc); hipFree(h->part);
    hipFree(h->tb_order);
    hipFree(h->sp.r); hipFree(h->sp.z[0]); hipFree(h->sp.z[1]); hipFree(h->sp.zpre); hipFree(h->sp.x);
    hipFree(h->x64n);
    for (int i = 2; i < MAXPH; ++i) hipFree(h->sp.ph[i]);
    hipFree(h->sp.ahist);
    hipFree(h->sp.code); hipFree(h->sp.rh); hipFree(h->sp.zh); hipFree(h->sp.zw); hipFree(h->sp.b);
    for (int l = 1; l < MAXL; ++l) { hipFree(h->sp.lcode[l]); hipFree(h->sp.zt[l]); }
    hipFree(h->sp.r1h); hipFree(h->sp.zp1h); hipFree(h->sp.zt1h);
    if (h->probe) {
        for (int i = 0; i < 2 * h->probe->max; ++i) hipEventDestroy(h->probe->ev[i]);
        free(h->probe);
    }
    if (h->ev) hipEventDestroy(h->ev);
    if (h->hsc) hipHostFree(h->hsc);
    if (h->hpart) hipHostFree(h->hpart);
    if (h->cap) hipStreamDestroy(h->cap);
    free(h);
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_set_smoothing(flub_mgpcg *h, int pre_post, int bottom)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    FLUB_REQUIRE(pre_post >= 0 && pre_post <= 64 && bottom >= 0 && bottom <= 256, FLUB_EINVAL,
                 "smoothing counts out of range");
    if (pre_post != h->pre_post || bottom != h->bottom) {
        FLUB_CHECK_HIP(hipStreamSynchronize(h->s));
        drop_graphs(h);
    }
    h->pre_post = pre_post;
    h->bottom = bottom;
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_set_coarse_scale(flub_mgpcg *h, float cs)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    FLUB_REQUIRE(cs > 0.0f && cs <= 4.0f, FLUB_EINVAL, "coarse-grid correction scale %g outside (0, 4]", (double)cs);
    if (h->L > 1 && cs != h->lv[1].cs) {
        FLUB_CHECK_HIP(hipStreamSynchronize(h->s));
        drop_graphs(h);   // captured restrictions hold the old scale
    }
    for (int l = 1; l < h->L; ++l) h->lv[l].cs = cs;
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_set_sweep_cap(flub_mgpcg *h, int cap)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    FLUB_REQUIRE(cap >= 0 && cap <= 1024, FLUB_EINVAL, "sweep cap %d outside [0, 1024]", cap);
    if (cap != h->sweep_cap) {
        FLUB_CHECK_HIP(hipStreamSynchronize(h->s));
        drop_graphs(h);
    }
    h->sweep_cap = cap;
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_set_tags(flub_mgpcg *h, const int32_t *tag)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    Level &L0 = h->lv[0];
    if (tag) {
        if (L0.n % 4 == 0 && ((uintptr_t)tag & 15u) == 0)
            k_pack_tags4<<<nb256(L0.n / 4), 256, 0, h->s>>>(tag, L0.tag, L0.n / 4);
        else
            k_pack_tags<<<nb256(L0.n), 256, 0, h->s>>>(tag, L0.tag, L0.n);
        FLUB_CHECK_LAUNCH();
    } else {
        FLUB_REQUIRE(h->tags_set, FLUB_ESTATE, "set_tags(None) before any tags were given");
    }
    for (int l = 0; l + 1 < h->L; ++l) {
        Level &C = h->lv[l + 1];
        k_restrict_tag<<<nb256(C.n), 256, 0, h->s>>>(h->lv[l].tag, C.tag, C.nx, C.ny, C.nz);
        FLUB_CHECK_LAUNCH();
    }
    for (int l = 0; l < h->L; ++l) TRY(make_codes(h->s, h->lv[l]));
    for (int l = 0; l < h->nsplit; ++l) TRY(split_convert(h->s, h->lv[l], h->lv[l].code, h->sp.lcode[l], 2, true));
    h->tags_set = true;
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_init(flub_mgpcg *h, const float *r, float k)
{
    FLUB_REQUIRE(h && r, FLUB_EINVAL, "flub_mgpcg_init: null argument");
    Level &L0 = h->lv[0];
    const uint32_t n = L0.n;
    k_scale_copy<<<nb256(n), 256, 0, h->s>>>(L0.r, r, k, n);
    FLUB_CHECK_LAUNCH();
    k_fill_f32<<<nb256(n), 256, 0, h->s>>>(L0.z, 0.0f, n);
    k_fill_f32<<<nb256(n), 256, 0, h->s>>>(h->Ap, 0.0f, n);
    k_fill_f32<<<nb256(n), 256, 0, h->s>>>(h->p, 0.0f, n);
    k_fill_f32<<<nb256(n), 256, 0, h->s>>>(h->x, 0.0f, n);
    FLUB_CHECK_LAUNCH();
    FLUB_CHECK_HIP(hipMemsetAsync(h->sc, 0, SC_NUM * sizeof(double), h->s));
    h->x64_valid = false;
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_mask_interior(flub_mgpcg *h)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    FLUB_REQUIRE(h->tags_set, FLUB_ESTATE, "mask_interior before set_tags");
    Level &L0 = h->lv[0];
    k_mask_interior<<<nb256(L0.n), 256, 0, h->s>>>(L0.tag, L0.r, L0.n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_zero_centered(flub_mgpcg *h)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    FLUB_REQUIRE(h->tags_set, FLUB_ESTATE, "zero_centered before set_tags");
    return mg_zero_centered(h);
}

static bool smoother_ok(int s) { return s >= 0 && s <= 2; }

FLUB_API int flub_mgpcg_apply_preconditioner(flub_mgpcg *h, int smoother)
{
    FLUB_REQUIRE(h && smoother_ok(smoother), FLUB_EINVAL, "bad handle or smoother");
    FLUB_REQUIRE(h->tags_set, FLUB_ESTATE, "apply_preconditioner before set_tags");
    if (h->prec != FLUB_PREC_FP32) TRY(mg_mixed_rscale(h));
    return mg_precondition(h, smoother, false);
}

FLUB_API int flub_mgpcg_graph_precondition(flub_mgpcg *h, int smoother)
{
    FLUB_REQUIRE(h && smoother_ok(smoother), FLUB_EINVAL, "bad handle or smoother");
    FLUB_REQUIRE(h->tags_set, FLUB_ESTATE, "graph_precondition before set_tags");
    if (h->prec != FLUB_PREC_FP32) TRY(mg_mixed_rscale(h));
    return mg_precondition(h, smoother, true);
}

static int compute_Ap_dot(flub_mgpcg *h, int smoother, int op, double eps)
{
    Level &L0 = h->lv[0];
    const int nb = red_blocks(L0.n);
    if (smoother == FLUB_SMOOTHER_DEFAULT)
        k_Ap_dot<false><<<nb, RBS, 0, h->s>>>(h->Ap, h->p, L0.tag, L0.nx, L0.ny, L0.nz, h->part);
    else
        k_Ap_dot<true><<<nb, RBS, 0, h->s>>>(h->Ap, h->p, L0.tag, L0.nx, L0.ny, L0.nz, h->part);
    FLUB_CHECK_LAUNCH();
    return launch_finish(h, nb, op, eps);
}

FLUB_API int flub_mgpcg_compute_Ap(flub_mgpcg *h, int smoother)
{
    FLUB_REQUIRE(h && smoother_ok(smoother), FLUB_EINVAL, "bad handle or smoother");
    FLUB_REQUIRE(h->tags_set || smoother == FLUB_SMOOTHER_DEFAULT, FLUB_ESTATE, "compute_Ap before set_tags");
    return compute_Ap_dot(h, smoother, FIN_SUM, 0.0);
}

FLUB_API int flub_mgpcg_update_p(flub_mgpcg *h)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    k_update_p<<<nb256(h->lv[0].n), 256, 0, h->s>>>(h->p, h->lv[0].z, h->sc, h->lv[0].n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_update_r(flub_mgpcg *h)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    k_update_r<<<nb256(h->lv[0].n), 256, 0, h->s>>>(h->lv[0].r, h->Ap, h->sc, h->lv[0].n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_update_x(flub_mgpcg *h)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    h->x64_valid = false;
    k_update_x<<<nb256(h->lv[0].n), 256, 0, h->s>>>(h->x, h->p, h->sc, h->lv[0].n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_update_alpha(flub_mgpcg *h, double eps)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    k_alpha<<<1, 1, 0, h->s>>>(h->sc, eps);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_update_beta(flub_mgpcg *h, double eps)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    k_beta<<<1, 1, 0, h->s>>>(h->sc, eps);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

static float *buf_ptr(flub_mgpcg *h, int b)
{
    switch (b) {
    case 0: return h->x;
    case 1: return h->p;
    case 2: return h->Ap;
    case 3: return h->lv[0].r;
    case 4: return h->lv[0].z;
    default: return nullptr;
    }
}

FLUB_API int flub_mgpcg_inner(flub_mgpcg *h, int a, int b, int scalar_out)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    float *pa = buf_ptr(h, a), *pb = buf_ptr(h, b);
    FLUB_REQUIRE(pa && pb && scalar_out >= 0 && scalar_out <= 3, FLUB_EINVAL, "bad buffer/scalar id");
    static const int ops[4] = {-1, -1, FIN_SUM, FIN_OLDZTR};
    const uint32_t n = h->lv[0].n;
    const int nb = red_blocks(n);
    k_dot<<<nb, RBS, 0, h->s>>>(pa, pb, n, h->part);
    FLUB_CHECK_LAUNCH();
    if (ops[scalar_out] >= 0) return launch_finish(h, nb, ops[scalar_out], 0.0);
    return launch_finish(h, nb, FIN_OUT, 0.0, h->sc + (scalar_out == 0 ? SC_ALPHA : SC_BETA));
}

static int read_scalar(flub_mgpcg *h, int idx, double *out)
{
    FLUB_CHECK_HIP(hipMemcpyAsync(h->hsc + idx, h->sc + idx, sizeof(double), hipMemcpyDeviceToHost, h->s));
    FLUB_CHECK_HIP(hipStreamSynchronize(h->s));
    *out = h->hsc[idx];
    return FLUB_OK;
}

// The FP64 true residual of a split solve's FP64 x: r = b - ndt(x) (FP64
// stencil, stored FP32) replaces the solve's r; returns ||r||^2.  The split
// solves stop when it is below the tolerance of the reference's test (the
// recursive residual passed it first).
int mg_split_true_residual(flub_mgpcg *h, float *R, double *rTrue)
{
    Level &L0 = h->lv[0];
    const bool hsum = h->hpart_dev != nullptr;
    int nbt = 0;
    TRY(split_true_residual(h->s, L0, R, h->sp.b, h->sp.x, h->sp.code, hsum ? h->hpart_dev : h->part, &nbt));
    if (hsum) FLUB_CHECK_HIP(hipEventRecord(h->ev, h->s));
    else TRY(finish_rtr_to_host(h, nbt));
    FLUB_CHECK_HIP(hipEventSynchronize(h->ev));
    *rTrue = hsum ? mg_host_partial_sum(h->hpart, nbt) : h->hsc[SC_RTR];
    return FLUB_OK;
}

int mg_split_enter(flub_mgpcg *h, bool &x_set)
{
    Level &L0 = h->lv[0];
    SplitBuf &S = h->sp;
    if (h->presplit) {   // the fused projection's divergence kernel wrote S.r and S.b; x = 0
        x_set = false;
        return FLUB_OK;
    }
    // r and x in the split layout; b = r - ndt(x), the right-hand side the
    // caller's r and x stand for (r itself when x = 0, as after init)
    TRY(split_convert(h->s, L0, L0.r, S.r, 4, true));
    TRY(split_x_from_natural(h->s, L0, h->x, S.x));
    int nbt = 0;
    TRY(split_true_residual(h->s, L0, S.b, S.r, S.x, S.code, h->part, &nbt));
    x_set = true;
    return FLUB_OK;
}

int mg_split_leave(flub_mgpcg *h, bool x_set)
{
    Level &L0 = h->lv[0];
    if (!x_set) FLUB_CHECK_HIP(hipMemsetAsync(h->sp.x, 0, (size_t)L0.n * sizeof(double), h->s));
    TRY(split_x_to_natural(h->s, L0, h->sp.x, h->x, h->want_x64n ? h->x64n : nullptr, h->p_out, h->p_h));
    h->p_written = h->p_out != nullptr;
    h->x64_valid = true;
    return FLUB_OK;
}

namespace {
__global__ void k_widen(const float *__restrict__ x, double *__restrict__ y, uint32_t n)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t < n) y[t] = (double)x[t];
}
}  // namespace

FLUB_API int flub_mgpcg_solution_f64(flub_mgpcg *h, double *x)
{
    FLUB_REQUIRE(h && x, FLUB_EINVAL, "flub_mgpcg_solution_f64: null argument");
    Level &L0 = h->lv[0];
    if (h->x64_valid) return split_x_to_natural(h->s, L0, h->sp.x, nullptr, x);
    k_widen<<<nb256(L0.n), 256, 0, h->s>>>(h->x, x, L0.n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

// mgpcg.py:378-462 for the "phases" smoother on the colour-split level 0:
// r, x, z, p and Ap of level 0 are colour-split for the whole solve and
// converted back at the end.  update_p then Ap; update_x (x += alpha p with
// the same alpha and p) is deferred until x is needed, every p kept in a ring
// of buffers meanwhile (PxState), so that update_r reads and writes only r and
// x (FP64, SplitBuf::x) moves once per flush.  The V-cycle is queued
// before the host reads the residual norm, into the z buffer the iteration
// does not use: the GPU keeps working while the host tests for convergence.
// If the test stops the solve, that speculative z is not used (x, r, z and
// the scalars are those of the reference sequence; the coarse levels hold the
// speculative cycle's values).  When the reference's test passes on the
// recursive residual, the FP64 true residual b - ndt(x) replaces r; the solve
// stops if it passes the same test, else it goes on (at most
// max_replacements times).  *rTr_out: that FP64 true residual.
static int solve_split(flub_mgpcg *h, int max_iters, double eps, double abs_tol, double rel_tol, int verbose,
                       int *iters, double *rTr_out)
{
    Level &L0 = h->lv[0];
    SplitBuf &S = h->sp;
    const uint32_t n = L0.n;
    const int nb = red_blocks(n);
    const int max_replacements = 3;
    int replacements = 0;
    int zs = 0;
    PxState px;
    px_init(h, px, S.x, n);
    TRY(mg_split_enter(h, px.x_set));
    float *const R = S.r;
    // the r.r partials go to the mapped host buffer and the host sums them
    // (no finish kernel on the stream)
    const bool hsum = h->hpart_dev != nullptr;
    k_dot<<<nb, RBS, 0, h->s>>>(R, R, n, hsum ? h->hpart_dev : h->part);
    FLUB_CHECK_LAUNCH();
    // the tolerance is first needed after iteration 1: the first V-cycle runs
    // while the initial r.r travels to the host
    if (hsum) FLUB_CHECK_HIP(hipEventRecord(h->ev, h->s));
    else TRY(finish_rtr_to_host(h, nb));
    TRY(mg_precondition_split(h, zs));
    TRY(launch_finish(h, h->vc_dot_nblk, FIN_OLDZTR, 0.0));
    FLUB_CHECK_HIP(hipEventSynchronize(h->ev));
    const double rTr0 = hsum ? mg_host_partial_sum(h->hpart, nb) : h->hsc[SC_RTR];
    const double tol = fmax(abs_tol, rTr0 * rel_tol);
    const int cap = max_iters < 0 ? SOLVE_SAFETY_CAP : max_iters;
    int it = 0;
    double rTr = rTr0, r_prev = rTr0, rate = 1.0;   // rate: last reduction factor of r.r
    int status = FLUB_OK;
    while (it < cap) {
        int nbp = 0;
        // speculate unless this test is predicted to stop the solve (a wasted
        // speculative cycle would delay the end of the solve)
        const bool speculate = it + 1 < cap && r_prev * rate > 16.0 * tol;
        // p = z + beta*p (with the pending x updates), then Ap = ndt(p)
        TRY(split_pstep_Ap(h->s, px, S.z[zs], nullptr, L0, h->Ap, S.code, h->part, &nbp, h->pap_ib));
        TRY(launch_finish(h, nbp, FIN_ALPHA, eps));
        // r only: x += alpha p waits for a later p step (or the end)
        k_update_rx_dot<false><<<nb, RBS, 0, h->s>>>(R, nullptr, h->Ap, nullptr, h->sc, n,
                                                      hsum ? h->hpart_dev : h->part);
        FLUB_CHECK_LAUNCH();
        px_after_r_update(px);
        if (hsum) FLUB_CHECK_HIP(hipEventRecord(h->ev, h->s));
        else TRY(finish_rtr_to_host(h, nb));
        if (speculate) TRY(mg_precondition_split(h, zs ^ 1));
        FLUB_CHECK_HIP(hipEventSynchronize(h->ev));
        rTr = hsum ? mg_host_partial_sum(h->hpart, nb) : h->hsc[SC_RTR];
        if (verbose) printf("after iter %d, |residual|_2=%.9g\n", it + 1, sqrt(rTr));
        if (!isfinite(rTr)) { status = FLUB_ENONFINITE; flub_set_error("non-finite residual at iteration %d", it + 1); break; }
        bool redo = false;   // the speculative cycle used an r that was then replaced
        if (rTr < tol) {
            TRY(px_flush(h->s, px));
            double rTrue;
            TRY(mg_split_true_residual(h, R, &rTrue));
            if (verbose) printf("  true |residual|_2=%.9g\n", sqrt(rTrue));
            rTr = rTrue;
            if (!isfinite(rTrue)) { status = FLUB_ENONFINITE; flub_set_error("non-finite true residual"); break; }
            if (rTrue < tol || replacements >= max_replacements) break;
            ++replacements;
            redo = true;
        }
        rate = r_prev > 0.0 ? fmin(rTr / r_prev, 1.0) : 1.0;
        r_prev = rTr;
        if (!speculate || redo) TRY(mg_precondition_split(h, zs ^ 1));
        zs ^= 1;
        TRY(launch_finish(h, h->vc_dot_nblk, FIN_BETA, eps));
        it += 1;
    }
    TRY(px_flush(h->s, px));
    if (hsum) TRY(mg_rtr_to_device(h, rTr));
    // back to the reference layout (sp.zpre is free scratch now)
    TRY(mg_split_leave(h, px.x_set));
    if (!h->x_only) {
        TRY(split_convert(h->s, L0, R, L0.r, 4, false));
        TRY(split_convert(h->s, L0, S.z[zs], L0.z, 4, false));
        TRY(split_convert(h->s, L0, h->Ap, S.zpre, 4, false));
        FLUB_CHECK_HIP(hipMemcpyAsync(h->Ap, S.zpre, (size_t)n * 4, hipMemcpyDeviceToDevice, h->s));
        TRY(split_convert(h->s, L0, px.p(), S.zpre, 4, false));
        FLUB_CHECK_HIP(hipMemcpyAsync(h->p, S.zpre, (size_t)n * 4, hipMemcpyDeviceToDevice, h->s));
    }
    if (status == FLUB_OK && it >= cap && max_iters < 0) {
        status = FLUB_ENOTCONV;
        flub_set_error("no convergence within %d iterations", cap);
    }
    if (verbose) printf("finished with  %d  iterations\n", it);
    if (iters) *iters = it;
    if (rTr_out) *rTr_out = rTr;
    return status;
}

// mgpcg.py:378-462 for the "phases" smoother on the fused kernels in the
// reference layout (grids the colour-split layout does not support): the z.r
// partials come out of the V-cycle's last sweep, update_p is folded into the
// Ap kernel (out of place, alternating p and p_alt) and update_x into
// update_r, all in FP32 as in the reference.
static int solve_fast(flub_mgpcg *h, int max_iters, double eps, double abs_tol, double rel_tol, int verbose,
                      int *iters, double *rTr_out)
{
    Level &L0 = h->lv[0];
    const uint32_t n = L0.n;
    const int nb = red_blocks(n);
    float *const p_home = h->p;
    float *R = L0.r, *X = h->x;
    k_dot<<<nb, RBS, 0, h->s>>>(R, R, n, h->part);
    FLUB_CHECK_LAUNCH();
    TRY(finish_rtr_to_host(h, nb));
    TRY(mg_precondition(h, FLUB_SMOOTHER_PHASES, true));
    TRY(launch_finish(h, h->vc_dot_nblk, FIN_OLDZTR, 0.0));
    FLUB_CHECK_HIP(hipEventSynchronize(h->ev));
    const double rTr0 = h->hsc[SC_RTR];
    const double tol = fmax(abs_tol, rTr0 * rel_tol);
    const int cap = max_iters < 0 ? SOLVE_SAFETY_CAP : max_iters;
    int it = 0;
    double rTr = rTr0;
    int status = FLUB_OK;
    while (it < cap) {
        int nbp = 0;
        TRY(fast_p_Ap(h->s, L0, h->p, h->p_alt, h->Ap, h->sc, h->part, &nbp));
        float *t = h->p; h->p = h->p_alt; h->p_alt = t;
        TRY(launch_finish(h, nbp, FIN_ALPHA, eps));
        k_update_rx_dot<true><<<nb, RBS, 0, h->s>>>(R, X, h->Ap, h->p, h->sc, n, h->part);
        FLUB_CHECK_LAUNCH();
        TRY(finish_rtr_to_host(h, nb));
        FLUB_CHECK_HIP(hipEventSynchronize(h->ev));
        rTr = h->hsc[SC_RTR];
        if (verbose) printf("after iter %d, |residual|_2=%.9g\n", it + 1, sqrt(rTr));
        if (!isfinite(rTr)) { status = FLUB_ENONFINITE; flub_set_error("non-finite residual at iteration %d", it + 1); break; }
        if (rTr < tol) break;
        TRY(mg_precondition(h, FLUB_SMOOTHER_PHASES, true));
        TRY(launch_finish(h, h->vc_dot_nblk, FIN_BETA, eps));
        it += 1;
    }
    if (h->p != p_home) {   // leave p in the buffer the handle exposes
        FLUB_CHECK_HIP(hipMemcpyAsync(p_home, h->p, (size_t)n * 4, hipMemcpyDeviceToDevice, h->s));
        h->p_alt = h->p;
        h->p = p_home;
    }
    if (status == FLUB_OK && it >= cap && max_iters < 0) {
        status = FLUB_ENOTCONV;
        flub_set_error("no convergence within %d iterations", cap);
    }
    if (verbose) printf("finished with  %d  iterations\n", it);
    if (iters) *iters = it;
    if (rTr_out) *rTr_out = rTr;
    return status;
}

static int solve_fp32(flub_mgpcg *h, int max_iters, double eps, double abs_tol, double rel_tol, int smoother,
                      int verbose, int *iters, double *rTr_out)
{
    Level &L0 = h->lv[0];
    const uint32_t n = L0.n;
    const int nb = red_blocks(n);
    const bool centred = smoother != FLUB_SMOOTHER_PHASES;
    if (centred) TRY(mg_zero_centered(h));
    TRY(mg_dot(h, L0.r, L0.r, FIN_RTR, 0.0));
    double rTr0;
    TRY(read_scalar(h, SC_RTR, &rTr0));
    const double tol = fmax(abs_tol, rTr0 * rel_tol);
    TRY(mg_precondition(h, smoother, true));
    k_update_p<<<nb256(n), 256, 0, h->s>>>(h->p, L0.z, h->sc, n);
    FLUB_CHECK_LAUNCH();
    TRY(mg_dot(h, L0.z, L0.r, FIN_OLDZTR, 0.0));
    const int cap = max_iters < 0 ? SOLVE_SAFETY_CAP : max_iters;
    int it = 0;
    double rTr = rTr0;
    int status = FLUB_OK;
    while (it < cap) {
        TRY(compute_Ap_dot(h, smoother, FIN_ALPHA, eps));
        if (centred) {
            k_update_r<<<nb256(n), 256, 0, h->s>>>(L0.r, h->Ap, h->sc, n);
            k_update_x<<<nb256(n), 256, 0, h->s>>>(h->x, h->p, h->sc, n);
            FLUB_CHECK_LAUNCH();
            TRY(mg_zero_centered(h));
            TRY(mg_dot(h, L0.r, L0.r, FIN_RTR, 0.0));
        } else {
            k_update_rx_dot<true><<<nb, RBS, 0, h->s>>>(L0.r, h->x, h->Ap, h->p, h->sc, n, h->part);
            FLUB_CHECK_LAUNCH();
            TRY(launch_finish(h, nb, FIN_RTR, 0.0));
        }
        TRY(read_scalar(h, SC_RTR, &rTr));
        if (verbose) printf("after iter %d, |residual|_2=%.9g\n", it + 1, sqrt(rTr));
        if (!isfinite(rTr)) { status = FLUB_ENONFINITE; flub_set_error("non-finite residual at iteration %d", it + 1); break; }
        if (rTr < tol) break;
        TRY(mg_precondition(h, smoother, true));
        TRY(mg_dot(h, L0.z, L0.r, FIN_BETA, eps));
        k_update_p<<<nb256(n), 256, 0, h->s>>>(h->p, L0.z, h->sc, n);
        FLUB_CHECK_LAUNCH();
        it += 1;
    }
    if (status == FLUB_OK && it >= cap && max_iters < 0) {
        status = FLUB_ENOTCONV;
        flub_set_error("no convergence within %d iterations", cap);
    }
    if (verbose) printf("finished with  %d  iterations\n", it);
    if (iters) *iters = it;
    if (rTr_out) *rTr_out = rTr;
    return status;
}

FLUB_API int flub_mgpcg_solve(flub_mgpcg *h, int max_iters, double eps, double abs_tol, double rel_tol,
                              int smoother, int verbose, int *iters, double *rTr)
{
    FLUB_REQUIRE(h && smoother_ok(smoother), FLUB_EINVAL, "bad handle or smoother");
    FLUB_REQUIRE(h->tags_set, FLUB_ESTATE, "solve before set_tags");
    h->x64_valid = false;   // a split solve sets it when it hands x back
    if (h->prec != FLUB_PREC_FP32)
        return mg_solve_mixed(h, max_iters, eps, abs_tol, rel_tol, smoother, verbose, iters, rTr);
    if (smoother == FLUB_SMOOTHER_PHASES && use_fast(h, smoother))
        return h->split ? solve_split(h, max_iters, eps, abs_tol, rel_tol, verbose, iters, rTr)
                        : solve_fast(h, max_iters, eps, abs_tol, rel_tol, verbose, iters, rTr);
    return solve_fp32(h, max_iters, eps, abs_tol, rel_tol, smoother, verbose, iters, rTr);
}

FLUB_API int flub_mgpcg_buffer(flub_mgpcg *h, int which, int level, void **ptr)
{
    FLUB_REQUIRE(h && ptr, FLUB_EINVAL, "null argument");
    FLUB_REQUIRE(level >= 0 && level < h->L, FLUB_EINVAL, "level %d out of range", level);
    switch (which) {
    case 0: *ptr = h->x; break;
    case 1: *ptr = h->p; break;
    case 2: *ptr = h->Ap; break;
    case 3: *ptr = h->lv[level].r; break;
    case 4: *ptr = h->lv[level].z; break;
    case 5: *ptr = h->lv[level].tag; break;
    case 6: *ptr = h->sc; break;
    default: FLUB_REQUIRE(false, FLUB_EINVAL, "unknown buffer id %d", which);
    }
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_compute_divd(flub_mgpcg *h, int level, float *divd)
{
    FLUB_REQUIRE(h && divd, FLUB_EINVAL, "null argument");
    FLUB_REQUIRE(level >= 0 && level < h->L, FLUB_EINVAL, "level out of range");
    Level &L = h->lv[level];
    k_divd<<<nb256(L.n), 256, 0, h->s>>>(L.tag, divd, L.nx, L.ny, L.nz);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_scalars(flub_mgpcg *h, double out[5])
{
    FLUB_REQUIRE(h && out, FLUB_EINVAL, "null argument");
    FLUB_CHECK_HIP(hipMemcpyAsync(h->hsc, h->sc, SC_NUM * sizeof(double), hipMemcpyDeviceToHost, h->s));
    FLUB_CHECK_HIP(hipStreamSynchronize(h->s));
    out[0] = h->hsc[SC_ALPHA];
    out[1] = h->hsc[SC_BETA];
    out[2] = h->hsc[SC_SUM];
    out[3] = h->hsc[SC_OLDZTR];
    out[4] = h->hsc[SC_RTR];
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_level_shape(flub_mgpcg *h, int level, int shape[3])
{
    FLUB_REQUIRE(h && shape, FLUB_EINVAL, "null argument");
    FLUB_REQUIRE(level >= 0 && level < h->L, FLUB_EINVAL, "level out of range");
    shape[0] = h->lv[level].nx;
    shape[1] = h->lv[level].ny;
    shape[2] = h->lv[level].nz;
    return FLUB_OK;
}

FLUB_API int flub_inner(const float *a, const float *b, size_t n, double *out_dev, void *stream)
{
    FLUB_REQUIRE(a && b && out_dev, FLUB_EINVAL, "flub_inner: null pointer");
    FLUB_REQUIRE(n < 2147483648ull, FLUB_ESHAPE, "flub_inner: n too large");
    static thread_local double *part = nullptr;
    if (!part) FLUB_CHECK_HIP(hipMalloc((void **)&part, MAXBLK * sizeof(double)));
    hipStream_t s = as_stream(stream);
    const int nb = red_blocks((uint32_t)n);
    if (n) {
        k_dot<<<nb, RBS, 0, s>>>(a, b, (uint32_t)n, part);
        FLUB_CHECK_LAUNCH();
    }
    k_finish<<<1, RBS, 0, s>>>(part, n ? nb : 0, out_dev, FIN_OUT, 0.0);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

// ---------------------------------------------------------------- inner product without allocation
// flub_inner_noalloc keeps its FP64 partials (at most MAXBLK, one per block of
// the reduction) in this file-scope device array instead of memory the
// library allocates: every GPU the library runs on holds one copy of it (16
// KiB) until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: the kernels receive the per-device address that
// hipGetSymbolAddress returns, resolved once per device on the entry point's
// first call there.  Calls on one GPU share its copy (the last call's partials
// stay in it): calls that overlap on one GPU, from several streams or
// threads, race.
static __device__ __align__(256) double scratch[MAXBLK];
#define FLUB_NOALLOC_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[FLUB_NOALLOC_MAX_DEVICES];

static int noalloc_scratch(double **out)
{
    int dev = 0;
    FLUB_CHECK_HIP(hipGetDevice(&dev));
    FLUB_REQUIRE(dev >= 0 && dev < FLUB_NOALLOC_MAX_DEVICES, FLUB_EINVAL,
                 "flub_inner_noalloc: device %d beyond the %d the scratch table holds", dev,
                 FLUB_NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            flub_set_error("flub_inner_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                           hipGetErrorString(e));
            return e != hipSuccess ? (int)e : FLUB_ESTATE;
        }
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return FLUB_OK;
}

FLUB_API int flub_inner_noalloc(const float *a, const float *b, size_t n, double *out_dev, void *stream)
{
    FLUB_REQUIRE(a && b && out_dev, FLUB_EINVAL, "flub_inner_noalloc: null pointer");
    FLUB_REQUIRE(n <= FLUB_INNER_NOALLOC_MAX_N, FLUB_ESHAPE,
                 "flub_inner_noalloc: n = %zu exceeds the maximum %llu", n,
                 (unsigned long long)FLUB_INNER_NOALLOC_MAX_N);
    double *part = nullptr;
    TRY(noalloc_scratch(&part));
    hipStream_t s = as_stream(stream);
    const int nb = red_blocks((uint32_t)n);   // <= MAXBLK: the scratch holds every partial
    if (n) {
        k_dot<<<nb, RBS, 0, s>>>(a, b, (uint32_t)n, part);
        FLUB_CHECK_LAUNCH();
    }
    k_finish<<<1, RBS, 0, s>>>(part, n ? nb : 0, out_dev, FIN_OUT, 0.0);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

// example_fluid.py:120-147 + water_3d.py:58-68.
FLUB_API int flub_project(flub_mgpcg *h, float *u, float *v, float *w, const float *ur, const float *vr,
                          const float *wr, const int32_t *tag, float h_cell, int max_iters, double abs_tol,
                          double rel_tol, int smoother, float *p, float *div, int *iters, double *rTr)
{
    FLUB_REQUIRE(h && u && v && w && ur && vr && wr && tag && p, FLUB_EINVAL, "flub_project: null argument");
    FLUB_REQUIRE(smoother_ok(smoother), FLUB_EINVAL, "bad smoother");
    FLUB_REQUIRE(h_cell > 0.0f, FLUB_EINVAL, "cell size must be positive");
    Level &L0 = h->lv[0];
    const int nx = L0.nx, ny = L0.ny, nz = L0.nz;
    TRY(flub_mgpcg_set_tags(h, tag));
    const float s_div = (float)(-1.0 / (double)h_cell);
    // The colour-split solve takes r from the divergence kernel directly, in
    // its layout: r = (FLUID ? d * (-h) : 0), the values init + mask_interior
    // give, and the copy of it that its FP64 true residual uses as b; the same
    // kernel zeroes p as init does (x starts at zero without being written:
    // the solve's first x update stores alpha p); the scalars are zeroed as
    // init does (z and Ap are written before they are read).
    const bool presplit = h->split && smoother == FLUB_SMOOTHER_PHASES && use_fast(h, smoother) &&
                          divergence_split_ok(u, v, div, L0.tag, nz);
    if (presplit) {
        TRY(divergence_fused_split(u, v, w, div, ur, vr, wr, L0.tag, nx, ny, nz, s_div, h->sp.r, -h_cell, h->s,
                                   h->sp.b, nullptr, h->p));
        FLUB_CHECK_HIP(hipMemsetAsync(h->sc, 0, SC_NUM * sizeof(double), h->s));
    } else {
        float *d = div ? div : h->Ap;   // Ap is rewritten by init
        TRY(flub_divergence_fused_u8(u, v, w, d, ur, vr, wr, L0.tag, nx, ny, nz, s_div, h->s));
        TRY(flub_mgpcg_init(h, d, -h_cell));
        TRY(flub_mgpcg_mask_interior(h));
    }
    // only the pressure x is read after the solve: r, z, p, Ap stay where the
    // solver keeps them (the next init resets them)
    h->x_only = true;
    h->presplit = presplit;
    // the gradient reads the split FP64 x directly when it can (no reference-layout copy)
    const bool grad3 = h->split && subtract_gradient3_split_ok(u, v, ur, vr, nz) && getenv("FLUB_GRAD_PER_AXIS") == nullptr;
    h->want_x64n = h->x64n != nullptr && !grad3;
    h->p_out = p;
    h->p_h = h_cell;
    h->p_written = false;
    int st = flub_mgpcg_solve(h, max_iters, 1e-12, abs_tol, rel_tol, smoother, 0, iters, rTr);
    h->x_only = false;
    h->presplit = false;
    h->want_x64n = false;
    h->p_out = nullptr;
    if (st != FLUB_OK && st != FLUB_ENOTCONV) return st;
    if (h->x64_valid && grad3) {
        // a split solve: the gradient of its FP64 x (the scale chain p = x h,
        // p / h is x itself), each face update rounded once; the projected
        // field's divergence is then the solve's residual
        TRY(subtract_gradient3_split(u, v, w, ur, vr, wr, h->sp.x, L0.tag, nx, ny, nz, h->s));
    } else if (h->x64_valid) {
        TRY(flub_subtract_gradient_x64_u8(u, ur, h->x64n, L0.tag, nx, ny, nz, 1, h->s));
        TRY(flub_subtract_gradient_x64_u8(v, vr, h->x64n, L0.tag, nx, ny, nz, 2, h->s));
        TRY(flub_subtract_gradient_x64_u8(w, wr, h->x64n, L0.tag, nx, ny, nz, 3, h->s));
    } else {
        TRY(flub_subtract_gradient_x_u8(u, ur, h->x, L0.tag, nx, ny, nz, 1, h_cell, h->s));
        TRY(flub_subtract_gradient_x_u8(v, vr, h->x, L0.tag, nx, ny, nz, 2, h_cell, h->s));
        TRY(flub_subtract_gradient_x_u8(w, wr, h->x, L0.tag, nx, ny, nz, 3, h_cell, h->s));
    }
    if (!h->p_written) {
        const float inv_h = (float)(1.0 / (double)h_cell);
        k_store_pressure<<<nb256(L0.n), 256, 0, h->s>>>(p, h->x, h_cell, inv_h, L0.n);
        FLUB_CHECK_LAUNCH();
    }
    h->p_written = false;
    return st;
}

// ---------------------------------------------------------------- in-solve timing
FLUB_API int flub_mgpcg_probe(flub_mgpcg *h, int kernel, int max_launches)
{
    FLUB_REQUIRE(h, FLUB_EINVAL, "null handle");
    FLUB_REQUIRE(kernel == 8, FLUB_EINVAL, "flub_mgpcg_probe: kernel %d not probed (only 8)", kernel);
    FLUB_REQUIRE(max_launches >= 1 && max_launches <= PROBE_MAX, FLUB_EINVAL,
                 "flub_mgpcg_probe: max_launches must be 1..%d", PROBE_MAX);
    FLUB_REQUIRE(h->split && h->pair_l0, FLUB_ESTATE, "flub_mgpcg_probe: no colour-split FP32 level 0");
    FLUB_REQUIRE(!h->probe, FLUB_ESTATE, "flub_mgpcg_probe: already armed");
    Probe *p = (Probe *)calloc(1, sizeof(Probe));
    FLUB_REQUIRE(p, FLUB_EINVAL, "flub_mgpcg_probe: out of host memory");
    for (int i = 0; i < 2 * max_launches; ++i) {
        const hipError_t e = hipEventCreate(&p->ev[i]);
        if (e != hipSuccess) {
            for (int j = 0; j < i; ++j) hipEventDestroy(p->ev[j]);
            free(p);
            flub_set_error("hipEventCreate: %s", hipGetErrorString(e));
            return (int)e;
        }
    }
    p->max = max_launches;
    h->probe = p;
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_probe_read(flub_mgpcg *h, float *avg_ms, float *min_ms, int *count)
{
    FLUB_REQUIRE(h && avg_ms && count, FLUB_EINVAL, "null argument");
    FLUB_REQUIRE(h->probe, FLUB_ESTATE, "flub_mgpcg_probe_read: no probe armed");
    Probe *p = h->probe;
    h->probe = nullptr;
    int rc = FLUB_OK;
    double sum = 0.0;
    float mn = 0.0f;
    if (p->n > 0) rc = (int)hipEventSynchronize(p->ev[2 * p->n - 1]);
    for (int i = 0; i < p->n && rc == FLUB_OK; ++i) {
        float ms = 0.0f;
        rc = (int)hipEventElapsedTime(&ms, p->ev[2 * i], p->ev[2 * i + 1]);
        sum += ms;
        mn = i == 0 ? ms : fminf(mn, ms);
    }
    if (rc != FLUB_OK) flub_set_error("flub_mgpcg_probe_read: %s", hipGetErrorString((hipError_t)rc));
    *count = p->n;
    *avg_ms = p->n ? (float)(sum / p->n) : 0.0f;
    if (min_ms) *min_ms = mn;
    for (int i = 0; i < 2 * p->max; ++i) hipEventDestroy(p->ev[i]);
    free(p);
    return rc;
}

// ---------------------------------------------------------------- measurement
FLUB_API int flub_mgpcg_profile(flub_mgpcg *h, int kernel, int reps, float *avg_ms, double *bytes)
{
    FLUB_REQUIRE(h && avg_ms && reps > 0, FLUB_EINVAL, "flub_mgpcg_profile: bad argument");
    FLUB_REQUIRE(h->tags_set, FLUB_ESTATE, "profile before set_tags");
    h->x64_valid = false;
    h->rh_ready = h->zh0_ready = -1;   // the replayed V-cycle converts r itself
    Level &L0 = h->lv[0];
    const double N = (double)L0.n, v = h->prec == FLUB_PREC_FP32 ? 4.0 : 2.0;
    double b = 0.0;
    hipEvent_t e0, e1;
    FLUB_CHECK_HIP(hipEventCreate(&e0));
    FLUB_CHECK_HIP(hipEventCreate(&e1));
    const int nb = red_blocks(L0.n);
    auto launch_once = [&]() -> int {
        switch (kernel) {
        case 0:
            if (h->fast) return fast_sweep(h->s, L0, L0.z, L0.z, nullptr, 0, SWEEP_PLAIN, nullptr, nullptr);
            return half_sweep(h, 0, 1, 0);
        case 1:
            if (h->fast) {
                int nbp = 0;
                TRY(fast_p_Ap(h->s, L0, h->p, h->p_alt, h->Ap, h->sc, h->part, &nbp));
                return launch_finish(h, nbp, FIN_SUM, 0.0);
            }
            return compute_Ap_dot(h, FLUB_SMOOTHER_PHASES, FIN_SUM, 0.0);
        case 2:
            k_update_rx_dot<true><<<nb, RBS, 0, h->s>>>(L0.r, h->x, h->Ap, h->p, h->sc, L0.n, h->part);
            FLUB_CHECK_LAUNCH();
            return launch_finish(h, nb, FIN_RTR, 0.0);
        case 3: return h->split ? mg_precondition_split(h, 0) : mg_precondition(h, FLUB_SMOOTHER_PHASES, true);
        case 4: {
            Level &C = h->lv[1];
            if (h->fast) return fast_restrict(h->s, L0, L0.zpre, C);
            k_restrict<true><<<nb256(C.n), 256, 0, h->s>>>(L0.r, L0.z, L0.tag, C.r, C.tag, L0.nx, L0.ny, L0.nz, C.cs);
            FLUB_CHECK_LAUNCH();
            return FLUB_OK;
        }
        case 5:
            if (h->fast) return fast_sweep(h->s, L0, L0.zpre, L0.z, h->lv[1].z, 1, SWEEP_PROLONG, nullptr, nullptr);
            k_prolong<<<nb256(L0.n), 256, 0, h->s>>>(L0.z, h->lv[1].z, L0.nx, L0.ny, L0.nz);
            FLUB_CHECK_LAUNCH();
            return FLUB_OK;
        case 6:
            k_update_p<<<nb256(L0.n), 256, 0, h->s>>>(h->p, L0.z, h->sc, L0.n);
            FLUB_CHECK_LAUNCH();
            return FLUB_OK;
        case 7: {   // one blocked launch of 4 level-0 half-sweeps (zpre -> z)
            const uint8_t ph[4] = {1, 0, 1, 0};
            FLUB_REQUIRE(h->tb_order && tb_supported(L0), FLUB_ESTATE, "no blocked smoothing for this grid");
            return tb_chain(h->s, h->tb_order, L0, FLUB_PREC_FP32, SWEEP_PLAIN, L0.zpre, nullptr, 1.0f, L0.r, ph, 4,
                            L0.z, L0.zpre, nullptr, nullptr, h->tb_g8);
        }
        case 8:   // level-0 half-sweep in the colour-split layout
            FLUB_REQUIRE(h->split, FLUB_ESTATE, "no colour-split level 0 for this handle");
            return split_sweep(h->s, L0, h->sp.z[0], h->sp.z[0], nullptr, h->sp.r, h->sp.code, 0, SWEEP_PLAIN,
                               nullptr, nullptr);
        case 9:
            FLUB_REQUIRE(h->split, FLUB_ESTATE, "no colour-split level 0 for this handle");
            return split_restrict(h->s, L0, h->sp.r, h->sp.zpre, h->sp.code, h->lv[1]);
        case 10: {   // the split solve's p step: p into a ring buffer, then Ap and the p.Ap partials
            FLUB_REQUIRE(h->split, FLUB_ESTATE, "no colour-split level 0 for this handle");
            int nbp = 0;
            PxState px;
            px_init(h, px, h->sp.x, L0.n);
            return split_pstep_Ap(h->s, px, h->sp.z[0], nullptr, L0, h->Ap, h->sp.code, h->part, &nbp, h->pap_ib);
        }
        case 14: {   // kernel 10 followed by the finish that sums the partials into sc[SC_SUM]
            FLUB_REQUIRE(h->split, FLUB_ESTATE, "no colour-split level 0 for this handle");
            int nbp = 0;
            PxState px;
            px_init(h, px, h->sp.x, L0.n);
            TRY(split_pstep_Ap(h->s, px, h->sp.z[0], nullptr, L0, h->Ap, h->sp.code, h->part, &nbp, h->pap_ib));
            return launch_finish(h, nbp, FIN_SUM, 0.0);
        }
        case 11:   // the split solve's r step: r -= alpha Ap and r.r
            k_update_rx_dot<false><<<nb, RBS, 0, h->s>>>(L0.r, nullptr, h->Ap, nullptr, h->sc, L0.n, h->part);
            FLUB_CHECK_LAUNCH();
            return launch_finish(h, nb, FIN_RTR, 0.0);
        case 12: {   // the mixed split solve's level-0 post-smoothing half-sweep: FP32 z, 16-bit r
            FLUB_REQUIRE(h->split && h->prec != FLUB_PREC_FP32, FLUB_ESTATE,
                         "no mixed-precision colour-split level 0 for this handle");
            SplitMixedSweep m{};
            m.kind = SMX_POST;
            m.phase = 0;
            m.zin = h->sp.z[0];
            m.zout = h->sp.z[0];
            m.rh = h->sp.rh;
            m.code = h->sp.code;
            m.r32 = h->sp.r;
            m.sc = h->sc;
            return split_sweep_mixed(h->s, L0, h->prec, m);
        }
        default: flub_set_error("unknown kernel id %d", kernel); return FLUB_EINVAL;
        }
    };
    switch (kernel) {
    case 0: b = N * (2 * v + 1); break;                      // z (7-pt), r (own colour), tags
    case 1: b = h->fast ? N * (4 * 4 + 1) : N * (2 * 4 + 1); break;   // (z, p_old ->) p, tags -> Ap
    case 2: b = N * 6 * 4; break;                            // r, x, Ap, p -> r, x
    case 3: b = N * 0; break;                                // filled by the caller's model
    case 4: b = N * (2 * v + 1) + N / 8 * (v + 1); break;
    case 5: b = h->fast ? N * (2.625 * v + 1) : N * 2 * v + N / 8 * v; break;   // fused: zpre, zc, r/2, tags -> z
    case 6: b = N * 3 * 4; break;
    case 7: b = 4 * N * (2 * v + 1); break;                  // four half-sweeps
    case 8: b = N * (2 * v + 1); break;
    case 9: b = N * (2 * v + 1) + N / 8 * (v + 1); break;
    case 10: case 14: b = N * (4 * 4 + 2); break;            // p, z, codes -> p', Ap
    case 11: b = N * 3 * 4; break;                           // r, Ap -> r
    case 12: b = N * 8; break;                               // as 8 with the 16-bit r (own colour)
    default: break;
    }
    int rc = launch_once();   // warm-up (captures the graph for id 3)
    if (rc == FLUB_OK) {
        FLUB_CHECK_HIP(hipStreamSynchronize(h->s));
        FLUB_CHECK_HIP(hipEventRecord(e0, h->s));
        for (int i = 0; i < reps && rc == FLUB_OK; ++i) rc = launch_once();
        FLUB_CHECK_HIP(hipEventRecord(e1, h->s));
        FLUB_CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        FLUB_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        *avg_ms = ms / reps;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (bytes) *bytes = b;
    return rc;
}

// ---------------------------------------------------------------- caller buffers
FLUB_API int flub_mgpcg_apply_A(flub_mgpcg *h, const float *x, float *y)
{
    FLUB_REQUIRE(h && x && y, FLUB_EINVAL, "flub_mgpcg_apply_A: null argument");
    FLUB_REQUIRE(h->tags_set, FLUB_ESTATE, "apply_A before set_tags");
    Level &L0 = h->lv[0];
    const int nb = red_blocks(L0.n);
    k_Ap_dot<true><<<nb, RBS, 0, h->s>>>(y, x, L0.tag, L0.nx, L0.ny, L0.nz, h->part);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_mgpcg_precondition_into(flub_mgpcg *h, const float *r, float *z, int smoother)
{
    FLUB_REQUIRE(h && r && z && smoother_ok(smoother), FLUB_EINVAL, "flub_mgpcg_precondition_into: bad argument");
    FLUB_REQUIRE(h->tags_set, FLUB_ESTATE, "precondition before set_tags");
    Level &L0 = h->lv[0];
    if (h->split && smoother == FLUB_SMOOTHER_PHASES) {
        // the cycle converts the caller's r: no BF16 r stored by an earlier
        // solve's r update (one that returned early) may stand in for it
        h->rh_ready = h->zh0_ready = -1;
        TRY(split_convert(h->s, L0, r, h->sp.r, 4, true));
        if (h->prec != FLUB_PREC_FP32) TRY(mg_mixed_rscale_of(h, h->sp.r));
        TRY(mg_precondition_split(h, 0));
        TRY(split_convert(h->s, L0, h->sp.z[0], z, 4, false));
        // the mixed cycle leaves z times its storage scale
        return h->prec != FLUB_PREC_FP32 ? scale_by_device(h->s, z, h->sc + SC_ZINV0, L0.n) : FLUB_OK;
    }
    FLUB_CHECK_HIP(hipMemcpyAsync(L0.r, r, (size_t)L0.n * 4, hipMemcpyDeviceToDevice, h->s));
    if (h->prec != FLUB_PREC_FP32) TRY(mg_mixed_rscale(h));
    TRY(mg_precondition(h, smoother, true));
    FLUB_CHECK_HIP(hipMemcpyAsync(z, L0.z, (size_t)L0.n * 4, hipMemcpyDeviceToDevice, h->s));
    return FLUB_OK;
}
