// Fast tagged multigrid kernels (levels with nz % 4 == 0).
//
// Each thread owns a "quad": 4 consecutive cells along k of one (i, j) row,
// loaded and stored as one 16-byte vector (tags as one 32-bit word), so a
// wavefront moves whole 256-byte runs of z, r and 64-byte runs of tags.  The
// i/j neighbours are the quads of the adjacent rows; the k neighbours beyond
// the quad are two scalar loads.  Per cell the arithmetic is exactly the
// reference's (mgpcg.py:72-101,118-142), so results are bitwise those of the
// one-cell-per-thread kernels and of the FP32 oracle.
//
// Fusions (each bitwise equal to the unfused launch sequence):
//  * ZERO: the first pre-smoothing half-sweep of a V-cycle level starts from
//    z = 0 (mgpcg.py:299,312): z is not read, all four cells are written.
//  * PROLONG: the first post-smoothing half-sweep reads z + z_coarse[parent]
//    (mgpcg.py:113-116,339) at every stencil point and writes the prolongated
//    value to the cells it does not update.
//  * DOT: the last half-sweep of the cycle also accumulates z.r (FP64) of its
//    final values (mgpcg.py:451).
#include "common.h"
#include "mgpcg_internal.h"
#include "smooth.h"

namespace {

enum { MODE_PLAIN = 0, MODE_ZERO = 1, MODE_PROLONG = 2 };

struct Quad {
    float v[4];
    int t[4];
};

__device__ __forceinline__ void unpack_tags(uint32_t w, int t[4])
{
    t[0] = w & 0xff; t[1] = (w >> 8) & 0xff; t[2] = (w >> 16) & 0xff; t[3] = w >> 24;
}

// The reference's neighbor_dot_tag for one cell from explicit neighbour
// values; an out-of-domain neighbour is given as the cell itself (clamp) and
// excluded from divd.
__device__ __forceinline__ float ndt7(float xI, int tI, float xip, int tip, bool ip, float xim, int tim,
                                      bool im, float xjp, int tjp, bool jp, float xjm, int tjm, bool jm,
                                      float xkp, int tkp, bool kp, float xkm, int tkm, bool km, int *divd)
{
    float ret = 0.0f;
    int d = 0;
    {
        const float a = (tip == 0 ? 0.0f : xip) - xI, b = (tim == 0 ? 0.0f : xim) - xI;
        ret = ret + (a * (tip < 2 ? 1.0f : 0.0f) + b * (tim < 2 ? 1.0f : 0.0f));
        d += (ip && tip < 2) + (im && tim < 2);
    }
    {
        const float a = (tjp == 0 ? 0.0f : xjp) - xI, b = (tjm == 0 ? 0.0f : xjm) - xI;
        ret = ret + (a * (tjp < 2 ? 1.0f : 0.0f) + b * (tjm < 2 ? 1.0f : 0.0f));
        d += (jp && tjp < 2) + (jm && tjm < 2);
    }
    {
        const float a = (tkp == 0 ? 0.0f : xkp) - xI, b = (tkm == 0 ? 0.0f : xkm) - xI;
        ret = ret + (a * (tkp < 2 ? 1.0f : 0.0f) + b * (tkm < 2 ? 1.0f : 0.0f));
        d += (kp && tkp < 2) + (km && tkm < 2);
    }
    if (tI != 1) ret = 0.0f;
    *divd = tI == 1 ? d : 0;
    return ret;
}

struct Geo {
    int nx, ny, nz;
    int cny, cnz;  // coarse dims (PROLONG)
};

// Load the quad of row (i, j) starting at k0 (z values, prolongated when
// PROLONG, zero when ZERO) and its tags.
template <int MODE>
__device__ __forceinline__ void load_quad(const float *z, const float *__restrict__ zc,
                                          const uint8_t *__restrict__ tag, const Geo &g, int i, int j,
                                          int k0, Quad &q)
{
    const uint32_t base = cidx(i, j, k0, g.ny, g.nz);
    unpack_tags(*reinterpret_cast<const uint32_t *>(tag + base), q.t);
    if (MODE == MODE_ZERO) {
        q.v[0] = q.v[1] = q.v[2] = q.v[3] = 0.0f;
        return;
    }
    const float4 a = *reinterpret_cast<const float4 *>(z + base);
    q.v[0] = a.x; q.v[1] = a.y; q.v[2] = a.z; q.v[3] = a.w;
    if (MODE == MODE_PROLONG) {
        const uint32_t cb = cidx(i >> 1, j >> 1, k0 >> 1, g.cny, g.cnz);
        const float2 c = *reinterpret_cast<const float2 *>(zc + cb);
        q.v[0] = q.v[0] + c.x; q.v[1] = q.v[1] + c.x;
        q.v[2] = q.v[2] + c.y; q.v[3] = q.v[3] + c.y;
    }
}

template <int MODE>
__device__ __forceinline__ float load_one(const float *z, const float *__restrict__ zc,
                                          const Geo &g, int i, int j, int k)
{
    if (MODE == MODE_ZERO) return 0.0f;
    float v = z[cidx(i, j, k, g.ny, g.nz)];
    if (MODE == MODE_PROLONG) v = v + zc[cidx(i >> 1, j >> 1, k >> 1, g.cny, g.cnz)];
    return v;
}

}  // namespace

// One red-black half-sweep (phase p updates (i+j+k)&1 == p^1) over quads.
// z is read from zin and the quad written to zout: in place (zin == zout) for
// PLAIN and ZERO, where a thread rewrites the cells it does not update with
// the bits it read; out of place for PROLONG, whose rewritten cells change.
template <int MODE, bool DOT>
__global__ void __launch_bounds__(256) k_rbgs4(const float *zin, float *zout, const float *__restrict__ r,
                                               const uint8_t *__restrict__ tag, const float *__restrict__ zc,
                                               int nx, int ny, int nz, int phase, double *__restrict__ part)
{
    __shared__ double lds[4];
    const Geo g{nx, ny, nz, ny >> 1, nz >> 1};
    const int nq = nz >> 2;
    const uint32_t total = (uint32_t)nx * ny * nq;
    double acc = 0.0;
    const Chunk ch = xcd_chunk(total);
    for (uint32_t t = ch.begin + threadIdx.x; t < ch.end; t += 256u) {
        const uint32_t row = t / nq;
        const int k0 = (int)(t - row * nq) * 4;
        const int i = (int)(row / ny), j = (int)(row - (uint32_t)i * ny);
        Quad c, xp, xm, yp, ym;
        load_quad<MODE>(zin, zc, tag, g, i, j, k0, c);
        const bool ip = i + 1 < nx, im = i > 0, jp = j + 1 < ny, jm = j > 0;
        if (ip) load_quad<MODE>(zin, zc, tag, g, i + 1, j, k0, xp); else xp = c;
        if (im) load_quad<MODE>(zin, zc, tag, g, i - 1, j, k0, xm); else xm = c;
        if (jp) load_quad<MODE>(zin, zc, tag, g, i, j + 1, k0, yp); else yp = c;
        if (jm) load_quad<MODE>(zin, zc, tag, g, i, j - 1, k0, ym); else ym = c;
        const bool kpo = k0 + 4 < nz, kmo = k0 > 0;
        const uint32_t base = cidx(i, j, k0, ny, nz);
        const float zk4 = kpo ? load_one<MODE>(zin, zc, g, i, j, k0 + 4) : 0.0f;
        const int tk4 = kpo ? tag[base + 4] : 0;
        const float zkm = kmo ? load_one<MODE>(zin, zc, g, i, j, k0 - 1) : 0.0f;
        const int tkm = kmo ? tag[base - 1] : 0;
        const float4 r4 = *reinterpret_cast<const float4 *>(r + base);
        const float rv[4] = {r4.x, r4.y, r4.z, r4.w};
        float out[4];
        const int upd = ((i + j + k0) & 1) ^ phase ^ 1;   // quad offset parity that is updated
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            out[q] = c.v[q];
            if ((q & 1) != upd) continue;
            const bool kp = q < 3 ? true : kpo, km = q > 0 ? true : kmo;
            const float xkp = q < 3 ? c.v[q + 1] : (kpo ? zk4 : c.v[q]);
            const int tkp = q < 3 ? c.t[q + 1] : (kpo ? tk4 : c.t[q]);
            const float xkm = q > 0 ? c.v[q - 1] : (kmo ? zkm : c.v[q]);
            const int tkmv = q > 0 ? c.t[q - 1] : (kmo ? tkm : c.t[q]);
            int divd;
            const float nd = ndt7(c.v[q], c.t[q], xp.v[q], xp.t[q], ip, xm.v[q], xm.t[q], im, yp.v[q], yp.t[q], jp,
                                  ym.v[q], ym.t[q], jm, xkp, tkp, kp, xkm, tkmv, km, &divd);
            out[q] = divd != 0 ? c.v[q] + (rv[q] + nd) / (float)divd : 0.0f;
        }
        *reinterpret_cast<float4 *>(zout + base) = make_float4(out[0], out[1], out[2], out[3]);
        if (DOT) {
#pragma unroll
            for (int q = 0; q < 4; ++q) acc += (double)out[q] * (double)rv[q];
        }
    }
    if (DOT) {
        const double s = block_sum<256>(acc, lds);
        if (threadIdx.x == 0) part[blockIdx.x] = s;
    }
}

// Gather restriction (mgpcg.py:94-101) for two coarse cells per thread: the
// fine quad k0..k0+3 of the four child rows (2I+a, 2J+b).  Residual
// r - (-ndt(z)) per fine cell; a FLUID parent sums its 8 children in the
// serial launch order (ii, jj, kk), others get 0.
__global__ void __launch_bounds__(256) k_restrict4(const float *__restrict__ r, const float *__restrict__ z,
                                                   const uint8_t *__restrict__ tag, float *__restrict__ rc,
                                                   const uint8_t *__restrict__ tagc, int nx, int ny, int nz)
{
    const Geo g{nx, ny, nz, 0, 0};
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
    const int nq = nz >> 2;   // coarse pairs per coarse row
    const uint32_t total = (uint32_t)cx * cy * nq;
    const Chunk ch = xcd_chunk(total);
    for (uint32_t t = ch.begin + threadIdx.x; t < ch.end; t += 256u) {
        const uint32_t crow = t / nq;
        const int kq = (int)(t - crow * nq);
        const int I = (int)(crow / cy), J = (int)(crow - (uint32_t)I * cy);
        const int k0 = kq * 4;
        const uint32_t cb = cidx(I, J, kq * 2, cy, cz);
        const uint32_t tc2 = *reinterpret_cast<const uint16_t *>(tagc + cb);
        const int tcp[2] = {(int)(tc2 & 0xff), (int)(tc2 >> 8)};
        float res[2][2][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int i = 2 * I + a, j = 2 * J + b;
                Quad c, xp, xm, yp, ym;
                load_quad<MODE_PLAIN>(z, nullptr, tag, g, i, j, k0, c);
                const bool ip = i + 1 < nx, im = i > 0, jp = j + 1 < ny, jm = j > 0;
                if (ip) load_quad<MODE_PLAIN>(z, nullptr, tag, g, i + 1, j, k0, xp); else xp = c;
                if (im) load_quad<MODE_PLAIN>(z, nullptr, tag, g, i - 1, j, k0, xm); else xm = c;
                if (jp) load_quad<MODE_PLAIN>(z, nullptr, tag, g, i, j + 1, k0, yp); else yp = c;
                if (jm) load_quad<MODE_PLAIN>(z, nullptr, tag, g, i, j - 1, k0, ym); else ym = c;
                const uint32_t base = cidx(i, j, k0, ny, nz);
                const bool kpo = k0 + 4 < nz, kmo = k0 > 0;
                const float zk4 = kpo ? z[base + 4] : 0.0f;
                const int tk4 = kpo ? tag[base + 4] : 0;
                const float zkm = kmo ? z[base - 1] : 0.0f;
                const int tkm = kmo ? tag[base - 1] : 0;
                const float4 r4 = *reinterpret_cast<const float4 *>(r + base);
                const float rv[4] = {r4.x, r4.y, r4.z, r4.w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool kp = q < 3 ? true : kpo, km = q > 0 ? true : kmo;
                    const float xkp = q < 3 ? c.v[q + 1] : (kpo ? zk4 : c.v[q]);
                    const int tkp = q < 3 ? c.t[q + 1] : (kpo ? tk4 : c.t[q]);
                    const float xkm = q > 0 ? c.v[q - 1] : (kmo ? zkm : c.v[q]);
                    const int tkmv = q > 0 ? c.t[q - 1] : (kmo ? tkm : c.t[q]);
                    int divd;
                    const float nd = ndt7(c.v[q], c.t[q], xp.v[q], xp.t[q], ip, xm.v[q], xm.t[q], im, yp.v[q],
                                          yp.t[q], jp, ym.v[q], ym.t[q], jm, xkp, tkp, kp, xkm, tkmv, km, &divd);
                    res[a][b][q] = rv[q] - (-nd);
                }
            }
        float outv[2];
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
            float s = 0.0f;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int kk = 0; kk < 2; ++kk) s += res[a][b][2 * cc + kk];
            outv[cc] = tcp[cc] == 1 ? s : 0.0f;
        }
        *reinterpret_cast<float2 *>(rc + cb) = make_float2(outv[0], outv[1]);
    }
}

// p_new = z + beta*p_old at every stencil point (mgpcg.py:229-232), then
// Ap = ndt(p_new) (mgpcg.py:36-39) and the FP64 partial of p_new.Ap; p_new is
// stored to pout (a different buffer from p: neighbours still read p_old).
__global__ void __launch_bounds__(256) k_p_Ap4(const float *__restrict__ p, float *__restrict__ pout,
                                               float *__restrict__ Ap,
                                               const float *__restrict__ z, const uint8_t *__restrict__ tag,
                                               const double *__restrict__ sc, int nx, int ny, int nz,
                                               double *__restrict__ part)
{
    __shared__ double lds[4];
    const float beta = (float)sc[SC_BETA];
    const Geo g{nx, ny, nz, 0, 0};
    const int nq = nz >> 2;
    const uint32_t total = (uint32_t)nx * ny * nq;
    double acc = 0.0;
    auto pq = [&](int i, int j, int k0, Quad &q) {
        const uint32_t base = cidx(i, j, k0, g.ny, g.nz);
        unpack_tags(*reinterpret_cast<const uint32_t *>(tag + base), q.t);
        const float4 a = *reinterpret_cast<const float4 *>(p + base);
        const float4 b = *reinterpret_cast<const float4 *>(z + base);
        q.v[0] = fmaf(beta, a.x, b.x); q.v[1] = fmaf(beta, a.y, b.y);
        q.v[2] = fmaf(beta, a.z, b.z); q.v[3] = fmaf(beta, a.w, b.w);
    };
    const Chunk ch = xcd_chunk(total);
    for (uint32_t t = ch.begin + threadIdx.x; t < ch.end; t += 256u) {
        const uint32_t row = t / nq;
        const int k0 = (int)(t - row * nq) * 4;
        const int i = (int)(row / ny), j = (int)(row - (uint32_t)i * ny);
        Quad c, xp, xm, yp, ym;
        pq(i, j, k0, c);
        const bool ip = i + 1 < nx, im = i > 0, jp = j + 1 < ny, jm = j > 0;
        if (ip) pq(i + 1, j, k0, xp); else xp = c;
        if (im) pq(i - 1, j, k0, xm); else xm = c;
        if (jp) pq(i, j + 1, k0, yp); else yp = c;
        if (jm) pq(i, j - 1, k0, ym); else ym = c;
        const uint32_t base = cidx(i, j, k0, ny, nz);
        const bool kpo = k0 + 4 < nz, kmo = k0 > 0;
        const float zk4 = kpo ? fmaf(beta, p[base + 4], z[base + 4]) : 0.0f;
        const int tk4 = kpo ? tag[base + 4] : 0;
        const float zkm = kmo ? fmaf(beta, p[base - 1], z[base - 1]) : 0.0f;
        const int tkm = kmo ? tag[base - 1] : 0;
        float out[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bool kp = q < 3 ? true : kpo, km = q > 0 ? true : kmo;
            const float xkp = q < 3 ? c.v[q + 1] : (kpo ? zk4 : c.v[q]);
            const int tkp = q < 3 ? c.t[q + 1] : (kpo ? tk4 : c.t[q]);
            const float xkm = q > 0 ? c.v[q - 1] : (kmo ? zkm : c.v[q]);
            const int tkmv = q > 0 ? c.t[q - 1] : (kmo ? tkm : c.t[q]);
            int divd;
            out[q] = ndt7(c.v[q], c.t[q], xp.v[q], xp.t[q], ip, xm.v[q], xm.t[q], im, yp.v[q], yp.t[q], jp,
                          ym.v[q], ym.t[q], jm, xkp, tkp, kp, xkm, tkmv, km, &divd);
            acc += (double)c.v[q] * (double)out[q];
        }
        *reinterpret_cast<float4 *>(pout + base) = make_float4(c.v[0], c.v[1], c.v[2], c.v[3]);
        *reinterpret_cast<float4 *>(Ap + base) = make_float4(out[0], out[1], out[2], out[3]);
    }
    const double s = block_sum<256>(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ------------------------------------------------------------------ host
// Workgroup count: one item per thread up to MAXBLK workgroups, rounded up to a
// multiple of 8 (xcd_chunk's remap is a bijection on multiples of 8).
static unsigned grid_for(uint32_t items)
{
    unsigned b = blocks_for(items, 256);
    b = b < 1 ? 1 : (b > MAXBLK ? MAXBLK : b);
    return (b + 7u) & ~7u;
}

int fast_sweep(hipStream_t s, const Level &L, const float *zin, float *zout, const float *zc, int phase, int mode,
               double *part, int *nblk)
{
    const uint32_t quads = L.n / 4;
    const unsigned nb = grid_for(quads);
    if (nblk) *nblk = (int)nb;
    const bool dot = part != nullptr;
#define SW(M, D) k_rbgs4<M, D><<<nb, 256, 0, s>>>(zin, zout, L.r, L.tag, zc, L.nx, L.ny, L.nz, phase, part)
    if (mode == MODE_ZERO) { if (dot) SW(MODE_ZERO, true); else SW(MODE_ZERO, false); }
    else if (mode == MODE_PROLONG) { if (dot) SW(MODE_PROLONG, true); else SW(MODE_PROLONG, false); }
    else { if (dot) SW(MODE_PLAIN, true); else SW(MODE_PLAIN, false); }
#undef SW
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

int fast_restrict(hipStream_t s, const Level &F, const float *zf, const Level &C)
{
    const uint32_t items = (uint32_t)C.nx * C.ny * (F.nz / 4);
    k_restrict4<<<grid_for(items), 256, 0, s>>>(F.r, zf, F.tag, C.r, C.tag, F.nx, F.ny, F.nz);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

int fast_p_Ap(hipStream_t s, const Level &L0, const float *p, float *pout, float *Ap, const double *sc,
              double *part, int *nblk)
{
    const unsigned nb = grid_for(L0.n / 4);
    *nblk = (int)nb;
    k_p_Ap4<<<nb, 256, 0, s>>>(p, pout, Ap, L0.z, L0.tag, sc, L0.nx, L0.ny, L0.nz, part);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}
