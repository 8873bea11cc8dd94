// Fast tagged multigrid kernels for levels with nz % 4 == 0, in FP32 (the
// reference arithmetic, bitwise) and in FP16/BF16 storage (the mixed solver).
//
// Each thread owns a "quad": 4 consecutive cells along k of one (i, j) row,
// moved as one 16-byte (FP32) or 8-byte (FP16/BF16) vector.  The i/j
// neighbours are the quads of the adjacent rows; the k neighbours beyond the
// quad come from the adjacent lanes (lane +-1 holds quad +-1 of the same row)
// and are loaded only at wavefront edges.
//
// Neighbour classes come from the per-cell 16-bit stencil code described in
// stencil.h, so results stay bitwise those of the reference kernels and of
// the FP32 oracle.
//
// Fusions (each equal to the unfused launch sequence):
//  * ZERO: the first pre-smoothing half-sweep of a level starts from z = 0
//    (mgpcg.py:299,312): z is not read, all four cells are written.
//  * PROLONG: the first post-smoothing half-sweep reads z + c*z_coarse[parent]
//    (mgpcg.py:113-116,339; c = 1 in FP32, 8 for the mixed level scaling) at
//    every stencil point and writes the prolongated value to the cells it does
//    not update.  Out of place: its neighbours still read the unprolongated z.
//  * DOT: the last half-sweep of the cycle also accumulates z.r (FP64).
//  * FINAL (mixed): the last level-0 sweep writes the FP32 z of the CG and the
//    FP64 partials of z.r and z.q.
#include "common.h"
#include "mgpcg_internal.h"
#include "smooth.h"
#include "stencil.h"

namespace {

struct Idx {
    int nx, ny, nz, nq, lq, lny;
    __device__ __forceinline__ void decode(uint32_t t, int &i, int &j, int &k0) const
    {
        uint32_t row, kq;
        if (lq >= 0) { row = t >> lq; kq = t & ((1u << lq) - 1u); }
        else { row = t / (uint32_t)nq; kq = t - row * (uint32_t)nq; }
        if (lny >= 0) { i = (int)(row >> lny); j = (int)(row & ((1u << lny) - 1u)); }
        else { i = (int)(row / (uint32_t)ny); j = (int)(row - (uint32_t)i * ny); }
        k0 = (int)kq * 4;
    }
};

}  // namespace

struct SweepArgs {
    const void *zin;      // T: z read (unused for ZERO)
    void *zout;           // T: z written (OUT 0/1)
    const void *zc;       // T: coarse z (PROLONG)
    float pfac;           // prolongation factor (1 FP32, 8 mixed)
    const void *rsrc;     // T: level r, or the FP32 CG residual when SRC32
    void *rh_out;         // T: SRC32 stores the scaled r here
    const uint16_t *code;
    const double *sc;     // device scalars (SC_RSCALE for SRC32 / FINAL)
    const float *r32, *q32;
    float *z32;           // FINAL output
    double *part;
    Idx g;
    int phase;
};

enum { OUT_STORE = 0, OUT_DOT = 1, OUT_FINAL = 2 };

template <typename T, int MODE>
__device__ __forceinline__ void row4(const SweepArgs &a, int i, int j, int k0, float v[4])
{
    if (MODE == SWEEP_ZERO) { v[0] = v[1] = v[2] = v[3] = 0.0f; return; }
    ld4<T>(reinterpret_cast<const T *>(a.zin) + cidx(i, j, k0, a.g.ny, a.g.nz), v);
    if (MODE == SWEEP_PROLONG) {
        float c[2];
        ld2<T>(reinterpret_cast<const T *>(a.zc) + cidx(i >> 1, j >> 1, k0 >> 1, a.g.ny >> 1, a.g.nz >> 1), c);
        v[0] = v[0] + a.pfac * c[0]; v[1] = v[1] + a.pfac * c[0];
        v[2] = v[2] + a.pfac * c[1]; v[3] = v[3] + a.pfac * c[1];
    }
}

template <typename T, int MODE>
__device__ __forceinline__ float one(const SweepArgs &a, int i, int j, int k)
{
    if (MODE == SWEEP_ZERO) return 0.0f;
    float v = ldv<T>(reinterpret_cast<const T *>(a.zin)[cidx(i, j, k, a.g.ny, a.g.nz)]);
    if (MODE == SWEEP_PROLONG)
        v = v + a.pfac * ldv<T>(reinterpret_cast<const T *>(a.zc)[cidx(i >> 1, j >> 1, k >> 1, a.g.ny >> 1, a.g.nz >> 1)]);
    return v;
}

// One red-black half-sweep (phase p updates (i+j+k)&1 == p^1, mgpcg.py:72-86).
template <typename T, int MODE, bool SRC32, int OUT>
__global__ void __launch_bounds__(256) k_sweep(SweepArgs a)
{
    __shared__ double lds[4];
    const Idx g = a.g;
    const uint32_t total = (uint32_t)g.nx * g.ny * g.nq;
    const int lane = threadIdx.x & 63;
    const float s0 = (SRC32 || OUT == OUT_FINAL) ? (float)a.sc[SC_RSCALE] : 1.0f;
    double acc1 = 0.0, acc2 = 0.0;
    const Chunk ch = xcd_chunk(total);
    for (uint32_t t = ch.first; t < ch.end; t += ch.stride) {
        int i, j, k0;
        g.decode(t, i, j, k0);
        const uint32_t base = cidx(i, j, k0, g.ny, g.nz);
        float c[4], xp[4], xm[4], yp[4], ym[4];
        // neighbour rows clamped to the domain (a wall neighbour reads the cell's
        // own row, as the reference's clamped sample does; its class is 0):
        // no branches, so all loads issue at once
        row4<T, MODE>(a, i, j, k0, c);
        row4<T, MODE>(a, min(i + 1, g.nx - 1), j, k0, xp);
        row4<T, MODE>(a, max(i - 1, 0), j, k0, xm);
        row4<T, MODE>(a, i, min(j + 1, g.ny - 1), k0, yp);
        row4<T, MODE>(a, i, max(j - 1, 0), k0, ym);
        // k neighbours from the adjacent lanes (same row), loaded at wavefront edges
        const float left = __shfl_up(c[3], 1, 64), right = __shfl_down(c[0], 1, 64);
        float zkm = c[0], zk4 = c[3];
        if (k0 > 0) zkm = lane > 0 ? left : one<T, MODE>(a, i, j, k0 - 1);
        if (k0 + 4 < g.nz) zk4 = (lane < 63 && t + 1 < ch.end) ? right : one<T, MODE>(a, i, j, k0 + 4);
        uint32_t cd[4];
        codes4(a.code, base, cd);
        float rv[4];
        if (SRC32) {
            const float4 r4 = *reinterpret_cast<const float4 *>(a.r32 + base);
            rv[0] = r4.x * s0; rv[1] = r4.y * s0; rv[2] = r4.z * s0; rv[3] = r4.w * s0;
            st4<T>(reinterpret_cast<T *>(a.rh_out) + base, rv);
            if constexpr (sizeof(T) != 4) {     // continue with the stored (rounded) values
                rv[0] = ldv<T>(stv<T>(rv[0])); rv[1] = ldv<T>(stv<T>(rv[1]));
                rv[2] = ldv<T>(stv<T>(rv[2])); rv[3] = ldv<T>(stv<T>(rv[3]));
            }
        } else {
            ld4<T>(reinterpret_cast<const T *>(a.rsrc) + base, rv);
        }
        float out[4];
        const int upd = ((i + j + k0) & 1) ^ a.phase ^ 1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            out[q] = c[q];
            if ((q & 1) != upd) continue;
            const float xkp = q < 3 ? c[q + 1] : zk4;
            const float xkm = q > 0 ? c[q - 1] : zkm;
            out[q] = gs_update(cd[q], c[q], rv[q], xp[q], xm[q], yp[q], ym[q], xkp, xkm);
        }
        if (OUT == OUT_FINAL) {
            const float inv = 1.0f / s0;
            const float4 r4 = *reinterpret_cast<const float4 *>(a.r32 + base);
            const float4 q4 = *reinterpret_cast<const float4 *>(a.q32 + base);
            const float o0 = out[0] * inv, o1 = out[1] * inv, o2 = out[2] * inv, o3 = out[3] * inv;
            *reinterpret_cast<float4 *>(a.z32 + base) = make_float4(o0, o1, o2, o3);
            acc1 += (double)o0 * r4.x + (double)o1 * r4.y + (double)o2 * r4.z + (double)o3 * r4.w;
            acc2 += (double)o0 * q4.x + (double)o1 * q4.y + (double)o2 * q4.z + (double)o3 * q4.w;
        } else {
            st4<T>(reinterpret_cast<T *>(a.zout) + base, out);
            if (OUT == OUT_DOT) {
#pragma unroll
                for (int q = 0; q < 4; ++q) acc1 += (double)out[q] * (double)rv[q];
            }
        }
    }
    if (OUT != OUT_STORE) {
        const double s1 = block_sum<256>(acc1, lds);
        if (OUT == OUT_FINAL) {
            const double s2 = block_sum<256>(acc2, lds);
            if (threadIdx.x == 0) a.part[gridDim.x + blockIdx.x] = s2;
        }
        if (threadIdx.x == 0) a.part[blockIdx.x] = s1;
    }
}

// Gather restriction (mgpcg.py:94-101), two coarse cells per thread: the fine
// quad of the four child rows (2I+a, 2J+b).  Residual r - (-ndt(z)) per fine
// cell; a FLUID parent sums its 8 children in the serial launch order
// (ii, jj, kk), others get 0.  The coarse value is stored times cscale
// (the coarse level's cs in FP32; cs/8 for the mixed level scaling).
template <typename T>
__global__ void __launch_bounds__(256) k_restrict(const T *__restrict__ r, const T *__restrict__ z,
                                                  const uint16_t *__restrict__ code, T *__restrict__ rc,
                                                  const uint8_t *__restrict__ tagc, Idx g, float cscale)
{
    // g describes the coarse grid in units of coarse pairs (nq = cz/2 = fine quads per row)
    const int nx = g.nx * 2, ny = g.ny * 2, nz = g.nz * 2;
    const int cy = g.ny, cz = g.nz;
    const uint32_t total = (uint32_t)g.nx * g.ny * g.nq;
    const int lane = threadIdx.x & 63;
    const Chunk ch = xcd_chunk(total);
    for (uint32_t t = ch.first; t < ch.end; t += ch.stride) {
        int I, J, kc;
        g.decode(t, I, J, kc);   // kc = 4 * pair index = fine k0
        const int k0 = kc;
        float res[2][2][4];
#pragma unroll
        for (int aa = 0; aa < 2; ++aa)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int i = 2 * I + aa, j = 2 * J + b;
                const uint32_t base = cidx(i, j, k0, ny, nz);
                float c[4], xp[4], xm[4], yp[4], ym[4];
                ld4<T>(z + base, c);   // neighbour rows clamped to the domain (walls: class 0)
                ld4<T>(z + cidx(min(i + 1, nx - 1), j, k0, ny, nz), xp);
                ld4<T>(z + cidx(max(i - 1, 0), j, k0, ny, nz), xm);
                ld4<T>(z + cidx(i, min(j + 1, ny - 1), k0, ny, nz), yp);
                ld4<T>(z + cidx(i, max(j - 1, 0), k0, ny, nz), ym);
                const float left = __shfl_up(c[3], 1, 64), right = __shfl_down(c[0], 1, 64);
                float zkm = c[0], zk4 = c[3];
                if (k0 > 0) zkm = lane > 0 ? left : ldv<T>(z[base - 1]);
                if (k0 + 4 < nz) zk4 = (lane < 63 && t + 1 < ch.end) ? right : ldv<T>(z[base + 4]);
                uint32_t cd[4];
                codes4(code, base, cd);
                float rv[4];
                ld4<T>(r + base, rv);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float xkp = q < 3 ? c[q + 1] : zk4;
                    const float xkm = q > 0 ? c[q - 1] : zkm;
                    float nd = nsum_code(cd[q], c[q], xp[q], xm[q], yp[q], ym[q], xkp, xkm);
                    if (!code_fluid(cd[q])) nd = 0.0f;
                    res[aa][b][q] = rv[q] - (-nd);
                }
            }
        const uint32_t cb = cidx(I, J, k0 >> 1, cy, cz);
        const uint32_t tc2 = *reinterpret_cast<const uint16_t *>(tagc + cb);
        float o[2];
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
            float s = 0.0f;
#pragma unroll
            for (int aa = 0; aa < 2; ++aa)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int kk = 0; kk < 2; ++kk) s += res[aa][b][2 * cc + kk];
            const int tp = cc ? (int)(tc2 >> 8) : (int)(tc2 & 0xff);
            o[cc] = tp == 1 ? s * cscale : 0.0f;
        }
        if constexpr (sizeof(T) == 4) {
            *reinterpret_cast<float2 *>(rc + cb) = make_float2(o[0], o[1]);
        } else {
            rc[cb] = stv<T>(o[0]);
            rc[cb + 1] = stv<T>(o[1]);
        }
    }
}

// p_new = z + beta*p_old at every stencil point (mgpcg.py:229-232), then
// Ap = ndt(p_new) (mgpcg.py:36-39) and the FP64 partial of p_new.Ap; p_new is
// stored to pout (not p: the neighbours still read p_old).
__global__ void __launch_bounds__(256) k_p_Ap(const float *__restrict__ p, float *__restrict__ pout,
                                              float *__restrict__ Ap, const float *__restrict__ z,
                                              const uint16_t *__restrict__ code, const double *__restrict__ sc,
                                              Idx g, double *__restrict__ part)
{
    __shared__ double lds[4];
    const float beta = (float)sc[SC_BETA];
    const uint32_t total = (uint32_t)g.nx * g.ny * g.nq;
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
    auto pq = [&](int i, int j, int k0, float v[4]) {
        const uint32_t b = cidx(i, j, k0, g.ny, g.nz);
        const float4 a = *reinterpret_cast<const float4 *>(p + b);
        const float4 zz = *reinterpret_cast<const float4 *>(z + b);
        v[0] = fmaf(beta, a.x, zz.x); v[1] = fmaf(beta, a.y, zz.y);
        v[2] = fmaf(beta, a.z, zz.z); v[3] = fmaf(beta, a.w, zz.w);
    };
    const Chunk ch = xcd_chunk(total);
    for (uint32_t t = ch.first; t < ch.end; t += ch.stride) {
        int i, j, k0;
        g.decode(t, i, j, k0);
        const uint32_t base = cidx(i, j, k0, g.ny, g.nz);
        float c[4], xp[4], xm[4], yp[4], ym[4];
        pq(i, j, k0, c);   // neighbour rows clamped to the domain (walls: class 0)
        pq(min(i + 1, g.nx - 1), j, k0, xp);
        pq(max(i - 1, 0), j, k0, xm);
        pq(i, min(j + 1, g.ny - 1), k0, yp);
        pq(i, max(j - 1, 0), k0, ym);
        const float left = __shfl_up(c[3], 1, 64), right = __shfl_down(c[0], 1, 64);
        float zkm = c[0], zk4 = c[3];
        if (k0 > 0) zkm = lane > 0 ? left : fmaf(beta, p[base - 1], z[base - 1]);
        if (k0 + 4 < g.nz) zk4 = (lane < 63 && t + 1 < ch.end) ? right : fmaf(beta, p[base + 4], z[base + 4]);
        uint32_t cd[4];
        codes4(code, base, cd);
        float out[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float xkp = q < 3 ? c[q + 1] : zk4;
            const float xkm = q > 0 ? c[q - 1] : zkm;
            const float s = nsum_code(cd[q], c[q], xp[q], xm[q], yp[q], ym[q], xkp, xkm);
            out[q] = code_fluid(cd[q]) ? s : 0.0f;
            acc += (double)c[q] * (double)out[q];
        }
        *reinterpret_cast<float4 *>(pout + base) = make_float4(c[0], c[1], c[2], c[3]);
        *reinterpret_cast<float4 *>(Ap + base) = make_float4(out[0], out[1], out[2], out[3]);
    }
    const double s = block_sum<256>(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Stencil codes of a level from its one-byte tags.
__global__ void k_make_codes(const uint8_t *__restrict__ tag, uint16_t *__restrict__ code, int nx, int ny, int nz)
{
    const uint32_t n = (uint32_t)nx * ny * nz;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = t % nz, j = (t / nz) % ny, i = t / ((uint32_t)nz * ny);
    const int self = tag[t];
    uint32_t cd = 0, divd = 0;
    const int ni[6] = {i + 1, i - 1, i, i, i, i}, nj[6] = {j, j, j + 1, j - 1, j, j}, nk[6] = {k, k, k, k, k + 1, k - 1};
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        uint32_t cls = 0;
        if (ni[d] >= 0 && ni[d] < nx && nj[d] >= 0 && nj[d] < ny && nk[d] >= 0 && nk[d] < nz) {
            const int tn = tag[cidx(ni[d], nj[d], nk[d], ny, nz)];
            cls = tn == 0 ? 1u : (tn == 1 ? 2u : 0u);
            divd += tn < 2;
        }
        cd |= cls << (2 * d);
    }
    if (self == 1) cd |= (divd << 12) | 0x8000u;
    code[t] = (uint16_t)cd;
}

// k_make_codes for 4 consecutive cells along k per thread (nz % 4 == 0): the
// tag rows of the cell and its i/j neighbours as 4-byte loads, the k-neighbours
// beyond the vector as single bytes, the 4 codes as one 8-byte store.
__global__ void __launch_bounds__(256) k_make_codes4(const uint8_t *__restrict__ tag, uint16_t *__restrict__ code,
                                                     int nx, int ny, int nz)
{
    const uint32_t nq = (uint32_t)nz >> 2;
    const uint32_t tq = blockIdx.x * 256u + threadIdx.x;
    if (tq >= (uint32_t)nx * ny * nq) return;
    const uint32_t row = tq / nq, q = tq - row * nq;
    const int i = (int)(row / (uint32_t)ny), j = (int)(row - (uint32_t)i * ny), k0 = 4 * (int)q;
    const uint32_t t = row * (uint32_t)nz + k0, plane = (uint32_t)ny * nz;
    auto row4 = [&](uint32_t b, int v[4]) {
        const uchar4 a = *reinterpret_cast<const uchar4 *>(tag + b);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    };
    // -1: outside the domain (class 0, not counted in divd)
    int c[4], xp[4] = {-1, -1, -1, -1}, xm[4] = {-1, -1, -1, -1}, yp[4] = {-1, -1, -1, -1}, ym[4] = {-1, -1, -1, -1};
    row4(t, c);
    if (i + 1 < nx) row4(t + plane, xp);
    if (i > 0) row4(t - plane, xm);
    if (j + 1 < ny) row4(t + nz, yp);
    if (j > 0) row4(t - nz, ym);
    const int km0 = k0 > 0 ? (int)tag[t - 1] : -1, kp3 = k0 + 4 < nz ? (int)tag[t + 4] : -1;
    uint32_t out[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int nb[6] = {xp[e], xm[e], yp[e], ym[e], e < 3 ? c[e + 1] : kp3, e > 0 ? c[e - 1] : km0};
        uint32_t cd = 0, divd = 0;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const int tn = nb[d];
            const uint32_t cls = tn == 0 ? 1u : (tn == 1 ? 2u : 0u);
            divd += (tn >= 0 && tn < 2) ? 1u : 0u;
            cd |= cls << (2 * d);
        }
        if (c[e] == 1) cd |= (divd << 12) | 0x8000u;
        out[e] = cd;
    }
    *reinterpret_cast<uint2 *>(code + t) = make_uint2(out[0] | (out[1] << 16), out[2] | (out[3] << 16));
}

// ------------------------------------------------------------------ host
static int ilog2_exact(int v)
{
    if (v <= 0 || (v & (v - 1))) return -1;
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

static Idx make_idx(int nx, int ny, int nz)
{
    Idx g;
    g.nx = nx; g.ny = ny; g.nz = nz; g.nq = nz / 4;
    g.lq = ilog2_exact(g.nq);
    g.lny = ilog2_exact(ny);
    return g;
}

// Workgroup count: one item per thread up to MAXBLK workgroups, rounded up to a
// multiple of 8 (xcd_chunk's remap needs it).
static unsigned grid_for(uint32_t items)
{
    unsigned b = blocks_for(items, 256);
    b = b < 1 ? 1 : (b > MAXBLK ? MAXBLK : b);
    return (b + 7u) & ~7u;
}

int make_codes(hipStream_t s, const Level &L)
{
    if (L.nz % 4 == 0)   // the level's buffers come from hipMalloc: aligned
        k_make_codes4<<<blocks_for(L.n / 4, 256), 256, 0, s>>>(L.tag, L.code, L.nx, L.ny, L.nz);
    else
        k_make_codes<<<blocks_for(L.n, 256), 256, 0, s>>>(L.tag, L.code, L.nx, L.ny, L.nz);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

template <typename T, int MODE, bool SRC32, int OUT>
static void launch_k(hipStream_t s, unsigned nb, const SweepArgs &a)
{
    k_sweep<T, MODE, SRC32, OUT><<<nb, 256, 0, s>>>(a);
}

template <typename T>
static int sweep_t(hipStream_t s, const Level &L, const SweepInfo &si, int *nblk)
{
    SweepArgs a{};
    a.zin = si.zin; a.zout = si.zout; a.zc = si.zc; a.pfac = si.pfac;
    a.rsrc = si.rsrc; a.rh_out = si.rh_out; a.code = L.code; a.sc = si.sc;
    a.r32 = si.r32; a.q32 = si.q32; a.z32 = si.z32; a.part = si.part;
    a.g = make_idx(L.nx, L.ny, L.nz);
    a.phase = si.phase;
    const unsigned nb = grid_for(L.n / 4);
    if (nblk) *nblk = (int)nb;
    const int m = si.mode, o = si.out;
    if (m == SWEEP_ZERO && si.src32) launch_k<T, SWEEP_ZERO, true, OUT_STORE>(s, nb, a);
    else if (m == SWEEP_ZERO) launch_k<T, SWEEP_ZERO, false, OUT_STORE>(s, nb, a);
    else if (m == SWEEP_PROLONG) launch_k<T, SWEEP_PROLONG, false, OUT_STORE>(s, nb, a);
    else if (o == OUT_DOT) launch_k<T, SWEEP_PLAIN, false, OUT_DOT>(s, nb, a);
    else if (o == OUT_FINAL) launch_k<T, SWEEP_PLAIN, false, OUT_FINAL>(s, nb, a);
    else launch_k<T, SWEEP_PLAIN, false, OUT_STORE>(s, nb, a);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

int sweep(hipStream_t s, const Level &L, int prec, const SweepInfo &si, int *nblk)
{
    if (prec == FLUB_PREC_FP16) return sweep_t<__half>(s, L, si, nblk);
    if (prec == FLUB_PREC_BF16) return sweep_t<__hip_bfloat16>(s, L, si, nblk);
    return sweep_t<float>(s, L, si, nblk);
}

int restrict_level(hipStream_t s, const Level &F, const void *rf, const void *zf, const Level &C, void *rc, int prec)
{
    const Idx g = make_idx(C.nx, C.ny, C.nz);   // nq = cz/2 = fine quads per row
    Idx gc = g;
    gc.nq = F.nz / 4;
    gc.lq = ilog2_exact(gc.nq);
    const uint32_t items = (uint32_t)C.nx * C.ny * gc.nq;
    const unsigned nb = grid_for(items);
    if (prec == FLUB_PREC_FP16)
        k_restrict<__half><<<nb, 256, 0, s>>>((const __half *)rf, (const __half *)zf, F.code, (__half *)rc, C.tag, gc, 0.125f * C.cs);
    else if (prec == FLUB_PREC_BF16)
        k_restrict<__hip_bfloat16><<<nb, 256, 0, s>>>((const __hip_bfloat16 *)rf, (const __hip_bfloat16 *)zf, F.code,
                                                      (__hip_bfloat16 *)rc, C.tag, gc, 0.125f * C.cs);
    else
        k_restrict<float><<<nb, 256, 0, s>>>((const float *)rf, (const float *)zf, F.code, (float *)rc, C.tag, gc, C.cs);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

int fast_sweep(hipStream_t s, const Level &L, const float *zin, float *zout, const float *zc, int phase, int mode,
               double *part, int *nblk)
{
    SweepInfo si{};
    si.zin = zin; si.zout = zout; si.zc = zc; si.pfac = 1.0f; si.rsrc = L.r;
    si.phase = phase; si.mode = mode; si.out = part ? OUT_DOT : OUT_STORE; si.part = part;
    return sweep(s, L, FLUB_PREC_FP32, si, nblk);
}

int fast_restrict(hipStream_t s, const Level &F, const float *zf, const Level &C)
{
    return restrict_level(s, F, F.r, zf, C, C.r, FLUB_PREC_FP32);
}

int fast_p_Ap(hipStream_t s, const Level &L0, const float *p, float *pout, float *Ap, const double *sc,
              double *part, int *nblk)
{
    const unsigned nb = grid_for(L0.n / 4);
    *nblk = (int)nb;
    k_p_Ap<<<nb, 256, 0, s>>>(p, pout, Ap, L0.z, L0.code, sc, make_idx(L0.nx, L0.ny, L0.nz), part);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}
