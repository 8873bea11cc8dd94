// The coarsest levels of a V-cycle in one workgroup.
//
// Below a few thousand cells a level's half-sweep is a few hundred cell
// updates, and a V-cycle that launches kernels for it (mgpcg.py:296-349: the
// pre-smoothing, the restriction, the bottom sweeps, the prolongation and the
// post-smoothing of every level) waits on launch and memory latency several
// microseconds per launch.  This kernel runs the whole lower part of the
// cycle from a level T (at most TAIL_MAX_CELLS cells) down to the bottom and
// back in one launch of one workgroup: z, r and the stencil codes of every
// level live in LDS, and a workgroup barrier separates the half-sweeps.
//
// Every step is the reference arithmetic in the reference order, so the z it
// leaves on level T is bitwise that of the launch-by-launch cycle:
// * half-sweep `phase` updates the cells with (i + j + k) & 1 == phase ^ 1
//   (mgpcg.py:72-86) by gs_update (IEEE division by divd, +0 off FLUID);
// * the restriction is the gather of mgpcg.py:94-101: a FLUID parent sums
//   r - (-ndt(z)) of its children in (ii, jj, kk) order from +0, times the
//   level's coarse-correction scale; any other parent gets 0;
// * the prolongation adds the parent's z to every cell (mgpcg.py:113-116);
// * pre-smoothing (0, 1) pairs from z = 0, bottom (0, 1) then (1, 0) pairs,
//   post-smoothing (1, 0) pairs.
#include "common.h"
#include "mgpcg_internal.h"
#include "smooth.h"
#include "stencil.h"

namespace {

constexpr int TAIL_NT = 1024;
constexpr int TAIL_LDS = TAIL_MAX_CELLS + TAIL_MAX_CELLS / 7 + 16;   // every level of the tail

struct TailArgs {
    const float *r0;        // r of level T (reference layout)
    float *zg[TAIL_MAXL];   // every level's z after the cycle (reference layout)
    float *rg[TAIL_MAXL];   // every level's r (the restricted ones written back)
    const uint16_t *code[TAIL_MAXL];
    int nx[TAIL_MAXL], ny[TAIL_MAXL], nz[TAIL_MAXL];
    uint32_t off[TAIL_MAXL];   // LDS offset of each level's cells
    int sweeps[TAIL_MAXL];     // pre/post red-black iterations (all levels but the last)
    float cs[TAIL_MAXL];       // scale of the restriction into each level (from index 1)
    int nl;                    // levels in the tail, the last one the bottom
    int nb;                    // bottom iterations per direction (bottom_smoothing / 2)
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// One half-sweep of a level held in LDS.
__device__ void tail_half(float *z, const float *r, const uint16_t *cd, int nx, int ny, int nz, int phase)
{
    const int nzh = nz >> 1;
    const uint32_t n = (uint32_t)nx * ny * nzh;
    for (uint32_t t = threadIdx.x; t < n; t += TAIL_NT) {
        const int kk = (int)(t % (uint32_t)nzh);
        const uint32_t row = t / (uint32_t)nzh;
        const int j = (int)(row % (uint32_t)ny), i = (int)(row / (uint32_t)ny);
        const int k = 2 * kk + ((((i + j) & 1) == phase) ? 1 : 0);
        const uint32_t I = cidx(i, j, k, ny, nz);
        const float xip = z[cidx(clampi(i + 1, 0, nx - 1), j, k, ny, nz)];
        const float xim = z[cidx(clampi(i - 1, 0, nx - 1), j, k, ny, nz)];
        const float xjp = z[cidx(i, clampi(j + 1, 0, ny - 1), k, ny, nz)];
        const float xjm = z[cidx(i, clampi(j - 1, 0, ny - 1), k, ny, nz)];
        const float xkp = z[cidx(i, j, clampi(k + 1, 0, nz - 1), ny, nz)];
        const float xkm = z[cidx(i, j, clampi(k - 1, 0, nz - 1), ny, nz)];
        z[I] = gs_update<false, false>(cd[I], z[I], r[I], xip, xim, xjp, xjm, xkp, xkm);
    }
}

__global__ void __launch_bounds__(TAIL_NT) k_tail(TailArgs a)
{
    __shared__ float zs[TAIL_LDS], rs[TAIL_LDS];
    __shared__ uint16_t cs[TAIL_LDS];
    // stage the codes of every level, r of level T; z of level T starts at 0
    for (int l = 0; l < a.nl; ++l) {
        const uint32_t n = (uint32_t)a.nx[l] * a.ny[l] * a.nz[l];
        for (uint32_t t = threadIdx.x; t < n; t += TAIL_NT) cs[a.off[l] + t] = a.code[l][t];
    }
    {
        const uint32_t n = (uint32_t)a.nx[0] * a.ny[0] * a.nz[0];
        for (uint32_t t = threadIdx.x; t < n; t += TAIL_NT) { rs[t] = a.r0[t]; zs[t] = 0.0f; }
    }
    __syncthreads();
    // down: pre-smoothing, restriction into the next level (its z zeroed)
    for (int l = 0; l + 1 < a.nl; ++l) {
        float *z = zs + a.off[l];
        const float *r = rs + a.off[l];
        const uint16_t *cd = cs + a.off[l];
        const int nx = a.nx[l], ny = a.ny[l], nz = a.nz[l];
        for (int it = 0; it < a.sweeps[l]; ++it) {
            tail_half(z, r, cd, nx, ny, nz, 0);
            __syncthreads();
            tail_half(z, r, cd, nx, ny, nz, 1);
            __syncthreads();
        }
        const int cy = a.ny[l + 1], cz = a.nz[l + 1];
        const uint32_t nc = (uint32_t)a.nx[l + 1] * cy * cz;
        float *zc = zs + a.off[l + 1], *rc = rs + a.off[l + 1];
        const uint16_t *cdc = cs + a.off[l + 1];
        const float scale = a.cs[l + 1];
        for (uint32_t t = threadIdx.x; t < nc; t += TAIL_NT) {
            const int K = (int)(t % (uint32_t)cz);
            const uint32_t row = t / (uint32_t)cz;
            const int J = (int)(row % (uint32_t)cy), I = (int)(row / (uint32_t)cy);
            float s = 0.0f;
            if (code_fluid(cdc[t])) {
                for (int ii = 0; ii < 2; ++ii)
                    for (int jj = 0; jj < 2; ++jj)
                        for (int kk = 0; kk < 2; ++kk) {
                            const int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                            const uint32_t c = cidx(i, j, k, ny, nz);
                            const uint32_t cd0 = cd[c];
                            float nd = 0.0f;
                            if (code_fluid(cd0))
                                nd = nsum_code(cd0, z[c], z[cidx(clampi(i + 1, 0, nx - 1), j, k, ny, nz)],
                                               z[cidx(clampi(i - 1, 0, nx - 1), j, k, ny, nz)],
                                               z[cidx(i, clampi(j + 1, 0, ny - 1), k, ny, nz)],
                                               z[cidx(i, clampi(j - 1, 0, ny - 1), k, ny, nz)],
                                               z[cidx(i, j, clampi(k + 1, 0, nz - 1), ny, nz)],
                                               z[cidx(i, j, clampi(k - 1, 0, nz - 1), ny, nz)]);
                            s += r[c] - (-nd);
                        }
                s = s * scale;
            }
            rc[t] = s;
            zc[t] = 0.0f;
        }
        __syncthreads();
    }
    // bottom: (0, 1) x nb, then (1, 0) x nb
    {
        const int l = a.nl - 1;
        float *z = zs + a.off[l];
        const float *r = rs + a.off[l];
        const uint16_t *cd = cs + a.off[l];
        for (int it = 0; it < 2 * a.nb; ++it) {
            const int first = it < a.nb ? 0 : 1;
            tail_half(z, r, cd, a.nx[l], a.ny[l], a.nz[l], first);
            __syncthreads();
            tail_half(z, r, cd, a.nx[l], a.ny[l], a.nz[l], first ^ 1);
            __syncthreads();
        }
    }
    // up: prolongation, post-smoothing
    for (int l = a.nl - 2; l >= 0; --l) {
        float *z = zs + a.off[l];
        const float *r = rs + a.off[l];
        const uint16_t *cd = cs + a.off[l];
        const float *zc = zs + a.off[l + 1];
        const int nx = a.nx[l], ny = a.ny[l], nz = a.nz[l];
        const uint32_t n = (uint32_t)nx * ny * nz;
        for (uint32_t t = threadIdx.x; t < n; t += TAIL_NT) {
            const int k = (int)(t % (uint32_t)nz);
            const uint32_t row = t / (uint32_t)nz;
            const int j = (int)(row % (uint32_t)ny), i = (int)(row / (uint32_t)ny);
            z[t] = z[t] + zc[cidx(i >> 1, j >> 1, k >> 1, ny >> 1, nz >> 1)];
        }
        __syncthreads();
        for (int it = 0; it < a.sweeps[l]; ++it) {
            tail_half(z, r, cd, nx, ny, nz, 1);
            __syncthreads();
            tail_half(z, r, cd, nx, ny, nz, 0);
            __syncthreads();
        }
    }
    // every level's z, and the r restricted into the levels below T, as the
    // launch-by-launch cycle leaves them in the level buffers
    for (int l = 0; l < a.nl; ++l) {
        const uint32_t n = (uint32_t)a.nx[l] * a.ny[l] * a.nz[l];
        for (uint32_t t = threadIdx.x; t < n; t += TAIL_NT) {
            a.zg[l][t] = zs[a.off[l] + t];
            if (l > 0) a.rg[l][t] = rs[a.off[l] + t];
        }
    }
}

}  // namespace

int tail_first_level(const flub_mgpcg *h, int lowest)
{
    // the tail starts at the first level (>= lowest, >= 1) of at most
    // TAIL_MAX_CELLS cells, if the levels from there down fit its tables
    if (h->no_tail || h->L < 2) return -1;
    for (int l = lowest < 1 ? 1 : lowest; l < h->L; ++l)
        if (h->lv[l].n <= TAIL_MAX_CELLS) return h->L - l <= TAIL_MAXL ? l : -1;
    return -1;
}

int tail_cycle(hipStream_t s, flub_mgpcg *h, int first)
{
    FLUB_REQUIRE(first >= 1 && first < h->L && h->L - first <= TAIL_MAXL && h->lv[first].n <= TAIL_MAX_CELLS,
                 FLUB_EINVAL, "tail cycle: level %d does not fit", first);
    TailArgs a{};
    a.nl = h->L - first;
    a.nb = h->bottom / 2;
    a.r0 = h->lv[first].r;
    uint32_t off = 0;
    for (int q = 0; q < a.nl; ++q) {
        const Level &L = h->lv[first + q];
        a.zg[q] = L.z;
        a.rg[q] = L.r;
        a.code[q] = L.code;
        a.nx[q] = L.nx; a.ny[q] = L.ny; a.nz[q] = L.nz;
        a.off[q] = off;
        a.sweeps[q] = mg_sweeps(h, first + q);
        a.cs[q] = L.cs;
        off += (L.n + 3u) & ~3u;
    }
    FLUB_REQUIRE(off <= (uint32_t)TAIL_LDS, FLUB_EINVAL, "tail cycle: %u cells exceed the LDS tables", off);
    k_tail<<<1, TAIL_NT, 0, s>>>(a);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}
