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
#include "div_small.h"

namespace {

constexpr int TAIL_NT = 1024;
constexpr int TAIL_LDS = TAIL_MAX_CELLS + TAIL_MAX_CELLS / 7 + 16;   // every level of the tail

struct TailArgs {
    const float *r0;        // r of level T (reference layout)
    float *zg[TAIL_MAXL];   // every level's z after the cycle (reference layout)
    float *rg[TAIL_MAXL];   // every level's r (the restricted ones written back)
    const uint16_t *code[TAIL_MAXL];
    int nx[TAIL_MAXL], ny[TAIL_MAXL], nz[TAIL_MAXL];
    int ly[TAIL_MAXL], lz[TAIL_MAXL];   // log2 of ny, nz (every extent a power of two)
    uint32_t off[TAIL_MAXL];   // LDS offset of each level's cells
    int sweeps[TAIL_MAXL];     // pre/post red-black iterations (all levels but the last)
    float cs[TAIL_MAXL];       // scale of the restriction into each level (from index 1)
    int nl;                    // levels in the tail, the last one the bottom
    int nb;                    // bottom iterations per direction (bottom_smoothing / 2)
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// One half-sweep of a level held in LDS (extents powers of two: ny = 2^ly,
// nz = 2^lz).  A thread takes up to TAIL_CPT cells of the colour, all their
// LDS reads issued before the first update.  The quotient by divd is
// div_small_wave with y = RN(1/divd), bitwise the IEEE division of gs_update.
constexpr int TAIL_CPT = (int)(TAIL_MAX_CELLS / 2 / TAIL_NT);

__device__ __forceinline__ void tail_half(float *z, const float *r, const uint16_t *cd, const float *ytab, int nx,
                                          int ny, int nz, int ly, int lz, int phase)
{
    const uint32_t n = (uint32_t)nx << (ly + lz - 1), last = 2u * n - 1u;
    const uint32_t si = 1u << (ly + lz), sj = 1u << lz;
    uint32_t I[TAIL_CPT], c[TAIL_CPT];
    float v[TAIL_CPT][8];
#pragma unroll
    for (int e = 0; e < TAIL_CPT; ++e) {
        const uint32_t t = threadIdx.x + (uint32_t)e * TAIL_NT;
        if (t >= n) break;
        const uint32_t row = t >> (lz - 1);   // (i, j) row; the cell's k = 2 kk + (1 if (i + j) & 1 == phase)
        const uint32_t ij = (row >> ly) + (row & (uint32_t)(ny - 1));
        const uint32_t x = (t << 1) + ((ij & 1u) == (uint32_t)phase ? 1u : 0u);
        I[e] = x;
        c[e] = cd[x];
        // neighbours outside the level read some cell of it: their class is 0
        // (masked), so any in-bounds value will do
        v[e][0] = z[x];
        v[e][1] = z[min(x + si, last)];
        v[e][2] = z[x >= si ? x - si : 0u];
        v[e][3] = z[min(x + sj, last)];
        v[e][4] = z[x >= sj ? x - sj : 0u];
        v[e][5] = z[min(x + 1u, last)];
        v[e][6] = z[x >= 1u ? x - 1u : 0u];
        v[e][7] = r[x];
    }
#pragma unroll
    for (int e = 0; e < TAIL_CPT; ++e) {
        const uint32_t t = threadIdx.x + (uint32_t)e * TAIL_NT;
        if (t >= n) break;
        const float s = nsum_code_air0(c[e], v[e][0], v[e][1], v[e][2], v[e][3], v[e][4], v[e][5], v[e][6]);
        const uint32_t divd = (c[e] >> 12) & 7u;
        const float q = div_small_wave(v[e][7] + s, (float)divd, ytab[divd]);
        z[I[e]] = keep_if(v[e][0] + q, 0u - (uint32_t)(divd != 0u));
    }
}

__global__ void __launch_bounds__(TAIL_NT) k_tail(TailArgs a)
{
    __shared__ float zs[TAIL_LDS], rs[TAIL_LDS];
    __shared__ uint16_t cs[TAIL_LDS];
    __shared__ float ytab[8];
    if (threadIdx.x < 8) {   // RN(1/d), d = 0..7 (d = 0: the quotient is discarded)
        const float t[8] = {0.0f, 1.0f, 0.5f, 1.0f / 3.0f, 0.25f, 0.2f, 1.0f / 6.0f, 1.0f / 7.0f};
        ytab[threadIdx.x] = t[threadIdx.x];
    }
    // stage the codes of every level, r of level T (z of level T starts at 0),
    // every global load of a thread issued before its first LDS store
    {
        constexpr int NL = (int)(TAIL_MAX_CELLS / TAIL_NT);
        const uint32_t n0 = (uint32_t)a.nx[0] * a.ny[0] * a.nz[0];
        float rv[NL];
        uint32_t cv[NL];
#pragma unroll
        for (int e = 0; e < NL; ++e) {
            const uint32_t t = threadIdx.x + (uint32_t)e * TAIL_NT;
            if (t < n0) { rv[e] = a.r0[t]; cv[e] = a.code[0][t]; }
        }
        static_assert(TAIL_MAX_CELLS / 8 <= (uint32_t)TAIL_NT, "one coarse cell per thread");
        uint32_t cw[TAIL_MAXL];
#pragma unroll
        for (int l = 1; l < TAIL_MAXL; ++l) {   // the coarser levels: at most TAIL_MAX_CELLS / 8 cells
            const uint32_t t = threadIdx.x;
            if (l < a.nl && t < (uint32_t)a.nx[l] * a.ny[l] * a.nz[l]) cw[l] = a.code[l][t];
        }
#pragma unroll
        for (int e = 0; e < NL; ++e) {
            const uint32_t t = threadIdx.x + (uint32_t)e * TAIL_NT;
            if (t < n0) { rs[t] = rv[e]; zs[t] = 0.0f; cs[t] = (uint16_t)cv[e]; }
        }
#pragma unroll
        for (int l = 1; l < TAIL_MAXL; ++l) {
            const uint32_t t = threadIdx.x;
            if (l < a.nl && t < (uint32_t)a.nx[l] * a.ny[l] * a.nz[l]) cs[a.off[l] + t] = (uint16_t)cw[l];
        }
    }
    __syncthreads();
    // down: pre-smoothing, restriction into the next level (its z zeroed)
    for (int l = 0; l + 1 < a.nl; ++l) {
        float *z = zs + a.off[l];
        const float *r = rs + a.off[l];
        const uint16_t *cd = cs + a.off[l];
        const int nx = a.nx[l], ny = a.ny[l], nz = a.nz[l];
        for (int it = 0; it < a.sweeps[l]; ++it) {
            tail_half(z, r, cd, ytab, nx, ny, nz, a.ly[l], a.lz[l], 0);
            __syncthreads();
            tail_half(z, r, cd, ytab, nx, ny, nz, a.ly[l], a.lz[l], 1);
            __syncthreads();
        }
        const int cy = a.ny[l + 1], cz = a.nz[l + 1];
        const uint32_t nc = (uint32_t)a.nx[l + 1] * cy * cz;
        float *zc = zs + a.off[l + 1], *rc = rs + a.off[l + 1];
        const uint16_t *cdc = cs + a.off[l + 1];
        const float scale = a.cs[l + 1];
        const int lcy = a.ly[l + 1], lcz = a.lz[l + 1];
        const uint32_t si = 1u << (a.ly[l] + a.lz[l]), sj = 1u << a.lz[l];
        const uint32_t last = (uint32_t)nx * ny * nz - 1u;
        for (uint32_t t = threadIdx.x; t < nc; t += TAIL_NT) {
            const int K = (int)(t & (uint32_t)(cz - 1));
            const int J = (int)((t >> lcz) & (uint32_t)(cy - 1)), I = (int)(t >> (lcy + lcz));
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
                                nd = nsum_code_air0(cd0, z[c], z[min(c + si, last)], z[c >= si ? c - si : 0u],
                                                    z[min(c + sj, last)], z[c >= sj ? c - sj : 0u],
                                                    z[min(c + 1u, last)], z[c >= 1u ? c - 1u : 0u]);
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
            tail_half(z, r, cd, ytab, a.nx[l], a.ny[l], a.nz[l], a.ly[l], a.lz[l], first);
            __syncthreads();
            tail_half(z, r, cd, ytab, a.nx[l], a.ny[l], a.nz[l], a.ly[l], a.lz[l], first ^ 1);
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
        const int ly = a.ly[l], lz = a.lz[l];
        // With post-smoothing the non-FLUID cells keep +0 here (the sweeps
        // write +0 to them anyway and no sweep reads their value), which keeps
        // the invariant of nsum_code_air0; without it they take the sum as in
        // the reference.
        const bool all = a.sweeps[l] == 0;
        for (uint32_t t = threadIdx.x; t < n; t += TAIL_NT) {
            const int k = (int)(t & (uint32_t)(nz - 1));
            const int j = (int)((t >> lz) & (uint32_t)(ny - 1)), i = (int)(t >> (ly + lz));
            if (all || code_fluid(cd[t]))
                z[t] = z[t] + zc[((uint32_t)(i >> 1) << (ly + lz - 2)) + ((uint32_t)(j >> 1) << (lz - 1)) +
                                 (uint32_t)(k >> 1)];
        }
        __syncthreads();
        for (int it = 0; it < a.sweeps[l]; ++it) {
            tail_half(z, r, cd, ytab, nx, ny, nz, a.ly[l], a.lz[l], 1);
            __syncthreads();
            tail_half(z, r, cd, ytab, nx, ny, nz, a.ly[l], a.lz[l], 0);
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
    auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
    for (int l = lowest < 1 ? 1 : lowest; l < h->L; ++l)
        if (h->lv[l].n <= TAIL_MAX_CELLS) {
            if (h->L - l > TAIL_MAXL) return -1;
            for (int q = l; q < h->L; ++q)
                if (!pow2(h->lv[q].nx) || !pow2(h->lv[q].ny) || !pow2(h->lv[q].nz)) return -1;
            return l;
        }
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
        a.ly[q] = __builtin_ctz((unsigned)L.ny); a.lz[q] = __builtin_ctz((unsigned)L.nz);
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
