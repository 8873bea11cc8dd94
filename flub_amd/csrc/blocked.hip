// Temporally blocked red-black smoothing for the small multigrid levels.
//
// On a level of a few million cells or less each half-sweep launch costs a
// few microseconds of dispatch and memory latency for very little data, so a
// V-cycle that launches one kernel per half-sweep (mgpcg.py:296-349: 4*s*2^l
// of them on level l) is bound by launch latency there.  This kernel runs up
// to G consecutive half-sweeps in one launch.  Each workgroup owns a core
// tile of CI x CJ x CK cells, loads z on the core plus a halo of G cells
// into LDS, and applies the half-sweeps in LDS: step s (of n) updates the
// cells of its colour within distance n-1-s of the core (the 7-point
// dependency cone), so after n steps the core holds exactly the values the
// launch-per-half-sweep sequence produces; the halo cells are recomputed
// redundantly by the neighbouring workgroups.  Every cell update is the
// reference arithmetic (gs_update), so the results are bitwise those of the
// sequential half-sweeps.  The launch reads z from one buffer and writes the
// core to another (the neighbours' halos still read the old values).
//
// The cells a thread may update, with their r and stencil codes, stay in
// registers for the whole launch (pairs of cells along k: one of each
// colour); LDS holds only z.  ZERO starts from z = 0 without reading it;
// PROLONG starts from z + pfac * z_coarse[parent] (mgpcg.py:113-116).
#include "common.h"
#include "div_small.h"
#include "mgpcg_internal.h"
#include "smooth.h"
#include "stencil.h"

#include <algorithm>
#include <vector>

namespace {

struct TbArgs {
    const void *zin;      // T: z at the start (unused for ZERO)
    void *zout;           // T: z of the core after the steps
    const void *zc;       // T: coarse z (PROLONG)
    const void *r;        // T: right-hand side of the level
    const uint16_t *code;
    const uint32_t *order;  // updatable pairs, nearest to the core first
    double *part;         // DOT: per-workgroup partials of z.r over the core
    float pfac;
    int nx, ny, nz;
    int ti, tj, tk;       // tiles per axis
    int nsteps;           // 1..G (the configuration's halo width)
    uint32_t phases;      // bit s: phase of step s (updates (i+j+k)&1 == phase^1)
    int zmasked;          // zin is +0 on every non-FLUID cell (an earlier launch of the chain wrote it)
    int cnt[TB_GMAX];     // cnt[d]: pairs within distance d of the core
};

// The red-black update of a cell (mgpcg.py:72-86) with the LDS invariant of
// this kernel: non-FLUID cells hold +0 there (zeroed at load; every update of
// a non-FLUID cell writes +0).  An AIR neighbour then reads 0, the
// reference's x' = 0, and only SOLID/outside neighbours are masked to +0
// (the reference's (x' - x_I) * 0.0; a zero term of either sign leaves the
// sum unchanged, see stencil.h).  Bitwise equal to gs_update.
__device__ __forceinline__ float tb_update(uint32_t cd, float z, float r, float xip, float xim, float xjp, float xjm,
                                           float xkp, float xkm, float y)
{
    const uint32_t nzb = (cd | (cd >> 1)) & 0x555u;   // bit 2n: neighbour n is AIR or FLUID
    const float xs[6] = {xip, xim, xjp, xjm, xkp, xkm};
    float s = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t mp = (uint32_t)((int32_t)(nzb << (31 - 4 * a)) >> 31);
        const uint32_t mm = (uint32_t)((int32_t)(nzb << (29 - 4 * a)) >> 31);
        s = s + (keep_if(xs[2 * a] - z, mp) + keep_if(xs[2 * a + 1] - z, mm));
    }
    const float q = div_small_wave(r + s, (float)((cd >> 12) & 7u), y);
    // +0 unless divd != 0: a FLUID cell with no AIR/FLUID neighbour has divd 0
    // and is set to 0 as well (mgpcg.py:80-83), whatever a prolongation added
    return keep_if(z + q, 0u - (uint32_t)(((cd >> 12) & 7u) != 0u));
}

// Kernel layout of a pair entry in `order`: LDS index of its even cell (bits
// 0-15) and the pair's distance to the core (bits 16-19).  Register layout of
// an owned pair: LDS index (0-15), parity of the even cell (16), distance of
// the even / odd cell (17-20 / 21-24; 15 for a pair outside the domain).
template <typename T, int CI, int CJ, int CK, int NT, int G, int MODE, bool DOT>
__global__ void __launch_bounds__(NT) k_tb(TbArgs a)
{
    constexpr int BI = CI + 2 * G, BJ = CJ + 2 * G, BK = CK + 2 * G;
    constexpr int SJ = BK, SI = BJ * BK;
    constexpr int BP = BK / 2;                  // cell pairs per LDS row
    constexpr int NOWN = (BI - 2) * (BJ - 2) * BP;
    constexpr int NU = (NOWN + NT - 1) / NT;
    constexpr int GS = (NU + 1) / 2;            // slots per gather group
    static_assert(G % 4 == 0 && CK % 4 == 0 && BI * BJ * BK < 65536, "tile layout");
    __shared__ float zs[BI * BJ * BK];
    __shared__ float ytab[8];
    __shared__ double red[NT / 64];

    if (threadIdx.x < 8) {   // RN(1/d), d = 0..7 (0 for the non-FLUID d = 0)
        const float t[8] = {0.0f, 1.0f, 0.5f, 1.0f / 3.0f, 0.25f, 0.2f, 1.0f / 6.0f, 1.0f / 7.0f};
        ytab[threadIdx.x] = t[threadIdx.x];
    }

    // XCD-aware tile order: each XCD takes a contiguous run of tiles
    const uint32_t ntile = (uint32_t)a.ti * a.tj * a.tk;
    uint32_t b = blockIdx.x;
    if ((ntile & 7u) == 0) b = (b & 7u) * (ntile >> 3) + (b >> 3);
    const int bk = (int)(b % (uint32_t)a.tk);
    b /= (uint32_t)a.tk;
    const int bj = (int)(b % (uint32_t)a.tj), bi = (int)(b / (uint32_t)a.tj);
    const int oi = bi * CI - G, oj = bj * CJ - G, ok = bk * CK - G;   // LDS origin (global)

    // Loads first, LDS stores after: z and the stencil codes on the core and
    // a halo of G as quads along k (the box starts at a multiple of 4 in k;
    // cells outside the domain hold 0), then r, codes and positions of the
    // pairs this thread may update.
    constexpr int BQ = BK / 4, NQ = BI * BJ * BQ, NL = (NQ + NT - 1) / NT;
    float zq[NL][4], cq[NL][2];
    uint32_t fq[NL];
    if (MODE != SWEEP_ZERO) {
#pragma unroll
        for (int u = 0; u < NL; ++u) {
            const int q = threadIdx.x + u * NT;
            zq[u][0] = zq[u][1] = zq[u][2] = zq[u][3] = 0.0f;
            cq[u][0] = cq[u][1] = 0.0f;
            fq[u] = 0u;
            const int kq = q % BQ, row = q / BQ;
            const int gi = oi + row / BJ, gj = oj + row % BJ, gk = ok + 4 * kq;
            if (q < NQ && gi >= 0 && gi < a.nx && gj >= 0 && gj < a.ny && gk >= 0 && gk < a.nz) {
                const uint32_t gx = cidx(gi, gj, gk, a.ny, a.nz);
                ld4<T>(reinterpret_cast<const T *>(a.zin) + gx, zq[u]);
                if (a.zmasked) {
                    fq[u] = 15u;   // masking is the identity: skip the codes
                } else {
                    const uint2 c4 = *reinterpret_cast<const uint2 *>(a.code + gx);
                    fq[u] = ((c4.x >> 15) & 1u) | ((c4.x >> 30) & 2u) | (((c4.y >> 15) & 1u) << 2) |
                            ((c4.y >> 28) & 8u);
                }
                if (MODE == SWEEP_PROLONG)
                    ld2<T>(reinterpret_cast<const T *>(a.zc) + cidx(gi >> 1, gj >> 1, gk >> 1, a.ny >> 1, a.nz >> 1),
                           cq[u]);
            }
        }
    }
    float r0[NU], r1[NU];
    uint32_t cw[NU], pk[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int q = threadIdx.x + u * NT;
        r0[u] = 0.0f; r1[u] = 0.0f; cw[u] = 0u; pk[u] = 0x1fe0000u;   // distances 15: never updated
        if (q < NOWN) {
            const uint32_t x0 = a.order[q] & 0xffffu;
            const int li = (int)(x0 / SI), lj = (int)((x0 / BK) % BJ), lk = (int)(x0 % BK);
            const int gi = oi + li, gj = oj + lj, gk = ok + lk;
            if (gi >= 0 && gi < a.nx && gj >= 0 && gj < a.ny && gk >= 0 && gk < a.nz) {
                const uint32_t gx = cidx(gi, gj, gk, a.ny, a.nz);
                float v[2];
                ld2<T>(reinterpret_cast<const T *>(a.r) + gx, v);
                r0[u] = v[0]; r1[u] = v[1];
                cw[u] = *reinterpret_cast<const uint32_t *>(a.code + gx);
                auto dist = [](int l, int c) { return l < G ? G - l : (l >= G + c ? l - (G + c - 1) : 0); };
                const int dij = max(dist(li, CI), dist(lj, CJ));
                const uint32_t d0 = (uint32_t)max(dij, dist(lk, CK)), d1 = (uint32_t)max(dij, dist(lk + 1, CK));
                pk[u] = x0 | ((uint32_t)((gi + gj + gk) & 1) << 16) | (d0 << 17) | (d1 << 21);
            }
        }
    }
    if (MODE == SWEEP_ZERO) {
        for (int q = threadIdx.x; q < NQ; q += NT)
            *reinterpret_cast<float4 *>(&zs[4 * q]) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {
#pragma unroll
        for (int u = 0; u < NL; ++u) {
            const int q = threadIdx.x + u * NT;
            if (q >= NQ) continue;
            float *v = zq[u];
            if (MODE == SWEEP_PROLONG) {
                v[0] = v[0] + a.pfac * cq[u][0]; v[1] = v[1] + a.pfac * cq[u][0];
                v[2] = v[2] + a.pfac * cq[u][1]; v[3] = v[3] + a.pfac * cq[u][1];
            }
            // non-FLUID cells hold +0 (the invariant of tb_update)
            *reinterpret_cast<float4 *>(&zs[4 * q]) =
                make_float4(keep_if(v[0], 0u - (fq[u] & 1u)), keep_if(v[1], 0u - ((fq[u] >> 1) & 1u)),
                            keep_if(v[2], 0u - ((fq[u] >> 2) & 1u)), keep_if(v[3], 0u - ((fq[u] >> 3) & 1u)));
        }
    }
    __syncthreads();

    // A step updates one colour and reads only the other (and each cell's
    // own value, read by its owner), so the reads of a step may precede its
    // writes.  Slots are gathered in groups, unconditionally (an inactive
    // lane reads a core cell), and pinned in registers: otherwise the
    // compiler sinks the reads into branches and they run one after another.
    // Slots whose pairs all lie beyond the step's distance are skipped
    // uniformly (the pairs are ordered by distance).
    constexpr int XDUMMY = (G * BJ + G) * BK + G;
    for (int s = 0; s < a.nsteps; ++s) {
        const uint32_t phase = (a.phases >> s) & 1u;
        const uint32_t lim = (uint32_t)(a.nsteps - 1 - s);
        const int cnt = a.cnt[lim];
#pragma unroll
        for (int g0 = 0; g0 < NU; g0 += GS) {
            if (g0 * NT >= cnt) break;
            int xs[GS];
            float nb[GS][8];
#pragma unroll
            for (int v = 0; v < GS; ++v) {
                const int u = g0 + v;
                xs[v] = -1;
                if (u >= NU || u * NT >= cnt) continue;
                const uint32_t p = pk[u];
                const uint32_t sel = ((p >> 16) & 1u) ^ phase ^ 1u;   // the pair's cell of colour phase^1
                const uint32_t d = (p >> (17 + 4 * sel)) & 15u;
                const bool act = d <= lim;
                const int x = act ? (int)(p & 0xffffu) + (int)sel : XDUMMY;
                const uint32_t cd = (cw[u] >> (16 * sel)) & 0xffffu;
                xs[v] = act ? x : -1;
                nb[v][0] = zs[x]; nb[v][1] = zs[x + SI]; nb[v][2] = zs[x - SI]; nb[v][3] = zs[x + SJ];
                nb[v][4] = zs[x - SJ]; nb[v][5] = zs[x + 1]; nb[v][6] = zs[x - 1];
                nb[v][7] = ytab[(cd >> 12) & 7u];
            }
#pragma unroll
            for (int v = 0; v < GS; ++v) {
                if (g0 + v >= NU || (g0 + v) * NT >= cnt) continue;
#pragma unroll
                for (int e = 0; e < 8; ++e) asm volatile("" : "+v"(nb[v][e]));
            }
#pragma unroll
            for (int v = 0; v < GS; ++v) {
                const int u = g0 + v;
                if (u >= NU || xs[v] < 0) continue;
                const uint32_t sel = ((pk[u] >> 16) & 1u) ^ phase ^ 1u;
                const uint32_t cd = (cw[u] >> (16 * sel)) & 0xffffu;
                zs[xs[v]] = tb_update(cd, nb[v][0], sel ? r1[u] : r0[u], nb[v][1], nb[v][2], nb[v][3], nb[v][4],
                                      nb[v][5], nb[v][6], nb[v][7]);
            }
        }
        __syncthreads();
    }

    if (DOT) {   // z.r over the core: the pairs at distance 0 (whole pairs: the core starts at an even k)
        double acc = 0.0;
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            if (u * NT >= a.cnt[0]) break;
            const uint32_t p = pk[u];
            if (((p >> 17) & 0xffu) != 0u) continue;
            const int x = (int)(p & 0xffffu);
            acc += (double)zs[x] * (double)r0[u] + (double)zs[x + 1] * (double)r1[u];
        }
        const double tot = block_sum<NT>(acc, red);
        if (threadIdx.x == 0) a.part[blockIdx.x] = tot;
    }

    // the core (always inside the domain: the level is a multiple of the tile)
    for (int q = threadIdx.x; q < CI * CJ * (CK / 4); q += NT) {
        const int kq = q % (CK / 4), row = q / (CK / 4);
        const int cj = row % CJ, ci = row / CJ;
        const int x = ((G + ci) * BJ + (G + cj)) * BK + G + 4 * kq;
        const float4 w = *reinterpret_cast<const float4 *>(&zs[x]);
        const float v[4] = {w.x, w.y, w.z, w.w};
        st4<T>(reinterpret_cast<T *>(a.zout) + cidx(oi + G + ci, oj + G + cj, ok + G + 4 * kq, a.ny, a.nz), v);
    }
}

// Tile configurations: core CI x CJ x CK, threads, halo width G (half-sweeps
// per launch).  The G = 8 configurations (8 half-sweeps per launch on levels
// of at most 64^3 cells) are optional: measured on a 256^3 solve, their larger
// halos cost more than the launches they save.
enum { TB_CFG_16 = 0, TB_CFG_8x16 = 1, TB_CFG_8 = 2, TB_CFG_8x16_G8 = 3, TB_CFG_8_G8 = 4, TB_CFG_4x8 = 5, TB_NCFG = 6 };
struct TbGeom {
    int ci, cj, ck, nt, g;
};
const TbGeom kGeom[TB_NCFG] = {{16, 16, 16, 1024, 4}, {8, 8, 16, 1024, 4}, {8, 8, 8, 1024, 4},
                               {8, 8, 16, 512, 8}, {8, 8, 8, 512, 8}, {4, 4, 8, 1024, 4}};
#define TB_G8_MAX_CELLS (1u << 18)
// Levels of at most 32^3 cells take 4 x 4 x 8 cores: on a 32^3 level they
// give 256 workgroups (32 with 8 x 8 x 16 cores), so the launch is spread
// over every CU, at three times the halo work per core cell.
#define TB_SMALL_MAX_CELLS (1u << 15)

bool fits(const Level &L, int c) { return L.nx % kGeom[c].ci == 0 && L.ny % kGeom[c].cj == 0 && L.nz % kGeom[c].ck == 0; }

int tb_cfg(const Level &L, bool allow_g8)
{
    if (allow_g8 && L.n <= TB_G8_MAX_CELLS) {
        if (fits(L, TB_CFG_8x16_G8)) return TB_CFG_8x16_G8;
        if (fits(L, TB_CFG_8_G8)) return TB_CFG_8_G8;
    }
    if (L.n >= (1u << 21) && fits(L, TB_CFG_16)) return TB_CFG_16;
    if (L.n <= TB_SMALL_MAX_CELLS && fits(L, TB_CFG_4x8)) return TB_CFG_4x8;
    if (fits(L, TB_CFG_8x16)) return TB_CFG_8x16;
    if (fits(L, TB_CFG_8)) return TB_CFG_8;
    return -1;
}

// Order tables (host copy): the updatable pairs of a tile sorted by their
// distance to the core (stable), and the counts within each distance.
struct TbTable {
    std::vector<uint32_t> order;
    int cnt[TB_GMAX];
    size_t offset;   // in the concatenated device table
};

const std::vector<TbTable> &tb_tables()
{
    static const std::vector<TbTable> t = [] {
        std::vector<TbTable> v(TB_NCFG);
        size_t off = 0;
        for (int c = 0; c < TB_NCFG; ++c) {
            const TbGeom &g = kGeom[c];
            const int G = g.g, BI = g.ci + 2 * G, BJ = g.cj + 2 * G, BK = g.ck + 2 * G;
            auto dist = [G](int l, int n) { return l < G ? G - l : (l >= G + n ? l - (G + n - 1) : 0); };
            std::vector<std::pair<int, uint32_t>> e;
            for (int li = 1; li < BI - 1; ++li)
                for (int lj = 1; lj < BJ - 1; ++lj)
                    for (int lm = 0; lm < BK / 2; ++lm) {
                        const int dij = std::max(dist(li, g.ci), dist(lj, g.cj));
                        const int d = std::max(dij, std::min(dist(2 * lm, g.ck), dist(2 * lm + 1, g.ck)));
                        e.push_back({d, (uint32_t)((li * BJ + lj) * BK + 2 * lm)});
                    }
            std::stable_sort(e.begin(), e.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
            TbTable &t = v[c];
            for (const auto &x : e) t.order.push_back(x.second | ((uint32_t)x.first << 16));
            for (int d = 0; d < G; ++d)
                t.cnt[d] = (int)std::count_if(e.begin(), e.end(), [d](const auto &x) { return x.first <= d; });
            t.offset = off;
            off += t.order.size();
        }
        return v;
    }();
    return t;
}

unsigned tb_blocks(int cfg, const Level &L)
{
    const TbGeom &g = kGeom[cfg];
    return (unsigned)((L.nx / g.ci) * (L.ny / g.cj) * (L.nz / g.ck));
}

template <typename T, int MODE, bool DOT>
void tb_launch(hipStream_t s, int cfg, TbArgs a)
{
    const TbGeom &g = kGeom[cfg];
    a.ti = a.nx / g.ci; a.tj = a.ny / g.cj; a.tk = a.nz / g.ck;
    const unsigned nb = (unsigned)(a.ti * a.tj * a.tk);
    if (cfg == TB_CFG_16) k_tb<T, 16, 16, 16, 1024, 4, MODE, DOT><<<nb, 1024, 0, s>>>(a);
    else if (cfg == TB_CFG_8x16) k_tb<T, 8, 8, 16, 1024, 4, MODE, DOT><<<nb, 1024, 0, s>>>(a);
    else if (cfg == TB_CFG_8) k_tb<T, 8, 8, 8, 1024, 4, MODE, DOT><<<nb, 1024, 0, s>>>(a);
    else if (cfg == TB_CFG_4x8) k_tb<T, 4, 4, 8, 1024, 4, MODE, DOT><<<nb, 1024, 0, s>>>(a);
    else if (cfg == TB_CFG_8x16_G8) k_tb<T, 8, 8, 16, 512, 8, MODE, DOT><<<nb, 512, 0, s>>>(a);
    else k_tb<T, 8, 8, 8, 512, 8, MODE, DOT><<<nb, 512, 0, s>>>(a);
}

template <typename T>
void tb_dispatch(hipStream_t s, int cfg, int mode, const TbArgs &a)
{
    if constexpr (sizeof(T) == 4) {
        if (a.part) {
            if (mode == SWEEP_PROLONG) tb_launch<T, SWEEP_PROLONG, true>(s, cfg, a);
            else tb_launch<T, SWEEP_PLAIN, true>(s, cfg, a);
            return;
        }
    }
    if (mode == SWEEP_ZERO) tb_launch<T, SWEEP_ZERO, false>(s, cfg, a);
    else if (mode == SWEEP_PROLONG) tb_launch<T, SWEEP_PROLONG, false>(s, cfg, a);
    else tb_launch<T, SWEEP_PLAIN, false>(s, cfg, a);
}

}  // namespace

bool tb_supported(const Level &L) { return tb_cfg(L, true) >= 0; }

int tb_launches(const Level &L, int nsteps, bool allow_g8)
{
    const int cfg = tb_cfg(L, allow_g8);
    return cfg < 0 ? 0 : (nsteps + kGeom[cfg].g - 1) / kGeom[cfg].g;
}

int tb_tables_create(uint32_t **dev)
{
    *dev = nullptr;
    const std::vector<TbTable> &t = tb_tables();
    std::vector<uint32_t> all;
    for (const TbTable &x : t) all.insert(all.end(), x.order.begin(), x.order.end());
    FLUB_CHECK_HIP(hipMalloc((void **)dev, all.size() * sizeof(uint32_t)));
    hipError_t e = hipMemcpy(*dev, all.data(), all.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(*dev);
        *dev = nullptr;
        FLUB_CHECK_HIP(e);
    }
    return FLUB_OK;
}

int tb_chain(hipStream_t s, const uint32_t *tables, const Level &L, int prec, int mode0, const void *zin,
             const void *zc, float pfac, const void *r, const uint8_t *phases, int nsteps, void *dst, void *other,
             double *part, int *nblk, bool allow_g8)
{
    const int cfg = tb_cfg(L, allow_g8);
    FLUB_REQUIRE(tables, FLUB_ESTATE, "blocked smoothing: order tables missing");
    FLUB_REQUIRE(cfg >= 0, FLUB_ESHAPE, "blocked smoothing: level %dx%dx%d is not a multiple of a tile", L.nx, L.ny,
                 L.nz);
    FLUB_REQUIRE(nsteps >= 1 && dst && other && dst != other && r, FLUB_EINVAL, "blocked smoothing: bad arguments");
    const int G = cfg < 0 ? 1 : kGeom[cfg].g;
    const int m = (nsteps + G - 1) / G;
    void *w = (m & 1) ? dst : other;    // the last launch writes dst
    FLUB_REQUIRE(mode0 == SWEEP_ZERO || (zin && zin != w), FLUB_EINVAL,
                 "blocked smoothing: the first launch must not write its input");
    FLUB_REQUIRE(mode0 != SWEEP_PROLONG || zc, FLUB_EINVAL, "blocked smoothing: PROLONG needs the coarse z");
    FLUB_REQUIRE(!part || (prec == FLUB_PREC_FP32 && (m >= 2 || mode0 != SWEEP_ZERO) &&
                           tb_blocks(cfg, L) <= MAXPART),
                 FLUB_EINVAL, "blocked smoothing: the fused z.r needs FP32, a last launch that reads z and at most "
                 "%d tiles", MAXPART);
    const TbTable &tab = tb_tables()[cfg];
    const void *in = zin;
    for (int c = 0; c < m; ++c) {
        const int st = c * G, cnt = nsteps - st < G ? nsteps - st : G;
        TbArgs a{};
        a.zin = in; a.zout = w; a.zc = zc; a.r = r; a.code = L.code; a.pfac = pfac;
        a.order = tables + tab.offset;
        for (int d = 0; d < G; ++d) a.cnt[d] = tab.cnt[d];
        a.nx = L.nx; a.ny = L.ny; a.nz = L.nz;
        a.nsteps = cnt;
        a.zmasked = c > 0;   // the previous launch stored the whole level with +0 on non-FLUID cells
        uint32_t seen = 0;
        for (int q = 0; q < cnt; ++q) {
            a.phases |= (uint32_t)(phases[st + q] & 1u) << q;
            seen |= 1u << (phases[st + q] & 1u);
        }
        // every launch must sweep both colours: its non-FLUID core cells then
        // end at 0, as in the sequential half-sweeps (the kernel zeroes them)
        FLUB_REQUIRE(seen == 3u, FLUB_EINVAL, "blocked smoothing: each launch needs both phases");
        const int mode = c == 0 ? mode0 : SWEEP_PLAIN;
        if (c + 1 == m && part) {
            a.part = part;
            if (nblk) *nblk = (int)tb_blocks(cfg, L);
        }
        if (prec == FLUB_PREC_FP16) tb_dispatch<__half>(s, cfg, mode, a);
        else if (prec == FLUB_PREC_BF16) tb_dispatch<__hip_bfloat16>(s, cfg, mode, a);
        else tb_dispatch<float>(s, cfg, mode, a);
        FLUB_CHECK_LAUNCH();
        in = w;
        w = (w == dst) ? other : dst;
    }
    return FLUB_OK;
}
