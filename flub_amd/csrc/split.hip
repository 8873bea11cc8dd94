// Colour-split storage of level 0 for the fused FP32 solve.
//
// A red-black half-sweep (mgpcg.py:72-86) updates the cells of one colour
// from the other colour's values.  In the reference's layout (k fastest)
// every cache line holds both colours, so a half-sweep streams all of z and
// r although it writes half of z.  Inside the solve, level 0 is therefore
// stored split by colour: an array of N values holds colour c = (i+j+k)&1 in
// its half c, cell (i, j, k) at index ((i*ny + j) * nz/2 + (k >> 1)) of that
// half.  Cell k of row (i, j) with colour c sits at k = 2m + o, o =
// (c - i - j) & 1; its k-neighbours (the other colour) are at m + o and
// m + o - 1 of the same row of the other half, its i/j neighbours at index m
// of the neighbouring rows of the other half.  A half-sweep then reads the
// updated colour's z, r and codes and the other colour's z: 9 bytes per cell
// of the level instead of the 14 its cache lines move in the reference
// layout.  The coarse levels (1..) keep the reference layout; the coarse
// index of a fine cell (i, j, k) is (i/2, j/2, k/2) = (i/2, j/2, m).
//
// The per-cell arithmetic and its order are those of the reference-layout
// kernels (gs_update, the residual of the restriction, p = z + beta*p and
// ndt(p)), so the results are bitwise the same; only the FP64 reduction
// order of the dot products differs.
//
// Threads own vectors of consecutive cells of one half-row (4, or 8 in the
// 16-bit pre-smoothing sweeps so that every load is 16 bytes); the k-neighbour beyond the
// vector comes from the adjacent lane (which holds the adjacent vector of the
// same half-row) or, at a wavefront edge, from memory.
#include "common.h"
#include "mgpcg_internal.h"
#include "smooth.h"
#include "stencil.h"

namespace {

// Work item t -> (i, j, m0): the vector index q is fastest, then the rows.
// With j-blocking (ljb >= 0) the rows run in blocks of JB = 2^ljb rows along
// j, i next, then the block: consecutive items sweep an i-plane of JB rows,
// so the rows i-1 and i+1 a work item reads were fetched at most two such
// planes earlier (a few hundred KB) and are still in L2 on large grids.
struct SGeo {
    int nx, ny, nh, nq;   // nh = nz/2 cells per half-row, nq = nh/cpt thread vectors per half-row
    int cpt;              // cells per thread (4, or 8 for 16-bit storage)
    int lq, lny, lnx;     // log2 of nq, ny and nx (-1: not a power of two)
    int ljb;              // log2 of the rows per j-block (-1: no blocking)
    uint32_t half;        // N/2
    __device__ __forceinline__ void decode(uint32_t t, int &i, int &j, int &m0) const
    {
        uint32_t row, q;
        if (lq >= 0) { row = t >> lq; q = t & ((1u << lq) - 1u); }
        else { row = t / (uint32_t)nq; q = t - row * (uint32_t)nq; }
        if (ljb >= 0) {
            const uint32_t jj = row & ((1u << ljb) - 1u), rest = row >> ljb;   // rest = block * nx + i
            uint32_t blk, ii;
            if (lnx >= 0) { blk = rest >> lnx; ii = rest & ((1u << lnx) - 1u); }
            else { blk = rest / (uint32_t)nx; ii = rest - blk * (uint32_t)nx; }
            i = (int)ii;
            j = (int)((blk << ljb) + jj);
        } else if (lny >= 0) {
            i = (int)(row >> lny); j = (int)(row & ((1u << lny) - 1u));
        } else {
            i = (int)(row / (uint32_t)ny); j = (int)(row - (uint32_t)i * ny);
        }
        m0 = cpt * (int)q;
    }
    __device__ __forceinline__ uint32_t row(int i, int j) const { return ((uint32_t)i * ny + j) * (uint32_t)nh; }
};

__device__ __forceinline__ void f4(const float *p, float v[4])
{
    const float4 a = *reinterpret_cast<const float4 *>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void st_f4(float *p, const float v[4])
{
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
// four consecutive values (16-byte aligned FP32 or 32-byte aligned FP64), widened to FP64
__device__ __forceinline__ void d4(const float *p, double v[4])
{
    const float4 a = *reinterpret_cast<const float4 *>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void d4(const double *p, double v[4])
{
    const double2 a = reinterpret_cast<const double2 *>(p)[0], b = reinterpret_cast<const double2 *>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}

struct SSweepArgs {
    const void *zin;       // TZI split z (PLAIN, PROLONG)
    void *zout;            // TZO split z written (STORE / DOT)
    const float *zc;       // PROLONG: z of level 1 (reference layout, FP32)
    void *r;               // TR split r of the V-cycle (SRC32: written)
    const uint16_t *code;  // split stencil codes
    const float *r32;      // SRC32: the FP32 CG residual (split)
    const void *rold;      // FINAL: TR split r of the previous V-cycle (r_old times its scale)
    const double *zinv_old;  // FINAL: 1/scale of rold
    double *zinv;          // FINAL: receives 1/scale (the z buffer holds z * scale)
    const double *sc;      // SC_RSCALE: storage scale of r (SRC32, FINAL)
    double *part;          // DOT: partials of z.r; FINAL: z.r then z.r_old
    SGeo g;
    int cu;                // colour updated (the half-sweep phase ^ 1)
};

// One red-black half-sweep of a level in the split layout, with z read as
// TZI, written as TZO and r stored as TR (float, __half or __hip_bfloat16;
// arithmetic in FP32).  ZERO starts from z = 0 and writes only the updated
// colour; the next sweep (PLAIN_OWN0) takes its own colour's old z as +0
// instead of reading zeros.  PROLONG reads z + z_coarse[parent] everywhere and
// also writes the other colour's prolongated values (out of place, zin !=
// zout).  SRC32 takes r of the updated colour from the FP32 CG residual times
// the storage scale and stores it (rounded to TR, and continues with the
// rounded values).  OUT: STORE; DOT accumulates z.r over both colours; FINAL
// stores like STORE (z stays in storage scale; 1/scale goes to zinv) and
// accumulates (z / scale).(r / scale) (r as stored, TR) and (z / scale).r_old
// over both colours; NATURAL stores both colours in the reference layout.
// Zeros standing in for the neighbour rows outside the domain in the direct-form
// sweeps (a row of at most ZROW_FLOATS values of the z type).
#define ZROW_FLOATS 4096
static __device__ __align__(16) float g_zero_row[ZROW_FLOATS];

template <int MODE, int OUT, typename TZI, typename TZO, typename TR, bool SRC32, int CPT>
__global__ void __launch_bounds__(256) k_ssweep(SSweepArgs a)
{
    // the reduced-precision pre-smoothing sweeps (16-bit z) update in direct
    // form (gs_direct): every neighbour holds +0 on non-FLUID cells, and the
    // rows outside the domain are read from a row of zeros.  The FP32 z of the
    // post-smoothing keeps the incremental form: there z holds the coarse
    // correction, and the direct form's sum of six such values rounds worse
    // (at 256^3 it left ten times the pressure error at the same residual).
    constexpr bool DIRECT = sizeof(TR) == 2 && sizeof(TZO) == 2 && MODE != SWEEP_PROLONG;
    __shared__ double lds[4];
    const SGeo g = a.g;
    const uint32_t total = (uint32_t)g.nx * g.ny * g.nq;
    const int lane = threadIdx.x & 63;
    const uint32_t own = (uint32_t)a.cu * g.half, oth = (uint32_t)(a.cu ^ 1) * g.half;
    const float s0 = (SRC32 || OUT == SWEEP_OUT_FINAL) ? (float)a.sc[SC_RSCALE] : 1.0f;
    if (OUT == SWEEP_OUT_FINAL && blockIdx.x == 0 && threadIdx.x == 0) a.zinv[0] = (double)(1.0f / s0);
    const TZI *zin = reinterpret_cast<const TZI *>(a.zin);
    double acc1 = 0.0, acc2 = 0.0;
    const Chunk ch = xcd_chunk(total);
    for (uint32_t t = ch.first; t < ch.end; t += ch.stride) {
        int i, j, m0;
        g.decode(t, i, j, m0);
        const int o = (a.cu - i - j) & 1;   // the updated colour sits at k = 2m + o
        auto zq = [&](uint32_t h, int ii, int jj, float v[CPT]) {
            if (MODE == SWEEP_ZERO) {
#pragma unroll
                for (int e = 0; e < CPT; ++e) v[e] = 0.0f;
                return;
            }
            ldn<TZI, CPT>(zin + h + g.row(ii, jj) + m0, v);
            if (MODE == SWEEP_PROLONG) {
                float p[CPT];
                ldn<float, CPT>(a.zc + cidx(ii >> 1, jj >> 1, m0, g.ny >> 1, g.nh), p);
#pragma unroll
                for (int e = 0; e < CPT; ++e) v[e] = v[e] + p[e];
            }
        };
        auto z1 = [&](uint32_t h, int ii, int jj, int m) -> float {
            if (MODE == SWEEP_ZERO) return 0.0f;
            float v = ldv<TZI>(zin[h + g.row(ii, jj) + m]);
            if (MODE == SWEEP_PROLONG) v = v + a.zc[cidx(ii >> 1, jj >> 1, m, g.ny >> 1, g.nh)];
            return v;
        };
        // neighbour rows clamped to the domain (a wall neighbour has stencil
        // class 0 and is never used): no branches, so all loads issue at once
        float c[CPT], w[CPT], xp[CPT], xm[CPT], yp[CPT], ym[CPT];
        if (MODE == SWEEP_PLAIN_OWN0 || DIRECT) {   // the direct form does not read the old value
#pragma unroll
            for (int e = 0; e < CPT; ++e) c[e] = 0.0f;
        } else {
            zq(own, i, j, c);
        }
        zq(oth, i, j, w);
        if constexpr (DIRECT && MODE != SWEEP_ZERO) {
            const TZI *zr = reinterpret_cast<const TZI *>(g_zero_row) + m0;
            auto rowp = [&](bool in, int ii, int jj) { return in ? zin + oth + g.row(ii, jj) + m0 : zr; };
            ldn<TZI, CPT>(rowp(i + 1 < g.nx, i + 1, j), xp);
            ldn<TZI, CPT>(rowp(i > 0, i - 1, j), xm);
            ldn<TZI, CPT>(rowp(j + 1 < g.ny, i, j + 1), yp);
            ldn<TZI, CPT>(rowp(j > 0, i, j - 1), ym);
        } else {
            zq(oth, min(i + 1, g.nx - 1), j, xp);
            zq(oth, max(i - 1, 0), j, xm);
            zq(oth, i, min(j + 1, g.ny - 1), yp);
            zq(oth, i, max(j - 1, 0), ym);
        }
        // the k-neighbour beyond the vector: other colour at m0-1 (o = 0) or m0+CPT (o = 1)
        const float left = __shfl_up(w[CPT - 1], 1, 64), right = __shfl_down(w[0], 1, 64);
        float edge = 0.0f;   // outside the domain: stencil class 0, unused
        if (o == 0) {
            if (m0 > 0) edge = lane > 0 ? left : z1(oth, i, j, m0 - 1);
        } else {
            if (m0 + CPT < g.nh) edge = (lane < 63 && t + 1 < ch.end) ? right : z1(oth, i, j, m0 + CPT);
        }
        const uint32_t base = g.row(i, j) + m0;
        uint32_t cd[CPT];
        codesn<CPT>(a.code, own + base, cd);
        float rv[CPT];
        if (SRC32) {
            ldn<float, CPT>(a.r32 + own + base, rv);
#pragma unroll
            for (int e = 0; e < CPT; ++e) rv[e] = rv[e] * s0;
            stn<TR, CPT>(reinterpret_cast<TR *>(a.r) + own + base, rv);
            if constexpr (sizeof(TR) != 4) {   // continue with the stored (rounded) values
#pragma unroll
                for (int e = 0; e < CPT; ++e) rv[e] = ldv<TR>(stv<TR>(rv[e]));
            }
        } else {
            ldn<TR, CPT>(reinterpret_cast<const TR *>(a.r) + own + base, rv);
        }
        float out[CPT];
#pragma unroll
        for (int e = 0; e < CPT; ++e) {
            float kp, km;
            if (o == 0) { kp = w[e]; km = e > 0 ? w[e - 1] : edge; }
            else { km = w[e]; kp = e < CPT - 1 ? w[e + 1] : edge; }
            // neighbours hold +0 on non-FLUID cells except in the prolongating
            // sweep, which adds the coarse correction to every cell on the fly
            if constexpr (DIRECT) out[e] = gs_direct(cd[e], rv[e], xp[e], xm[e], yp[e], ym[e], kp, km);
            else out[e] = gs_update<MODE != SWEEP_PROLONG, sizeof(TR) == 2>(cd[e], c[e], rv[e], xp[e], xm[e], yp[e],
                                                                            ym[e], kp, km);
        }
        if (OUT == SWEEP_OUT_FINAL) {
            // z and r over the storage scale s0 (a power of two: exact); the
            // products scale by s0^2, applied once to the thread's sums
            float rw[CPT], qo[CPT], qw[CPT];
            stn<TZO, CPT>(reinterpret_cast<TZO *>(a.zout) + own + base, out);
            // r from the sweep's stored r (r times the storage scale, rounded to
            // TR): the dot's rounding errors average out over the level, and the
            // other colour costs sizeof(TR) bytes a cell
            ldn<TR, CPT>(reinterpret_cast<const TR *>(a.r) + oth + base, rw);
            // r_old as the previous V-cycle stored it (over its own scale)
            ldn<TR, CPT>(reinterpret_cast<const TR *>(a.rold) + own + base, qo);
            ldn<TR, CPT>(reinterpret_cast<const TR *>(a.rold) + oth + base, qw);
            // the vector's terms summed in FP32, the sums accumulated in FP64
            // (this is the flexible CG's beta of a reduced-precision V-cycle:
            // an FP32 rounding per term is far below the V-cycle's own error)
            float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
            for (int e = 0; e < CPT; ++e) {
                s1 = fmaf(out[e], rv[e], s1);
                s1 = fmaf(w[e], rw[e], s1);
                s2 = fmaf(out[e], qo[e], s2);
                s2 = fmaf(w[e], qw[e], s2);
            }
            acc1 += (double)s1;
            acc2 += (double)s2;
        } else if (OUT == SWEEP_OUT_NATURAL) {
            // cells k = 2 m0 .. 2 m0 + 2 CPT - 1 of row (i, j): this colour at
            // k = 2m + o, the other one (unchanged, +0 on non-FLUID cells) at 2m + 1 - o
            float nat[2 * CPT];
#pragma unroll
            for (int e = 0; e < CPT; ++e) { nat[2 * e + o] = out[e]; nat[2 * e + 1 - o] = w[e]; }
            stn<float, 2 * CPT>(reinterpret_cast<float *>(a.zout) + cidx(i, j, 2 * m0, g.ny, 2 * g.nh), nat);
        } else {
            stn<TZO, CPT>(reinterpret_cast<TZO *>(a.zout) + own + base, out);
            if (MODE == SWEEP_PROLONG) {
                // the other colour's prolongated values, +0 on its non-FLUID cells:
                // the next sweep writes 0 there and never reads them as values, so
                // this keeps every later read and the cycle's result unchanged while
                // the sweeps after it see +0 on all non-FLUID neighbours.  That
                // cell is this cell's k+ (o = 0) or k- (o = 1) neighbour.
                float wm[CPT];
#pragma unroll
                for (int e = 0; e < CPT; ++e)
                    wm[e] = keep_if(w[e], 0u - ((cd[e] >> (o == 0 ? 9 : 11)) & 1u));
                stn<TZO, CPT>(reinterpret_cast<TZO *>(a.zout) + oth + base, wm);
            }
            if (OUT == SWEEP_OUT_DOT) {
                float ro[CPT];
                ldn<TR, CPT>(reinterpret_cast<const TR *>(a.r) + oth + base, ro);
#pragma unroll
                for (int e = 0; e < CPT; ++e) acc1 += (double)out[e] * (double)rv[e] + (double)w[e] * (double)ro[e];
            }
        }
    }
    if (OUT == SWEEP_OUT_FINAL) {   // (z / s0).(r / s0) and (z / s0).(r_old / s_old): powers of two
        const double inv = 1.0 / (double)s0;
        acc1 *= inv * inv;
        acc2 *= inv * a.zinv_old[0];
    }
    if (OUT == SWEEP_OUT_DOT || OUT == SWEEP_OUT_FINAL) {   // the only modes with partials (part != nullptr)
        const double s1 = block_sum<256>(acc1, lds);
        if (OUT == SWEEP_OUT_FINAL) {
            const double s2 = block_sum<256>(acc2, lds);
            if (threadIdx.x == 0) a.part[gridDim.x + blockIdx.x] = s2;
        }
        if (threadIdx.x == 0) a.part[blockIdx.x] = s1;
    }
}

// Two half-sweeps of a split level in one launch (FP32): colour cuA (the
// first sweep's, in mode MODE1: SWEEP_ZERO, SWEEP_PLAIN or SWEEP_PROLONG), then
// colour cuA ^ 1 (SWEEP_PLAIN_OWN0 after a ZERO one, else SWEEP_PLAIN), read
// from zin and written to zout (zout != zin: both colours; NAT: the whole z in
// the reference layout, like SWEEP_OUT_NATURAL).  A workgroup owns a tile of
// P2_TI x P2_TJ rows (i, j) over the full k range.  The first half-sweep runs
// on the tile and a one-row ring around it (clamped to the domain) into LDS;
// the second reads its neighbours of the first colour from there.  Each cell
// update is gs_update on the same operands as in the two k_ssweep launches,
// so zout is bitwise what they leave.  Used on the coarse split levels, where
// a half-sweep is short and launches, not bytes, set the time, and for the
// first pre- and post-smoothing pairs of the FP32 level 0 (the second sweep of
// each reads the first one's colour from LDS instead of HBM).
#define P2_TI 8
#define P2_TJ 4
#define P2_NHMAX 128
template <int MODE1, bool NAT, int NT, typename TZI = float, typename TZO = float, typename TR = float>
__global__ void __launch_bounds__(NT) k_ssweep2(const TZI *__restrict__ zin, TZO *__restrict__ zout,
                                                 const float *__restrict__ zc, const TR *__restrict__ r,
                                                 const uint16_t *__restrict__ code, SGeo g, int cuA)
{
    constexpr int HJ = P2_TJ + 2, HR = (P2_TI + 2) * HJ;
    __shared__ float4 A4[HR * P2_NHMAX / 4];   // first colour after its half-sweep, tile + ring rows
    float *const A = reinterpret_cast<float *>(A4);
    const int tpr = g.ny / P2_TJ, ntiles = (g.nx / P2_TI) * tpr;
    int b = (int)blockIdx.x;
    if ((ntiles & 7) == 0) b = (b & 7) * (ntiles >> 3) + (b >> 3);   // an XCD's blocks take adjacent tiles
    const int i0 = (b / tpr) * P2_TI, j0 = (b % tpr) * P2_TJ;
    const int cuB = cuA ^ 1, nh = g.nh, nq = g.nq, lq = g.lq;
    const uint32_t hA = (uint32_t)cuA * g.half, hB = (uint32_t)cuB * g.half;
    auto zq = [&](uint32_t h, int ii, int jj, int m0, float v[4]) {
        if (MODE1 == SWEEP_ZERO) {
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = 0.0f;
            return;
        }
        ldn<TZI, 4>(zin + h + g.row(ii, jj) + m0, v);
        if (MODE1 == SWEEP_PROLONG) {
            float pc[4];
            ldn<float, 4>(zc + cidx(ii >> 1, jj >> 1, m0, g.ny >> 1, nh), pc);
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = v[e] + pc[e];
        }
    };
    // first half-sweep: colour cuA on the tile and its ring
    for (int u = (int)threadIdx.x; u < HR * nq; u += NT) {
        const int hr = u >> lq, m0 = 4 * (u & (nq - 1));
        const int hi = hr / HJ, hj = hr - hi * HJ;
        const int i = i0 - 1 + hi, j = j0 - 1 + hj;
        if (i < 0 || i >= g.nx || j < 0 || j >= g.ny) continue;
        const int o = (cuA - i - j) & 1;
        float c[4], w[4], xp[4], xm[4], yp[4], ym[4];
        zq(hA, i, j, m0, c);
        zq(hB, i, j, m0, w);
        zq(hB, min(i + 1, g.nx - 1), j, m0, xp);
        zq(hB, max(i - 1, 0), j, m0, xm);
        zq(hB, i, min(j + 1, g.ny - 1), m0, yp);
        zq(hB, i, max(j - 1, 0), m0, ym);
        float edge = 0.0f;   // outside the domain: stencil class 0, unused
        if (MODE1 != SWEEP_ZERO) {
            const int me = o == 0 ? m0 - 1 : m0 + 4;
            if (me >= 0 && me < nh) {
                edge = ldv<TZI>(zin[hB + g.row(i, j) + me]);
                if (MODE1 == SWEEP_PROLONG) edge = edge + zc[cidx(i >> 1, j >> 1, me, g.ny >> 1, nh)];
            }
        }
        const uint32_t base = g.row(i, j) + m0;
        uint32_t cd[4];
        codesn<4>(code, hA + base, cd);
        float rv[4], out[4];
        ldn<TR, 4>(r + hA + base, rv);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float kp, km;
            if (o == 0) { kp = w[e]; km = e > 0 ? w[e - 1] : edge; }
            else { km = w[e]; kp = e < 3 ? w[e + 1] : edge; }
            out[e] = gs_update<MODE1 != SWEEP_PROLONG, sizeof(TR) == 2>(cd[e], c[e], rv[e], xp[e], xm[e], yp[e], ym[e], kp, km);
        }
        if constexpr (sizeof(TZO) == 2) {   // the second sweep reads the first's stored (rounded) values
#pragma unroll
            for (int e = 0; e < 4; ++e) out[e] = ldv<TZO>(stv<TZO>(out[e]));
        }
        A4[(hr * nh + m0) >> 2] = make_float4(out[0], out[1], out[2], out[3]);
    }
    __syncthreads();
    // second half-sweep: colour cuB on the tile, its cuA neighbours from LDS
    for (int u = (int)threadIdx.x; u < P2_TI * P2_TJ * nq; u += NT) {
        const int tr = u >> lq, m0 = 4 * (u & (nq - 1));
        const int ti = tr / P2_TJ, tj = tr - ti * P2_TJ;
        const int i = i0 + ti, j = j0 + tj;
        const int o = (cuB - i - j) & 1;
        auto arow = [&](int ii, int jj) -> const float * { return A + ((ii - i0 + 1) * HJ + (jj - j0 + 1)) * nh; };
        auto a4 = [&](const float *rowp, float v[4]) {
            const float4 t = *reinterpret_cast<const float4 *>(rowp + m0);
            v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
        };
        const float *Ac = arow(i, j);
        float c[4], w[4], xp[4], xm[4], yp[4], ym[4];
        a4(Ac, w);
        a4(arow(min(i + 1, g.nx - 1), j), xp);
        a4(arow(max(i - 1, 0), j), xm);
        a4(arow(i, min(j + 1, g.ny - 1)), yp);
        a4(arow(i, max(j - 1, 0)), ym);
        const int me = o == 0 ? m0 - 1 : m0 + 4;
        const float edge = (me >= 0 && me < nh) ? Ac[me] : 0.0f;
        if (MODE1 == SWEEP_ZERO) {   // SWEEP_PLAIN_OWN0
#pragma unroll
            for (int e = 0; e < 4; ++e) c[e] = 0.0f;
        } else {
            zq(hB, i, j, m0, c);
        }
        const uint32_t base = g.row(i, j) + m0;
        uint32_t cd[4];
        codesn<4>(code, hB + base, cd);
        float rv[4], out[4];
        ldn<TR, 4>(r + hB + base, rv);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float kp, km;
            if (o == 0) { kp = w[e]; km = e > 0 ? w[e - 1] : edge; }
            else { km = w[e]; kp = e < 3 ? w[e + 1] : edge; }
            out[e] = gs_update<true, sizeof(TR) == 2>(cd[e], c[e], rv[e], xp[e], xm[e], yp[e], ym[e], kp, km);
        }
        if (NAT) {
            float nat[8];
#pragma unroll
            for (int e = 0; e < 4; ++e) { nat[2 * e + o] = out[e]; nat[2 * e + 1 - o] = w[e]; }
            stn<TZO, 8>(zout + cidx(i, j, 2 * m0, g.ny, 2 * nh), nat);
        } else {
            stn<TZO, 4>(zout + hB + base, out);
            stn<TZO, 4>(zout + hA + base, w);
        }
    }
}

// Gather restriction (mgpcg.py:94-101) from split level 0 into level 1
// (reference layout), four coarse cells per thread: the children of coarse
// (I, J, K) are, per child row (2I+a, 2J+b), the cells m = K of both halves.
// Residual r - (-ndt(z)) per fine cell (0 unless FLUID); a FLUID parent sums
// its children in the serial launch order (a, b, kk), others get 0.
template <typename TR, typename TZ, int CPT, typename TO = float>
__global__ void __launch_bounds__(256) k_srestrict(const TR *__restrict__ r, const TZ *__restrict__ z,
                                                   const uint16_t *__restrict__ code, TO *__restrict__ rc,
                                                   const uint8_t *__restrict__ tagc, SGeo g, SGeo cg, int csplit,
                                                   float cs)
{
    // cg: the coarse grid in thread vectors of CPT cells (nx, ny coarse; nh = coarse nz = fine nz/2)
    const uint32_t total = (uint32_t)cg.nx * cg.ny * cg.nq;
    const int lane = threadIdx.x & 63;
    const Chunk ch = xcd_chunk(total);
    for (uint32_t t = ch.first; t < ch.end; t += ch.stride) {
        int I, J, K0;
        cg.decode(t, I, J, K0);
        const int i0 = 2 * I, j0 = 2 * J;
        // both colours of fine row (ii, jj) at m = K0 ..
        auto ld2 = [&](int ii, int jj, float v[2][CPT]) {
            const uint32_t b = g.row(ii, jj) + K0;
            ldn<TZ, CPT>(z + b, v[0]);
            ldn<TZ, CPT>(z + g.half + b, v[1]);
        };
        // Each row is loaded once: the rows i-1, i, i+1 of the two child j
        // rotate through P, C, N as the child i advances; the y-halo rows
        // (i, j0-1) and (i, j0+2) are loaded per child i.  Rows clamped to
        // the domain (a wall neighbour has stencil class 0 and is never used).
        float P[2][2][CPT], C[2][2][CPT], N[2][2][CPT];   // [b][colour][e]
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            ld2(max(i0 - 1, 0), j0 + b, P[b]);
            ld2(i0, j0 + b, C[b]);
        }
        float s[CPT];
#pragma unroll
        for (int e = 0; e < CPT; ++e) s[e] = 0.0f;
#pragma unroll
        for (int aa = 0; aa < 2; ++aa) {
            const int i = i0 + aa;
#pragma unroll
            for (int b = 0; b < 2; ++b) ld2(min(i + 1, g.nx - 1), j0 + b, N[b]);
            float YM[2][CPT], YP[2][CPT];
            ld2(i, max(j0 - 1, 0), YM);
            ld2(i, min(j0 + 2, g.ny - 1), YP);
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int j = j0 + b;
                const uint32_t base = g.row(i, j) + K0;
                float res[2][CPT];   // [kk][e]: child (i, j, 2(K0+e)+kk)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                    const uint32_t own = (uint32_t)cc * g.half, oth = (uint32_t)(cc ^ 1) * g.half;
                    const int o = (cc - i - j) & 1;   // this colour's child is kk = o
                    const float *w = C[b][cc ^ 1];
                    const float *yp = b == 0 ? C[1][cc ^ 1] : YP[cc ^ 1];   // row j+1
                    const float *ym = b == 0 ? YM[cc ^ 1] : C[0][cc ^ 1];   // row j-1
                    const float left = __shfl_up(w[CPT - 1], 1, 64), right = __shfl_down(w[0], 1, 64);
                    float edge = 0.0f;
                    if (o == 0) {
                        if (K0 > 0) edge = lane > 0 ? left : ldv<TZ>(z[oth + base - 1]);
                    } else {
                        if (K0 + CPT < g.nh) edge = (lane < 63 && t + 1 < ch.end) ? right : ldv<TZ>(z[oth + base + CPT]);
                    }
                    uint32_t cd[CPT];
                    codesn<CPT>(code, own + base, cd);
                    float rv[CPT];
                    ldn<TR, CPT>(r + own + base, rv);
#pragma unroll
                    for (int e = 0; e < CPT; ++e) {
                        float kp, km;
                        if (o == 0) { kp = w[e]; km = e > 0 ? w[e - 1] : edge; }
                        else { km = w[e]; kp = e < CPT - 1 ? w[e + 1] : edge; }
                        // the pre-smoothed z holds +0 on non-FLUID cells
                        float nd = nsum_code_air0(cd[e], C[b][cc][e], N[b][cc ^ 1][e], P[b][cc ^ 1][e], yp[e], ym[e],
                                                  kp, km);
                        if (!code_fluid(cd[e])) nd = 0.0f;
                        res[o][e] = rv[e] - (-nd);
                    }
                }
                // the children in the serial order (a, b, kk)
#pragma unroll
                for (int e = 0; e < CPT; ++e) s[e] = (s[e] + res[0][e]) + res[1][e];
            }
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int e = 0; e < CPT; ++e) { P[b][c][e] = C[b][c][e]; C[b][c][e] = N[b][c][e]; }
        }
        const uint32_t cb = cidx(I, J, K0, cg.ny, cg.nh);
        uint32_t tg[CPT / 4];   // the coarse tags, 4 to a word
#pragma unroll
        for (int q = 0; q < CPT / 4; ++q) tg[q] = reinterpret_cast<const uint32_t *>(tagc + cb)[q];
        float out[CPT];
#pragma unroll
        for (int e = 0; e < CPT; ++e) out[e] = ((tg[e >> 2] >> (8 * (e & 3))) & 0xffu) == 1u ? s[e] * cs : 0.0f;
        if (csplit) {   // colour-split coarse r: even e to half (I + J) & 1, odd e to the other (K0 even)
            const uint32_t nhc = (uint32_t)cg.nh >> 1, halfc = (uint32_t)cg.nx * (uint32_t)cg.ny * nhc;
            const uint32_t b = ((uint32_t)I * cg.ny + J) * nhc + (uint32_t)(K0 >> 1);
            const uint32_t c0 = (uint32_t)((I + J) & 1);
            float ev[CPT / 2], od[CPT / 2];
#pragma unroll
            for (int e = 0; e < CPT / 2; ++e) { ev[e] = out[2 * e]; od[e] = out[2 * e + 1]; }
            stn<TO, CPT / 2>(rc + c0 * halfc + b, ev);
            stn<TO, CPT / 2>(rc + (c0 ^ 1u) * halfc + b, od);
        } else {
            stn<TO, CPT>(rc + cb, out);
        }
    }
}

// Ap = ndt(p) (mgpcg.py:36-39) and the FP64 partial of p.Ap, in the split
// layout; each thread does the quads of both colours at one position.  (The
// p update runs before, elementwise: fused into this kernel it costs the
// neighbour reads of z as well, and the kernel is then bound by the L1 fills
// of the neighbour rows.)
__global__ void __launch_bounds__(256) k_sp_Ap(const float *__restrict__ p, float *__restrict__ Ap,
                                               const uint16_t *__restrict__ code, SGeo g, double *__restrict__ part)
{
    __shared__ double lds[4];
    const uint32_t total = (uint32_t)g.nx * g.ny * g.nq;
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
    const Chunk ch = xcd_chunk(total);
    for (uint32_t t = ch.first; t < ch.end; t += ch.stride) {
        int i, j, m0;
        g.decode(t, i, j, m0);
        auto pq = [&](uint32_t h, int ii, int jj, float v[4]) { f4(p + h + g.row(ii, jj) + m0, v); };
        auto p1 = [&](uint32_t b) -> float { return p[b]; };
        const uint32_t base = g.row(i, j) + m0;
        float pn[2][4];
        pq(0, i, j, pn[0]);
        pq(g.half, i, j, pn[1]);
        const float l0 = __shfl_up(pn[0][3], 1, 64), r0 = __shfl_down(pn[0][0], 1, 64);
        const float l1 = __shfl_up(pn[1][3], 1, 64), r1 = __shfl_down(pn[1][0], 1, 64);
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
            const uint32_t own = (uint32_t)cc * g.half, oth = (uint32_t)(cc ^ 1) * g.half;
            const int o = (cc - i - j) & 1;
            const float *w = pn[cc ^ 1];
            float xp[4], xm[4], yp[4], ym[4];   // rows clamped to the domain (walls: class 0)
            pq(oth, min(i + 1, g.nx - 1), j, xp);
            pq(oth, max(i - 1, 0), j, xm);
            pq(oth, i, min(j + 1, g.ny - 1), yp);
            pq(oth, i, max(j - 1, 0), ym);
            float edge = 0.0f;
            if (o == 0) {
                if (m0 > 0) edge = lane > 0 ? (cc ? l0 : l1) : p1(oth + base - 1);
            } else {
                if (m0 + 4 < g.nh) edge = (lane < 63 && t + 1 < ch.end) ? (cc ? r0 : r1) : p1(oth + base + 4);
            }
            uint32_t cd[4];
            codes4(code, own + base, cd);
            float out[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float kp, km;
                if (o == 0) { kp = w[e]; km = e > 0 ? w[e - 1] : edge; }
                else { km = w[e]; kp = e < 3 ? w[e + 1] : edge; }
                const float s = nsum_code(cd[e], pn[cc][e], xp[e], xm[e], yp[e], ym[e], kp, km);
                out[e] = code_fluid(cd[e]) ? s : 0.0f;
                acc += (double)pn[cc][e] * (double)out[e];
            }
            st_f4(Ap + own + base, out);
        }
    }
    const double s = block_sum<256>(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// The p step and Ap of the split solves in one pass: p = beta p_prev + z (z
// times zinv[0] when set; k_update_pp's arithmetic) and Ap = ndt(p) with the
// FP64 partials of p.Ap (k_sp_Ap's arithmetic, so p and Ap are bitwise those
// of the two-pass step).  A workgroup owns a tile of jb = 256 / nq rows along
// j, each row whole along k (nq vectors of 4 cells per half-row), and marches
// along i over the rows [i0, i1) of its segment.  Each step computes p of row
// i + 1 (both colours) from the loads issued one step earlier, keeps rows
// i - 1, i, i + 1 of its own (j, vector) in registers and puts row i + 1 into
// an LDS plane together with the tile's two halo rows along j (computed again
// by helper threads: 4 nq vectors); the i and j neighbours of row i then come
// from registers and LDS, and the k neighbour beyond a vector from the
// adjacent lane (a half-row's vectors sit in consecutive lanes of one
// wavefront).  Every value of p_prev and z is thus loaded about
// (ib + 2) / ib * (jb + 2) / jb times (mostly from L2) instead of five times
// from L1.  Rows i0 - 1 and i1 are computed for their neighbours only; row
// indices are clamped to the grid as in k_sp_Ap.  Three planes in rotation
// need one barrier a step: a step writes plane (i + 1) % 3 and reads plane
// i % 3, and plane (i + 1) % 3 was last read two steps earlier.  Block b runs
// tile (b & 7) * tpx + (b >> 3), so that the segments of a tile (which share
// their halo rows along i) run on one XCD; blocks past the last tile write a
// zero partial.
struct PApArgs {
    float *p_out;
    const float *p_prev, *z;
    const double *sc, *zinv;
    double *asave;   // receives sc[SC_ALPHA] (block 0), as in k_update_pp
    float *Ap;
    const uint16_t *code;
    double *part;
    SGeo g;
    int lq;          // log2 of nq (2..6)
    int jb;          // rows of a tile along j: 256 / nq
    int ib;          // rows of a segment along i
    int nseg;        // segments along i
    uint32_t ntiles, tpx;   // tiles, tiles per XCD
};

__global__ void __launch_bounds__(256) k_sp_pAp(PApArgs a)
{
    extern __shared__ float4 pl[];   // 3 planes of (jb + 2) rows x 2 colours x nq vectors
    __shared__ double lds[4];
    const SGeo &g = a.g;
    const int tid = threadIdx.x;
    const uint32_t T = (blockIdx.x & 7u) * a.tpx + (blockIdx.x >> 3);
    if (T >= a.ntiles) {
        if (tid == 0) a.part[blockIdx.x] = 0.0;
        return;
    }
    if (a.asave && blockIdx.x == 0 && tid == 0) a.asave[0] = a.sc[SC_ALPHA];   // read by a later launch only
    const int nq = g.nq, lq = a.lq;
    const int seg = (int)(T % (uint32_t)a.nseg), jt = (int)(T / (uint32_t)a.nseg);
    const int i0 = seg * a.ib, i1 = min(i0 + a.ib, g.nx);
    const int j0 = jt * a.jb;
    const int q = tid & (nq - 1), jl = tid >> lq, j = j0 + jl, m0 = 4 * q;
    // helper threads: halo row j0 - 1 (hr 0) or j0 + jb (hr 1), clamped, colour hc
    const bool helper = tid < 4 * nq;
    const int hr = (tid >> (lq + 1)) & 1, hc = (tid >> lq) & 1;
    const int hj = hr == 0 ? max(j0 - 1, 0) : min(j0 + a.jb, g.ny - 1);
    const uint32_t plane = (uint32_t)(a.jb + 2) * 2u * (uint32_t)nq;
    const uint32_t own_slot = (uint32_t)((jl + 1) * 2) * nq + q;   // + c * nq
    const uint32_t halo_slot = (uint32_t)(((hr == 0 ? 0 : a.jb + 1) * 2 + hc) * nq + q);
    const float beta = (float)a.sc[SC_BETA];
    const bool scaled = a.zinv != nullptr;
    const float zi = scaled ? (float)a.zinv[0] : 1.0f;
    auto ld = [&](const float *b, int c, int ii, int jj) -> float4 {
        return *reinterpret_cast<const float4 *>(b + (uint32_t)c * g.half + g.row(ii, jj) + m0);
    };
    auto pval = [&](float4 pp, float4 zz) -> float4 {
        if (scaled) { zz.x *= zi; zz.y *= zi; zz.z *= zi; zz.w *= zi; }
        return make_float4(fmaf(beta, pp.x, zz.x), fmaf(beta, pp.y, zz.y), fmaf(beta, pp.z, zz.z),
                           fmaf(beta, pp.w, zz.w));
    };
    float4 P0[2], P1[2], P2[2];   // p of rows i - 1, i, i + 1 (clamped) at (j, m0), both colours
    float4 PP[2], ZZ[2], HP, HZ;  // loads of the next row: own, helper
    uint2 CD[2];                  // stencil codes of row i (loaded one step ahead, like p_prev and z:
                                  // a wait for a load also waits for every load issued before it)
    auto ldc = [&](int c, int ii) -> uint2 {
        return *reinterpret_cast<const uint2 *>(a.code + (uint32_t)c * g.half + g.row(ii, j) + m0);
    };
    {
        const int im = max(i0 - 1, 0);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            P0[c] = pval(ld(a.p_prev, c, im, j), ld(a.z, c, im, j));
            P1[c] = pval(ld(a.p_prev, c, i0, j), ld(a.z, c, i0, j));
            pl[(uint32_t)(i0 % 3) * plane + own_slot + (uint32_t)c * nq] = P1[c];
        }
        if (helper) pl[(uint32_t)(i0 % 3) * plane + halo_slot] = pval(ld(a.p_prev, hc, i0, hj), ld(a.z, hc, i0, hj));
        const int in = min(i0 + 1, g.nx - 1);
#pragma unroll
        for (int c = 0; c < 2; ++c) { CD[c] = ldc(c, i0); PP[c] = ld(a.p_prev, c, in, j); ZZ[c] = ld(a.z, c, in, j); }
        HP = ld(a.p_prev, hc, in, hj);
        HZ = ld(a.z, hc, in, hj);
    }
    double acc = 0.0;
    for (int i = i0; i < i1; ++i) {
        const uint32_t b1 = (uint32_t)((i + 1) % 3) * plane, b0 = (uint32_t)(i % 3) * plane;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            P2[c] = pval(PP[c], ZZ[c]);
            pl[b1 + own_slot + (uint32_t)c * nq] = P2[c];
        }
        if (helper) pl[b1 + halo_slot] = pval(HP, HZ);
        uint32_t cd[2][4];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            cd[c][0] = CD[c].x & 0xffffu; cd[c][1] = CD[c].x >> 16; cd[c][2] = CD[c].y & 0xffffu; cd[c][3] = CD[c].y >> 16;
        }
        if (i + 1 < i1) {   // the next step's codes (row i + 1) and rows i + 2 (clamped)
            const int in = min(i + 2, g.nx - 1);
#pragma unroll
            for (int c = 0; c < 2; ++c) CD[c] = ldc(c, i + 1);
#pragma unroll
            for (int c = 0; c < 2; ++c) { PP[c] = ld(a.p_prev, c, in, j); ZZ[c] = ld(a.z, c, in, j); }
            // every thread loads a halo vector (hj, hc are in the grid for all): with no
            // branch around the loads the waits for them are exact
            HP = ld(a.p_prev, hc, in, hj);
            HZ = ld(a.z, hc, in, hj);
        }
        const uint32_t base = g.row(i, j) + m0;
        __syncthreads();
        float4 YP[2], YM[2];   // rows j + 1 and j - 1 (clamped through the halo rows)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            YP[c] = pl[b0 + own_slot + (uint32_t)(2 * nq) + (uint32_t)c * nq];
            YM[c] = pl[b0 + own_slot - (uint32_t)(2 * nq) + (uint32_t)c * nq];
        }
        const float l0 = __shfl_up(P1[0].w, 1, 64), r0 = __shfl_down(P1[0].x, 1, 64);
        const float l1 = __shfl_up(P1[1].w, 1, 64), r1 = __shfl_down(P1[1].x, 1, 64);
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
            const int o = (cc - i - j) & 1;
            const float4 wv = P1[cc ^ 1], own4 = P1[cc];
            const float w[4] = {wv.x, wv.y, wv.z, wv.w}, pn[4] = {own4.x, own4.y, own4.z, own4.w};
            const float xp[4] = {P2[cc ^ 1].x, P2[cc ^ 1].y, P2[cc ^ 1].z, P2[cc ^ 1].w};
            const float xm[4] = {P0[cc ^ 1].x, P0[cc ^ 1].y, P0[cc ^ 1].z, P0[cc ^ 1].w};
            const float yp[4] = {YP[cc ^ 1].x, YP[cc ^ 1].y, YP[cc ^ 1].z, YP[cc ^ 1].w};
            const float ym[4] = {YM[cc ^ 1].x, YM[cc ^ 1].y, YM[cc ^ 1].z, YM[cc ^ 1].w};
            float edge = 0.0f;
            if (o == 0) {
                if (m0 > 0) edge = cc ? l0 : l1;
            } else {
                if (m0 + 4 < g.nh) edge = cc ? r0 : r1;
            }
            float out[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float kp, km;
                if (o == 0) { kp = w[e]; km = e > 0 ? w[e - 1] : edge; }
                else { km = w[e]; kp = e < 3 ? w[e + 1] : edge; }
                const float s = nsum_code(cd[cc][e], pn[e], xp[e], xm[e], yp[e], ym[e], kp, km);
                out[e] = code_fluid(cd[cc][e]) ? s : 0.0f;
                acc += (double)pn[e] * (double)out[e];
            }
            const uint32_t at = (uint32_t)cc * g.half + base;
            st_f4(a.Ap + at, out);
            *reinterpret_cast<float4 *>(a.p_out + at) = own4;
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) { P0[c] = P1[c]; P1[c] = P2[c]; }
    }
    const double s = block_sum<256>(acc, lds);
    if (tid == 0) a.part[blockIdx.x] = s;
}

// True residual r = b - ndt(x) with the stencil evaluated in FP64 (from the
// FP32 or FP64 x), stored as FP32 r; 0 off FLUID cells.  Partials: sum of r^2
// (FP64 values) in part[block], max |r| in part[nblk + block].  Split layout,
// both colours at each position.
template <typename XT>
__global__ void __launch_bounds__(256) k_strue_residual(float *__restrict__ r, const float *__restrict__ b,
                                                        const XT *__restrict__ x,
                                                        const uint16_t *__restrict__ code, SGeo g,
                                                        double *__restrict__ part)
{
    __shared__ double lds[4];
    __shared__ double lm[4];
    const uint32_t total = (uint32_t)g.nx * g.ny * g.nq;
    const int lane = threadIdx.x & 63;
    double acc = 0.0, mx = 0.0;
    const Chunk ch = xcd_chunk(total);
    for (uint32_t t = ch.first; t < ch.end; t += ch.stride) {
        int i, j, m0;
        g.decode(t, i, j, m0);
        const uint32_t base = g.row(i, j) + m0;
        double xs[2][4];
        d4(x + base, xs[0]);
        d4(x + g.half + base, xs[1]);
        const double l0 = __shfl_up(xs[0][3], 1, 64), r0 = __shfl_down(xs[0][0], 1, 64);
        const double l1 = __shfl_up(xs[1][3], 1, 64), r1 = __shfl_down(xs[1][0], 1, 64);
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
            const uint32_t own = (uint32_t)cc * g.half, oth = (uint32_t)(cc ^ 1) * g.half;
            const int o = (cc - i - j) & 1;
            const double *w = xs[cc ^ 1];
            double xp[4], xm[4], yp[4], ym[4];   // rows clamped to the domain (walls: class 0)
            d4(x + oth + g.row(min(i + 1, g.nx - 1), j) + m0, xp);
            d4(x + oth + g.row(max(i - 1, 0), j) + m0, xm);
            d4(x + oth + g.row(i, min(j + 1, g.ny - 1)) + m0, yp);
            d4(x + oth + g.row(i, max(j - 1, 0)) + m0, ym);
            double edge = 0.0;
            if (o == 0) {
                if (m0 > 0) edge = lane > 0 ? (cc ? l0 : l1) : (double)x[oth + base - 1];
            } else {
                if (m0 + 4 < g.nh) edge = (lane < 63 && t + 1 < ch.end) ? (cc ? r0 : r1) : (double)x[oth + base + 4];
            }
            uint32_t cd[4];
            codes4(code, own + base, cd);
            float bv[4];
            f4(b + own + base, bv);
            float out[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                double kp, km;
                if (o == 0) { kp = w[e]; km = e > 0 ? w[e - 1] : edge; }
                else { km = w[e]; kp = e < 3 ? w[e + 1] : edge; }
                const double xI = xs[cc][e];
                const double nbv[6] = {xp[e], xm[e], yp[e], ym[e], kp, km};
                double sacc = 0.0;
#pragma unroll
                for (int d = 0; d < 6; ++d) {
                    const uint32_t cls = (cd[e] >> (2 * d)) & 3u;
                    if (cls != 0u) sacc += (cls == 2u ? nbv[d] : 0.0) - xI;
                }
                const double rv = code_fluid(cd[e]) ? (double)bv[e] - sacc : 0.0;
                out[e] = (float)rv;
                acc += rv * rv;
                mx = fmax(mx, fabs(rv));
            }
            st_f4(r + own + base, out);
        }
    }
    const double s = block_sum<256>(acc, lds);
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
    if (lane == 0) lm[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s;
        part[gridDim.x + blockIdx.x] = fmax(fmax(lm[0], lm[1]), fmax(lm[2], lm[3]));
    }
}

// Reference layout <-> split layout, one reference quad (4 cells along k)
// per thread: cells k0, k0+2 go to one half, k0+1, k0+3 to the other, both
// at m = k0/2, k0/2 + 1.
template <typename TS, typename TD, bool TO_SPLIT>
__global__ void __launch_bounds__(256) k_split_conv(const TS *__restrict__ src, TD *__restrict__ dst, SGeo g)
{
    const uint32_t nzq = (uint32_t)g.nh / 2;   // reference quads per row
    const uint32_t total = (uint32_t)g.nx * g.ny * nzq;
    for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < total; t += gridDim.x * 256u) {
        const uint32_t row = t / nzq, kq = t - row * nzq;
        const int i = (int)(row / (uint32_t)g.ny), j = (int)(row - (uint32_t)i * g.ny);
        const int k0 = 4 * (int)kq;
        const uint32_t nat = row * (uint32_t)(2 * g.nh) + (uint32_t)k0;
        const uint32_t c0 = (uint32_t)((i + j + k0) & 1);   // colour of cells k0, k0+2
        const uint32_t s0 = c0 * g.half + row * (uint32_t)g.nh + (uint32_t)(k0 >> 1);
        const uint32_t s1 = (c0 ^ 1u) * g.half + row * (uint32_t)g.nh + (uint32_t)(k0 >> 1);
        if (TO_SPLIT) {
            const TS v0 = src[nat], v1 = src[nat + 1], v2 = src[nat + 2], v3 = src[nat + 3];
            dst[s0] = (TD)v0; dst[s0 + 1] = (TD)v2;
            dst[s1] = (TD)v1; dst[s1 + 1] = (TD)v3;
        } else {
            const TS v0 = src[s0], v2 = src[s0 + 1], v1 = src[s1], v3 = src[s1 + 1];
            dst[nat] = (TD)v0; dst[nat + 1] = (TD)v1; dst[nat + 2] = (TD)v2; dst[nat + 3] = (TD)v3;
        }
    }
}

// The split FP64 solution to the reference layout, rounded to FP32 into x32
// and as it is into x64 (k_split_conv's mapping, both outputs in one pass;
// x64 may be null); with pout also the projection's pressure from the FP32 x
// through the reference's scale chain, ((x h) / h) h as (x h) * inv_h * h.
__global__ void __launch_bounds__(256) k_split_x_out(const double *__restrict__ src, float *__restrict__ x32,
                                                     double *__restrict__ x64, SGeo g, float *__restrict__ pout,
                                                     float hh, float inv_h)
{
    const uint32_t nzq = (uint32_t)g.nh / 2;
    const uint32_t total = (uint32_t)g.nx * g.ny * nzq;
    for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < total; t += gridDim.x * 256u) {
        const uint32_t row = t / nzq, kq = t - row * nzq;
        const int i = (int)(row / (uint32_t)g.ny), j = (int)(row - (uint32_t)i * g.ny);
        const int k0 = 4 * (int)kq;
        const uint32_t nat = row * (uint32_t)(2 * g.nh) + (uint32_t)k0;
        const uint32_t c0 = (uint32_t)((i + j + k0) & 1);
        const uint32_t s0 = c0 * g.half + row * (uint32_t)g.nh + (uint32_t)(k0 >> 1);
        const uint32_t s1 = (c0 ^ 1u) * g.half + row * (uint32_t)g.nh + (uint32_t)(k0 >> 1);
        const double v0 = src[s0], v2 = src[s0 + 1], v1 = src[s1], v3 = src[s1 + 1];
        const float f0 = (float)v0, f1 = (float)v1, f2 = (float)v2, f3 = (float)v3;
        *reinterpret_cast<float4 *>(x32 + nat) = make_float4(f0, f1, f2, f3);
        if (x64) {
            reinterpret_cast<double2 *>(x64 + nat)[0] = make_double2(v0, v1);
            reinterpret_cast<double2 *>(x64 + nat)[1] = make_double2(v2, v3);
        }
        if (pout)
            *reinterpret_cast<float4 *>(pout + nat) =
                make_float4(((f0 * hh) * inv_h) * hh, ((f1 * hh) * inv_h) * hh, ((f2 * hh) * inv_h) * hh,
                            ((f3 * hh) * inv_h) * hh);
    }
}

int ilog2_or_neg(int v)
{
    if (v <= 0 || (v & (v - 1))) return -1;
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

SGeo make_sgeo(int nx, int ny, int nz, int cpt = 4)
{
    SGeo g;
    g.nx = nx; g.ny = ny; g.nh = nz / 2; g.cpt = cpt; g.nq = g.nh / cpt;
    g.lq = ilog2_or_neg(g.nq);
    g.lny = ilog2_or_neg(ny);
    g.lnx = ilog2_or_neg(nx);
    // 32-row blocks: at 256^3 the half-sweep runs 24.0 us instead of 24.9, at
    // 512^3 261 us instead of 285 (the whole solve 5% faster there)
    const int jb = 32;
    g.ljb = (jb < ny && ny % jb == 0) ? ilog2_or_neg(jb) : -1;
    g.half = (uint32_t)nx * ny * (uint32_t)g.nh;
    return g;
}

unsigned grid8(uint32_t items)
{
    unsigned b = blocks_for(items, 256);
    b = b < 1 ? 1 : (b > MAXBLK ? MAXBLK : b);
    return (b + 7u) & ~7u;   // xcd_chunk needs a multiple of 8
}

template <int MODE, int OUT, typename TZI, typename TZO, typename TR, bool SRC32, int CPT = 4>
void sweep_launch(hipStream_t s, unsigned nb, const SSweepArgs &a)
{
    k_ssweep<MODE, OUT, TZI, TZO, TR, SRC32, CPT><<<nb, 256, 0, s>>>(a);
}

// The mixed-precision level-0 sweeps: pre-smoothing with T storage of z and
// r (the first two convert r from the FP32 CG residual), post-smoothing with
// FP32 z and T r, the last one FINAL.
template <typename T>
int split_sweep_mixed_t(hipStream_t s, const Level &L0, const SplitMixedSweep &m)
{
    SSweepArgs a{};
    a.zin = m.zin; a.zout = m.zout; a.zc = m.zc; a.r = m.rh; a.code = m.code;
    a.r32 = m.r32; a.rold = m.rold; a.zinv_old = m.zinv_old; a.zinv = m.zinv; a.sc = m.sc; a.part = m.part;
    a.cu = m.phase ^ 1;
    // the pre-smoothing sweeps (16-bit z and r) move 8 cells per thread, in
    // 16-byte vectors, when the half-rows allow it
    const bool pre = m.kind == SMX_PRE_ZERO_SRC || m.kind == SMX_PRE_SRC || m.kind == SMX_PRE ||
                     m.kind == SMX_PRE_ZERO_RH || m.kind == SMX_PRE_OWN0_RH;
    const int cpt = pre && (L0.nz / 2) % 8 == 0 ? 8 : 4;
    FLUB_REQUIRE(L0.nz / 2 <= ZROW_FLOATS, FLUB_ESHAPE, "mixed sweep: nz = %d above %d", L0.nz, 2 * ZROW_FLOATS);
    a.g = make_sgeo(L0.nx, L0.ny, L0.nz, cpt);
    const unsigned nb = grid8((uint32_t)L0.nx * L0.ny * a.g.nq);
    if (m.nblk) *m.nblk = (int)nb;
    switch (m.kind) {
    case SMX_PRE_ZERO_SRC:
        if (cpt == 8) sweep_launch<SWEEP_ZERO, SWEEP_OUT_STORE, T, T, T, true, 8>(s, nb, a);
        else sweep_launch<SWEEP_ZERO, SWEEP_OUT_STORE, T, T, T, true>(s, nb, a);
        break;
    case SMX_PRE_SRC:   // follows the ZERO sweep: this colour starts at 0
        if (cpt == 8) sweep_launch<SWEEP_PLAIN_OWN0, SWEEP_OUT_STORE, T, T, T, true, 8>(s, nb, a);
        else sweep_launch<SWEEP_PLAIN_OWN0, SWEEP_OUT_STORE, T, T, T, true>(s, nb, a);
        break;
    case SMX_PRE_ZERO_RH:
        if (cpt == 8) sweep_launch<SWEEP_ZERO, SWEEP_OUT_STORE, T, T, T, false, 8>(s, nb, a);
        else sweep_launch<SWEEP_ZERO, SWEEP_OUT_STORE, T, T, T, false>(s, nb, a);
        break;
    case SMX_PRE_OWN0_RH:
        if (cpt == 8) sweep_launch<SWEEP_PLAIN_OWN0, SWEEP_OUT_STORE, T, T, T, false, 8>(s, nb, a);
        else sweep_launch<SWEEP_PLAIN_OWN0, SWEEP_OUT_STORE, T, T, T, false>(s, nb, a);
        break;
    case SMX_PRE:
        if (cpt == 8) sweep_launch<SWEEP_PLAIN, SWEEP_OUT_STORE, T, T, T, false, 8>(s, nb, a);
        else sweep_launch<SWEEP_PLAIN, SWEEP_OUT_STORE, T, T, T, false>(s, nb, a);
        break;
    case SMX_POST_PROLONG: sweep_launch<SWEEP_PROLONG, SWEEP_OUT_STORE, T, float, T, false>(s, nb, a); break;
    case SMX_POST: sweep_launch<SWEEP_PLAIN, SWEEP_OUT_STORE, float, float, T, false>(s, nb, a); break;
    default: sweep_launch<SWEEP_PLAIN, SWEEP_OUT_FINAL, float, float, T, false>(s, nb, a); break;
    }
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

}  // namespace

bool split_supported(const Level &L0) { return L0.nz % 8 == 0 && L0.nx % 2 == 0 && L0.ny % 2 == 0; }

int split_convert(hipStream_t s, const Level &L0, const void *src, void *dst, int elem_bytes, bool to_split)
{
    const SGeo g = make_sgeo(L0.nx, L0.ny, L0.nz);
    const unsigned nb = blocks_for(L0.n / 4, 256) < 4096 ? blocks_for(L0.n / 4, 256) : 4096;
    if (elem_bytes == 4) {
        if (to_split) k_split_conv<float, float, true><<<nb, 256, 0, s>>>((const float *)src, (float *)dst, g);
        else k_split_conv<float, float, false><<<nb, 256, 0, s>>>((const float *)src, (float *)dst, g);
    } else {
        if (to_split) k_split_conv<uint16_t, uint16_t, true><<<nb, 256, 0, s>>>((const uint16_t *)src, (uint16_t *)dst, g);
        else k_split_conv<uint16_t, uint16_t, false><<<nb, 256, 0, s>>>((const uint16_t *)src, (uint16_t *)dst, g);
    }
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

int split_x_from_natural(hipStream_t s, const Level &L0, const float *x, double *x64)
{
    const SGeo g = make_sgeo(L0.nx, L0.ny, L0.nz);
    const unsigned nb = blocks_for(L0.n / 4, 256) < 4096 ? blocks_for(L0.n / 4, 256) : 4096;
    k_split_conv<float, double, true><<<nb, 256, 0, s>>>(x, x64, g);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

int split_x_to_natural(hipStream_t s, const Level &L0, const double *x64, float *x, double *x64_natural, float *p,
                       float hh)
{
    const SGeo g = make_sgeo(L0.nx, L0.ny, L0.nz);
    const unsigned nb = blocks_for(L0.n / 4, 256) < 4096 ? blocks_for(L0.n / 4, 256) : 4096;
    if (x && (x64_natural || p))
        k_split_x_out<<<nb, 256, 0, s>>>(x64, x, x64_natural, g, p, hh, (float)(1.0 / (double)hh));
    else if (x) k_split_conv<double, float, false><<<nb, 256, 0, s>>>(x64, x, g);
    else if (x64_natural) k_split_conv<double, double, false><<<nb, 256, 0, s>>>(x64, x64_natural, g);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

int split_sweep(hipStream_t s, const Level &L0, const float *zin, float *zout, const float *zc, const float *r,
                const uint16_t *code, int phase, int mode, double *part, int *nblk, bool natural_out)
{
    SSweepArgs a{};
    a.zin = zin; a.zout = zout; a.zc = zc; a.r = const_cast<float *>(r); a.code = code; a.part = part;
    a.g = make_sgeo(L0.nx, L0.ny, L0.nz);
    a.cu = phase ^ 1;
    const unsigned nb = grid8((uint32_t)L0.nx * L0.ny * a.g.nq);
    if (nblk) *nblk = (int)nb;
    if (mode == SWEEP_ZERO) sweep_launch<SWEEP_ZERO, SWEEP_OUT_STORE, float, float, float, false>(s, nb, a);
    else if (mode == SWEEP_PROLONG) sweep_launch<SWEEP_PROLONG, SWEEP_OUT_STORE, float, float, float, false>(s, nb, a);
    else if (mode == SWEEP_PLAIN_OWN0) sweep_launch<SWEEP_PLAIN_OWN0, SWEEP_OUT_STORE, float, float, float, false>(s, nb, a);
    else if (natural_out) sweep_launch<SWEEP_PLAIN, SWEEP_OUT_NATURAL, float, float, float, false>(s, nb, a);
    else if (part) sweep_launch<SWEEP_PLAIN, SWEEP_OUT_DOT, float, float, float, false>(s, nb, a);
    else sweep_launch<SWEEP_PLAIN, SWEEP_OUT_STORE, float, float, float, false>(s, nb, a);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

template <int NT, typename TZI, typename TZO, typename TR>
static void pair_launch(hipStream_t s, unsigned nb, int mode, bool nat, const void *zin_, void *zout_, const float *zc,
                        const void *r_, const uint16_t *code, const SGeo &g, int cu)
{
    const TZI *zin = static_cast<const TZI *>(zin_);
    TZO *zout = static_cast<TZO *>(zout_);
    const TR *r = static_cast<const TR *>(r_);
    if (mode == SWEEP_ZERO && !nat)
        k_ssweep2<SWEEP_ZERO, false, NT, TZI, TZO, TR><<<nb, NT, 0, s>>>(zin, zout, zc, r, code, g, cu);
    else if (mode == SWEEP_PROLONG && !nat)
        k_ssweep2<SWEEP_PROLONG, false, NT, TZI, TZO, TR><<<nb, NT, 0, s>>>(zin, zout, zc, r, code, g, cu);
    else if (mode == SWEEP_PLAIN && !nat)
        k_ssweep2<SWEEP_PLAIN, false, NT, TZI, TZO, TR><<<nb, NT, 0, s>>>(zin, zout, zc, r, code, g, cu);
    else if (mode == SWEEP_PLAIN)
        k_ssweep2<SWEEP_PLAIN, true, NT, TZI, TZO, TR><<<nb, NT, 0, s>>>(zin, zout, zc, r, code, g, cu);
    else if (mode == SWEEP_PROLONG)
        k_ssweep2<SWEEP_PROLONG, true, NT, TZI, TZO, TR><<<nb, NT, 0, s>>>(zin, zout, zc, r, code, g, cu);
    else k_ssweep2<SWEEP_ZERO, true, NT, TZI, TZO, TR><<<nb, NT, 0, s>>>(zin, zout, zc, r, code, g, cu);
}

bool split_pair_supported(const Level &L)
{
    const SGeo g = make_sgeo(L.nx, L.ny, L.nz);
    return split_supported(L) && L.nx % P2_TI == 0 && L.ny % P2_TJ == 0 && g.lq >= 0 && g.nh <= P2_NHMAX;
}

int split_sweep_pair(hipStream_t s, const Level &L, const void *zin, void *zout, const float *zc, const void *r,
                     const uint16_t *code, int phase, int mode, bool natural_out, int types)
{
    FLUB_REQUIRE(split_pair_supported(L), FLUB_ESHAPE, "sweep pair: level %dx%dx%d not supported", L.nx, L.ny, L.nz);
    FLUB_REQUIRE(zin != zout || mode == SWEEP_ZERO, FLUB_EINVAL, "sweep pair: zin and zout must differ");
    FLUB_REQUIRE(mode == SWEEP_ZERO || mode == SWEEP_PLAIN || mode == SWEEP_PROLONG, FLUB_EINVAL,
                 "sweep pair: bad mode %d", mode);
    FLUB_REQUIRE(types == PAIR_F32 || types == PAIR_BF16_R || !natural_out, FLUB_EINVAL,
                 "sweep pair: the reference-layout output is FP32");
    const SGeo g = make_sgeo(L.nx, L.ny, L.nz);
    const unsigned nb = (unsigned)((L.nx / P2_TI) * (L.ny / P2_TJ));
    const int cu = phase ^ 1;
    using BF = __hip_bfloat16;
    if (types == PAIR_BF16) {   // the BF16 pre-smoothing (1024 threads)
        pair_launch<1024, BF, BF, BF>(s, nb, mode, false, zin, zout, zc, r, code, g, cu);
    } else if (types == PAIR_BF16_IN) {   // BF16 z and r in, FP32 z out (the prolongating pair)
        pair_launch<1024, BF, float, BF>(s, nb, mode, false, zin, zout, zc, r, code, g, cu);
    } else if (types == PAIR_BF16_R) {   // FP32 z, BF16 r
        if (natural_out) pair_launch<1024, float, float, BF>(s, nb, mode, true, zin, zout, zc, r, code, g, cu);
        else pair_launch<1024, float, float, BF>(s, nb, mode, false, zin, zout, zc, r, code, g, cu);
    } else {
        static const int nt = getenv("FLUB_PAIR_NT") ? atoi(getenv("FLUB_PAIR_NT")) : 1024;
        if (nt == 256) pair_launch<256, float, float, float>(s, nb, mode, natural_out, zin, zout, zc, r, code, g, cu);
        else if (nt == 512) pair_launch<512, float, float, float>(s, nb, mode, natural_out, zin, zout, zc, r, code, g, cu);
        else pair_launch<1024, float, float, float>(s, nb, mode, natural_out, zin, zout, zc, r, code, g, cu);
    }
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

int split_sweep_mixed(hipStream_t s, const Level &L0, int prec, const SplitMixedSweep &m)
{
    if (prec == FLUB_PREC_BF16) return split_sweep_mixed_t<__hip_bfloat16>(s, L0, m);
    return split_sweep_mixed_t<__half>(s, L0, m);
}

int split_restrict(hipStream_t s, const Level &F, const float *r, const float *z, const uint16_t *code,
                   const Level &C, bool csplit)
{
    const SGeo g = make_sgeo(F.nx, F.ny, F.nz);
    SGeo cg = make_sgeo(C.nx, C.ny, 2 * C.nz);   // nh = C.nz
    const unsigned nb = grid8((uint32_t)C.nx * C.ny * cg.nq);
    k_srestrict<float, float, 4><<<nb, 256, 0, s>>>(r, z, code, C.r, C.tag, g, cg, (int)csplit, C.cs);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

int split_restrict_mixed(hipStream_t s, const Level &F, int prec, const void *rh, const void *zh,
                         const uint16_t *code, const Level &C, bool csplit, void *rc16)
{
    // 4 coarse cells per thread: at 8 (16-byte loads) the kernel needs 256
    // VGPRs and runs 53 us instead of 40 us at 256^3
    const SGeo g = make_sgeo(F.nx, F.ny, F.nz);
    SGeo cg = make_sgeo(C.nx, C.ny, 2 * C.nz);
    const unsigned nb = grid8((uint32_t)C.nx * C.ny * cg.nq);
    using BF = __hip_bfloat16;
    if (prec == FLUB_PREC_BF16 && rc16)   // the coarse r in BF16 (the BF16 level-1 pre-smoothing)
        k_srestrict<BF, BF, 4, BF><<<nb, 256, 0, s>>>((const BF *)rh, (const BF *)zh, code, (BF *)rc16, C.tag, g, cg,
                                                      (int)csplit, C.cs);
    else if (prec == FLUB_PREC_BF16)
        k_srestrict<BF, BF, 4><<<nb, 256, 0, s>>>((const BF *)rh, (const BF *)zh, code, C.r, C.tag, g, cg,
                                                  (int)csplit, C.cs);
    else
        k_srestrict<__half, __half, 4><<<nb, 256, 0, s>>>((const __half *)rh, (const __half *)zh, code, C.r, C.tag,
                                                          g, cg, (int)csplit, C.cs);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

int split_Ap(hipStream_t s, const Level &L0, const float *p, float *Ap, const uint16_t *code, double *part, int *nblk)
{
    const SGeo g = make_sgeo(L0.nx, L0.ny, L0.nz);
    const unsigned nb = grid8((uint32_t)L0.nx * L0.ny * g.nq);
    *nblk = (int)nb;
    k_sp_Ap<<<nb, 256, 0, s>>>(p, Ap, code, g, part);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

// Launch plan of k_sp_pAp: nq a power of two in [4, 64] (256 threads hold
// whole half-rows of jb = 256 / nq rows), ny a multiple of jb, at most MAXBLK
// blocks.  ib <= 0 picks the rows per segment: from 16, doubled while the
// tiles exceed MAXBLK, halved (to 2) while they are fewer than 512.  (At 256^3,
// 8 x 32-row tiles: the step with its finish takes 81 us in two passes, 67,
// 60, 54.7 and 58.8 us with 4, 8, 16 and 32 rows per segment.)
struct PApPlan {
    int lq, jb, ib, nseg;
    uint32_t ntiles, tpx, nblocks;
    size_t shmem;
};
static bool pap_plan(const Level &L0, int ib, PApPlan &pl)
{
    const SGeo g = make_sgeo(L0.nx, L0.ny, L0.nz);
    if (g.nq < 4 || g.nq > 64 || g.lq < 0 || g.nh != 4 * g.nq || L0.nx < 2) return false;
    pl.lq = g.lq;
    pl.jb = 256 / g.nq;
    if (L0.ny % pl.jb != 0) return false;
    const uint64_t njb = (uint64_t)(L0.ny / pl.jb);
    auto tiles = [&](int b) { return (uint64_t)((L0.nx + b - 1) / b) * njb; };
    if (ib <= 0) {
        ib = 16;
        while (ib < L0.nx && tiles(ib) > MAXBLK) ib *= 2;
        while (ib > 2 && tiles(ib) < 512) ib /= 2;
    }
    pl.ib = ib;
    pl.nseg = (L0.nx + ib - 1) / ib;
    pl.ntiles = (uint32_t)tiles(ib);
    pl.tpx = (pl.ntiles + 7u) / 8u;
    pl.nblocks = 8u * pl.tpx;
    if (pl.nblocks > MAXBLK) return false;
    pl.shmem = 3u * (size_t)(pl.jb + 2) * 2u * (size_t)g.nq * sizeof(float4);
    return true;
}

int split_pstep_Ap(hipStream_t s, PxState &st, const float *z, const double *zinv, const Level &L0, float *Ap,
                   const uint16_t *code, double *part, int *nblk, int ib)
{
    PApPlan pl;
    if (ib < 0 || !pap_plan(L0, ib, pl)) {
        const int rc = px_step(s, st, z, zinv);
        if (rc) return rc;
        return split_Ap(s, L0, st.p(), Ap, code, part, nblk);
    }
    int out;
    bool save;
    const int rc = px_step_slot(s, st, &out, &save);
    if (rc) return rc;
    PApArgs a{};
    a.p_out = st.ring[out];
    a.p_prev = st.ring[st.cur];
    a.z = z;
    a.sc = st.sc;
    a.zinv = zinv;
    a.asave = save ? st.ahist + st.cur : nullptr;
    a.Ap = Ap;
    a.code = code;
    a.part = part;
    a.g = make_sgeo(L0.nx, L0.ny, L0.nz);
    a.lq = pl.lq;
    a.jb = pl.jb;
    a.ib = pl.ib;
    a.nseg = pl.nseg;
    a.ntiles = pl.ntiles;
    a.tpx = pl.tpx;
    k_sp_pAp<<<pl.nblocks, 256, pl.shmem, s>>>(a);
    FLUB_CHECK_LAUNCH();
    px_step_done(st, out, save);
    *nblk = (int)pl.nblocks;
    return FLUB_OK;
}

namespace {
__global__ void k_scale_by(float *__restrict__ x, const double *__restrict__ f, uint32_t n)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t < n) x[t] = x[t] * (float)f[0];
}
}  // namespace

int scale_by_device(hipStream_t s, float *x, const double *f, uint32_t n)
{
    k_scale_by<<<blocks_for(n, 256), 256, 0, s>>>(x, f, n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

int split_true_residual(hipStream_t s, const Level &L0, float *r, const float *b, const double *x,
                        const uint16_t *code, double *part, int *nblk)
{
    const SGeo g = make_sgeo(L0.nx, L0.ny, L0.nz);
    const unsigned nb = grid8((uint32_t)L0.nx * L0.ny * g.nq);
    *nblk = (int)nb;
    k_strue_residual<double><<<nb, 256, 0, s>>>(r, b, x, code, g, part);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}
