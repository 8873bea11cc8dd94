// Fast tagged multigrid kernels (smooth.hip), used by the FP32 and the
// mixed-precision solvers.
#pragma once
#include "mgpcg_internal.h"

// SWEEP_PLAIN_OWN0 (split sweeps only): PLAIN with the updated colour's old
// z taken as +0 (the sweep after a SWEEP_ZERO one, which then need not store
// the other colour's zeros).
enum { SWEEP_PLAIN = 0, SWEEP_ZERO = 1, SWEEP_PROLONG = 2, SWEEP_PLAIN_OWN0 = 3 };
// SWEEP_OUT_NATURAL (split sweeps only): the whole z after the half-sweep, both
// colours, in the reference layout into zout (a buffer other than zin).
enum { SWEEP_OUT_STORE = 0, SWEEP_OUT_DOT = 1, SWEEP_OUT_FINAL = 2, SWEEP_OUT_NATURAL = 3 };

// One half-sweep of a level in the storage type of `prec`.
struct SweepInfo {
    const void *zin;     // z read (unused for SWEEP_ZERO)
    void *zout;          // z written (SWEEP_OUT_STORE / SWEEP_OUT_DOT)
    const void *zc;      // coarse z (SWEEP_PROLONG)
    float pfac;          // prolongation factor
    const void *rsrc;    // level r (storage type), or the FP32 residual when src32
    void *rh_out;        // src32: scaled r stored here
    const double *sc;    // device scalars (scale of r: src32 / final)
    const float *r32, *q32;
    float *z32;          // SWEEP_OUT_FINAL output
    double *part;        // partials (SWEEP_OUT_DOT: z.r; FINAL: z.r then z.q)
    int phase, mode, out;
    bool src32;
};

int sweep(hipStream_t s, const Level &L, int prec, const SweepInfo &si, int *nblk);
// Gather restriction of level F (r = rf, z = zf) into rc of level C, both in
// the storage type of prec (coarse values times 1/8 for reduced precision).
int restrict_level(hipStream_t s, const Level &F, const void *rf, const void *zf, const Level &C, void *rc, int prec);
// 16-bit stencil codes of a level from its tags.
int make_codes(hipStream_t s, const Level &L);

// FP32 conveniences.
int fast_sweep(hipStream_t s, const Level &L, const float *zin, float *zout, const float *zc, int phase, int mode,
               double *part, int *nblk);
int fast_restrict(hipStream_t s, const Level &F, const float *zf, const Level &C);
// pout = z + beta*p; Ap = ndt(pout); partials of pout.Ap.
int fast_p_Ap(hipStream_t s, const Level &L0, const float *p, float *pout, float *Ap, const double *sc,
              double *part, int *nblk);

// Temporally blocked smoothing (blocked.hip): up to 4 or 8 half-sweeps per
// launch, depending on the level's tile configuration.
#define TB_GMAX 8
// The level's shape is a multiple of one of the tile shapes.
bool tb_supported(const Level &L);
// Launches needed for nsteps half-sweeps of level L.
int tb_launches(const Level &L, int nsteps, bool allow_g8);
// nsteps half-sweeps (phase of step s: phases[s]) of level L in the storage
// type of prec; the first starts as mode0 (SWEEP_ZERO, SWEEP_PLAIN from zin,
// SWEEP_PROLONG from zin + pfac * zc[parent]).  The launches alternate
// between the buffers dst and other, the last one writing dst; the first one
// must not write zin (zin != the first buffer written, which is dst when
// tb_launches(L, nsteps, allow_g8) is odd, else other), and each launch
// must contain both phases.  allow_g8 admits the 8-step configurations.
// Device copy of the tiles' update-order tables (one per handle).
int tb_tables_create(uint32_t **dev);
// With part (FP32; the last launch not a ZERO one) the last launch also writes
// per-workgroup FP64 partials of z.r over the level into part and their count
// into *nblk.
int tb_chain(hipStream_t s, const uint32_t *tables, const Level &L, int prec, int mode0, const void *zin,
             const void *zc, float pfac, const void *r, const uint8_t *phases, int nsteps, void *dst, void *other,
             double *part = nullptr, int *nblk = nullptr, bool allow_g8 = true);

#define MG_MAX_PHASES (4 * 4096)
// Levels above this many cells are bandwidth-bound: one launch per
// half-sweep streams them better than the blocked kernel, whose halo
// recomputation then costs more than the launches it saves.
#define TB_MAX_CELLS (1u << 21)
// Blocked smoothing for nsteps half-sweeps of level l (level 0 only when the
// caller allows it and it is not switched off).
static inline bool mg_use_tb(const flub_mgpcg *h, int l, int nsteps, bool level0 = false)
{
    return !h->no_tb && h->tb_order && (l >= 1 || (level0 && !h->no_tb_l0)) && nsteps >= 1 &&
           nsteps <= MG_MAX_PHASES && h->lv[l].n <= TB_MAX_CELLS && tb_supported(h->lv[l]);
}
// ph = first, 1-first repeated `pairs` times (pre-smoothing mgpcg.py:303-309
// with first 0, post-smoothing mgpcg.py:342-348 with first 1); returns the
// step count.
static inline int mg_phase_pairs(uint8_t *ph, int pairs, int first)
{
    for (int i = 0; i < pairs; ++i) { ph[2 * i] = (uint8_t)first; ph[2 * i + 1] = (uint8_t)(1 - first); }
    return 2 * pairs;
}
// Bottom schedule (mgpcg.py:319-336): (0, 1) x nb, then (1, 0) x nb.
static inline int mg_bottom_phases(uint8_t *ph, int nb)
{
    mg_phase_pairs(ph, nb, 0);
    mg_phase_pairs(ph + 2 * nb, nb, 1);
    return 4 * nb;
}

// Colour-split levels (split.hip): level 0 inside the fused solves, and the
// next levels of the split V-cycle while they are large enough to stream.
bool split_supported(const Level &L0);
// Reference layout <-> split layout of a level's array of 4- or 2-byte values.
int split_convert(hipStream_t s, const Level &L0, const void *src, void *dst, int elem_bytes, bool to_split);
// One half-sweep of a split level (mode SWEEP_ZERO / PLAIN / PLAIN_OWN0 /
// PROLONG from zin + zc[parent], zc in the reference layout); with part, also
// the partials of z.r (count in *nblk); with natural_out, zout receives the
// whole z in the reference layout (SWEEP_OUT_NATURAL).
int split_sweep(hipStream_t s, const Level &L0, const float *zin, float *zout, const float *zc, const float *r,
                const uint16_t *code, int phase, int mode, double *part, int *nblk, bool natural_out = false);
// Two half-sweeps of a split level in one launch, phase `phase` then the
// other: mode SWEEP_ZERO (then SWEEP_PLAIN_OWN0), SWEEP_PLAIN or
// SWEEP_PROLONG (from zin + zc[parent]) for the first, SWEEP_PLAIN for the
// second; from zin into zout (a different buffer; both colours, or with
// natural_out the whole z in the reference layout).  Bitwise the two
// split_sweep launches.  `types` selects the storage: PAIR_F32 (z and r FP32),
// PAIR_BF16 (z in and out and r BF16), PAIR_BF16_IN (BF16 z in and r, FP32 z
// out), PAIR_BF16_R (FP32 z, BF16 r); the 16-bit forms divide through the
// hardware reciprocal like the level-0 16-bit sweeps.
enum { PAIR_F32 = 0, PAIR_BF16 = 1, PAIR_BF16_IN = 2, PAIR_BF16_R = 3 };
bool split_pair_supported(const Level &L);
int split_sweep_pair(hipStream_t s, const Level &L, const void *zin, void *zout, const float *zc, const void *r,
                     const uint16_t *code, int phase, int mode, bool natural_out = false, int types = PAIR_F32);
// Mixed-precision level-0 sweeps on the split layout (T = FP16 or BF16 by
// prec): pre-smoothing stores z and r as T (the first two sweeps take r from
// the FP32 CG residual times the storage scale), post-smoothing keeps z in
// FP32; the last one writes the CG's z = z / scale and the partials of z.r and
// z.r_old (r_old: the r the previous V-cycle stored).  SMX_PRE_ZERO_RH and
// SMX_PRE_OWN0_RH are the first two sweeps with the T r already stored (BF16
// at storage scale 1, written by the r update).
enum { SMX_PRE_ZERO_SRC = 0, SMX_PRE_SRC, SMX_PRE, SMX_POST_PROLONG, SMX_POST, SMX_FINAL, SMX_PRE_ZERO_RH,
       SMX_PRE_OWN0_RH };
struct SplitMixedSweep {
    int kind, phase;
    const void *zin;       // T (pre, post-prolong) or FP32 (post)
    void *zout;            // T (pre) or FP32 (post)
    const float *zc;       // level-1 z in the reference layout (post-prolong)
    void *rh;              // T split r
    const uint16_t *code;
    const float *r32;
    const void *rold;      // FINAL: T split r of the previous V-cycle (its r times its scale)
    const double *zinv_old;  // FINAL: 1/scale of rold
    double *zinv;          // FINAL: receives 1/scale (zout keeps z * scale)
    const double *sc;
    double *part;
    int *nblk;
};
int split_sweep_mixed(hipStream_t s, const Level &L0, int prec, const SplitMixedSweep &m);
// Restriction of a split level (r, z stored as T) into the next level's FP32
// r, colour-split when csplit; with rc16 (BF16 only) into that BF16 buffer.
int split_restrict_mixed(hipStream_t s, const Level &F, int prec, const void *rh, const void *zh,
                         const uint16_t *code, const Level &C, bool csplit = false, void *rc16 = nullptr);
// Restriction of a split level (r, z) into the next level's r, colour-split
// when csplit.
int split_restrict(hipStream_t s, const Level &F, const float *r, const float *z, const uint16_t *code,
                   const Level &C, bool csplit = false);
// Ap = ndt(p) and the partials of p.Ap (both split); *nblk: partial count.
int split_Ap(hipStream_t s, const Level &L0, const float *p, float *Ap, const uint16_t *code, double *part, int *nblk);
// x *= f[0] (a device scalar).
int scale_by_device(hipStream_t s, float *x, const double *f, uint32_t n);
// r = b - ndt(x) in FP64 from the FP64 x (stored FP32), partials of r.r and
// max|r| (split; r, b and x are distinct buffers).
int split_true_residual(hipStream_t s, const Level &L0, float *r, const float *b, const double *x,
                        const uint16_t *code, double *part, int *nblk);
// The FP64 solution between the reference layout and the split layout:
// x (FP32, reference layout) -> x64 (split); x64 (split) -> x (FP32, rounded
// once) and/or x64_natural (FP64, reference layout), either may be null.
int split_x_from_natural(hipStream_t s, const Level &L0, const float *x, double *x64);
int split_x_to_natural(hipStream_t s, const Level &L0, const double *x64, float *x, double *x64_natural,
                       float *p = nullptr, float hh = 1.0f);
