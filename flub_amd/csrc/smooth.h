// Fast tagged multigrid kernels (smooth.hip), used by the FP32 and the
// mixed-precision solvers.
#pragma once
#include "mgpcg_internal.h"

enum { SWEEP_PLAIN = 0, SWEEP_ZERO = 1, SWEEP_PROLONG = 2 };
enum { SWEEP_OUT_STORE = 0, SWEEP_OUT_DOT = 1, SWEEP_OUT_FINAL = 2 };

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
