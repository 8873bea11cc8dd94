// Fast tagged multigrid kernels (smooth.hip), called by the FP32 solver.
#pragma once
#include "mgpcg_internal.h"

enum { SWEEP_PLAIN = 0, SWEEP_ZERO = 1, SWEEP_PROLONG = 2 };

// One red-black half-sweep of level L over quads: z read from zin, written to
// zout (zin == zout except for SWEEP_PROLONG, which adds zc[parent]).  With
// part != NULL the block partials of z.r are written and *nblk set.
int fast_sweep(hipStream_t s, const Level &L, const float *zin, float *zout, const float *zc, int phase, int mode,
               double *part, int *nblk);
// Gather restriction of level F (z = zf) into level C.
int fast_restrict(hipStream_t s, const Level &F, const float *zf, const Level &C);
// pout = z + beta*p; Ap = ndt(pout); partials of pout.Ap.
int fast_p_Ap(hipStream_t s, const Level &L0, const float *p, float *pout, float *Ap, const double *sc,
              double *part, int *nblk);
