/*
 * TEST INFRASTRUCTURE -- CPU oracle for the pressure-projection path
 * (divergence -> MGPCG -> gradient subtraction) of the reference Warp code:
 *   divergence.py:44-99,174-180, multigrid/mgpcg.py:9-499,
 *   subtract_gradient.py:38-68, utils/util.py:20-28,162-165,
 *   geometry.py:335-344,390-397, example_fluid.py:84-147, water_3d.py:58-68.
 *
 * It is the checker and the CPU baseline, never part of the shipped path:
 * only tests/, the smoke check and the CPU-baseline leg of bench.py load it.
 * Built with -O3 -ffp-contract=off so every expression rounds exactly as
 * written; the one fused multiply-add (the CG axpys) is spelled fma()/fmaf().
 *
 * Parity pins (tests/test_oracle_*.py): restrict_tag against min over all 3^8
 * child combinations; divd and the operator against an independently assembled
 * sparse graph Laplacian; the smoother against textbook Gauss-Seidel on that
 * matrix; restriction/prolongation against reshape sums and R = P^T; the V-cycle
 * against symmetry and positivity; the solve against scipy spsolve in FP64;
 * divergence/subtract-gradient against adjointness, telescoping and the
 * post-projection-divergence == residual invariant.  The Warp FP32 summation
 * order of array_inner cannot be pinned (Warp is not installed): parity
 * unpinned for bitwise FP32 iteration histories, pinned to tolerance instead.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXL 12
#define ORC_SAFETY_CAP 20000 /* iteration cap when max_iters == -1 */
#define IDX(i, j, k) ((((size_t)(i)) * (size_t)ny + (size_t)(j)) * (size_t)nz + (size_t)(k))

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

#define REAL float
#define SUF _f32
#define FMA fmaf
#include "oracle_body.h"
#undef REAL
#undef SUF
#undef FMA

#define REAL double
#define SUF _f64
#define FMA fma
#include "oracle_body.h"
#undef REAL
#undef SUF
#undef FMA

/* geometry.py:335-344 _tag_dirichlet: phi > 0 -> AIR (0), else FLUID (1). */
void orc_tag_dirichlet(const float *phi, int32_t *tag, size_t n)
{
    for (size_t I = 0; I < n; ++I) tag[I] = phi[I] > 0.0f ? 0 : 1;
}

/* geometry.py:390-397 _tag_neumann: phi_rigid < 0 -> SOLID (2), else unchanged. */
void orc_tag_neumann(const float *phi_rigid, int32_t *tag, size_t n)
{
    for (size_t I = 0; I < n; ++I)
        if (phi_rigid[I] < 0.0f) tag[I] = 2;
}

int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
