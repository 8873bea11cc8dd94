/*
 * TEST INFRASTRUCTURE -- CPU oracle body, included twice by flub_oracle.c with
 *   REAL = float  (SUF = _f32)  : the FP32 restatement of the reference path
 *   REAL = double (SUF = _f64)  : its FP64 twin, used as the accuracy judge.
 * Nothing in the product links or calls this code; only tests/, the smoke check
 * and the CPU-baseline leg of bench.py do.
 *
 * Every routine restates one Warp kernel of the reference, in the reference's
 * launch order and floating-point operation order, serially (as Warp's CPU
 * backend runs a launch), with optional OpenMP over cells where the result does
 * not depend on the order (red-black colours, per-cell stencils, gathers).
 * Layout: C-contiguous [i][j][k], k fastest, idx = ((i*ny)+j)*nz+k.
 * Tags: int32, 0 = AIR, 1 = FLUID, 2 = SOLID.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

/* mgpcg.py:118-142 neighbor_dot_tag.  Returns sum over axes of
 * (x'_N+ - x_I)*[tag_N+<2] + (x'_N- - x_I)*[tag_N-<2] with clamped neighbours,
 * x'_N = 0 where tag_N == 0; 0 when tag_I != 1. */
static inline REAL FN(ndt)(const REAL *x, const int32_t *tag, int nx, int ny, int nz,
                           int i, int j, int k)
{
    const size_t I = IDX(i, j, k);
    const REAL xI = x[I];
    REAL ret = (REAL)0.0;
    for (int a = 0; a < 3; ++a) {
        int pi = i, pj = j, pk = k, mi = i, mj = j, mk = k;
        if (a == 0) { pi = imin(i + 1, nx - 1); mi = imax(i - 1, 0); }
        if (a == 1) { pj = imin(j + 1, ny - 1); mj = imax(j - 1, 0); }
        if (a == 2) { pk = imin(k + 1, nz - 1); mk = imax(k - 1, 0); }
        const size_t P = IDX(pi, pj, pk), M = IDX(mi, mj, mk);
        const REAL fp = tag[P] < 2 ? (REAL)1.0 : (REAL)0.0;
        const REAL fm = tag[M] < 2 ? (REAL)1.0 : (REAL)0.0;
        const REAL xp = tag[P] == 0 ? (REAL)0.0 : x[P];
        const REAL xm = tag[M] == 0 ? (REAL)0.0 : x[M];
        ret += (xp - xI) * fp + (xm - xI) * fm;
    }
    if (tag[I] != 1) ret = (REAL)0.0;
    return ret;
}

/* mgpcg.py:15-28 neighbor_sum (tag-free operator of the "default" smoother). */
static inline REAL FN(nsum)(const REAL *x, int nx, int ny, int nz, int i, int j, int k)
{
    const size_t I = IDX(i, j, k);
    REAL ret = (REAL)0.0;
    for (int a = 0; a < 3; ++a) {
        int pi = i, pj = j, pk = k, mi = i, mj = j, mk = k;
        if (a == 0) { pi = imin(i + 1, nx - 1); mi = imax(i - 1, 0); }
        if (a == 1) { pj = imin(j + 1, ny - 1); mj = imax(j - 1, 0); }
        if (a == 2) { pk = imin(k + 1, nz - 1); mk = imax(k - 1, 0); }
        ret += x[IDX(pi, pj, pk)] + x[IDX(mi, mj, mk)] - x[I] * (REAL)2.0;
    }
    return ret;
}

/* mgpcg.py:51-58 _compute_divd with neighbor_sum_tag (mgpcg.py:144-160). */
void FN(orc_compute_divd)(const int32_t *tag, REAL *divd, int nx, int ny, int nz)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const size_t I = IDX(i, j, k);
                if (tag[I] == 1) {
                    int c = 0;
                    if (i + 1 < nx && tag[IDX(i + 1, j, k)] < 2) c++;
                    if (i - 1 >= 0 && tag[IDX(i - 1, j, k)] < 2) c++;
                    if (j + 1 < ny && tag[IDX(i, j + 1, k)] < 2) c++;
                    if (j - 1 >= 0 && tag[IDX(i, j - 1, k)] < 2) c++;
                    if (k + 1 < nz && tag[IDX(i, j, k + 1)] < 2) c++;
                    if (k - 1 >= 0 && tag[IDX(i, j, k - 1)] < 2) c++;
                    divd[I] = (REAL)c;
                } else {
                    divd[I] = (REAL)0.0;
                }
            }
}

/* mgpcg.py:36-39 compute_Ap_mcadams: Ap = neighbor_dot_tag(p) (= -A p). */
void FN(orc_compute_Ap_mcadams)(REAL *Ap, const REAL *p, const int32_t *tag, int nx, int ny, int nz)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
                Ap[IDX(i, j, k)] = FN(ndt)(p, tag, nx, ny, nz, i, j, k);
}

/* mgpcg.py:30-33 compute_Ap (tag-free). */
void FN(orc_compute_Ap_default)(REAL *Ap, const REAL *p, int nx, int ny, int nz)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
                Ap[IDX(i, j, k)] = FN(nsum)(p, nx, ny, nz, i, j, k);
}

/* mgpcg.py:72-86 _smooth_mcadams, launch dim (nx, ny, nz/2): phase p updates
 * the cells with (i+j+k)&1 == p^1.  Cells of one colour read only the other
 * colour and themselves, so the result is independent of the visiting order. */
void FN(orc_smooth_mcadams)(REAL *z, const REAL *r, const int32_t *tag, const REAL *divd,
                            int nx, int ny, int nz, int phase)
{
    const int nzh = nz / 2;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int kk = 0; kk < nzh; ++kk) {
                const int k = (((i + j) & 1) == phase) ? kk * 2 + 1 : kk * 2;
                const size_t I = IDX(i, j, k);
                const REAL d = divd[I];
                if (d != (REAL)0.0)
                    z[I] = z[I] + (r[I] + FN(ndt)(z, tag, nx, ny, nz, i, j, k)) / d;
                else
                    z[I] = (REAL)0.0;
            }
}

/* mgpcg.py:41-49 _smooth (tag-free "default" smoother).  Note the reference uses
 * res = r.shape[0] on all three axes and updates (i+j+k)&1 == phase. */
void FN(orc_smooth_default)(REAL *z, const REAL *r, int nx, int ny, int nz, int phase)
{
    const int res = nx;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                if (((i + j + k) & 1) != phase) continue;
                const REAL dx = (i == 0 || i == res - 1) ? (REAL)1.0 : (REAL)2.0;
                const REAL dy = (j == 0 || j == res - 1) ? (REAL)1.0 : (REAL)2.0;
                const REAL dz = (k == 0 || k == res - 1) ? (REAL)1.0 : (REAL)2.0;
                const size_t I = IDX(i, j, k);
                z[I] = z[I] + (r[I] + FN(nsum)(z, nx, ny, nz, i, j, k)) / (dx + dy + dz);
            }
}

/* mgpcg.py:94-101 _restrict_mcadams, restated literally: a serial launch over
 * the fine grid in row-major order, atomic_add into a zeroed coarse array when
 * the parent is FLUID, plain store of 0 otherwise.  nx,ny,nz are FINE dims. */
void FN(orc_restrict_mcadams_serial)(const REAL *r_h, const REAL *z_h, const int32_t *tag_h,
                                     REAL *r_2h, const int32_t *tag_2h, int nx, int ny, int nz)
{
    const int cy = ny / 2, cz = nz / 2;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const REAL res = r_h[IDX(i, j, k)] - (-FN(ndt)(z_h, tag_h, nx, ny, nz, i, j, k));
                const size_t P = ((size_t)(i / 2) * cy + (j / 2)) * cz + (k / 2);
                if (tag_2h[P] == 1) r_2h[P] += res;
                else r_2h[P] = (REAL)0.0;
            }
}

/* Same result as the serial restatement above, computed as a gather: each
 * parent sums its 8 children in the order the serial launch visits them
 * ((ii,jj,kk) lexicographic, kk fastest), starting from the zeroed value. */
void FN(orc_restrict_mcadams)(const REAL *r_h, const REAL *z_h, const int32_t *tag_h,
                              REAL *r_2h, const int32_t *tag_2h, int nx, int ny, int nz)
{
    const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < cx; ++I)
        for (int J = 0; J < cy; ++J)
            for (int K = 0; K < cz; ++K) {
                const size_t P = ((size_t)I * cy + J) * cz + K;
                if (tag_2h[P] != 1) { r_2h[P] = (REAL)0.0; continue; }
                REAL s = (REAL)0.0;
                for (int ii = 0; ii < 2; ++ii)
                    for (int jj = 0; jj < 2; ++jj)
                        for (int kk = 0; kk < 2; ++kk) {
                            const int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                            s += r_h[IDX(i, j, k)] - (-FN(ndt)(z_h, tag_h, nx, ny, nz, i, j, k));
                        }
                r_2h[P] = s;
            }
}

/* mgpcg.py:88-92 _restrict (tag-free), as a gather in serial order. */
void FN(orc_restrict_default)(const REAL *r_h, const REAL *z_h, REAL *r_2h, int nx, int ny, int nz)
{
    const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < cx; ++I)
        for (int J = 0; J < cy; ++J)
            for (int K = 0; K < cz; ++K) {
                REAL s = (REAL)0.0;
                for (int ii = 0; ii < 2; ++ii)
                    for (int jj = 0; jj < 2; ++jj)
                        for (int kk = 0; kk < 2; ++kk) {
                            const int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                            s += r_h[IDX(i, j, k)] - (-FN(nsum)(z_h, nx, ny, nz, i, j, k));
                        }
                r_2h[((size_t)I * cy + J) * cz + K] = s;
            }
}

/* mgpcg.py:113-116 _prolongate.  nx,ny,nz are FINE dims. */
void FN(orc_prolongate)(REAL *z_h, const REAL *z_2h, int nx, int ny, int nz)
{
    const int cy = ny / 2, cz = nz / 2;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
                z_h[IDX(i, j, k)] = z_h[IDX(i, j, k)] + z_2h[((size_t)(i / 2) * cy + j / 2) * cz + k / 2];
}

/* divergence.py:44-99 divergence_tag.  The reference takes res = shape[0] for
 * every axis (cubic grids only); this restatement uses each axis' own extent,
 * which is identical on cubic grids. */
void FN(orc_divergence_tag)(const REAL *u, const REAL *v, const REAL *w, REAL *div,
                            const REAL *ur, const REAL *vr, const REAL *wr,
                            const int32_t *tag, int nx, int ny, int nz)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const size_t iu0 = ((size_t)i * ny + j) * nz + k;          /* u (nx+1,ny,nz) */
                const size_t iu1 = ((size_t)(i + 1) * ny + j) * nz + k;
                const size_t iv0 = ((size_t)i * (ny + 1) + j) * nz + k;    /* v (nx,ny+1,nz) */
                const size_t iv1 = ((size_t)i * (ny + 1) + j + 1) * nz + k;
                const size_t iw0 = ((size_t)i * ny + j) * (nz + 1) + k;    /* w (nx,ny,nz+1) */
                const size_t iw1 = iw0 + 1;
                REAL vl = u[iu0], vr_ = u[iu1], vb = v[iv0], vt = v[iv1], vf = w[iw0], vn = w[iw1];
                if (i == 0) vl = (REAL)0.0;
                else if (tag[IDX(i - 1, j, k)] == 2) vl = ur[iu0];
                if (i == nx - 1) vr_ = (REAL)0.0;
                else if (tag[IDX(i + 1, j, k)] == 2) vr_ = ur[iu1];
                if (j == 0) vb = (REAL)0.0;
                else if (tag[IDX(i, j - 1, k)] == 2) vb = vr[iv0];
                if (j == ny - 1) vt = (REAL)0.0;
                else if (tag[IDX(i, j + 1, k)] == 2) vt = vr[iv1];
                if (k == 0) vf = (REAL)0.0;
                else if (tag[IDX(i, j, k - 1)] == 2) vf = wr[iw0];
                if (k == nz - 1) vn = (REAL)0.0;
                else if (tag[IDX(i, j, k + 1)] == 2) vn = wr[iw1];
                div[IDX(i, j, k)] = (vr_ - vl + vt - vb + vn - vf);
            }
}

/* divergence.py:174-180 mask_with_tag. */
void FN(orc_mask_with_tag)(REAL *div, const int32_t *tag, size_t n)
{
#pragma omp parallel for schedule(static)
    for (long long I = 0; I < (long long)n; ++I)
        if (tag[I] != 1) div[I] = (REAL)0.0;
}

/* utils/util.py:162-165 scale_by_3d and mgpcg.py:9-12 _scale_by_3d. */
void FN(orc_scale)(REAL *q, REAL s, size_t n)
{
#pragma omp parallel for schedule(static)
    for (long long I = 0; I < (long long)n; ++I) q[I] = q[I] * s;
}

/* mgpcg.py:492-499 _mask_interior. */
void FN(orc_mask_interior)(const int32_t *tag, REAL *r, size_t n)
{
#pragma omp parallel for schedule(static)
    for (long long I = 0; I < (long long)n; ++I)
        if (tag[I] != 1) r[I] = (REAL)0.0;
}

/* subtract_gradient.py:38-68 subtract_gradient_tag over the face grid of axis d
 * (d = 1, 2, 3 for u, v, w); uf and u_rigid have the face shape of that axis. */
void FN(orc_subtract_gradient_tag)(REAL *uf, const REAL *u_rigid, const REAL *p, const int32_t *tag,
                                   int nx, int ny, int nz, int d)
{
    const int fx = nx + (d == 1), fy = ny + (d == 2), fz = nz + (d == 3);
    const int n_d = d == 1 ? nx : (d == 2 ? ny : nz);   /* uf.shape[d-1] - 1 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < fx; ++i)
        for (int j = 0; j < fy; ++j)
            for (int k = 0; k < fz; ++k) {
                const size_t F = ((size_t)i * fy + j) * fz + k;
                const int ii = d == 1 ? i : (d == 2 ? j : k);
                if (0 < ii && ii < n_d) {
                    const int li = i - (d == 1), lj = j - (d == 2), lk = k - (d == 3);
                    const REAL pl = p[IDX(li, lj, lk)], pr = p[IDX(i, j, k)];
                    const int tl = tag[IDX(li, lj, lk)], tr = tag[IDX(i, j, k)];
                    if (tl == 2 || tr == 2) uf[F] = u_rigid[F];
                    else if (tl == 1 || tr == 1) uf[F] = uf[F] - (pr - pl);
                    else uf[F] = (REAL)0.0;
                } else {
                    uf[F] = (REAL)0.0;
                }
            }
}

/* mgpcg.py:191-205 _mean then _sub (zero_centered, mgpcg.py:270-274): float
 * atomic sum in serial launch order, int count, then q -= mean on FLUID and
 * q = 0 elsewhere. */
void FN(orc_zero_centered)(REAL *q, const int32_t *tag, size_t n)
{
    REAL sum = (REAL)0.0;
    int cnt = 0;
    for (size_t I = 0; I < n; ++I)
        if (tag[I] == 1) { sum += q[I]; cnt += 1; }
    const REAL mean = sum / (REAL)cnt;
#pragma omp parallel for schedule(static)
    for (long long I = 0; I < (long long)n; ++I)
        q[I] = tag[I] == 1 ? q[I] - mean : (REAL)0.0;
}

/* zero_centered with the sum of the mean accumulated in double (mean_f64):
 * the same serial order and subtraction; mean = (REAL)sum / (REAL)count, as
 * the HIP path forms it. */
static void FN(zero_centered_f64)(REAL *q, const int32_t *tag, size_t n)
{
    double sum = 0.0;
    int cnt = 0;
    for (size_t I = 0; I < n; ++I)
        if (tag[I] == 1) { sum += (double)q[I]; cnt += 1; }
    const REAL mean = (REAL)sum / (REAL)cnt;
#pragma omp parallel for schedule(static)
    for (long long I = 0; I < (long long)n; ++I)
        q[I] = tag[I] == 1 ? q[I] - mean : (REAL)0.0;
}

/* warp.utils.array_inner (mgpcg.py:5): sum of a*b.  The accumulation type is a
 * parameter: dot_f64 = 1 accumulates in double (the design this build ships),
 * 0 in REAL, both in index order. */
double FN(orc_inner)(const REAL *a, const REAL *b, size_t n, int dot_f64)
{
    if (dot_f64) {
        double s = 0.0;
        for (size_t I = 0; I < n; ++I) s += (double)a[I] * (double)b[I];
        return s;
    }
    REAL s = (REAL)0.0;
    for (size_t I = 0; I < n; ++I) s += a[I] * b[I];
    return (double)s;
}

static double FN(inner_par)(const REAL *a, const REAL *b, size_t n)
{
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (long long I = 0; I < (long long)n; ++I) s += (double)a[I] * (double)b[I];
    return s;
}

/* -------------------------------------------------------------------------
 * MGPCG (mgpcg.py:234-490), same levels, buffers and control flow.
 * ------------------------------------------------------------------------- */
typedef struct {
    int L, pre_post, bottom, growth;   /* sweeps at level l: pre_post << (growth * l) */
    int cap;           /* > 0: at most cap sweeps on any level (no reference counterpart) */
    REAL cs;           /* restricted residuals stored times cs (1: the reference) */
    int nx[ORC_MAXL], ny[ORC_MAXL], nz[ORC_MAXL];
    size_t n[ORC_MAXL];
    REAL *r[ORC_MAXL], *z[ORC_MAXL], *divd[ORC_MAXL];
    int32_t *tag[ORC_MAXL];
    REAL *x, *p, *Ap;
    double alpha, beta, sum, old_zTr;   /* device scalars of the reference */
    int dot_f64;       /* 1: inner products accumulate in double */
    int parallel_dot;  /* 1: OpenMP reduction (order depends on threads) */
    int mean_f64;      /* 1: zero_centered sums the mean in double (0: the reference's REAL sum) */
} FN(orc_mgpcg);

FN(orc_mgpcg) *FN(orc_mgpcg_create)(int nx, int ny, int nz, int levels)
{
    if (levels < 1 || levels > ORC_MAXL) return NULL;
    FN(orc_mgpcg) *m = (FN(orc_mgpcg) *)calloc(1, sizeof(*m));
    m->L = levels;
    m->pre_post = 6;
    m->bottom = 6;
    m->growth = 1;
    m->cs = (REAL)1.0;
    m->dot_f64 = 1;
    for (int l = 0; l < levels; ++l) {
        m->nx[l] = nx >> l; m->ny[l] = ny >> l; m->nz[l] = nz >> l;
        m->n[l] = (size_t)m->nx[l] * m->ny[l] * m->nz[l];
        m->r[l] = (REAL *)calloc(m->n[l], sizeof(REAL));
        m->z[l] = (REAL *)calloc(m->n[l], sizeof(REAL));
        m->divd[l] = (REAL *)calloc(m->n[l], sizeof(REAL));
        m->tag[l] = (int32_t *)calloc(m->n[l], sizeof(int32_t));
    }
    m->x = (REAL *)calloc(m->n[0], sizeof(REAL));
    m->p = (REAL *)calloc(m->n[0], sizeof(REAL));
    m->Ap = (REAL *)calloc(m->n[0], sizeof(REAL));
    return m;
}

void FN(orc_mgpcg_destroy)(FN(orc_mgpcg) *m)
{
    if (!m) return;
    for (int l = 0; l < m->L; ++l) { free(m->r[l]); free(m->z[l]); free(m->divd[l]); free(m->tag[l]); }
    free(m->x); free(m->p); free(m->Ap); free(m);
}

void FN(orc_mgpcg_config)(FN(orc_mgpcg) *m, int pre_post, int bottom, int dot_f64, int parallel_dot)
{
    m->pre_post = pre_post; m->bottom = bottom; m->dot_f64 = dot_f64; m->parallel_dot = parallel_dot;
}

/* The mean of zero_centered inside solve: 0 (default) sums it in REAL in
 * serial launch order as the reference's atomic sum (mgpcg.py:191-196), 1 in
 * double (the accumulation the HIP path uses). */
void FN(orc_mgpcg_mean_f64)(FN(orc_mgpcg) *m, int mean_f64) { m->mean_f64 = mean_f64; }

/* Smoothing growth per level: 1 = the reference's pre_and_post_smoothing << l
 * (mgpcg.py:304,343), 0 = the same count on every level. */
void FN(orc_mgpcg_growth)(FN(orc_mgpcg) *m, int growth) { m->growth = growth; }

/* No reference counterpart: the V-cycle variant of the HIP path's
 * flub_mgpcg_set_coarse_scale / flub_mgpcg_set_sweep_cap.  Each restricted
 * residual is stored times cs (one rounding of the parent's sum); level l
 * smooths min(pre_post << (growth * l), cap) iterations when cap > 0.
 * cs = 1, cap = 0 is the reference's V-cycle. */
void FN(orc_mgpcg_cycle)(FN(orc_mgpcg) *m, double cs, int cap) { m->cs = (REAL)cs; m->cap = cap; }

static int FN(sweeps)(const FN(orc_mgpcg) *m, int l)
{
    const int s = m->pre_post << (m->growth * l);
    return m->cap > 0 && s > m->cap ? m->cap : s;
}

REAL *FN(orc_mgpcg_buf)(FN(orc_mgpcg) *m, int which, int level)
{
    switch (which) {
    case 0: return m->x;
    case 1: return m->p;
    case 2: return m->Ap;
    case 3: return m->r[level];
    case 4: return m->z[level];
    case 5: return m->divd[level];
    default: return NULL;
    }
}
int32_t *FN(orc_mgpcg_tag)(FN(orc_mgpcg) *m, int level) { return m->tag[level]; }

/* mgpcg.py:162-174 restrict_tag (coarse dims cx,cy,cz). */
void FN(orc_restrict_tag)(const int32_t *tag_h, int32_t *tag_2h, int cx, int cy, int cz)
{
    const int ny = cy * 2, nz = cz * 2;
    for (int i = 0; i < cx; ++i)
        for (int j = 0; j < cy; ++j)
            for (int k = 0; k < cz; ++k) {
                int t = 2;
                for (int ii = 0; ii < 2; ++ii)
                    for (int jj = 0; jj < 2; ++jj)
                        for (int kk = 0; kk < 2; ++kk) {
                            const int tg = tag_h[IDX(ii + i * 2, jj + j * 2, kk + k * 2)];
                            if (tg == 0) t = 0;
                            else if (t > 0 && tg == 1) t = 1;
                        }
                tag_2h[((size_t)i * cy + j) * cz + k] = t;
            }
}

/* mgpcg.py:360-369 set_tags. */
void FN(orc_mgpcg_set_tags)(FN(orc_mgpcg) *m, const int32_t *tag)
{
    if (tag) memcpy(m->tag[0], tag, m->n[0] * sizeof(int32_t));
    FN(orc_compute_divd)(m->tag[0], m->divd[0], m->nx[0], m->ny[0], m->nz[0]);
    for (int l = 0; l + 1 < m->L; ++l) {
        FN(orc_restrict_tag)(m->tag[l], m->tag[l + 1], m->nx[l + 1], m->ny[l + 1], m->nz[l + 1]);
        FN(orc_compute_divd)(m->tag[l + 1], m->divd[l + 1], m->nx[l + 1], m->ny[l + 1], m->nz[l + 1]);
    }
}

/* mgpcg.py:276-284 init. */
void FN(orc_mgpcg_init)(FN(orc_mgpcg) *m, const REAL *r, REAL k)
{
    memset(m->z[0], 0, m->n[0] * sizeof(REAL));
    memset(m->Ap, 0, m->n[0] * sizeof(REAL));
    memset(m->p, 0, m->n[0] * sizeof(REAL));
    memset(m->x, 0, m->n[0] * sizeof(REAL));
    memcpy(m->r[0], r, m->n[0] * sizeof(REAL));
    FN(orc_scale)(m->r[0], k, m->n[0]);
    m->beta = 0.0;
}

void FN(orc_mgpcg_mask_interior)(FN(orc_mgpcg) *m)
{
    FN(orc_mask_interior)(m->tag[0], m->r[0], m->n[0]);
}

/* smoother: 0 = "default", 1 = "phases", 2 = "fraction". */
static void FN(half_sweep)(FN(orc_mgpcg) *m, int l, int smoother, int phase)
{
    if (smoother == 0) FN(orc_smooth_default)(m->z[l], m->r[l], m->nx[l], m->ny[l], m->nz[l], phase);
    else FN(orc_smooth_mcadams)(m->z[l], m->r[l], m->tag[l], m->divd[l], m->nx[l], m->ny[l], m->nz[l], phase);
}

/* mgpcg.py:296-349 apply_preconditioner: symmetric V-cycle into z[0]. */
void FN(orc_mgpcg_apply_preconditioner)(FN(orc_mgpcg) *m, int smoother)
{
    memset(m->z[0], 0, m->n[0] * sizeof(REAL));
    for (int l = 0; l + 1 < m->L; ++l) {
        for (int it = 0; it < FN(sweeps)(m, l); ++it) {
            FN(half_sweep)(m, l, smoother, 0);
            FN(half_sweep)(m, l, smoother, 1);
        }
        memset(m->z[l + 1], 0, m->n[l + 1] * sizeof(REAL));
        memset(m->r[l + 1], 0, m->n[l + 1] * sizeof(REAL));
        if (smoother == 0)
            FN(orc_restrict_default)(m->r[l], m->z[l], m->r[l + 1], m->nx[l], m->ny[l], m->nz[l]);
        else
            FN(orc_restrict_mcadams)(m->r[l], m->z[l], m->tag[l], m->r[l + 1], m->tag[l + 1],
                                     m->nx[l], m->ny[l], m->nz[l]);
        if (m->cs != (REAL)1.0)
            for (size_t I = 0; I < m->n[l + 1]; ++I) m->r[l + 1][I] = m->r[l + 1][I] * m->cs;
    }
    const int lb = m->L - 1;
    for (int it = 0; it < m->bottom / 2; ++it) {
        FN(half_sweep)(m, lb, smoother, 0);
        FN(half_sweep)(m, lb, smoother, 1);
    }
    for (int it = 0; it < m->bottom / 2; ++it) {
        FN(half_sweep)(m, lb, smoother, 1);
        FN(half_sweep)(m, lb, smoother, 0);
    }
    for (int l = m->L - 2; l >= 0; --l) {
        FN(orc_prolongate)(m->z[l], m->z[l + 1], m->nx[l], m->ny[l], m->nz[l]);
        for (int it = 0; it < FN(sweeps)(m, l); ++it) {
            FN(half_sweep)(m, l, smoother, 1);
            FN(half_sweep)(m, l, smoother, 0);
        }
    }
}

static double FN(dot)(FN(orc_mgpcg) *m, const REAL *a, const REAL *b)
{
    if (m->parallel_dot) return FN(inner_par)(a, b, m->n[0]);
    return FN(orc_inner)(a, b, m->n[0], m->dot_f64);
}

/* Scalars: with dot_f64 the reference's 1-element device scalars are kept in
 * double and rounded to REAL where an axpy applies them; without it every
 * scalar operation rounds to REAL as the FP32 reference does. */
static inline double FN(rs)(FN(orc_mgpcg) *m, double v) { return m->dot_f64 ? v : (double)(REAL)v; }

/* mgpcg.py:207-232 axpys; x + a*p is evaluated as one fused multiply-add. */
static void FN(axpy)(REAL *y, REAL a, const REAL *xv, size_t n)   /* y = y + a*x */
{
#pragma omp parallel for schedule(static)
    for (long long I = 0; I < (long long)n; ++I) y[I] = FMA(a, xv[I], y[I]);
}
static void FN(xpay)(REAL *p, const REAL *z, REAL b, size_t n)    /* p = z + b*p */
{
#pragma omp parallel for schedule(static)
    for (long long I = 0; I < (long long)n; ++I) p[I] = FMA(b, p[I], z[I]);
}

/* mgpcg.py:378-462 solve.  Returns the completed-iteration count `it`;
 * *rTr_out receives the last squared residual norm.  Deviations from the
 * reference (both are bugs there): beta starts at 0 instead of uninitialised
 * memory, and the first preconditioner call executes the V-cycle instead of
 * only capturing it. */
int FN(orc_mgpcg_solve)(FN(orc_mgpcg) *m, int max_iters, double eps, double abs_tol, double rel_tol,
                        int smoother, int verbose, double *rTr_out)
{
    const size_t n = m->n[0];
    if (smoother != 1) {
        if (m->mean_f64) FN(zero_centered_f64)(m->r[0], m->tag[0], n);
        else FN(orc_zero_centered)(m->r[0], m->tag[0], n);
    }
    const double rTr0 = FN(rs)(m, FN(dot)(m, m->r[0], m->r[0]));
    const double tol = abs_tol > rTr0 * rel_tol ? abs_tol : rTr0 * rel_tol;
    FN(orc_mgpcg_apply_preconditioner)(m, smoother);
    FN(xpay)(m->p, m->z[0], (REAL)m->beta, n);
    m->old_zTr = FN(rs)(m, FN(dot)(m, m->z[0], m->r[0]));
    int it = 0;
    double rTr = rTr0;
    const int cap = max_iters < 0 ? ORC_SAFETY_CAP : max_iters;
    while (it < cap) {
        if (smoother == 0) FN(orc_compute_Ap_default)(m->Ap, m->p, m->nx[0], m->ny[0], m->nz[0]);
        else FN(orc_compute_Ap_mcadams)(m->Ap, m->p, m->tag[0], m->nx[0], m->ny[0], m->nz[0]);
        m->sum = FN(rs)(m, FN(dot)(m, m->p, m->Ap));
        m->alpha = FN(rs)(m, m->old_zTr / FN(rs)(m, m->sum + FN(rs)(m, eps)));
        FN(axpy)(m->r[0], (REAL)(-(REAL)m->alpha), m->Ap, n);
        if (smoother != 1) {
            if (m->mean_f64) FN(zero_centered_f64)(m->r[0], m->tag[0], n);
            else FN(orc_zero_centered)(m->r[0], m->tag[0], n);
        }
        rTr = FN(rs)(m, FN(dot)(m, m->r[0], m->r[0]));
        if (verbose) printf("after iter %d, |residual|_2=%.9g\n", it + 1, sqrt(rTr));
        if (rTr < tol || !isfinite(rTr)) {
            FN(axpy)(m->x, (REAL)m->alpha, m->p, n);
            break;
        }
        FN(orc_mgpcg_apply_preconditioner)(m, smoother);
        m->sum = FN(rs)(m, FN(dot)(m, m->z[0], m->r[0]));
        m->beta = FN(rs)(m, m->sum / FN(rs)(m, m->old_zTr + FN(rs)(m, eps)));
        m->old_zTr = m->sum;
        FN(axpy)(m->x, (REAL)m->alpha, m->p, n);
        FN(xpay)(m->p, m->z[0], (REAL)m->beta, n);
        it += 1;
    }
    if (rTr_out) *rTr_out = rTr;
    return it;
}

/* example_fluid.py:120-147 project() with the water_3d.py:58-68 call_mgpcg,
 * minus the diagnostic divergence()/print and the rigid-velocity precompute:
 * div = -F/h (masked), r0 = -h*div, solve, p0 = x*h, u -= grad(p0/h), p0 *= h.
 * u,v,w are updated in place; p0 receives the stored pressure x*h.  eps is the
 * solve's regulariser (the reference passes 1e-12; the FP64 judge passes 0 so
 * that squared residuals below 1e-12 stay reachable). */
int FN(orc_project)(FN(orc_mgpcg) *m, REAL *u, REAL *v, REAL *w,
                    const REAL *ur, const REAL *vr, const REAL *wr, const int32_t *tag,
                    REAL *div, REAL *p0, double h, int max_iters, double eps, double abs_tol,
                    double rel_tol, int smoother, double *rTr_out)
{
    const int nx = m->nx[0], ny = m->ny[0], nz = m->nz[0];
    const size_t n = m->n[0];
    const REAL h0 = (REAL)h;
    FN(orc_divergence_tag)(u, v, w, div, ur, vr, wr, tag, nx, ny, nz);
    FN(orc_mask_with_tag)(div, tag, n);
    FN(orc_scale)(div, (REAL)(-1.0 / h), n);
    FN(orc_mgpcg_set_tags)(m, tag);
    FN(orc_mgpcg_init)(m, div, (REAL)(-h));
    FN(orc_mgpcg_mask_interior)(m);
    const int it = FN(orc_mgpcg_solve)(m, max_iters, eps, abs_tol, rel_tol, smoother, 0, rTr_out);
    memcpy(p0, m->x, n * sizeof(REAL));
    FN(orc_scale)(p0, h0, n);
    FN(orc_scale)(p0, (REAL)(1.0 / h), n);
    FN(orc_subtract_gradient_tag)(u, ur, p0, tag, nx, ny, nz, 1);
    FN(orc_subtract_gradient_tag)(v, vr, p0, tag, nx, ny, nz, 2);
    FN(orc_subtract_gradient_tag)(w, wr, p0, tag, nx, ny, nz, 3);
    FN(orc_scale)(p0, h0, n);
    return it;
}

#undef FN
#undef CAT
#undef CAT_
