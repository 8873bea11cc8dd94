/*
 * TEST INFRASTRUCTURE -- a driver for the AddressSanitizer / UBSan build of
 * the CPU oracle (make -C oracle sanitize): runs every oracle routine family
 * on small cubic and non-cubic grids whose arrays are allocated to their exact
 * sizes, so an out-of-range index or undefined behaviour in the oracle aborts
 * the run.  Nothing in the product uses it.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

typedef struct orc_mgpcg_f32 orc_mgpcg_f32;
typedef struct orc_mgpcg_f64 orc_mgpcg_f64;
orc_mgpcg_f32 *orc_mgpcg_create_f32(int nx, int ny, int nz, int levels);
orc_mgpcg_f64 *orc_mgpcg_create_f64(int nx, int ny, int nz, int levels);
void orc_mgpcg_destroy_f32(orc_mgpcg_f32 *m);
void orc_mgpcg_destroy_f64(orc_mgpcg_f64 *m);
void orc_mgpcg_config_f32(orc_mgpcg_f32 *m, int pre_post, int bottom, int dot_f64, int parallel_dot);
int orc_project_f32(orc_mgpcg_f32 *m, float *u, float *v, float *w, const float *ur, const float *vr,
                    const float *wr, const int32_t *tag, float *div, float *p0, double h, int max_iters, double eps,
                    double abs_tol, double rel_tol, int smoother, double *rTr_out);
int orc_project_f64(orc_mgpcg_f64 *m, double *u, double *v, double *w, const double *ur, const double *vr,
                    const double *wr, const int32_t *tag, double *div, double *p0, double h, int max_iters,
                    double eps, double abs_tol, double rel_tol, int smoother, double *rTr_out);
void orc_phi_prims(float *phi, int32_t *owner, int sx, int sy, int sz, float stx, float sty, float stz, float h,
                   int n, const float *x, const float *q, const int32_t *gtype, const float *radius);
void orc_phi_to_fraction(const float *phi, float *frac, int nx, int ny, int nz, int dim);
void orc_solid_vel(float *vel, const int32_t *owner, int sx, int sy, int sz, int d, float stx, float sty, float stz,
                   float h0, const float *x, const float *v, const float *omega);
void orc_divergence_fractional(const float *u, const float *v, const float *w, float *div, const float *fu,
                               const float *fv, const float *fw, const float *ur, const float *vr, const float *wr,
                               int nx, int ny, int nz);
void orc_set_unknown(const float *phi, int32_t *tu, int32_t *tv, int32_t *tw, int nx, int ny, int nz);
void orc_extrapolate_3d(float *q, int32_t *unknown, int sx, int sy, int sz);
void orc_flush_1(int32_t *tag, size_t n);
void orc_advect_mc(const float *u, const float *v, const float *w, int nx, int ny, int nz, const float *q_in,
                   float *q_out, float *q_tmp, int gx, int gy, int gz, const float *st, float dt, float h0);
void orc_advect_phi(const float *u, const float *v, const float *w, int nx, int ny, int nz, const float *phi,
                    float *new_phi, const float *st, float dt, float h0);
void orc_advect_simple(const float *u, const float *v, const float *w, int nx, int ny, int nz, const float *q,
                       float *new_q, float dt, float h);

static uint64_t rng_state = 0x2545F4914F6CDD1Dull;
static double urand(void)   /* xorshift64*: uniform [0, 1) */
{
    rng_state ^= rng_state >> 12; rng_state ^= rng_state << 25; rng_state ^= rng_state >> 27;
    return (double)((rng_state * 0x2545F4914F6CDD1Dull) >> 11) / 9007199254740992.0;
}
static float *frand(size_t n, float lo, float hi)
{
    float *a = malloc(n * sizeof(float));
    for (size_t i = 0; i < n; ++i) a[i] = (float)(lo + (hi - lo) * urand());
    return a;
}

static int projection(int nx, int ny, int nz, int levels)
{
    const size_t n = (size_t)nx * ny * nz;
    const size_t nu = (size_t)(nx + 1) * ny * nz, nv = (size_t)nx * (ny + 1) * nz, nw = (size_t)nx * ny * (nz + 1);
    int32_t *tag = malloc(n * sizeof(int32_t));
    for (size_t i = 0; i < n; ++i) { const double r = urand(); tag[i] = r < 0.2 ? 0 : (r < 0.9 ? 1 : 2); }
    int iters = 0;
    for (int smoother = 0; smoother < 3; ++smoother) {
        float *u = frand(nu, -1, 1), *v = frand(nv, -1, 1), *w = frand(nw, -1, 1);
        float *ur = frand(nu, -1, 1), *vr = frand(nv, -1, 1), *wr = frand(nw, -1, 1);
        float *div = malloc(n * sizeof(float)), *p0 = malloc(n * sizeof(float));
        double rtr = 0.0;
        orc_mgpcg_f32 *m = orc_mgpcg_create_f32(nx, ny, nz, levels);
        orc_mgpcg_config_f32(m, 2, 6, 1, 0);
        iters += orc_project_f32(m, u, v, w, ur, vr, wr, tag, div, p0, 1.0 / nx, 50, 1e-12, 0.0, 1e-12, smoother, &rtr);
        orc_mgpcg_destroy_f32(m);
        double *u64 = malloc(nu * 8), *v64 = malloc(nv * 8), *w64 = malloc(nw * 8);
        double *ur64 = malloc(nu * 8), *vr64 = malloc(nv * 8), *wr64 = malloc(nw * 8);
        double *div64 = malloc(n * 8), *p64 = malloc(n * 8);
        for (size_t i = 0; i < nu; ++i) { u64[i] = u[i]; ur64[i] = ur[i]; }
        for (size_t i = 0; i < nv; ++i) { v64[i] = v[i]; vr64[i] = vr[i]; }
        for (size_t i = 0; i < nw; ++i) { w64[i] = w[i]; wr64[i] = wr[i]; }
        orc_mgpcg_f64 *m64 = orc_mgpcg_create_f64(nx, ny, nz, levels);
        iters += orc_project_f64(m64, u64, v64, w64, ur64, vr64, wr64, tag, div64, p64, 1.0 / nx, 50, 0.0, 0.0,
                                 1e-20, smoother, &rtr);
        orc_mgpcg_destroy_f64(m64);
        free(u); free(v); free(w); free(ur); free(vr); free(wr); free(div); free(p0);
        free(u64); free(v64); free(w64); free(ur64); free(vr64); free(wr64); free(div64); free(p64);
    }
    free(tag);
    return iters;
}

static void geometry_and_steps(int nx, int ny, int nz)
{
    const size_t n = (size_t)nx * ny * nz, nc = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
    const size_t nf[3] = {(size_t)(nx + 1) * ny * nz, (size_t)nx * (ny + 1) * nz, (size_t)nx * ny * (nz + 1)};
    const int fs[3][3] = {{nx + 1, ny, nz}, {nx, ny + 1, nz}, {nx, ny, nz + 1}};
    const float st[3][3] = {{0.0f, 0.5f, 0.5f}, {0.5f, 0.0f, 0.5f}, {0.5f, 0.5f, 0.0f}};
    const float h = 1.0f / (float)nx;
    const float x[6] = {0.5f, 0.4f, 0.45f, 0.3f, 0.6f, 0.5f}, q[8] = {0.1f, 0.2f, 0.3f, 0.927f, 0, 0, 0, 1};
    const int32_t gt[2] = {3, 2};
    const float rad[6] = {0.1f, 0.15f, 0.12f, 0.2f, 0.2f, 0.2f}, vel[6] = {0.3f, -0.2f, 0.1f, 0, 0, 0.5f};
    const float om[6] = {0.1f, 0.2f, -0.3f, 0, 0, 0};
    float *corner = malloc(nc * sizeof(float));
    orc_phi_prims(corner, NULL, nx + 1, ny + 1, nz + 1, 0, 0, 0, h, 2, x, q, gt, rad);
    float *frac[3], *vs[3], *uvw[3];
    int32_t *own[3], *mark[3];
    for (int d = 0; d < 3; ++d) {
        frac[d] = malloc(nf[d] * sizeof(float));
        own[d] = malloc(nf[d] * sizeof(int32_t));
        orc_phi_prims(frac[d], own[d], fs[d][0], fs[d][1], fs[d][2], st[d][0], st[d][1], st[d][2], h, 2, x, q, gt, rad);
        orc_phi_to_fraction(corner, frac[d], nx, ny, nz, d + 1);
        vs[d] = malloc(nf[d] * sizeof(float));
        orc_solid_vel(vs[d], own[d], fs[d][0], fs[d][1], fs[d][2], d + 1, st[d][0], st[d][1], st[d][2], h, x, vel, om);
        uvw[d] = frand(nf[d], -1, 1);
        mark[d] = malloc(nf[d] * sizeof(int32_t));
        for (size_t i = 0; i < nf[d]; ++i) mark[d][i] = urand() < 0.5 ? -1 : 0;
    }
    float *div = malloc(n * sizeof(float));
    orc_divergence_fractional(uvw[0], uvw[1], uvw[2], div, frac[0], frac[1], frac[2], vs[0], vs[1], vs[2], nx, ny, nz);
    float *phi = frand(n, -1, 1);
    orc_set_unknown(phi, mark[0], mark[1], mark[2], nx, ny, nz);
    for (int pass = 0; pass < 3; ++pass)
        for (int d = 0; d < 3; ++d) {
            orc_extrapolate_3d(uvw[d], mark[d], fs[d][0], fs[d][1], fs[d][2]);
            orc_flush_1(mark[d], nf[d]);
        }
    for (int d = 0; d < 3; ++d) {
        float *out = malloc(nf[d] * sizeof(float)), *tmp = malloc(nf[d] * sizeof(float));
        orc_advect_mc(uvw[0], uvw[1], uvw[2], nx, ny, nz, uvw[d], out, tmp, fs[d][0], fs[d][1], fs[d][2], st[d],
                      0.5f, h);
        free(out); free(tmp);
    }
    const float cst[3] = {0.5f, 0.5f, 0.5f};
    float *q2 = malloc(n * sizeof(float)), *q3 = malloc(n * sizeof(float));
    orc_advect_phi(uvw[0], uvw[1], uvw[2], nx, ny, nz, phi, q2, cst, 0.5f, h);
    orc_advect_simple(uvw[0], uvw[1], uvw[2], nx, ny, nz, phi, q3, 0.5f, h);
    free(q2); free(q3); free(phi); free(div); free(corner);
    for (int d = 0; d < 3; ++d) { free(frac[d]); free(own[d]); free(vs[d]); free(uvw[d]); free(mark[d]); }
}

int main(void)
{
    int it = 0;
    it += projection(16, 16, 16, 3);
    it += projection(16, 8, 24, 3);
    it += projection(8, 12, 20, 2);
    geometry_and_steps(12, 12, 12);
    geometry_and_steps(9, 14, 7);
    printf("sanitized oracle run complete (%d solver iterations)\n", it);
    return 0;
}
