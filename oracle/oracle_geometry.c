/*
 * TEST INFRASTRUCTURE -- CPU oracle for the solid-geometry inputs of the
 * projection: primitive signed distances, face fractions, rigid face
 * velocities and the fraction-weighted divergence of the reference Warp code:
 *   geometry.py:17-32 (_sdf_cube), :82-113 (compute_sdf_atom_prims),
 *   :148-169 (sdf_argmin_prims), :225-239 (precompute_grid_point_phi_prims),
 *   :365-388 (update_argmin: cached level sets, sampled by the tricubic
 *   interpolation of oracle_advect.c),
 *   :264-290 (phi_to_fraction), :401-462 (PrecomputeFaceFraction);
 *   utils/util.py:168-228 (fraction_edge / fraction_triangle / fraction_quad),
 *   :231-235 (compute_v_solid_3d);
 *   rigid_coupling.py:51-70 (precompute_solid_vel);
 *   divergence.py:5-41 (divergence_fractional_3d);
 *   extrapolate.py:9-47 (extrapolate_3d, flush_1_3d, extrapolate_by_1),
 *   utils/util.py:424-437 (set_unknown).
 *
 * FP32 throughout, one point at a time in the reference's expression order;
 * built with -ffp-contract=off so every operation rounds as written.  Only
 * tests/, the smoke check and the CPU-baseline leg of bench.py load it.
 *
 * Parity pins (tests/test_oracle_geometry.py): the box and sphere distances
 * against closed forms on points rotated by scipy's Rotation; the quaternion
 * rotation and product against scipy; fraction_quad against the exact area of
 * a unit square cut by a plane (Sutherland-Hodgman clipping); phi_to_fraction
 * per axis against that exact area on every face of a grid cut by a plane;
 * the solid velocity against v + omega x (p - x); the fractional divergence
 * against the plain flux divergence (all fractions 1), the rigid flux
 * divergence (all fractions 0) and u == u_rigid (any fractions); the
 * cached-level-set resampling against a level set that is a polynomial of degree
 * <= 3 per axis (reproduced exactly by the tricubic interpolation away from
 * the clamped border), evaluated in FP64 at the rotated, translated point; the
 * extrapolation against a vectorised mean over the known 27-neighbourhood,
 * constant fields (kept exactly), idempotence and unchanged tags.
 *
 * Rigid bodies are structure-of-arrays: x[3n], q[4n] (x, y, z, w),
 * gtype[n] (0 none, 1 quad, 2 circle, 3 rect, 4 "x", 5 mesh), radius[3n].
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#define GIDX(i, j, k, sy, sz) ((((size_t)(i)) * (size_t)(sy) + (size_t)(j)) * (size_t)(sz) + (size_t)(k))

/* Rotation of v by the inverse of the unit quaternion q = (x, y, z, w):
 * v (2w^2 - 1) - 2w (q_v x v) + 2 q_v (q_v . v). */
static void quat_rotate_inv(const float *q, const float *v, float *out)
{
    const float qx = q[0], qy = q[1], qz = q[2], qw = q[3];
    const float c = 2.0f * qw * qw - 1.0f;
    const float t0 = qy * v[2] - qz * v[1];
    const float t1 = qz * v[0] - qx * v[2];
    const float t2 = qx * v[1] - qy * v[0];
    const float d = qx * v[0] + qy * v[1] + qz * v[2];
    out[0] = v[0] * c - t0 * qw * 2.0f + qx * d * 2.0f;
    out[1] = v[1] * c - t1 * qw * 2.0f + qy * d * 2.0f;
    out[2] = v[2] * c - t2 * qw * 2.0f + qz * d * 2.0f;
}

/* Hamilton product a * b of quaternions (x, y, z, w). */
static void quat_mul(const float *a, const float *b, float *o)
{
    o[0] = a[3] * b[0] + b[3] * a[0] + (a[1] * b[2] - a[2] * b[1]);
    o[1] = a[3] * b[1] + b[3] * a[1] + (a[2] * b[0] - a[0] * b[2]);
    o[2] = a[3] * b[2] + b[3] * a[2] + (a[0] * b[1] - a[1] * b[0]);
    o[3] = a[3] * b[3] - (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
}

/* geometry.py:17-32: signed distance of a box of half-extents radius, rotated by q. */
float orc_sdf_cube(const float *p_in, const float *radius, const float *q)
{
    float p[3], m[3];
    quat_rotate_inv(q, p_in, p);
    for (int a = 0; a < 3; ++a) m[a] = fmaxf(p[a], -p[a]) - radius[a];
    const float mx = fmaxf(fmaxf(m[0], m[1]), m[2]);
    if (mx < 0.0f) return mx;   /* -(min(-m)) = max(m) */
    const float a0 = fmaxf(m[0], 0.0f), a1 = fmaxf(m[1], 0.0f), a2 = fmaxf(m[2], 0.0f);
    return sqrtf(a0 * a0 + a1 * a1 + a2 * a2);
}

/* The rotation by pi/2 about z used by the "x" primitive: (0, 0, cos(pi/4), sin(pi/4)). */
static const float kZ2[4] = {0.0f, 0.0f, 0.70710677f, 0.70710677f};

/* geometry.py:82-113 compute_sdf_atom_prims at point (u, v, w). */
float orc_sdf_atom_prims(float u, float v, float w, const float *x, const float *q, int gtype,
                         const float *radius)
{
    const float p[3] = {u - x[0], v - x[1], w - x[2]};
    switch (gtype) {
    case 1:
    case 3:
        return orc_sdf_cube(p, radius, q);
    case 2:
        return sqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - radius[0];
    case 4: {
        const float zinv[4] = {-kZ2[0], -kZ2[1], -kZ2[2], kZ2[3]};
        float t[4], q2[4];
        quat_mul(kZ2, q, t);
        quat_mul(t, zinv, q2);
        return fminf(orc_sdf_cube(p, radius, q), orc_sdf_cube(p, radius, q2));
    }
    case 5:
        return 2.0f;   /* meshes are resolved from cached level sets later */
    default:
        return 1.0f;
    }
}

/* geometry.py:148-169 sdf_argmin_prims: the smallest distance below h0 (2.0 if none)
 * and the index of its body (-1 if none); the first minimum wins ties. */
float orc_sdf_argmin_prims(float u, float v, float w, int n, const float *x, const float *q,
                           const int32_t *gtype, const float *radius, float h0, int32_t *argmin)
{
    float interior = 2.0f;
    int32_t am = -1;
    for (int b = 0; b < n; ++b) {
        const float s = orc_sdf_atom_prims(u, v, w, x + 3 * b, q + 4 * b, gtype[b], radius + 3 * b);
        if (s < fminf(h0, interior)) { interior = s; am = b; }
    }
    if (argmin) *argmin = am;
    return interior;
}

/* geometry.py:225-239 precompute_grid_point_phi_prims over a (sx, sy, sz) array:
 * point ((i + st_x) h, (j + st_y) h, (k + st_z) h) written as i*h + st_x*h; the owner
 * index is written only for a staggered (face) array. */
void orc_phi_prims(float *phi, int32_t *owner, int sx, int sy, int sz, float stx, float sty, float stz,
                   float h, int n, const float *x, const float *q, const int32_t *gtype, const float *radius)
{
    const int staggered = stx != 0.0f || sty != 0.0f || stz != 0.0f;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < sx; ++i)
        for (int j = 0; j < sy; ++j)
            for (int k = 0; k < sz; ++k) {
                const float u = (float)i * h + stx * h;
                const float v = (float)j * h + sty * h;
                const float w = (float)k * h + stz * h;
                int32_t am;
                const size_t t = GIDX(i, j, k, sy, sz);
                phi[t] = orc_sdf_argmin_prims(u, v, w, n, x, q, gtype, radius, h, &am);
                if (staggered && owner) owner[t] = am;
            }
}

void orc_tricubic(const float *q, int gx, int gy, int gz, const float *str, float *out, int n);

/* geometry.py:365-388 update_argmin over an (sx, sy, sz) array: the point
 * (i h + st_x h, ...) less the body position x, rotated by q^-1, divided by
 * h_ls, less ijk, samples the level set (lx, ly, lz) tricubically
 * (trilerp_stagger with stagger 0, utils/util.py:98-104); a sample below phi
 * replaces it and, on a staggered array, sets owner to id. */
void orc_update_argmin(float *phi, int32_t *owner, int sx, int sy, int sz, const float *ls, int lx, int ly, int lz,
                       int id, const float *x, const float *q, float h_ls, const float *st, const float *ijk,
                       float h)
{
    const int staggered = st[0] != 0.0f || st[1] != 0.0f || st[2] != 0.0f;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < sx; ++i)
        for (int j = 0; j < sy; ++j)
            for (int k = 0; k < sz; ++k) {
                float p[3] = {(float)i * h, (float)j * h, (float)k * h};
                for (int a = 0; a < 3; ++a) p[a] = p[a] + st[a] * h;
                for (int a = 0; a < 3; ++a) p[a] = p[a] - x[a];
                float b[3];
                quat_rotate_inv(q, p, b);
                float s3[3];
                for (int a = 0; a < 3; ++a) s3[a] = b[a] / h_ls - ijk[a];
                for (int a = 0; a < 3; ++a) s3[a] = s3[a] - 0.0f;
                float m;
                orc_tricubic(ls, lx, ly, lz, s3, &m, 1);
                const size_t t = GIDX(i, j, k, sy, sz);
                if (m < phi[t]) {
                    phi[t] = m;
                    if (staggered && owner) owner[t] = id;
                }
            }
}

/* utils/util.py:168-187: fraction of an edge with sdf > 0 (0 below 0.1). */
float orc_fraction_edge(float yl, float yr)
{
    float ret;
    if (yl > 0.0f && yr > 0.0f) ret = 1.0f;
    else if (yl > 0.0f) ret = yl / (yl - yr);
    else if (yr > 0.0f) ret = yr / (yr - yl);
    else ret = 0.0f;
    if (ret < 0.1f) ret = 0.0f;
    else if (ret > 1.0f) ret = 1.0f;
    return ret;
}

/* utils/util.py:189-217: fraction of a triangle with sdf > 0. */
float orc_fraction_triangle(float ya, float yb, float yc)
{
    const float fab = orc_fraction_edge(ya, yb), fbc = orc_fraction_edge(yb, yc), fca = orc_fraction_edge(yc, ya);
    const int a = ya > 0.0f, b = yb > 0.0f, c = yc > 0.0f;
    float ret;
    if (a && b && c) ret = 1.0f;
    else if (a && b) ret = 1.0f - (1.0f - fca) * (1.0f - fbc);
    else if (a && c) ret = 1.0f - (1.0f - fab) * (1.0f - fbc);
    else if (b && c) ret = 1.0f - (1.0f - fab) * (1.0f - fca);
    else if (a) ret = fca * fab;
    else if (b) ret = fab * fbc;
    else if (c) ret = fbc * fca;
    else ret = 0.0f;
    return fminf(fmaxf(ret, 0.0f), 1.0f);
}

/* utils/util.py:219-228: a face as four triangles around its centre. */
float orc_fraction_quad(float ya, float yb, float yc, float yd, float ycen)
{
    float f = orc_fraction_triangle(ya, yb, ycen) + orc_fraction_triangle(yb, yc, ycen) +
              orc_fraction_triangle(yc, yd, ycen) + orc_fraction_triangle(yd, ya, ycen);
    f = f * 0.25f;
    if (f < 0.1f) f = 0.0f;
    else if (f > 1.0f) f = 1.0f;
    return f;
}

/* geometry.py:264-290 phi_to_fraction for axis dim (1 u, 2 v, 3 w) of an nx x ny x nz
 * cell grid: phi holds the corner values ((nx+1) x (ny+1) x (nz+1)); frac holds the
 * face-centre values on entry and the face fractions on return. */
void orc_phi_to_fraction(const float *phi, float *frac, int nx, int ny, int nz, int dim)
{
    const int fx = nx + (dim == 1), fy = ny + (dim == 2), fz = nz + (dim == 3);
    const int py = ny + 1, pz = nz + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < fx; ++i)
        for (int j = 0; j < fy; ++j)
            for (int k = 0; k < fz; ++k) {
                const float ya = phi[GIDX(i, j, k, py, pz)];
                float yb, yc, yd;
                if (dim == 1) {
                    yb = phi[GIDX(i, j + 1, k, py, pz)];
                    yc = phi[GIDX(i, j + 1, k + 1, py, pz)];
                    yd = phi[GIDX(i, j, k + 1, py, pz)];
                } else if (dim == 2) {
                    yb = phi[GIDX(i + 1, j, k, py, pz)];
                    yc = phi[GIDX(i + 1, j, k + 1, py, pz)];
                    yd = phi[GIDX(i, j, k + 1, py, pz)];
                } else {
                    yb = phi[GIDX(i + 1, j, k, py, pz)];
                    yc = phi[GIDX(i + 1, j + 1, k, py, pz)];
                    yd = phi[GIDX(i, j + 1, k, py, pz)];
                }
                const size_t t = GIDX(i, j, k, fy, fz);
                frac[t] = orc_fraction_quad(ya, yb, yc, yd, frac[t]);
            }
}

/* rigid_coupling.py:51-70 precompute_solid_vel for axis d over a face array
 * (sx, sy, sz) with stagger (st): the component d of v + omega x (pos - x) of
 * the owning body where h0/100 < i*h0 < 1 - h0/100 and an owner exists, else 0.
 * (The interior test reads i for every axis, as the reference does.) */
void orc_solid_vel(float *vel, const int32_t *owner, int sx, int sy, int sz, int d, float stx, float sty,
                   float stz, float h0, const float *x, const float *v, const float *omega)
{
    const float eps = h0 * 0.01f;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < sx; ++i)
        for (int j = 0; j < sy; ++j)
            for (int k = 0; k < sz; ++k) {
                const size_t t = GIDX(i, j, k, sy, sz);
                const int32_t o = owner[t];
                const float tmp = (float)i * h0;
                if (eps < tmp && tmp < 1.0f - eps && o != -1) {
                    const float p[3] = {((float)i + stx) * h0, ((float)j + sty) * h0, ((float)k + stz) * h0};
                    const float r[3] = {p[0] - x[3 * o], p[1] - x[3 * o + 1], p[2] - x[3 * o + 2]};
                    const float *om = omega + 3 * o, *vo = v + 3 * o;
                    float c;
                    if (d == 1) c = om[1] * r[2] - om[2] * r[1];
                    else if (d == 2) c = om[2] * r[0] - om[0] * r[2];
                    else c = om[0] * r[1] - om[1] * r[0];
                    vel[t] = vo[d - 1] + c;
                } else {
                    vel[t] = 0.0f;
                }
            }
}

/* divergence.py:5-41 divergence_fractional_3d over the nx x ny x nz cells
 * (sample_3d clamps are no-ops on MAC arrays of the matching shapes). */
void orc_divergence_fractional(const float *u, const float *v, const float *w, float *div, const float *fu,
                               const float *fv, const float *fw, const float *ur, const float *vr,
                               const float *wr, int nx, int ny, int nz)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const size_t ul = GIDX(i, j, k, ny, nz), uh = GIDX(i + 1, j, k, ny, nz);
                const size_t vl = GIDX(i, j, k, ny + 1, nz), vh = GIDX(i, j + 1, k, ny + 1, nz);
                const size_t wl = GIDX(i, j, k, ny, nz + 1), wh = GIDX(i, j, k + 1, ny, nz + 1);
                const float l = fu[ul], r = fu[uh], b = fv[vl], t = fv[vh], f = fw[wl], n = fw[wh];
                const float fluid = u[uh] * r - u[ul] * l + v[vh] * t - v[vl] * b + w[wh] * n - w[wl] * f;
                const float solid = (1.0f - r) * ur[uh] - (1.0f - l) * ur[ul] + (1.0f - t) * vr[vh] -
                                    (1.0f - b) * vr[vl] + (1.0f - n) * wr[wh] - (1.0f - f) * wr[wl];
                div[GIDX(i, j, k, ny, nz)] = fluid + solid;
            }
}

/* ---------------------------------------------------------------------------
 * Velocity extrapolation (extrapolate.py:9-47, utils/util.py:424-437).
 * ------------------------------------------------------------------------- */

/* utils/util.py:424-437 set_unknown over the nx x ny x nz cells: face arrays
 * tu/tv/tw (their own shapes) get 1 where phi >= 0, else 0, at the cell index;
 * the last face layer of each axis is left as it is. */
void orc_set_unknown(const float *phi, int32_t *tu, int32_t *tv, int32_t *tw, int nx, int ny, int nz)
{
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int32_t f = phi[GIDX(i, j, k, ny, nz)] >= 0.0f ? 1 : 0;
                tu[GIDX(i, j, k, ny, nz)] = f;
                tv[GIDX(i, j, k, ny + 1, nz)] = f;
                tw[GIDX(i, j, k, ny, nz + 1)] = f;
            }
}

/* extrapolate.py:9-31 extrapolate_3d on one face array (sx, sy, sz), in place
 * and serially: a face with unknown == 1 takes the mean of its in-range
 * 27-neighbourhood faces with unknown == 0 (summed ii, jj, kk from -1 to 1,
 * ii slowest) and is marked -1, or 0.0 when there is none.  Known faces are
 * never written and -1 / 1 are both "not known", so the serial order gives the
 * parallel launch's result. */
void orc_extrapolate_3d(float *q, int32_t *unknown, int sx, int sy, int sz)
{
    for (int i = 0; i < sx; ++i)
        for (int j = 0; j < sy; ++j)
            for (int k = 0; k < sz; ++k) {
                const size_t t = GIDX(i, j, k, sy, sz);
                if (unknown[t] != 1) continue;
                int n = 0;
                float s = 0.0f;
                for (int a = -1; a <= 1; ++a)
                    for (int b = -1; b <= 1; ++b)
                        for (int c = -1; c <= 1; ++c) {
                            const int ii = i + a, jj = j + b, kk = k + c;
                            if (ii < 0 || ii >= sx || jj < 0 || jj >= sy || kk < 0 || kk >= sz) continue;
                            const size_t u = GIDX(ii, jj, kk, sy, sz);
                            if (unknown[u] == 0) { n += 1; s += q[u]; }
                        }
                if (n > 0) { q[t] = s / (float)n; unknown[t] = -1; }
                else q[t] = 0.0f;
            }
}

/* extrapolate.py:33-37 flush_1_3d: -1 -> 1. */
void orc_flush_1(int32_t *tag, size_t n)
{
    for (size_t t = 0; t < n; ++t)
        if (tag[t] == -1) tag[t] = 1;
}
