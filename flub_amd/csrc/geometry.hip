// Solid-geometry inputs of the projection: signed distances of primitive
// rigid bodies, face fractions, rigid face velocities and the
// fraction-weighted divergence (geometry.py:17-290,401-462,
// utils/util.py:168-235, rigid_coupling.py:51-70, divergence.py:5-41).
//
// One thread per point, k fastest across consecutive lanes.  The bodies are
// few (a scene's rigid list), read as wavefront-uniform loads inside the
// kernels.  Every expression is evaluated in the order the reference writes
// it, rounded as written (-ffp-contract=off; IEEE division and square root),
// so the results are bitwise those of the CPU restatement.
//
// Rigid bodies are structure-of-arrays in device memory: x[3n], q[4n]
// (x, y, z, w), gtype[n] (0 none, 1 quad, 2 circle, 3 rect, 4 "x", 5 mesh),
// radius[3n], v[3n], omega[3n].
#include "common.h"
#include "interp.h"

namespace {

struct Bodies {
    const float *x, *q, *radius;
    const int32_t *gtype;
    int n;
};

// v rotated by the inverse of the unit quaternion q: v (2w^2 - 1) - 2w (q_v x v) + 2 q_v (q_v . v).
__device__ __forceinline__ void quat_rotate_inv(const float q[4], const float v[3], float o[3])
{
    const float c = 2.0f * q[3] * q[3] - 1.0f;
    const float t0 = q[1] * v[2] - q[2] * v[1];
    const float t1 = q[2] * v[0] - q[0] * v[2];
    const float t2 = q[0] * v[1] - q[1] * v[0];
    const float d = q[0] * v[0] + q[1] * v[1] + q[2] * v[2];
    o[0] = v[0] * c - t0 * q[3] * 2.0f + q[0] * d * 2.0f;
    o[1] = v[1] * c - t1 * q[3] * 2.0f + q[1] * d * 2.0f;
    o[2] = v[2] * c - t2 * q[3] * 2.0f + q[2] * d * 2.0f;
}

// Hamilton product a * b, quaternions as (x, y, z, w).
__device__ __forceinline__ void quat_mul(const float a[4], const float b[4], float o[4])
{
    o[0] = a[3] * b[0] + b[3] * a[0] + (a[1] * b[2] - a[2] * b[1]);
    o[1] = a[3] * b[1] + b[3] * a[1] + (a[2] * b[0] - a[0] * b[2]);
    o[2] = a[3] * b[2] + b[3] * a[2] + (a[0] * b[1] - a[1] * b[0]);
    o[3] = a[3] * b[3] - (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
}

// geometry.py:17-32: a box of half-extents r rotated by q.
__device__ float sdf_cube(const float p_in[3], const float r[3], const float q[4])
{
    float p[3];
    quat_rotate_inv(q, p_in, p);
    const float m0 = fmaxf(p[0], -p[0]) - r[0], m1 = fmaxf(p[1], -p[1]) - r[1], m2 = fmaxf(p[2], -p[2]) - r[2];
    const float mx = fmaxf(fmaxf(m0, m1), m2);
    if (mx < 0.0f) return mx;
    const float a0 = fmaxf(m0, 0.0f), a1 = fmaxf(m1, 0.0f), a2 = fmaxf(m2, 0.0f);
    return sqrtf(a0 * a0 + a1 * a1 + a2 * a2);
}

// geometry.py:82-113 compute_sdf_atom_prims.
__device__ float sdf_atom(float u, float v, float w, const Bodies &B, int b)
{
    const float p[3] = {u - B.x[3 * b], v - B.x[3 * b + 1], w - B.x[3 * b + 2]};
    const float q[4] = {B.q[4 * b], B.q[4 * b + 1], B.q[4 * b + 2], B.q[4 * b + 3]};
    const float r[3] = {B.radius[3 * b], B.radius[3 * b + 1], B.radius[3 * b + 2]};
    switch (B.gtype[b]) {
    case 1:
    case 3:
        return sdf_cube(p, r, q);
    case 2:
        return sqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - r[0];
    case 4: {
        // the same box turned by pi/2 about z: z q z^-1 with z = (0, 0, cos(pi/4), sin(pi/4))
        const float z[4] = {0.0f, 0.0f, 0.70710677f, 0.70710677f};
        const float zi[4] = {-0.0f, -0.0f, -0.70710677f, 0.70710677f};
        float t[4], q2[4];
        quat_mul(z, q, t);
        quat_mul(t, zi, q2);
        return fminf(sdf_cube(p, r, q), sdf_cube(p, r, q2));
    }
    case 5:
        return 2.0f;   // meshes are resolved from cached level sets afterwards
    default:
        return 1.0f;
    }
}

// geometry.py:148-169 sdf_argmin_prims: the smallest distance below h0 (2.0
// when none is) and its body (-1), the first minimum winning ties.
__device__ float sdf_argmin(float u, float v, float w, const Bodies &B, float h0, int32_t &am)
{
    float interior = 2.0f;
    am = -1;
    for (int b = 0; b < B.n; ++b) {
        const float s = sdf_atom(u, v, w, B, b);
        if (s < fminf(h0, interior)) { interior = s; am = b; }
    }
    return interior;
}

// utils/util.py:168-228.
__device__ __forceinline__ float fraction_edge(float yl, float yr)
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

__device__ float fraction_triangle(float ya, float yb, float yc)
{
    const float fab = fraction_edge(ya, yb), fbc = fraction_edge(yb, yc), fca = fraction_edge(yc, ya);
    const bool a = ya > 0.0f, b = yb > 0.0f, c = yc > 0.0f;
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

__device__ float fraction_quad(float ya, float yb, float yc, float yd, float ycen)
{
    float f = fraction_triangle(ya, yb, ycen) + fraction_triangle(yb, yc, ycen) + fraction_triangle(yc, yd, ycen) +
              fraction_triangle(yd, ya, ycen);
    f = f * 0.25f;
    if (f < 0.1f) f = 0.0f;
    else if (f > 1.0f) f = 1.0f;
    return f;
}

__device__ __forceinline__ void decode3(uint32_t t, int sy, int sz, int &i, int &j, int &k)
{
    const uint32_t row = t / (uint32_t)sz;
    k = (int)(t - row * (uint32_t)sz);
    i = (int)(row / (uint32_t)sy);
    j = (int)(row - (uint32_t)i * sy);
}

// geometry.py:225-239 precompute_grid_point_phi_prims.
__global__ void __launch_bounds__(256) k_phi_prims(float *__restrict__ phi, int32_t *__restrict__ owner, int sx,
                                                   int sy, int sz, float stx, float sty, float stz, float h,
                                                   Bodies B)
{
    const uint32_t n = (uint32_t)sx * sy * sz;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    decode3(t, sy, sz, i, j, k);
    int32_t am;
    phi[t] = sdf_argmin((float)i * h + stx * h, (float)j * h + sty * h, (float)k * h + stz * h, B, h, am);
    if (owner) owner[t] = am;
}

// geometry.py:264-290 phi_to_fraction for axis dim: corner phi of the
// (nx+1) x (ny+1) x (nz+1) points, frac holds the face-centre phi on entry.
__global__ void __launch_bounds__(256) k_phi_to_fraction(const float *__restrict__ phi, float *__restrict__ frac,
                                                         int nx, int ny, int nz, int dim)
{
    const int fy = ny + (dim == 2), fz = nz + (dim == 3);
    const uint32_t n = (uint32_t)(nx + (dim == 1)) * fy * fz;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    decode3(t, fy, fz, i, j, k);
    const uint32_t py = (uint32_t)ny + 1, pz = (uint32_t)nz + 1;
    auto P = [&](int a, int b, int c) { return phi[((uint32_t)a * py + (uint32_t)b) * pz + (uint32_t)c]; };
    const float ya = P(i, j, k);
    float yb, yc, yd;
    if (dim == 1) { yb = P(i, j + 1, k); yc = P(i, j + 1, k + 1); yd = P(i, j, k + 1); }
    else if (dim == 2) { yb = P(i + 1, j, k); yc = P(i + 1, j, k + 1); yd = P(i, j, k + 1); }
    else { yb = P(i + 1, j, k); yc = P(i + 1, j + 1, k); yd = P(i, j + 1, k); }
    frac[t] = fraction_quad(ya, yb, yc, yd, frac[t]);
}

// geometry.py:401-462 PrecomputeFaceFraction (precompute_fij_with_cache) for
// all three face arrays in one launch: face (i, j, k) of axis d takes the
// fraction_quad of its four corner distances and the distance at its centre,
// and the owner of the centre.  A workgroup covers a tile of FTI x FTJ x FTK
// grid points (i <= nx, j <= ny, k <= nz), each point the corner (i, j, k) of
// the u, v and w faces that exist there; it first evaluates the tile's corner
// distances once into LDS (a one-point halo on the high sides), so a point
// costs about 1.6 corner evaluations instead of 12.  The distances are the
// same function of the same arguments as one evaluation per face corner.
constexpr int FTI = 4, FTJ = 4, FTK = 64;
constexpr int FCI = FTI + 1, FCJ = FTJ + 1, FCK = FTK + 1, FC = FCI * FCJ * FCK;
__global__ void __launch_bounds__(256) k_face_fractions3(float *__restrict__ fu, float *__restrict__ fv,
                                                         float *__restrict__ fw, int32_t *__restrict__ ou,
                                                         int32_t *__restrict__ ov, int32_t *__restrict__ ow, int nx,
                                                         int ny, int nz, float h, Bodies B, int tiles_j, int tiles_k)
{
    __shared__ float cphi[FC];
    const int tk = blockIdx.x % tiles_k, rest = blockIdx.x / tiles_k;
    const int tj = rest % tiles_j, ti = rest / tiles_j;
    const int i0 = ti * FTI, j0 = tj * FTJ, k0 = tk * FTK;
    for (int idx = threadIdx.x; idx < FC; idx += 256) {
        const int ci = idx / (FCJ * FCK), cj = (idx / FCK) % FCJ, ck = idx % FCK;
        const int a = i0 + ci, b = j0 + cj, c = k0 + ck;
        float v = 0.0f;
        if (a <= nx && b <= ny && c <= nz) {
            int32_t unused;
            v = sdf_argmin((float)a * h + 0.0f * h, (float)b * h + 0.0f * h, (float)c * h + 0.0f * h, B, h, unused);
        }
        cphi[idx] = v;
    }
    __syncthreads();
    const int lk = threadIdx.x & 63;
    for (int f = threadIdx.x >> 6; f < FTI * FTJ; f += 4) {
        const int li = f / FTJ, lj = f % FTJ;
        const int i = i0 + li, j = j0 + lj, k = k0 + lk;
        if (i > nx || j > ny || k > nz) continue;
        auto C = [&](int a, int b, int c) { return cphi[((li + a) * FCJ + (lj + b)) * FCK + (lk + c)]; };
        const float ya = C(0, 0, 0);
        int32_t am;
        if (j < ny && k < nz) {   // u face (i, j, k), centre stagger (0, 0.5, 0.5)
            const float ycen = sdf_argmin((float)i * h + 0.0f * h, (float)j * h + 0.5f * h, (float)k * h + 0.5f * h,
                                          B, h, am);
            const uint32_t t = ((uint32_t)i * ny + j) * (uint32_t)nz + k;
            fu[t] = fraction_quad(ya, C(0, 1, 0), C(0, 1, 1), C(0, 0, 1), ycen);
            ou[t] = am;
        }
        if (i < nx && k < nz) {   // v face, stagger (0.5, 0, 0.5)
            const float ycen = sdf_argmin((float)i * h + 0.5f * h, (float)j * h + 0.0f * h, (float)k * h + 0.5f * h,
                                          B, h, am);
            const uint32_t t = ((uint32_t)i * (ny + 1) + j) * (uint32_t)nz + k;
            fv[t] = fraction_quad(ya, C(1, 0, 0), C(1, 0, 1), C(0, 0, 1), ycen);
            ov[t] = am;
        }
        if (i < nx && j < ny) {   // w face, stagger (0.5, 0.5, 0)
            const float ycen = sdf_argmin((float)i * h + 0.5f * h, (float)j * h + 0.5f * h, (float)k * h + 0.0f * h,
                                          B, h, am);
            const uint32_t t = ((uint32_t)i * ny + j) * (uint32_t)(nz + 1) + k;
            fw[t] = fraction_quad(ya, C(1, 0, 0), C(1, 1, 0), C(0, 1, 0), ycen);
            ow[t] = am;
        }
    }
}

// geometry.py:365-388 update_argmin: a body given as a cached level set
// (grid spacing h_ls, its cell ijk at the body origin) resampled at point
// (i h + st h, ...) in the body frame -- (p - x) rotated by q^-1, divided by
// h_ls, less ijk -- by the tricubic interpolation of utils/util.py:98-104
// (trilerp_stagger with stagger 0); the point takes the level set's value,
// and the body's id as its owner (staggered arrays only), where it is below
// the value already there.
struct Pose {
    float x[3], q[4], st[3], ijk[3];
};

__global__ void __launch_bounds__(256) k_update_argmin(float *__restrict__ phi, int32_t *__restrict__ owner, int sx,
                                                       int sy, int sz, Arr ls, int id, Pose P, float h_ls, float h)
{
    const uint32_t n = (uint32_t)sx * sy * sz;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    decode3(t, sy, sz, i, j, k);
    const float p[3] = {(float)i * h + P.st[0] * h - P.x[0], (float)j * h + P.st[1] * h - P.x[1],
                        (float)k * h + P.st[2] * h - P.x[2]};
    float b[3];
    quat_rotate_inv(P.q, p, b);
    const float s = b[0] / h_ls - P.ijk[0], tt = b[1] / h_ls - P.ijk[1], r = b[2] / h_ls - P.ijk[2];
    const float m = tricubic(ls, s - 0.0f, tt - 0.0f, r - 0.0f);
    if (m < phi[t]) {
        phi[t] = m;
        if (owner) owner[t] = id;
    }
}

// rigid_coupling.py:51-70 precompute_solid_vel for axis d: component d of
// v + omega x (p - x) of the owning body on faces with h0/100 < i*h0 <
// 1 - h0/100 (the test reads i for every axis, as the reference does), else 0.
__global__ void __launch_bounds__(256) k_solid_vel(float *__restrict__ vel, const int32_t *__restrict__ owner,
                                                   int sx, int sy, int sz, int d, float h0,
                                                   const float *__restrict__ x, const float *__restrict__ v,
                                                   const float *__restrict__ omega)
{
    const uint32_t n = (uint32_t)sx * sy * sz;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    decode3(t, sy, sz, i, j, k);
    const float eps = h0 * 0.01f;
    const int32_t o = owner[t];
    const float tmp = (float)i * h0;
    float out = 0.0f;
    if (eps < tmp && tmp < 1.0f - eps && o != -1) {
        const float stx = d == 1 ? 0.0f : 0.5f, sty = d == 2 ? 0.0f : 0.5f, stz = d == 3 ? 0.0f : 0.5f;
        const float r0 = ((float)i + stx) * h0 - x[3 * o];
        const float r1 = ((float)j + sty) * h0 - x[3 * o + 1];
        const float r2 = ((float)k + stz) * h0 - x[3 * o + 2];
        const float *om = omega + 3 * o;
        float c;
        if (d == 1) c = om[1] * r2 - om[2] * r1;
        else if (d == 2) c = om[2] * r0 - om[0] * r2;
        else c = om[0] * r1 - om[1] * r0;
        out = v[3 * o + d - 1] + c;
    }
    vel[t] = out;
}

// divergence.py:5-41 divergence_fractional_3d (the sample_3d clamps never
// bind on MAC arrays of the matching shapes).
__global__ void __launch_bounds__(256) k_divergence_fractional(
    const float *__restrict__ u, const float *__restrict__ v, const float *__restrict__ w, float *__restrict__ div,
    const float *__restrict__ fu, const float *__restrict__ fv, const float *__restrict__ fw,
    const float *__restrict__ ur, const float *__restrict__ vr, const float *__restrict__ wr, int nx, int ny, int nz)
{
    const uint32_t n = (uint32_t)nx * ny * nz;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    decode3(t, ny, nz, i, j, k);
    const uint32_t ul = t, uh = t + (uint32_t)ny * nz;
    const uint32_t vl = ((uint32_t)i * (ny + 1) + j) * nz + k, vh = vl + nz;
    const uint32_t wl = ((uint32_t)i * ny + j) * (nz + 1) + k, wh = wl + 1;
    const float l = fu[ul], r = fu[uh], b = fv[vl], tt = fv[vh], f = fw[wl], nn = fw[wh];
    const float fluid = u[uh] * r - u[ul] * l + v[vh] * tt - v[vl] * b + w[wh] * nn - w[wl] * f;
    const float solid = (1.0f - r) * ur[uh] - (1.0f - l) * ur[ul] + (1.0f - tt) * vr[vh] - (1.0f - b) * vr[vl] +
                        (1.0f - nn) * wr[wh] - (1.0f - f) * wr[wl];
    div[t] = fluid + solid;
}

bool dims_ok(int a, int b, int c)
{
    return a > 0 && b > 0 && c > 0 && (uint64_t)a * b * c < (1ull << 32) - 256;
}

bool bodies_ok(int n, const float *x, const float *q, const int32_t *gtype, const float *radius)
{
    return n == 0 || (n > 0 && x && q && gtype && radius);
}

}  // namespace

FLUB_API int flub_phi_prims(float *phi, int32_t *owner, int sx, int sy, int sz, float stx, float sty, float stz,
                            float h, int n_bodies, const float *x, const float *q, const int32_t *gtype,
                            const float *radius, void *stream)
{
    FLUB_REQUIRE(phi && dims_ok(sx, sy, sz), FLUB_EINVAL, "flub_phi_prims: bad array");
    FLUB_REQUIRE(bodies_ok(n_bodies, x, q, gtype, radius), FLUB_EINVAL, "flub_phi_prims: bad body arrays");
    const uint32_t n = (uint32_t)sx * sy * sz;
    const Bodies B{x, q, radius, gtype, n_bodies};
    k_phi_prims<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(phi, owner, sx, sy, sz, stx, sty, stz, h, B);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_phi_to_fraction(const float *phi, float *frac, int nx, int ny, int nz, int dim, void *stream)
{
    FLUB_REQUIRE(phi && frac && dim >= 1 && dim <= 3, FLUB_EINVAL, "flub_phi_to_fraction: bad arguments");
    FLUB_REQUIRE(dims_ok(nx + 1, ny + 1, nz + 1), FLUB_EINVAL, "flub_phi_to_fraction: bad grid");
    const uint32_t n = (uint32_t)(nx + (dim == 1)) * (ny + (dim == 2)) * (nz + (dim == 3));
    k_phi_to_fraction<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(phi, frac, nx, ny, nz, dim);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_face_fractions(float *fu, float *fv, float *fw, int32_t *owner_u, int32_t *owner_v,
                                 int32_t *owner_w, int nx, int ny, int nz, float h, int n_bodies, const float *x,
                                 const float *q, const int32_t *gtype, const float *radius, void *stream)
{
    FLUB_REQUIRE(fu && fv && fw && owner_u && owner_v && owner_w, FLUB_EINVAL, "flub_face_fractions: null array");
    FLUB_REQUIRE(dims_ok(nx + 1, ny + 1, nz + 1), FLUB_EINVAL, "flub_face_fractions: bad grid");
    FLUB_REQUIRE(bodies_ok(n_bodies, x, q, gtype, radius), FLUB_EINVAL, "flub_face_fractions: bad body arrays");
    const Bodies B{x, q, radius, gtype, n_bodies};
    const int ti = (nx + 1 + FTI - 1) / FTI, tj = (ny + 1 + FTJ - 1) / FTJ, tk = (nz + 1 + FTK - 1) / FTK;
    FLUB_REQUIRE((uint64_t)ti * tj * tk < (1ull << 31), FLUB_EINVAL, "flub_face_fractions: grid too large");
    k_face_fractions3<<<(unsigned)(ti * tj * tk), 256, 0, as_stream(stream)>>>(fu, fv, fw, owner_u, owner_v, owner_w,
                                                                               nx, ny, nz, h, B, tj, tk);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_update_argmin(float *phi, int32_t *owner, int sx, int sy, int sz, const float *levelset, int lx,
                                int ly, int lz, int id, const float *x, const float *q, float h_levelset,
                                const float *stagger, const float *ijk, float h, void *stream)
{
    FLUB_REQUIRE(phi && dims_ok(sx, sy, sz), FLUB_EINVAL, "flub_update_argmin: bad array");
    FLUB_REQUIRE(levelset && dims_ok(lx, ly, lz) && (uint64_t)lx * ly * lz * 4u < (1ull << 32), FLUB_EINVAL,
                 "flub_update_argmin: bad level-set array (at most 2^30 values)");
    FLUB_REQUIRE(x && q && stagger && ijk, FLUB_EINVAL, "flub_update_argmin: null pose");
    Pose P;
    for (int a = 0; a < 3; ++a) { P.x[a] = x[a]; P.st[a] = stagger[a]; P.ijk[a] = ijk[a]; }
    for (int a = 0; a < 4; ++a) P.q[a] = q[a];
    const bool staggered = stagger[0] != 0.0f || stagger[1] != 0.0f || stagger[2] != 0.0f;
    const uint32_t n = (uint32_t)sx * sy * sz;
    k_update_argmin<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(phi, staggered ? owner : nullptr, sx, sy, sz,
                                                                       Arr{levelset, lx, ly, lz}, id, P, h_levelset, h);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_solid_vel(float *vel, const int32_t *owner, int sx, int sy, int sz, int d, float h0,
                            const float *x, const float *v, const float *omega, void *stream)
{
    FLUB_REQUIRE(vel && owner && dims_ok(sx, sy, sz) && d >= 1 && d <= 3, FLUB_EINVAL, "flub_solid_vel: bad array");
    FLUB_REQUIRE(x && v && omega, FLUB_EINVAL, "flub_solid_vel: null body array");
    const uint32_t n = (uint32_t)sx * sy * sz;
    k_solid_vel<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(vel, owner, sx, sy, sz, d, h0, x, v, omega);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_divergence_fractional_3d(const float *u, const float *v, const float *w, float *div,
                                           const float *fu, const float *fv, const float *fw, const float *u_rigid,
                                           const float *v_rigid, const float *w_rigid, int nx, int ny, int nz,
                                           void *stream)
{
    FLUB_REQUIRE(u && v && w && div && fu && fv && fw && u_rigid && v_rigid && w_rigid, FLUB_EINVAL,
                 "flub_divergence_fractional_3d: null array");
    FLUB_REQUIRE(dims_ok(nx + 1, ny + 1, nz + 1), FLUB_EINVAL, "flub_divergence_fractional_3d: bad grid");
    const uint32_t n = (uint32_t)nx * ny * nz;
    k_divergence_fractional<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(u, v, w, div, fu, fv, fw, u_rigid,
                                                                               v_rigid, w_rigid, nx, ny, nz);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}
