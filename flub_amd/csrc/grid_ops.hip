// Cell classification, MAC divergence, masking, scaling and gradient
// subtraction: the grid kernels the reference launches around its pressure
// solve (geometry.py:335-397, divergence.py:44-180, utils/util.py:162-165,
// subtract_gradient.py:38-68).  One thread per cell or face, k fastest across
// consecutive lanes so every global access of a wavefront is one contiguous
// 256-byte run.
#include "common.h"
#include "grid_ops.h"

__global__ void k_tag_dirichlet(const float *__restrict__ phi, int32_t *__restrict__ tag, size_t n)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) tag[t] = phi[t] > 0.0f ? 0 : 1;
}

__global__ void k_tag_neumann(const float *__restrict__ phi, int32_t *__restrict__ tag, size_t n)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && phi[t] < 0.0f) tag[t] = 2;
}

FLUB_API int flub_tag_dirichlet(const float *phi, int32_t *tag, size_t n, void *stream)
{
    if (!n) return FLUB_OK;
    FLUB_REQUIRE(phi && tag, FLUB_EINVAL, "flub_tag_dirichlet: null pointer");
    k_tag_dirichlet<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(phi, tag, n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_tag_neumann(const float *phi_rigid, int32_t *tag, size_t n, void *stream)
{
    if (!n) return FLUB_OK;
    FLUB_REQUIRE(phi_rigid && tag, FLUB_EINVAL, "flub_tag_neumann: null pointer");
    k_tag_neumann<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(phi_rigid, tag, n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

// divergence.py:44-99.  FUSE additionally applies mask_with_tag (:174-180) and
// scale_by_3d(div, s) in registers: (tag==FLUID ? F : 0) * s, which is the
// value the three separate launches store.
template <typename TagT, bool FUSE>
__global__ void k_divergence(const float *__restrict__ u, const float *__restrict__ v,
                             const float *__restrict__ w, float *__restrict__ div,
                             const float *__restrict__ ur, const float *__restrict__ vr,
                             const float *__restrict__ wr, const TagT *__restrict__ tag, int nx,
                             int ny, int nz, float s)
{
    const uint32_t n = (uint32_t)nx * ny * nz;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = t % nz;
    const int j = (t / nz) % ny;
    const int i = t / ((uint32_t)nz * ny);
    const uint32_t plane = (uint32_t)ny * nz;
    const uint32_t iu0 = t, iu1 = t + plane;
    const uint32_t iv0 = ((uint32_t)i * (ny + 1) + j) * nz + k, iv1 = iv0 + nz;
    const uint32_t iw0 = ((uint32_t)i * ny + j) * (nz + 1) + k, iw1 = iw0 + 1;
    float vl = u[iu0], vrr = u[iu1], vb = v[iv0], vt = v[iv1], vf = w[iw0], vn = w[iw1];
    if (i == 0) vl = 0.0f;
    else if (tag[t - plane] == 2) vl = ur[iu0];
    if (i == nx - 1) vrr = 0.0f;
    else if (tag[t + plane] == 2) vrr = ur[iu1];
    if (j == 0) vb = 0.0f;
    else if (tag[t - nz] == 2) vb = vr[iv0];
    if (j == ny - 1) vt = 0.0f;
    else if (tag[t + nz] == 2) vt = vr[iv1];
    if (k == 0) vf = 0.0f;
    else if (tag[t - 1] == 2) vf = wr[iw0];
    if (k == nz - 1) vn = 0.0f;
    else if (tag[t + 1] == 2) vn = wr[iw1];
    float F = vrr - vl + vt - vb + vn - vf;
    if (FUSE) F = (tag[t] == 1 ? F : 0.0f) * s;
    div[t] = F;
}

// 4 consecutive values (16-byte or 4-byte aligned vector loads and stores)
__device__ __forceinline__ void ldf4(const float *p, float v[4])
{
    const float4 a = *reinterpret_cast<const float4 *>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void ldf4(const double *p, double v[4])
{
    const double2 a = reinterpret_cast<const double2 *>(p)[0], b = reinterpret_cast<const double2 *>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void stf4(float *p, const float v[4])
{
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void ldt4(const uint8_t *p, int v[4])
{
    const uchar4 a = *reinterpret_cast<const uchar4 *>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void ldt4(const int32_t *p, int v[4])
{
    const int4 a = *reinterpret_cast<const int4 *>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}

// k_divergence for 4 consecutive cells along k per thread (nz % 4 == 0,
// 16-byte aligned arrays): u, v, the tag rows and div move as vectors, w
// (rows of nz + 1) and the rigid values as scalars; the per-cell arithmetic
// is k_divergence's.
//
// RSPLIT (the fused projection; FUSE): instead of, or (div != nullptr) as
// well as, div it writes the solver's right-hand side r = (FLUID ? d * k : +0)
// with d the fused value (the scale copy and mask_interior of the solve's
// init), in the colour-split layout: cell (i, j, k) in half (i + j + k) & 1
// at (i * ny + j) * nz/2 + k/2; the same values into rcopy and +0 into zero0
// and zero1 at the same positions when they are set (the mixed solve's copy of
// the right-hand side; the solver's x and p).
template <typename TagT, bool FUSE, bool RSPLIT = false>
__global__ void __launch_bounds__(256) k_divergence4(const float *__restrict__ u, const float *__restrict__ v,
                                                     const float *__restrict__ w, float *__restrict__ div,
                                                     const float *__restrict__ ur, const float *__restrict__ vr,
                                                     const float *__restrict__ wr, const TagT *__restrict__ tag,
                                                     int nx, int ny, int nz, float s, float *__restrict__ rsplit = nullptr,
                                                     float k_r = 0.0f, float *__restrict__ rcopy = nullptr,
                                                     float *__restrict__ zero0 = nullptr,
                                                     float *__restrict__ zero1 = nullptr)
{
    const uint32_t nq = (uint32_t)nz >> 2;
    const uint32_t tq = blockIdx.x * 256u + threadIdx.x;
    if (tq >= (uint32_t)nx * ny * nq) return;
    const uint32_t row = tq / nq, q = tq - row * nq;
    const int i = (int)(row / (uint32_t)ny), j = (int)(row - (uint32_t)i * ny);
    const int k0 = 4 * (int)q;
    const uint32_t plane = (uint32_t)ny * nz;
    const uint32_t t = row * (uint32_t)nz + k0;
    const uint32_t iv0 = ((uint32_t)i * (ny + 1) + j) * nz + k0;
    const uint32_t iw0 = ((uint32_t)i * ny + j) * (nz + 1) + k0;
    float U0[4], U1[4], V0[4], V1[4], W[5];
    ldf4(u + t, U0);
    ldf4(u + t + plane, U1);
    ldf4(v + iv0, V0);
    ldf4(v + iv0 + nz, V1);
#pragma unroll
    for (int e = 0; e < 5; ++e) W[e] = w[iw0 + e];
    int T[4], Txm[4] = {0, 0, 0, 0}, Txp[4] = {0, 0, 0, 0}, Tym[4] = {0, 0, 0, 0}, Typ[4] = {0, 0, 0, 0};
    ldt4(tag + t, T);
    if (i > 0) ldt4(tag + t - plane, Txm);
    if (i < nx - 1) ldt4(tag + t + plane, Txp);
    if (j > 0) ldt4(tag + t - nz, Tym);
    if (j < ny - 1) ldt4(tag + t + nz, Typ);
    const int tk0 = k0 > 0 ? (int)tag[t - 1] : 0, tk4 = k0 + 4 < nz ? (int)tag[t + 4] : 0;
    float out[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int k = k0 + e;
        float vl = U0[e], vrr = U1[e], vb = V0[e], vt = V1[e], vf = W[e], vn = W[e + 1];
        if (i == 0) vl = 0.0f;
        else if (Txm[e] == 2) vl = ur[t + e];
        if (i == nx - 1) vrr = 0.0f;
        else if (Txp[e] == 2) vrr = ur[t + e + plane];
        if (j == 0) vb = 0.0f;
        else if (Tym[e] == 2) vb = vr[iv0 + e];
        if (j == ny - 1) vt = 0.0f;
        else if (Typ[e] == 2) vt = vr[iv0 + e + nz];
        const int tm = e > 0 ? T[e - 1] : tk0, tp = e < 3 ? T[e + 1] : tk4;
        if (k == 0) vf = 0.0f;
        else if (tm == 2) vf = wr[iw0 + e];
        if (k == nz - 1) vn = 0.0f;
        else if (tp == 2) vn = wr[iw0 + e + 1];
        float F = vrr - vl + vt - vb + vn - vf;
        if (FUSE) F = (T[e] == 1 ? F : 0.0f) * s;
        out[e] = F;
    }
    if (!RSPLIT || div) stf4(div + t, out);
    if (RSPLIT) {
        // k0 is even: cells k0, k0 + 2 go to half (i + j) & 1, k0 + 1, k0 + 3 to the other
        float rv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) rv[e] = T[e] == 1 ? out[e] * k_r : 0.0f;
        const uint32_t c0 = (uint32_t)((i + j) & 1), half = (uint32_t)nx * ny * (uint32_t)(nz >> 1);
        const uint32_t b = row * (uint32_t)(nz >> 1) + (uint32_t)(k0 >> 1);
        *reinterpret_cast<float2 *>(rsplit + c0 * half + b) = make_float2(rv[0], rv[2]);
        *reinterpret_cast<float2 *>(rsplit + (c0 ^ 1u) * half + b) = make_float2(rv[1], rv[3]);
        if (rcopy) {
            *reinterpret_cast<float2 *>(rcopy + c0 * half + b) = make_float2(rv[0], rv[2]);
            *reinterpret_cast<float2 *>(rcopy + (c0 ^ 1u) * half + b) = make_float2(rv[1], rv[3]);
        }
        const float2 z2 = make_float2(0.0f, 0.0f);
        if (zero0) {
            *reinterpret_cast<float2 *>(zero0 + c0 * half + b) = z2;
            *reinterpret_cast<float2 *>(zero0 + (c0 ^ 1u) * half + b) = z2;
        }
        if (zero1) {
            *reinterpret_cast<float2 *>(zero1 + c0 * half + b) = z2;
            *reinterpret_cast<float2 *>(zero1 + (c0 ^ 1u) * half + b) = z2;
        }
    }
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

static bool grid_ok(int nx, int ny, int nz)
{
    if (nx < 1 || ny < 1 || nz < 1) return false;
    const double cells = (double)(nx + 1) * (ny + 1) * (nz + 1);
    return cells < 2147483647.0;
}

template <typename TagT, bool FUSE>
static int divergence_impl(const float *u, const float *v, const float *w, float *div, const float *ur,
                           const float *vr, const float *wr, const TagT *tag, int nx, int ny, int nz,
                           float s, void *stream)
{
    FLUB_REQUIRE(u && v && w && div && ur && vr && wr && tag, FLUB_EINVAL, "divergence: null pointer");
    FLUB_REQUIRE(grid_ok(nx, ny, nz), FLUB_ESHAPE, "divergence: bad grid %dx%dx%d", nx, ny, nz);
    const size_t n = (size_t)nx * ny * nz;
    if (nz % 4 == 0 && aligned16(u) && aligned16(v) && aligned16(div) && aligned16(tag))
        k_divergence4<TagT, FUSE><<<blocks_for(n / 4, 256), 256, 0, as_stream(stream)>>>(u, v, w, div, ur, vr, wr, tag,
                                                                                     nx, ny, nz, s);
    else
        k_divergence<TagT, FUSE><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(u, v, w, div, ur, vr, wr, tag,
                                                                                  nx, ny, nz, s);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_divergence_tag(const float *u, const float *v, const float *w, float *div,
                                 const float *u_rigid, const float *v_rigid, const float *w_rigid,
                                 const int32_t *tag, int nx, int ny, int nz, void *stream)
{
    return divergence_impl<int32_t, false>(u, v, w, div, u_rigid, v_rigid, w_rigid, tag, nx, ny, nz,
                                           1.0f, stream);
}

int flub_divergence_fused_u8(const float *u, const float *v, const float *w, float *div,
                             const float *ur, const float *vr, const float *wr, const uint8_t *tag,
                             int nx, int ny, int nz, float s, void *stream)
{
    return divergence_impl<uint8_t, true>(u, v, w, div, ur, vr, wr, tag, nx, ny, nz, s, stream);
}

bool divergence_split_ok(const float *u, const float *v, const float *div, const uint8_t *tag, int nz)
{
    return nz % 8 == 0 && aligned16(u) && aligned16(v) && (!div || aligned16(div)) && aligned16(tag);
}

int divergence_fused_split(const float *u, const float *v, const float *w, float *div, const float *ur,
                           const float *vr, const float *wr, const uint8_t *tag, int nx, int ny, int nz, float s,
                           float *rsplit, float k_r, void *stream, float *rcopy, float *zero0, float *zero1)
{
    FLUB_REQUIRE(u && v && w && ur && vr && wr && tag && rsplit, FLUB_EINVAL, "divergence: null pointer");
    FLUB_REQUIRE(grid_ok(nx, ny, nz) && divergence_split_ok(u, v, div, tag, nz), FLUB_ESHAPE,
                 "divergence: the split right-hand side needs nz %% 8 == 0 and aligned arrays");
    const size_t n = (size_t)nx * ny * nz;
    k_divergence4<uint8_t, true, true><<<blocks_for(n / 4, 256), 256, 0, as_stream(stream)>>>(
        u, v, w, div, ur, vr, wr, tag, nx, ny, nz, s, rsplit, k_r, rcopy, zero0, zero1);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

__global__ void k_mask_with_tag(float *__restrict__ div, const int32_t *__restrict__ tag, size_t n)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && tag[t] != 1) div[t] = 0.0f;
}

FLUB_API int flub_mask_with_tag(float *div, const int32_t *tag, size_t n, void *stream)
{
    if (!n) return FLUB_OK;
    FLUB_REQUIRE(div && tag, FLUB_EINVAL, "flub_mask_with_tag: null pointer");
    k_mask_with_tag<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(div, tag, n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

__global__ void k_scale(float *__restrict__ q, float s, size_t n)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) q[t] = q[t] * s;
}

FLUB_API int flub_scale(float *q, float s, size_t n, void *stream)
{
    if (!n) return FLUB_OK;
    FLUB_REQUIRE(q, FLUB_EINVAL, "flub_scale: null pointer");
    k_scale<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(q, s, n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

// subtract_gradient.py:38-68 over the face grid of axis d.  PS != 0 reads the
// pressure as (x * h) * (1/h) -- the value the reference's scale chain
// (water_3d.py:67-68, example_fluid.py:134) hands to the kernel -- so the
// fused projection can read x directly.  PT: the pressure's type; with FP64
// (the fused projection's FP64 solution, PS false) the face update
// u - (p_r - p_l) is computed in FP64 and rounded once.
template <typename TagT, bool PS, typename PT = float>
__global__ void k_subtract_gradient(float *__restrict__ uf, const float *__restrict__ urig,
                                    const PT *__restrict__ p, const TagT *__restrict__ tag, int nx,
                                    int ny, int nz, int d, float h, float inv_h)
{
    const int fx = nx + (d == 1), fy = ny + (d == 2), fz = nz + (d == 3);
    const uint32_t nf = (uint32_t)fx * fy * fz;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf) return;
    const int k = t % fz;
    const int j = (t / fz) % fy;
    const int i = t / ((uint32_t)fz * fy);
    const int ii = d == 1 ? i : (d == 2 ? j : k);
    const int n_d = d == 1 ? nx : (d == 2 ? ny : nz);
    float out;
    if (0 < ii && ii < n_d) {
        const uint32_t I = cidx(i, j, k, ny, nz);
        const uint32_t off = d == 1 ? (uint32_t)ny * nz : (d == 2 ? (uint32_t)nz : 1u);
        PT pl = p[I - off], pr = p[I];
        if (PS) { pl = (pl * h) * inv_h; pr = (pr * h) * inv_h; }
        const int tl = tag[I - off], tr = tag[I];
        if (tl == 2 || tr == 2) out = urig[t];
        else if (tl == 1 || tr == 1) out = (float)((PT)uf[t] - (pr - pl));
        else out = 0.0f;
    } else {
        out = 0.0f;
    }
    uf[t] = out;
}

// k_subtract_gradient for the u and v faces (rows of nz faces), 4 consecutive
// faces along k per thread (nz % 4 == 0, 16-byte aligned arrays); the rigid
// velocities are read only by vectors that have a SOLID side.
template <typename TagT, bool PS, typename PT = float>
__global__ void __launch_bounds__(256) k_subtract_gradient4(float *__restrict__ uf, const float *__restrict__ urig,
                                                            const PT *__restrict__ p,
                                                            const TagT *__restrict__ tag, int nx, int ny, int nz,
                                                            int d, float h, float inv_h)
{
    const int fx = nx + (d == 1), fy = ny + (d == 2);
    const uint32_t nq = (uint32_t)nz >> 2;
    const uint32_t tq = blockIdx.x * 256u + threadIdx.x;
    if (tq >= (uint32_t)fx * fy * nq) return;
    const uint32_t row = tq / nq, q = tq - row * nq;
    const int i = (int)(row / (uint32_t)fy), j = (int)(row - (uint32_t)i * fy);
    const int k0 = 4 * (int)q;
    const uint32_t t = row * (uint32_t)nz + k0;
    const int ii = d == 1 ? i : j, n_d = d == 1 ? nx : ny;
    float out[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (0 < ii && ii < n_d) {
        const uint32_t I = cidx(i, j, k0, ny, nz);
        const uint32_t off = d == 1 ? (uint32_t)ny * nz : (uint32_t)nz;
        PT PL[4], PR[4];
        float UF[4], UR[4];
        int TL[4], TR[4];
        ldf4(p + I - off, PL);
        ldf4(p + I, PR);
        ldt4(tag + I - off, TL);
        ldt4(tag + I, TR);
        ldf4(uf + t, UF);
        bool solid = false;
#pragma unroll
        for (int e = 0; e < 4; ++e) solid = solid || TL[e] == 2 || TR[e] == 2;
        if (solid) ldf4(urig + t, UR);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            PT pl = PL[e], pr = PR[e];
            if (PS) { pl = (pl * h) * inv_h; pr = (pr * h) * inv_h; }
            if (TL[e] == 2 || TR[e] == 2) out[e] = UR[e];
            else if (TL[e] == 1 || TR[e] == 1) out[e] = (float)((PT)UF[e] - (pr - pl));
        }
    }
    stf4(uf + t, out);
}

// k_subtract_gradient for the w faces (rows of nz + 1 faces): face k lies
// between cells k - 1 and k, so the 4 faces k0 .. k0 + 3 of a thread read the
// cells k0 .. k0 + 3 as one vector and cell k0 - 1 as a scalar; the face
// values (unaligned rows) are scalars.  The thread of the last vector also
// writes the boundary face k = nz.
template <typename TagT, bool PS, typename PT = float>
__global__ void __launch_bounds__(256) k_subtract_gradient4w(float *__restrict__ uf, const float *__restrict__ urig,
                                                             const PT *__restrict__ p,
                                                             const TagT *__restrict__ tag, int nx, int ny, int nz,
                                                             float h, float inv_h)
{
    const uint32_t nq = (uint32_t)nz >> 2;
    const uint32_t tq = blockIdx.x * 256u + threadIdx.x;
    if (tq >= (uint32_t)nx * ny * nq) return;
    const uint32_t row = tq / nq, q = tq - row * nq;
    const int k0 = 4 * (int)q;
    const uint32_t I = row * (uint32_t)nz + k0, t = row * (uint32_t)(nz + 1) + k0;
    PT PR[4];
    int TR[4];
    ldf4(p + I, PR);
    ldt4(tag + I, TR);
    const PT p_m = k0 > 0 ? p[I - 1] : (PT)0;
    const int t_m = k0 > 0 ? (int)tag[I - 1] : 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int k = k0 + e;
        float out = 0.0f;
        if (k > 0) {
            PT pl = e > 0 ? PR[e - 1] : p_m, pr = PR[e];
            if (PS) { pl = (pl * h) * inv_h; pr = (pr * h) * inv_h; }
            const int tl = e > 0 ? TR[e - 1] : t_m, tr = TR[e];
            if (tl == 2 || tr == 2) out = urig[t + e];
            else if (tl == 1 || tr == 1) out = (float)((PT)uf[t + e] - (pr - pl));
        }
        uf[t + e] = out;
    }
    if (q == nq - 1) uf[t + 4] = 0.0f;   // k = nz
}

// The three face updates of subtract_gradient_tag (subtract_gradient.py:38-68,
// d = 1, 2, 3) in one pass from the fused projection's FP64 solution in the
// colour-split layout (cell (i, j, k) of colour c = (i + j + k) & 1 at
// c * half + (i * ny + j) * nz / 2 + k / 2).  A thread takes 4 cells along k
// and updates the u, v and w faces on their low sides (each face rounded once
// from the FP64 difference, as flub_subtract_gradient_x64_u8 does); the
// threads of the last cells of a row / column / plane also zero the
// boundary faces i = nx, j = ny, k = nz.  x is read once instead of once per
// axis, and no reference-layout copy of it is written.
__global__ void __launch_bounds__(256) k_subtract_gradient3_split(float *__restrict__ u, float *__restrict__ v,
                                                                  float *__restrict__ w,
                                                                  const float *__restrict__ ur,
                                                                  const float *__restrict__ vr,
                                                                  const float *__restrict__ wr,
                                                                  const double *__restrict__ xs,
                                                                  const uint8_t *__restrict__ tag, int nx, int ny,
                                                                  int nz)
{
    const uint32_t nq = (uint32_t)nz >> 2, nh = (uint32_t)nz >> 1;
    const uint32_t half = (uint32_t)nx * (uint32_t)ny * nh;
    const uint32_t tq = blockIdx.x * 256u + threadIdx.x;
    if (tq >= (uint32_t)nx * ny * nq) return;
    const uint32_t row = tq / nq, q = tq - row * nq;
    const int i = (int)(row / (uint32_t)ny), j = (int)(row - (uint32_t)i * ny);
    const int k0 = 4 * (int)q;
    const uint32_t I = row * (uint32_t)nz + (uint32_t)k0;
    // the 4 cells k0 .. k0 + 3 of row r: colours c0, c0 ^ 1 alternate
    auto ldx = [&](uint32_t r, int ci, int cj, double x[4]) {
        const uint32_t c0 = (uint32_t)((ci + cj) & 1);
        const uint32_t b = r * nh + (uint32_t)(k0 >> 1);
        const double2 a = *reinterpret_cast<const double2 *>(xs + c0 * half + b);
        const double2 o = *reinterpret_cast<const double2 *>(xs + (c0 ^ 1u) * half + b);
        x[0] = a.x; x[1] = o.x; x[2] = a.y; x[3] = o.y;
    };
    double X[4];
    int T[4];
    ldx(row, i, j, X);
    ldt4(tag + I, T);
    // u faces (i, j, k0..k0+3): between cells i - 1 and i
    {
        float out[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const uint32_t f = I;   // u is (nx+1) x ny x nz: face (i, j, k) at (i * ny + j) * nz + k
        if (i > 0) {
            double XL[4];
            int TL[4];
            ldx(row - (uint32_t)ny, i - 1, j, XL);
            ldt4(tag + I - (uint32_t)ny * nz, TL);
            float UF[4], UR[4];
            ldf4(u + f, UF);
            bool solid = false;
#pragma unroll
            for (int e = 0; e < 4; ++e) solid = solid || TL[e] == 2 || T[e] == 2;
            if (solid) ldf4(ur + f, UR);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (TL[e] == 2 || T[e] == 2) out[e] = UR[e];
                else if (TL[e] == 1 || T[e] == 1) out[e] = (float)((double)UF[e] - (X[e] - XL[e]));
            }
        }
        stf4(u + f, out);
        if (i == nx - 1) {
            const float zero[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            stf4(u + f + (uint32_t)ny * nz, zero);   // face i = nx
        }
    }
    // v faces (i, j, k0..k0+3): between cells j - 1 and j; v is nx x (ny+1) x nz
    {
        float out[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const uint32_t f = ((uint32_t)i * (ny + 1) + (uint32_t)j) * nz + (uint32_t)k0;
        if (j > 0) {
            double XL[4];
            int TL[4];
            ldx(row - 1u, i, j - 1, XL);
            ldt4(tag + I - (uint32_t)nz, TL);
            float VF[4], VR[4];
            ldf4(v + f, VF);
            bool solid = false;
#pragma unroll
            for (int e = 0; e < 4; ++e) solid = solid || TL[e] == 2 || T[e] == 2;
            if (solid) ldf4(vr + f, VR);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (TL[e] == 2 || T[e] == 2) out[e] = VR[e];
                else if (TL[e] == 1 || T[e] == 1) out[e] = (float)((double)VF[e] - (X[e] - XL[e]));
            }
        }
        stf4(v + f, out);
        if (j == ny - 1) {
            const float zero[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            stf4(v + f + (uint32_t)nz, zero);   // face j = ny
        }
    }
    // w faces (i, j, k0..k0+3): between cells k - 1 and k; w rows hold nz + 1 faces
    {
        const uint32_t f = row * (uint32_t)(nz + 1) + (uint32_t)k0;
        double xm = 0.0;
        int tm = 0;
        if (k0 > 0) {
            const uint32_t c = (uint32_t)((i + j + k0 - 1) & 1);
            xm = xs[c * half + row * nh + (uint32_t)((k0 - 1) >> 1)];
            tm = tag[I - 1];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int k = k0 + e;
            float out = 0.0f;
            if (k > 0) {
                const double pl = e > 0 ? X[e - 1] : xm;
                const int tl = e > 0 ? T[e - 1] : tm, tr = T[e];
                if (tl == 2 || tr == 2) out = wr[f + e];
                else if (tl == 1 || tr == 1) out = (float)((double)w[f + e] - (X[e] - pl));
            }
            w[f + e] = out;
        }
        if (q == nq - 1) w[f + 4] = 0.0f;   // face k = nz
    }
}

int subtract_gradient3_split(float *u, float *v, float *w, const float *ur, const float *vr, const float *wr,
                             const double *xs, const uint8_t *tag, int nx, int ny, int nz, void *stream)
{
    FLUB_REQUIRE(u && v && w && ur && vr && wr && xs && tag, FLUB_EINVAL, "subtract_gradient3_split: null pointer");
    FLUB_REQUIRE(grid_ok(nx, ny, nz) && nz % 4 == 0, FLUB_ESHAPE, "subtract_gradient3_split: bad grid");
    const size_t n4 = (size_t)nx * ny * (nz / 4);
    k_subtract_gradient3_split<<<blocks_for(n4, 256), 256, 0, as_stream(stream)>>>(u, v, w, ur, vr, wr, xs, tag, nx,
                                                                                   ny, nz);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

bool subtract_gradient3_split_ok(const float *u, const float *v, const float *ur, const float *vr, int nz)
{
    return nz % 4 == 0 && aligned16(u) && aligned16(v) && aligned16(ur) && aligned16(vr);
}

template <typename TagT, bool PS, typename PT = float>
static int subgrad_impl(float *uf, const float *urig, const PT *p, const TagT *tag, int nx, int ny,
                        int nz, int d, float h, void *stream)
{
    FLUB_REQUIRE(uf && urig && p && tag, FLUB_EINVAL, "subtract_gradient_tag: null pointer");
    FLUB_REQUIRE(d >= 1 && d <= 3, FLUB_EINVAL, "subtract_gradient_tag: d must be 1, 2 or 3 (got %d)", d);
    FLUB_REQUIRE(grid_ok(nx, ny, nz), FLUB_ESHAPE, "subtract_gradient_tag: bad grid");
    const size_t nf = (size_t)(nx + (d == 1)) * (ny + (d == 2)) * (nz + (d == 3));
    const float inv_h = (float)(1.0 / (double)h);
    if (d != 3 && nz % 4 == 0 && aligned16(uf) && aligned16(urig) && aligned16(p) && aligned16(tag))
        k_subtract_gradient4<TagT, PS, PT><<<blocks_for(nf / 4, 256), 256, 0, as_stream(stream)>>>(uf, urig, p, tag,
                                                                                                 nx, ny, nz, d, h,
                                                                                                 inv_h);
    else if (d == 3 && nz % 4 == 0 && aligned16(p) && aligned16(tag))
        k_subtract_gradient4w<TagT, PS, PT><<<blocks_for((size_t)nx * ny * (nz / 4), 256), 256, 0,
                                              as_stream(stream)>>>(uf, urig, p, tag, nx, ny, nz, h, inv_h);
    else
        k_subtract_gradient<TagT, PS, PT><<<blocks_for(nf, 256), 256, 0, as_stream(stream)>>>(uf, urig, p, tag, nx,
                                                                                            ny, nz, d, h, inv_h);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_subtract_gradient_tag(float *uf, const float *u_rigid, const float *p,
                                        const int32_t *tag, int nx, int ny, int nz, int d, void *stream)
{
    return subgrad_impl<int32_t, false>(uf, u_rigid, p, tag, nx, ny, nz, d, 1.0f, stream);
}

int flub_subtract_gradient_x_u8(float *uf, const float *u_rigid, const float *x, const uint8_t *tag,
                                int nx, int ny, int nz, int d, float h, void *stream)
{
    return subgrad_impl<uint8_t, true>(uf, u_rigid, x, tag, nx, ny, nz, d, h, stream);
}

FLUB_API int flub_subtract_gradient_tag_f64(float *uf, const float *u_rigid, const double *p, const int32_t *tag,
                                            int nx, int ny, int nz, int d, void *stream)
{
    return subgrad_impl<int32_t, false, double>(uf, u_rigid, p, tag, nx, ny, nz, d, 1.0f, stream);
}

int flub_subtract_gradient_x64_u8(float *uf, const float *u_rigid, const double *x, const uint8_t *tag,
                                  int nx, int ny, int nz, int d, void *stream)
{
    return subgrad_impl<uint8_t, false, double>(uf, u_rigid, x, tag, nx, ny, nz, d, 1.0f, stream);
}
