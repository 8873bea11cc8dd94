// Velocity extrapolation into the air (extrapolate.py:9-47) and its marks
// (utils/util.py:424-437 set_unknown), the step after the projection in the
// reference's water loop (water_3d.py:70-73).
//
// A face marked 1 takes the mean of the faces marked 0 in its in-range
// 27-neighbourhood (summed with the offsets ii, jj, kk from -1 to 1, ii
// slowest, as the reference loops) and is marked -1; with none it gets 0.0.
// The launch runs in place: threads read only faces marked 0, which no thread
// writes, and write only faces marked 1, so the result does not depend on the
// order in which faces are processed.  flush_1_3d then turns every -1 into 1,
// including -1s that were there before the pass: set_unknown writes only the
// first nx x ny x nz marks of each face array, and the reference reuses the
// rigid-owner arrays for the marks, so their last face layer still holds owner
// values (-1: no owner).  After one extrapolate_by_1 the marks are in {0, 1}
// plus owner indices >= 2 (never read or written); the second pass also fills
// the faces that were -1, and from then on every pass recomputes the same
// faces from the same known values.
#include "common.h"

namespace {

__device__ __forceinline__ void decode3(uint32_t t, int sy, int sz, int &i, int &j, int &k)
{
    const uint32_t row = t / (uint32_t)sz;
    k = (int)(t - row * (uint32_t)sz);
    i = (int)(row / (uint32_t)sy);
    j = (int)(row - (uint32_t)i * sy);
}

__global__ void __launch_bounds__(256) k_set_unknown(const float *__restrict__ phi, int32_t *__restrict__ tu,
                                                     int32_t *__restrict__ tv, int32_t *__restrict__ tw, int nx,
                                                     int ny, int nz)
{
    const uint32_t n = (uint32_t)nx * ny * nz;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    decode3(t, ny, nz, i, j, k);
    const int32_t f = phi[t] >= 0.0f ? 1 : 0;
    tu[t] = f;
    tv[((uint32_t)i * (ny + 1) + j) * nz + k] = f;
    tw[((uint32_t)i * ny + j) * (nz + 1) + k] = f;
}

// One extrapolate_3d launch on a face array (MODE 0), also applying
// flush_1_3d (MODE 1: the extrapolate + flush pair in one pass: extrapolated
// faces keep mark 1 and faces marked -1 get 1; else extrapolated faces are
// marked -1), or (MODE 2) two or more extrapolate + flush passes in one: the
// faces marked 1 or -1 are all filled and marked 1.  MODE 2 equals the passes
// because only faces marked 0 are read and none is ever written: the second
// pass recomputes the first pass's faces from the same values and adds the
// faces the first pass flushed from -1, and later passes repeat the second.
// Faces marked -1 are never read by other threads (only marks 0 are), so
// changing their marks in the same launch changes nothing the extrapolation
// sees.
template <int MODE>
__global__ void __launch_bounds__(256) k_extrapolate(float *__restrict__ q, int32_t *__restrict__ unknown, int sx,
                                                     int sy, int sz)
{
    const uint32_t n = (uint32_t)sx * sy * sz;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    const int32_t mark = unknown[t];
    if (MODE >= 1 && mark == -1) unknown[t] = 1;
    if (!(mark == 1 || (MODE == 2 && mark == -1))) return;
    int i, j, k;
    decode3(t, sy, sz, i, j, k);
    int cnt = 0;
    float s = 0.0f;
    for (int a = -1; a <= 1; ++a) {
        const int ii = i + a;
        if (ii < 0 || ii >= sx) continue;
        for (int b = -1; b <= 1; ++b) {
            const int jj = j + b;
            if (jj < 0 || jj >= sy) continue;
            const uint32_t row = ((uint32_t)ii * sy + jj) * (uint32_t)sz;
            for (int c = -1; c <= 1; ++c) {
                const int kk = k + c;
                if (kk < 0 || kk >= sz) continue;
                if (unknown[row + kk] == 0) {
                    cnt += 1;
                    s += q[row + kk];
                }
            }
        }
    }
    if (cnt > 0) {
        q[t] = s / (float)cnt;
        if (MODE == 0) unknown[t] = -1;
    } else {
        q[t] = 0.0f;
    }
}

__global__ void __launch_bounds__(256) k_flush_1(int32_t *__restrict__ tag, size_t n)
{
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n && tag[t] == -1) tag[t] = 1;
}

bool dims_ok(int a, int b, int c) { return a > 0 && b > 0 && c > 0 && (uint64_t)a * b * c < (1ull << 32) - 256; }

}  // namespace

FLUB_API int flub_set_unknown(const float *phi, int32_t *tu, int32_t *tv, int32_t *tw, int nx, int ny, int nz,
                              void *stream)
{
    FLUB_REQUIRE(phi && tu && tv && tw && dims_ok(nx + 1, ny + 1, nz + 1), FLUB_EINVAL, "flub_set_unknown: bad argument");
    const uint32_t n = (uint32_t)nx * ny * nz;
    k_set_unknown<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(phi, tu, tv, tw, nx, ny, nz);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_extrapolate_3d(float *q, int32_t *unknown, int sx, int sy, int sz, void *stream)
{
    FLUB_REQUIRE(q && unknown && dims_ok(sx, sy, sz), FLUB_EINVAL, "flub_extrapolate_3d: bad argument");
    const uint32_t n = (uint32_t)sx * sy * sz;
    k_extrapolate<0><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(q, unknown, sx, sy, sz);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_flush_1_3d(int32_t *tag, size_t n, void *stream)
{
    FLUB_REQUIRE(tag, FLUB_EINVAL, "flub_flush_1_3d: null pointer");
    if (!n) return FLUB_OK;
    k_flush_1<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(tag, n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

namespace {
// extrapolate_by_1 (MODE 1), or the result of two or more of them (MODE 2)
int extrapolate_faces(float *u, float *v, float *w, int32_t *tu, int32_t *tv, int32_t *tw, int nx, int ny, int nz,
                      bool several, void *stream)
{
    hipStream_t s = as_stream(stream);
    float *q[3] = {u, v, w};
    int32_t *t[3] = {tu, tv, tw};
    for (int d = 0; d < 3; ++d) {
        const int sx = nx + (d == 0), sy = ny + (d == 1), sz = nz + (d == 2);
        const uint32_t n = (uint32_t)sx * sy * sz;
        if (several) k_extrapolate<2><<<blocks_for(n, 256), 256, 0, s>>>(q[d], t[d], sx, sy, sz);
        else k_extrapolate<1><<<blocks_for(n, 256), 256, 0, s>>>(q[d], t[d], sx, sy, sz);
        FLUB_CHECK_LAUNCH();
    }
    return FLUB_OK;
}
}  // namespace

FLUB_API int flub_extrapolate_by_1(float *u, float *v, float *w, int32_t *tu, int32_t *tv, int32_t *tw, int nx, int ny,
                                   int nz, void *stream)
{
    FLUB_REQUIRE(u && v && w && tu && tv && tw && dims_ok(nx + 1, ny + 1, nz + 1), FLUB_EINVAL,
                 "flub_extrapolate_by_1: bad argument");
    return extrapolate_faces(u, v, w, tu, tv, tw, nx, ny, nz, false, stream);
}

FLUB_API int flub_extrapolate(float *u, float *v, float *w, int32_t *tu, int32_t *tv, int32_t *tw, const float *phi,
                              int nx, int ny, int nz, int n_passes, void *stream)
{
    FLUB_REQUIRE(phi && n_passes >= 0, FLUB_EINVAL, "flub_extrapolate: bad argument");
    FLUB_REQUIRE(u && v && w && tu && tv && tw, FLUB_EINVAL, "flub_extrapolate: bad argument");
    int rc = flub_set_unknown(phi, tu, tv, tw, nx, ny, nz, stream);
    if (rc != FLUB_OK || n_passes == 0) return rc;
    // any n_passes >= 2 gives the result of two, which one launch per face
    // array computes (see k_extrapolate)
    return extrapolate_faces(u, v, w, tu, tv, tw, nx, ny, nz, n_passes >= 2, stream);
}
