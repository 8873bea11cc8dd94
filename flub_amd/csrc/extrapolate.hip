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
//
// A workgroup fills a tile of TI x TJ x TK faces: it first stages the tile
// and a one-face halo in LDS (a class byte per face -- 0 known, 1 marked 1,
// 3 marked -1, 2 anything else -- and the value of the known ones), then
// each face sums its known neighbours from LDS in the reference's order.
constexpr int ETI = 8, ETJ = 8, ETK = 62;   // halo rows of 64 faces: one wave per row
constexpr int EHI = ETI + 2, EHJ = ETJ + 2, EHK = ETK + 2, EH = EHI * EHJ * EHK;
static_assert(EHK == 64, "staging maps one wave to a halo row");
template <int MODE>
__global__ void __launch_bounds__(256) k_extrapolate(float *__restrict__ q, int32_t *__restrict__ unknown, int sx,
                                                     int sy, int sz, int tiles_j, int tiles_k)
{
    __shared__ float qs[EH];
    __shared__ uint8_t cls[EH];
    const int tk = blockIdx.x % tiles_k, rest = blockIdx.x / tiles_k;
    const int tj = rest % tiles_j, ti = rest / tiles_j;
    const int i0 = ti * ETI - 1, j0 = tj * ETJ - 1, k0 = tk * ETK - 1;   // halo origin
    // staging: every load of the thread is issued before any is used
    constexpr int PER = (EH + 255) / 256;
    int32_t mk[PER];
    float qv[PER];
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int idx = threadIdx.x + 256 * r;
        const int row = idx >> 6, lk = idx & 63;   // EHK == 64
        const int li = row / EHJ, lj = row - li * EHJ;
        const int i = i0 + li, j = j0 + lj, k = k0 + lk;
        mk[r] = 2;
        qv[r] = 0.0f;
        if (idx < EH && i >= 0 && i < sx && j >= 0 && j < sy && k >= 0 && k < sz) {
            const uint32_t g = ((uint32_t)i * sy + j) * (uint32_t)sz + k;
            mk[r] = unknown[g];
            qv[r] = q[g];
        }
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int idx = threadIdx.x + 256 * r;
        if (idx < EH) {
            const int32_t m = mk[r];
            cls[idx] = m == 0 ? 0 : (m == 1 ? 1 : (m == -1 ? 3 : 2));
            qs[idx] = m == 0 ? qv[r] : 0.0f;   // +0 on every face that is not known
        }
    }
    __syncthreads();
    const int lk = (threadIdx.x & 63) + 1;
    if (lk > ETK) return;   // no barrier follows
    for (int f = threadIdx.x >> 6; f < ETI * ETJ; f += 4) {
        const int li = f / ETJ + 1, lj = f % ETJ + 1;
        const int i = i0 + li, j = j0 + lj, k = k0 + lk;
        if (i >= sx || j >= sy || k >= sz) continue;
        const int c = cls[(li * EHJ + lj) * EHK + lk];
        const bool fill = c == 1 || (MODE == 2 && c == 3);
        const uint32_t g = ((uint32_t)i * sy + j) * (uint32_t)sz + k;
        if (MODE >= 1 && c == 3) unknown[g] = 1;
        if (!fill) continue;
        // The reference sums the known neighbours inside the array in this
        // order.  Every other neighbour (unknown, or the halo beyond the
        // array: class 2) is staged as +0 and adding it changes nothing: the
        // running sum starts at +0 and is never -0 (an exact zero sum rounds
        // to +0), so s + (+0) == s bitwise.
        int cnt = 0;
        float s = 0.0f;
#pragma unroll
        for (int a = -1; a <= 1; ++a)
#pragma unroll
            for (int b = -1; b <= 1; ++b) {
                const int row = ((li + a) * EHJ + (lj + b)) * EHK + lk;
#pragma unroll
                for (int cc = -1; cc <= 1; ++cc) {
                    cnt += cls[row + cc] == 0 ? 1 : 0;
                    s = s + qs[row + cc];
                }
            }
        if (cnt > 0) {
            q[g] = s / (float)cnt;
            if (MODE == 0) unknown[g] = -1;
        } else {
            q[g] = 0.0f;
        }
    }
}

template <int MODE>
void launch_extrapolate(float *q, int32_t *unknown, int sx, int sy, int sz, hipStream_t s)
{
    const int ti = (sx + ETI - 1) / ETI, tj = (sy + ETJ - 1) / ETJ, tk = (sz + ETK - 1) / ETK;
    k_extrapolate<MODE><<<(unsigned)(ti * tj * tk), 256, 0, s>>>(q, unknown, sx, sy, sz, tj, tk);
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
    launch_extrapolate<0>(q, unknown, sx, sy, sz, as_stream(stream));
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
        if (several) launch_extrapolate<2>(q[d], t[d], sx, sy, sz, s);
        else launch_extrapolate<1>(q[d], t[d], sx, sy, sz, s);
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
