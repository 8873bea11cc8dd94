// BF16 3x3x3 convolutions of the learned preconditioner (no counterpart in
// the reference) as implicit GEMMs on the gfx950 matrix cores, FP32
// accumulation.  Tensors are channels-last: X[n][d][h][w][c] (BF16), weights
// Wt[cout][27*cin] with the 27 taps ordered (dz, dy, dx) row-major and cin
// fastest (BF16; FP32 master copy kept by the caller).  "Same" padding: taps
// outside the volume read zero.
//
// Forward / data gradient (k_conv3):  M = output voxels, N = cout,
// K = 27*cin, one v_mfma_f32_16x16x32_bf16 per 16 voxels x 16 cout x 32 K.
// Each wavefront owns MB 16-voxel tiles and all cout tiles, so one weight
// fragment feeds MB matrix instructions.  Operand layout of the instruction
// (wave64): lane l holds A[l%16][8*(l/16)+0..7] and B[8*(l/16)+0..7][l%16];
// the result D[4*(l/16)+i][l%16], i = 0..3.  cin is a multiple of 8, so the
// 8 K-values of a lane always lie in one tap and are one 16-byte load.
// The data gradient is the same kernel on dY with the taps flipped and the
// channel roles swapped (prepared by the caller).
//
// Weight gradient (k_conv3_wgrad): per tap a GEMM M = cout, N = cin, K =
// voxels, partial sums per (voxel chunk, wavefront) reduced in a fixed order
// by k_wgrad_reduce (deterministic).
#include "common.h"

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

namespace {

__device__ __forceinline__ bf16x8_t zero8()
{
    bf16x8_t z;
#pragma unroll
    for (int i = 0; i < 8; ++i) z[i] = (__bf16)0.0f;
    return z;
}

__device__ __forceinline__ bf16x8_t load8(const __hip_bfloat16 *p)
{
    return *reinterpret_cast<const bf16x8_t *>(p);
}

}  // namespace

template <int CIN, int COUT, int MB>
__global__ void __launch_bounds__(256) k_conv3(const __hip_bfloat16 *__restrict__ X,
                                               const __hip_bfloat16 *__restrict__ Wt,
                                               const float *__restrict__ bias, __hip_bfloat16 *__restrict__ Y,
                                               float *__restrict__ Y32, int nb, int D, int H, int W, int relu)
{
    constexpr int K = 27 * CIN;
    constexpr int KC = (K + 31) / 32;
    constexpr int NT = COUT / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const long nvox = (long)nb * D * H * W;
    const long m0 = ((long)blockIdx.x * 4 + wave) * (16 * MB);
    if (m0 >= nvox) return;
    int vb[MB], vd[MB], vh[MB], vw[MB];
    bool vok[MB];
#pragma unroll
    for (int mt = 0; mt < MB; ++mt) {
        long m = m0 + mt * 16 + r16;
        vok[mt] = m < nvox;
        if (!vok[mt]) m = 0;
        vw[mt] = (int)(m % W); m /= W;
        vh[mt] = (int)(m % H); m /= H;
        vd[mt] = (int)(m % D);
        vb[mt] = (int)(m / D);
    }
    f32x4_t acc[MB][NT];
#pragma unroll
    for (int mt = 0; mt < MB; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = (f32x4_t){0.0f, 0.0f, 0.0f, 0.0f};

    for (int kc = 0; kc < KC; ++kc) {
        const int k = kc * 32 + g * 8;
        const bool kvalid = k < K;
        const int tap = kvalid ? k / CIN : 0;
        const int ci0 = k - tap * CIN;
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
        bf16x8_t bfr[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
            bfr[nt] = kvalid ? load8(Wt + (long)(nt * 16 + r16) * K + k) : zero8();
#pragma unroll
        for (int mt = 0; mt < MB; ++mt) {
            const int zz = vd[mt] + dz, yy = vh[mt] + dy, xx = vw[mt] + dx;
            bf16x8_t a = zero8();
            if (vok[mt] && kvalid && zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 && xx < W)
                a = load8(X + ((((long)vb[mt] * D + zz) * H + yy) * W + xx) * CIN + ci0);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bfr[nt], acc[mt][nt], 0, 0, 0);
        }
    }
#pragma unroll
    for (int mt = 0; mt < MB; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = nt * 16 + r16;
            const float b = bias ? bias[co] : 0.0f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const long m = m0 + mt * 16 + 4 * g + i;
                if (m >= nvox) continue;
                float v = acc[mt][nt][i] + b;
                if (relu) v = fmaxf(v, 0.0f);
                if (Y32) Y32[m * COUT + co] = v;
                else Y[m * COUT + co] = __float2bfloat16(v);
            }
        }
}

// dW partials: part[slot][co][tap][ci] for the voxel chunk of this
// (block, wavefront); grid (27 taps, chunks).
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv3_wgrad(const __hip_bfloat16 *__restrict__ X,
                                                     const __hip_bfloat16 *__restrict__ dY, float *__restrict__ part,
                                                     int nb, int D, int H, int W, long vox_per_wave)
{
    constexpr int MT = COUT / 16, NT = CIN / 16;
    const int tap = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const long nvox = (long)nb * D * H * W;
    const long slot = (long)blockIdx.y * 4 + wave;
    const long v_begin = slot * vox_per_wave;
    const long v_end = v_begin + vox_per_wave < nvox ? v_begin + vox_per_wave : nvox;
    f32x4_t acc[MT][NT];
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) acc[a][b] = (f32x4_t){0.0f, 0.0f, 0.0f, 0.0f};
    for (long kb = v_begin; kb < v_end; kb += 32) {
        // this lane's 8 voxels: kb + 8g .. +7
        bf16x8_t afr[MT], bfr[NT];
#pragma unroll
        for (int a = 0; a < MT; ++a) afr[a] = zero8();
#pragma unroll
        for (int b = 0; b < NT; ++b) bfr[b] = zero8();
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const long v = kb + 8 * g + e;
            if (v >= v_end) break;
            long m = v;
            const int w = (int)(m % W); m /= W;
            const int h = (int)(m % H); m /= H;
            const int d = (int)(m % D);
            const int bb = (int)(m / D);
#pragma unroll
            for (int a = 0; a < MT; ++a) afr[a][e] = *reinterpret_cast<const __bf16 *>(dY + v * COUT + a * 16 + r16);
            const int zz = d + dz, yy = h + dy, xx = w + dx;
            if (zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 && xx < W) {
                const long src = (((long)bb * D + zz) * H + yy) * W + xx;
#pragma unroll
                for (int b = 0; b < NT; ++b) bfr[b][e] = *reinterpret_cast<const __bf16 *>(X + src * CIN + b * 16 + r16);
            }
        }
#pragma unroll
        for (int a = 0; a < MT; ++a)
#pragma unroll
            for (int b = 0; b < NT; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a], bfr[b], acc[a][b], 0, 0, 0);
    }
    // D[4g+i][r16]: co = a*16 + 4g + i, ci = b*16 + r16
    float *dst = part + slot * (long)COUT * 27 * CIN;
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int co = a * 16 + 4 * g + i, ci = b * 16 + r16;
                dst[((long)co * 27 + tap) * CIN + ci] = acc[a][b][i];
            }
}

__global__ void k_wgrad_reduce(const float *__restrict__ part, float *__restrict__ dw, long per, int slots,
                               float scale)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per) return;
    float s = 0.0f;
    for (int k = 0; k < slots; ++k) s += part[(long)k * per + t];
    dw[t] = s * scale;
}

// 1x1 head with one output channel, FP32: y[v] = b + sum_c x[v][c] w[c].
__global__ void k_head_fwd(const __hip_bfloat16 *__restrict__ X, const float *__restrict__ w, float b,
                           float *__restrict__ y, long nvox, int C)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    float s = b;
    for (int c = 0; c < C; ++c) s += __bfloat162float(X[v * C + c]) * w[c];
    y[v] = s;
}

// ------------------------------------------------------------------ host
namespace {

template <int CIN, int COUT>
int conv_launch(const void *x, const void *wt, const float *bias, void *y, float *y32, int nb, int d, int h, int w,
                int relu, hipStream_t s)
{
    constexpr int MB = COUT >= 64 ? 2 : 4;
    const long nvox = (long)nb * d * h * w;
    const long per_block = 4L * 16 * MB;
    const long nblk = (nvox + per_block - 1) / per_block;
    FLUB_REQUIRE(nblk < (1L << 31), FLUB_ESHAPE, "conv: volume too large");
    k_conv3<CIN, COUT, MB><<<(unsigned)nblk, 256, 0, s>>>((const __hip_bfloat16 *)x, (const __hip_bfloat16 *)wt,
                                                        bias, (__hip_bfloat16 *)y, y32, nb, d, h, w, relu);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

template <int CIN, int COUT>
int wgrad_launch(const void *x, const void *dy, float *dw, float *work, long work_floats, int nb, int d, int h, int w,
                 hipStream_t s)
{
    const long nvox = (long)nb * d * h * w;
    const long per = (long)COUT * 27 * CIN;
    int chunks = 64;
    while (chunks > 1 && (long)chunks * 4 * per > work_floats) chunks >>= 1;
    FLUB_REQUIRE((long)chunks * 4 * per <= work_floats, FLUB_EINVAL, "wgrad: workspace too small");
    long vpw = (nvox + chunks * 4 - 1) / (chunks * 4);
    vpw = (vpw + 31) / 32 * 32;
    const int slots = (int)((nvox + vpw - 1) / vpw);
    dim3 grid(27, (unsigned)((slots + 3) / 4));
    k_conv3_wgrad<CIN, COUT><<<grid, 256, 0, s>>>((const __hip_bfloat16 *)x, (const __hip_bfloat16 *)dy, work, nb, d,
                                                  h, w, vpw);
    FLUB_CHECK_LAUNCH();
    // slots beyond the volume were not written by any wavefront
    k_wgrad_reduce<<<(unsigned)((per + 255) / 256), 256, 0, s>>>(work, dw, per, slots, 1.0f);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

}  // namespace

#define FLUB_CONV_PAIRS(X) \
    X(16, 16) X(16, 32) X(32, 32) X(32, 64) X(64, 64) X(96, 32) X(48, 16) \
    X(32, 16) X(64, 32) X(32, 96) X(16, 48)

FLUB_API int flub_conv3d_bf16(const void *x, const void *wt, const float *bias, void *y, float *y32, int nb, int d,
                              int h, int w, int cin, int cout, int relu, void *stream)
{
    FLUB_REQUIRE(x && wt && (y || y32), FLUB_EINVAL, "flub_conv3d_bf16: null pointer");
    FLUB_REQUIRE(nb > 0 && d > 0 && h > 0 && w > 0, FLUB_ESHAPE, "flub_conv3d_bf16: empty volume");
    hipStream_t s = as_stream(stream);
#define CASE(a, b) \
    if (cin == a && cout == b) return conv_launch<a, b>(x, wt, bias, y, y32, nb, d, h, w, relu, s);
    FLUB_CONV_PAIRS(CASE)
#undef CASE
    flub_set_error("flub_conv3d_bf16: unsupported channels %d -> %d", cin, cout);
    return FLUB_EINVAL;
}

FLUB_API int flub_conv3d_wgrad_bf16(const void *x, const void *dy, float *dw, float *work, size_t work_floats, int nb,
                                    int d, int h, int w, int cin, int cout, void *stream)
{
    FLUB_REQUIRE(x && dy && dw && work, FLUB_EINVAL, "flub_conv3d_wgrad_bf16: null pointer");
    FLUB_REQUIRE(nb > 0 && d > 0 && h > 0 && w > 0, FLUB_ESHAPE, "flub_conv3d_wgrad_bf16: empty volume");
    hipStream_t s = as_stream(stream);
#define CASE(a, b) \
    if (cin == a && cout == b) return wgrad_launch<a, b>(x, dy, dw, work, (long)work_floats, nb, d, h, w, s);
    FLUB_CONV_PAIRS(CASE)
#undef CASE
    flub_set_error("flub_conv3d_wgrad_bf16: unsupported channels %d -> %d", cin, cout);
    return FLUB_EINVAL;
}

FLUB_API int flub_head_fwd(const void *x, const float *w, float b, float *y, size_t nvox, int c, void *stream)
{
    FLUB_REQUIRE(x && w && y && c > 0, FLUB_EINVAL, "flub_head_fwd: bad argument");
    if (!nvox) return FLUB_OK;
    k_head_fwd<<<(unsigned)((nvox + 255) / 256), 256, 0, as_stream(stream)>>>((const __hip_bfloat16 *)x, w, b, y,
                                                                            (long)nvox, c);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}
