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

// Tiled forward convolution.  A workgroup (4 wavefronts) computes an output
// tile of TZ x TY x 16 voxels and all COUT channels: it loads the input tile
// with a one-voxel halo, (TZ+2) x (TY+2) x 18 voxels x CIN channels, into LDS
// once (each input voxel is then read from HBM about (TZ+2)(TY+2)18 /
// (16 TZ TY) times instead of 27), and every wavefront takes its matrix
// fragments for all 27 taps from LDS.  Each wavefront owns RB = TZ*TY/4 rows
// of 16 voxels.  The matrix instruction computes D[cout][voxel] = W x X
// (weights as the A operand), so a lane ends with 4 consecutive output
// channels of one voxel: 8-byte stores, contiguous across the wavefront.
//
// With CA > 0 the input is the concatenation of the nearest 2x upsampling of
// Xa (CA channels at half resolution) and Xb (CIN - CA channels): the U-Net
// decoder's up2 + cat, gathered while the tile is loaded instead of being
// materialised.  Requires W % 16 == 0, H % TY == 0, D % TZ == 0 (and even
// sizes with CA > 0).  Tiles go to the XCDs in contiguous ranges (workgroup b
// runs on XCD b % 8), so the halos a tile shares with its neighbours are
// mostly read from the same L2.
template <int CIN, int COUT, int TZ, int TY, int CA>
__global__ void __launch_bounds__(256) k_conv3t(const __hip_bfloat16 *__restrict__ Xa,
                                                const __hip_bfloat16 *__restrict__ Xb,
                                                const __hip_bfloat16 *__restrict__ Wt,
                                                const float *__restrict__ bias, __hip_bfloat16 *__restrict__ Y,
                                                float *__restrict__ Y32, int nb, int D, int H, int W, int relu)
{
    constexpr int HZ = TZ + 2, HY = TY + 2, HX = 18;
    constexpr int NV = HZ * HY * HX;
    constexpr int C8 = CIN / 8;
    constexpr int VS = CIN;   // LDS values per voxel
    constexpr int K = 27 * CIN;
    constexpr int KC = (K + 31) / 32;
    constexpr int NT = COUT / 16;
    constexpr int RB = TZ * TY / 4;
    constexpr int CB = CIN - CA;
    __shared__ __align__(16) __hip_bfloat16 tile[NV * VS];

    const int txn = W / 16, tyn = H / TY, tzn = D / TZ;
    const uint32_t ntiles = (uint32_t)nb * tzn * tyn * txn;
    const uint32_t per = (ntiles + 7u) / 8u;
    const uint32_t t = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (t >= ntiles) return;
    uint32_t rest = t;
    const int tx = (int)(rest % txn); rest /= txn;
    const int ty = (int)(rest % tyn); rest /= tyn;
    const int tz = (int)(rest % tzn);
    const int bi = (int)(rest / tzn);
    const int x0 = tx * 16, y0 = ty * TY, z0 = tz * TZ;

    // the tile load: every load of a thread issued before the first LDS store
    constexpr int NLD = (NV * C8 + 255) / 256;
    uint4 buf[NLD];
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
        const int v = threadIdx.x + i * 256;
        buf[i] = make_uint4(0u, 0u, 0u, 0u);
        if (v < NV * C8) {
            const int vox = v / C8, c = (v - vox * C8) * 8;
            const int hx = vox % HX, hy = (vox / HX) % HY, hz = vox / (HX * HY);
            const int z = z0 + hz - 1, y = y0 + hy - 1, x = x0 + hx - 1;
            if (z >= 0 && z < D && y >= 0 && y < H && x >= 0 && x < W) {
                if (CA > 0 && c < CA) {
                    const long src = ((((long)bi * (D >> 1) + (z >> 1)) * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1));
                    buf[i] = *reinterpret_cast<const uint4 *>(Xa + src * CA + c);
                } else {
                    const long src = ((((long)bi * D + z) * H + y) * W + x);
                    buf[i] = *reinterpret_cast<const uint4 *>(Xb + src * CB + (c - CA));
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
        const int v = threadIdx.x + i * 256;
        if (v < NV * C8) {
            const int vox = v / C8, c = (v - vox * C8) * 8;
            *reinterpret_cast<uint4 *>(tile + vox * VS + c) = buf[i];
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    f32x4_t acc[RB][NT];
#pragma unroll
    for (int mt = 0; mt < RB; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = (f32x4_t){0.0f, 0.0f, 0.0f, 0.0f};

    // Weight fragments: KB steps of 32 K-values per group, the next group's
    // global loads issued before the current group's matrix instructions.
    constexpr int KB = 4;
    auto wload = [&](int kc, bf16x8_t w[NT]) {
        const int k = kc * 32 + g * 8;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) w[nt] = k < K ? load8(Wt + (long)(nt * 16 + r16) * K + k) : zero8();
    };
    bf16x8_t wcur[KB][NT], wnxt[KB][NT];
#pragma unroll
    for (int j = 0; j < KB; ++j) wload(j, wcur[j]);
#pragma unroll
    for (int kg = 0; kg < KC; kg += KB) {
        if (kg + KB < KC) {
#pragma unroll
            for (int j = 0; j < KB; ++j) wload(kg + KB + j, wnxt[j]);
        }
#pragma unroll
        for (int j = 0; j < KB; ++j) {
            const int kc = kg + j;
            if (kc >= KC) break;
            const int k = kc * 32 + g * 8;
            const bool kvalid = k < K;
            const int tap = kvalid ? k / CIN : 0;
            const int ci0 = k - tap * CIN;
            const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;   // offsets + 1 into the halo tile
#pragma unroll
            for (int mt = 0; mt < RB; ++mt) {
                const int row = wave * RB + mt;
                const int rz = row / TY, ry = row - rz * TY;
                const int vox = ((rz + dz) * HY + (ry + dy)) * HX + (r16 + dx);
                const bf16x8_t a = kvalid ? *reinterpret_cast<const bf16x8_t *>(tile + vox * VS + ci0) : zero8();
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wcur[j][nt], a, acc[mt][nt], 0, 0, 0);
            }
        }
#pragma unroll
        for (int j = 0; j < KB; ++j)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) wcur[j][nt] = wnxt[j][nt];
    }
    // D[cout = 4g + i][voxel = r16]
#pragma unroll
    for (int mt = 0; mt < RB; ++mt) {
        const int row = wave * RB + mt;
        const int rz = row / TY, ry = row - rz * TY;
        const long m = (((long)bi * D + (z0 + rz)) * H + (y0 + ry)) * W + (x0 + r16);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = nt * 16 + 4 * g;
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                v[i] = acc[mt][nt][i] + (bias ? bias[co + i] : 0.0f);
                if (relu) v[i] = fmaxf(v[i], 0.0f);
            }
            if (Y32) {
                *reinterpret_cast<float4 *>(Y32 + m * COUT + co) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
                uint2 w;
                __hip_bfloat16 *e = reinterpret_cast<__hip_bfloat16 *>(&w);
#pragma unroll
                for (int i = 0; i < 4; ++i) e[i] = __float2bfloat16(v[i]);
                *reinterpret_cast<uint2 *>(Y + m * COUT + co) = w;
            }
        }
    }
}

// 2x2x2 average pooling of a channels-last BF16 tensor (FP32 sum of the
// eight values in (dz, dy, dx) order, times 1/8, rounded to BF16); 8
// channels per thread.
__global__ void __launch_bounds__(256) k_pool2(const __hip_bfloat16 *__restrict__ X, __hip_bfloat16 *__restrict__ Y,
                                               int nb, int D, int H, int W, int C)
{
    const int C8 = C / 8, d2 = D / 2, h2 = H / 2, w2 = W / 2;
    const long n = (long)nb * d2 * h2 * w2 * C8;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    long r = t;
    const int c = (int)(r % C8) * 8; r /= C8;
    const int x = (int)(r % w2); r /= w2;
    const int y = (int)(r % h2); r /= h2;
    const int z = (int)(r % d2);
    const int b = (int)(r / d2);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const long src = ((((long)b * D + 2 * z + (q >> 2)) * H + 2 * y + ((q >> 1) & 1)) * W + 2 * x + (q & 1));
        const bf16x8_t v = load8(X + src * C + c);
#pragma unroll
        for (int e = 0; e < 8; ++e) s[e] += (float)v[e];
    }
    bf16x8_t o;
#pragma unroll
    for (int e = 0; e < 8; ++e) o[e] = (__bf16)(s[e] * 0.125f);
    *reinterpret_cast<bf16x8_t *>(Y + t * 8) = o;
}

// The network input: channel 0 = r / s on FLUID cells (0 elsewhere), channel
// 1 = the FLUID indicator, channels 2..15 zero (the matrix-core K width).
__global__ void __launch_bounds__(256) k_pack_input(const float *__restrict__ r, const uint8_t *__restrict__ fluid,
                                                    float s, __hip_bfloat16 *__restrict__ X, long n)
{
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    if (v >= n) return;
    const bool f = fluid[v] != 0;
    bf16x8_t a = zero8();
    a[0] = (__bf16)(f ? r[v] / s : 0.0f);
    a[1] = (__bf16)(f ? 1.0f : 0.0f);
    *reinterpret_cast<bf16x8_t *>(X + v * 16) = a;
    *reinterpret_cast<bf16x8_t *>(X + v * 16 + 8) = zero8();
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

// 1x1 head with one output channel, FP32: y[v] = b + sum_c x[v][c] w[c];
// with a mask, y[v] = mask[v] ? (that) * scale : 0.
__global__ void k_head_fwd(const __hip_bfloat16 *__restrict__ X, const float *__restrict__ w, float b,
                           float *__restrict__ y, long nvox, int C, const uint8_t *__restrict__ mask, float scale)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    float s = b;
    if (C % 8 == 0) {
        for (int c = 0; c < C; c += 8) {
            const bf16x8_t x8 = load8(X + v * C + c);
#pragma unroll
            for (int e = 0; e < 8; ++e) s += (float)x8[e] * w[c + e];
        }
    } else {
        for (int c = 0; c < C; ++c) s += __bfloat162float(X[v * C + c]) * w[c];
    }
    if (mask) s = mask[v] ? s * scale : 0.0f;
    y[v] = s;
}

// ------------------------------------------------------------------ host
namespace {

// Tile shapes of the LDS kernel per channel pair: TZ x TY x 16 output voxels,
// LDS (TZ+2)(TY+2)18 * CIN * 2 bytes (35-83 KB), TZ*TY/4 rows of 16 voxels
// per wavefront.
template <int CIN, int COUT> struct TileCfg {
    static constexpr int TZ = CIN <= 16 ? 4 : 2;
    static constexpr int TY = (CIN >= 64 || COUT >= 64) ? 4 : 8;
};

template <int CIN, int COUT, int CA>
bool conv_tiled(const void *xa, const void *xb, const void *wt, const float *bias, void *y, float *y32, int nb, int d,
                int h, int w, int relu, hipStream_t s)
{
    constexpr int TZ = TileCfg<CIN, COUT>::TZ, TY = TileCfg<CIN, COUT>::TY;
    if (w % 16 || h % TY || d % TZ) return false;
    if (CA > 0 && ((d | h | w) & 1)) return false;
    const long ntiles = (long)nb * (d / TZ) * (h / TY) * (w / 16);
    if (ntiles >= (1L << 28)) return false;
    const unsigned per = (unsigned)((ntiles + 7) / 8);
    k_conv3t<CIN, COUT, TZ, TY, CA><<<per * 8, 256, 0, s>>>((const __hip_bfloat16 *)xa, (const __hip_bfloat16 *)xb,
                                                          (const __hip_bfloat16 *)wt, bias, (__hip_bfloat16 *)y, y32,
                                                          nb, d, h, w, relu);
    return true;
}

template <int CIN, int COUT>
int conv_launch(const void *x, const void *wt, const float *bias, void *y, float *y32, int nb, int d, int h, int w,
                int relu, hipStream_t s)
{
    if (conv_tiled<CIN, COUT, 0>(nullptr, x, wt, bias, y, y32, nb, d, h, w, relu, s)) {
        FLUB_CHECK_LAUNCH();
        return FLUB_OK;
    }
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

FLUB_API int flub_conv3d_upcat_bf16(const void *xa, int ca, const void *xb, const void *wt, const float *bias, void *y,
                                    float *y32, int nb, int d, int h, int w, int cin, int cout, int relu, void *stream)
{
    FLUB_REQUIRE(xa && xb && wt && (y || y32), FLUB_EINVAL, "flub_conv3d_upcat_bf16: null pointer");
    FLUB_REQUIRE(nb > 0 && d > 0 && h > 0 && w > 0, FLUB_ESHAPE, "flub_conv3d_upcat_bf16: empty volume");
    hipStream_t s = as_stream(stream);
    bool done = false;
    if (cin == 48 && cout == 16 && ca == 32)
        done = conv_tiled<48, 16, 32>(xa, xb, wt, bias, y, y32, nb, d, h, w, relu, s);
    else if (cin == 96 && cout == 32 && ca == 64)
        done = conv_tiled<96, 32, 64>(xa, xb, wt, bias, y, y32, nb, d, h, w, relu, s);
    else {
        flub_set_error("flub_conv3d_upcat_bf16: unsupported channels %d (%d upsampled) -> %d", cin, ca, cout);
        return FLUB_EINVAL;
    }
    FLUB_REQUIRE(done, FLUB_ESHAPE, "flub_conv3d_upcat_bf16: volume %dx%dx%d not tileable", d, h, w);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_pool2_bf16(const void *x, void *y, int nb, int d, int h, int w, int c, void *stream)
{
    FLUB_REQUIRE(x && y && c % 8 == 0 && c > 0, FLUB_EINVAL, "flub_pool2_bf16: bad argument");
    FLUB_REQUIRE(nb > 0 && d > 1 && h > 1 && w > 1 && !((d | h | w) & 1), FLUB_ESHAPE, "flub_pool2_bf16: odd volume");
    const long n = (long)nb * (d / 2) * (h / 2) * (w / 2) * (c / 8);
    k_pool2<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>((const __hip_bfloat16 *)x,
                                                                       (__hip_bfloat16 *)y, nb, d, h, w, c);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_pack_input_bf16(const float *r, const uint8_t *fluid, float s, void *x, size_t nvox, void *stream)
{
    FLUB_REQUIRE(r && fluid && x, FLUB_EINVAL, "flub_pack_input_bf16: null pointer");
    if (!nvox) return FLUB_OK;
    k_pack_input<<<(unsigned)((nvox + 255) / 256), 256, 0, as_stream(stream)>>>(r, fluid, s, (__hip_bfloat16 *)x,
                                                                              (long)nvox);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
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
                                                                            (long)nvox, c, nullptr, 1.0f);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_head_fwd_masked(const void *x, const float *w, float b, float *y, size_t nvox, int c,
                                  const uint8_t *mask, float scale, void *stream)
{
    FLUB_REQUIRE(x && w && y && mask && c > 0, FLUB_EINVAL, "flub_head_fwd_masked: bad argument");
    if (!nvox) return FLUB_OK;
    k_head_fwd<<<(unsigned)((nvox + 255) / 256), 256, 0, as_stream(stream)>>>((const __hip_bfloat16 *)x, w, b, y,
                                                                            (long)nvox, c, mask, scale);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}
