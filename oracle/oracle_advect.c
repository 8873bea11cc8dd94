/*
 * TEST INFRASTRUCTURE -- CPU oracle for the MacCormack advection of the
 * reference Warp code: advect.py:11-56 (in_bound, backtrace, advect_phi),
 * :58-137 (_advect_mc_1/2/3, sample_minmax, advect_mc, advect_uv_mc,
 * advect_3d) and the interpolation it uses, utils/util.py:20-28 (sample_3d),
 * :33-86 (sample_velocity_3d), :90-159 (trilerp, trilerp_stagger, _trilerp,
 * cubic, _cubic, _bicubic, _tricubic).
 *
 * FP32, one point at a time in the reference's expression order, built with
 * -ffp-contract=off; wp.lerp(a, b, t) is a*(1-t) + b*t.  The three MacCormack
 * launches are restated literally (the GPU fuses them).  Only tests/, the
 * smoke check and the CPU-baseline leg of bench.py load it.
 *
 * Parity pins (tests/test_oracle_advect.py): sample_velocity_3d reproduces
 * linear velocity fields; the tricubic interpolation reproduces polynomials of
 * degree 3 per axis (it is the 4-point Lagrange interpolant); the RK3
 * backtrace moves by -dt/h U in a constant field and matches an independent
 * FP64 RK3 on an analytic linear field; sample_minmax against an independent
 * NumPy gather of the same (clamped) corners; MacCormack keeps constant fields
 * (to the rounding of the weights) and, with zero velocity, the field itself.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#define AIDX(i, j, k, sy, sz) ((((size_t)(i)) * (size_t)(sy) + (size_t)(j)) * (size_t)(sz) + (size_t)(k))

typedef struct {
    const float *q;
    int gx, gy, gz;
} Arr;

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* utils/util.py:20-28 sample_3d */
static inline float smp(Arr a, int u, int v, int w)
{
    return a.q[AIDX(clampi(u, 0, a.gx - 1), clampi(v, 0, a.gy - 1), clampi(w, 0, a.gz - 1), a.gy, a.gz)];
}

static inline float lerpf_(float a, float b, float t) { return a * (1.0f - t) + b * t; }

/* utils/util.py:33-86 sample_velocity_3d at p (grid units) */
static void sample_velocity(Arr u, Arr v, Arr w, const float p[3], float out[3])
{
    const float pu = p[0], pv = p[1], pw = p[2];
    const int iu = (int)floorf(pu), iv = (int)floorf(pv), iw = (int)floorf(pw);
    const float fu = pu - (float)iu, fv = pv - (float)iv, fw = pw - (float)iw;
    const float s = pu - 0.5f, t = pv - 0.5f, r = pw - 0.5f;
    const int is = (int)floorf(s), it = (int)floorf(t), ir = (int)floorf(r);
    const float fs = s - (float)is, ft = t - (float)it, fr = r - (float)ir;
    float a, b, c, d, e, f, g, h;
    a = smp(u, iu, it, ir); b = smp(u, iu + 1, it, ir); c = smp(u, iu, it + 1, ir); d = smp(u, iu + 1, it + 1, ir);
    e = smp(u, iu, it, ir + 1); f = smp(u, iu + 1, it, ir + 1); g = smp(u, iu, it + 1, ir + 1);
    h = smp(u, iu + 1, it + 1, ir + 1);
    out[0] = lerpf_(lerpf_(lerpf_(a, b, fu), lerpf_(c, d, fu), ft), lerpf_(lerpf_(e, f, fu), lerpf_(g, h, fu), ft), fr);
    a = smp(v, is, iv, ir); b = smp(v, is + 1, iv, ir); c = smp(v, is, iv + 1, ir); d = smp(v, is + 1, iv + 1, ir);
    e = smp(v, is, iv, ir + 1); f = smp(v, is + 1, iv, ir + 1); g = smp(v, is, iv + 1, ir + 1);
    h = smp(v, is + 1, iv + 1, ir + 1);
    out[1] = lerpf_(lerpf_(lerpf_(a, b, fs), lerpf_(c, d, fs), fv), lerpf_(lerpf_(e, f, fs), lerpf_(g, h, fs), fv), fr);
    a = smp(w, is, it, iw); b = smp(w, is + 1, it, iw); c = smp(w, is, it + 1, iw); d = smp(w, is + 1, it + 1, iw);
    e = smp(w, is, it, iw + 1); f = smp(w, is + 1, it, iw + 1); g = smp(w, is, it + 1, iw + 1);
    h = smp(w, is + 1, it + 1, iw + 1);
    out[2] = lerpf_(lerpf_(lerpf_(a, b, fs), lerpf_(c, d, fs), ft), lerpf_(lerpf_(e, f, fs), lerpf_(g, h, fs), ft), fw);
}

/* advect.py:11-22 backtrace (RK3) of p (grid units) */
static void backtrace(Arr u, Arr v, Arr w, float p[3], float dt, float h)
{
    float v1[3], v2[3], v3[3], p1[3], p2[3];
    sample_velocity(u, v, w, p, v1);
    const float c1 = 0.5f * dt / h, c2 = 0.75f * dt / h, c3 = dt / h;
    for (int a = 0; a < 3; ++a) p1[a] = p[a] - c1 * v1[a];
    sample_velocity(u, v, w, p1, v2);
    for (int a = 0; a < 3; ++a) p2[a] = p[a] - c2 * v2[a];
    sample_velocity(u, v, w, p2, v3);
    const float k1 = 2.0f / 9.0f, k2 = 1.0f / 3.0f, k3 = 4.0f / 9.0f;
    for (int a = 0; a < 3; ++a) p[a] = p[a] - c3 * (k1 * v1[a] + k2 * v2[a] + k3 * v3[a]);
}

/* utils/util.py:121-128 cubic: the 4-point weights at s, dotted with (a, b, c, d) */
static float cubic(float a, float b, float c, float d, float s)
{
    const float t0 = -1.0f / 3.0f * s + 0.5f * s * s - 1.0f / 6.0f * s * s * s;
    const float t1 = 1.0f - s * s + 0.5f * (s * s * s - s);
    const float t2 = s + 0.5f * (s * s - s * s * s);
    const float t3 = 1.0f / 6.0f * (s * s * s - s);
    return a * t0 + b * t1 + c * t2 + d * t3;
}

/* utils/util.py:130-159 _tricubic at (s, t, r) in array index units */
static float tricubic(Arr q, float s, float t, float r)
{
    const int iu = (int)floorf(s), iv = (int)floorf(t), iw = (int)floorf(r);
    const float fu = s - (float)iu, fv = t - (float)iv, fw = r - (float)iw;
    float zs[4];
    for (int c = 0; c < 4; ++c) {
        float ys[4];
        for (int b = 0; b < 4; ++b)
            ys[b] = cubic(smp(q, iu - 1, iv + b - 1, iw + c - 1), smp(q, iu, iv + b - 1, iw + c - 1),
                          smp(q, iu + 1, iv + b - 1, iw + c - 1), smp(q, iu + 2, iv + b - 1, iw + c - 1), fu);
        zs[c] = cubic(ys[0], ys[1], ys[2], ys[3], fv);
    }
    return cubic(zs[0], zs[1], zs[2], zs[3], fw);
}

/* utils/util.py:103-112 _trilerp */
static float trilerp_(Arr q, float s, float t, float r)
{
    const int iu = (int)floorf(s), iv = (int)floorf(t), iw = (int)floorf(r);
    const float fu = s - (float)iu, fv = t - (float)iv, fw = r - (float)iw;
    const float a = smp(q, iu, iv, iw), b = smp(q, iu + 1, iv, iw), c = smp(q, iu, iv + 1, iw),
                d = smp(q, iu + 1, iv + 1, iw), e = smp(q, iu, iv, iw + 1), f = smp(q, iu + 1, iv, iw + 1),
                g = smp(q, iu, iv + 1, iw + 1), h = smp(q, iu + 1, iv + 1, iw + 1);
    return lerpf_(lerpf_(lerpf_(a, b, fu), lerpf_(c, d, fu), fv), lerpf_(lerpf_(e, f, fu), lerpf_(g, h, fu), fv), fw);
}

/* advect.py:114-131 sample_minmax: p is divided by h (the reference passes h0
 * while p is already in grid units), then the 8 clamped corner samples. */
static void sample_minmax(Arr q, const float p[3], float h, const float st[3], float *mn, float *mx)
{
    const float rs = p[0] / h - st[0], rt = p[1] / h - st[1], rr = p[2] / h - st[2];
    const int i = (int)floorf(rs), j = (int)floorf(rt), k = (int)floorf(rr);
    const float a = smp(q, i, j, k), b = smp(q, i + 1, j, k), c = smp(q, i, j + 1, k), d = smp(q, i + 1, j + 1, k);
    const float e = smp(q, i, j, k + 1), f = smp(q, i + 1, j, k + 1), g = smp(q, i, j + 1, k + 1),
                hh = smp(q, i + 1, j + 1, k + 1);
    *mn = fminf(fminf(fminf(fminf(a, b), c), d), fminf(fminf(fminf(e, f), g), hh));
    *mx = fmaxf(fmaxf(fmaxf(fmaxf(a, b), c), d), fmaxf(fmaxf(fmaxf(e, f), g), hh));
}

/* ---------------------------------------------------------------- exports */
void orc_sample_velocity(const float *u, const float *v, const float *w, int nx, int ny, int nz, const float *p,
                         float *out, int n)
{
    const Arr U = {u, nx + 1, ny, nz}, V = {v, nx, ny + 1, nz}, W = {w, nx, ny, nz + 1};
    for (int t = 0; t < n; ++t) sample_velocity(U, V, W, p + 3 * t, out + 3 * t);
}

void orc_backtrace(const float *u, const float *v, const float *w, int nx, int ny, int nz, float *p, int n, float dt,
                   float h)
{
    const Arr U = {u, nx + 1, ny, nz}, V = {v, nx, ny + 1, nz}, W = {w, nx, ny, nz + 1};
    for (int t = 0; t < n; ++t) backtrace(U, V, W, p + 3 * t, dt, h);
}

void orc_tricubic(const float *q, int gx, int gy, int gz, const float *str, float *out, int n)
{
    const Arr Q = {q, gx, gy, gz};
    for (int t = 0; t < n; ++t) out[t] = tricubic(Q, str[3 * t], str[3 * t + 1], str[3 * t + 2]);
}

void orc_sample_minmax(const float *q, int gx, int gy, int gz, const float *p, float h, const float *st, float *mn,
                       float *mx, int n)
{
    const Arr Q = {q, gx, gy, gz};
    for (int t = 0; t < n; ++t) sample_minmax(Q, p + 3 * t, h, st, mn + t, mx + t);
}

/* advect.py:58-112,133-137 advect_mc of one field q_in (gx, gy, gz) with stagger st
 * through the MAC velocity (u, v, w) of an nx x ny x nz grid: the three launches,
 * literally (q_tmp holds the backward-traced values). */
void orc_advect_mc(const float *u, const float *v, const float *w, int nx, int ny, int nz, const float *q_in,
                   float *q_out, float *q_tmp, int gx, int gy, int gz, const float *st, float dt, float h0)
{
    const Arr U = {u, nx + 1, ny, nz}, V = {v, nx, ny + 1, nz}, W = {w, nx, ny, nz + 1}, Q = {q_in, gx, gy, gz};
    for (int pass = 1; pass <= 3; ++pass) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < gx; ++i)
            for (int j = 0; j < gy; ++j)
                for (int k = 0; k < gz; ++k) {
                    const size_t t = AIDX(i, j, k, gy, gz);
                    float p[3] = {(float)i + st[0], (float)j + st[1], (float)k + st[2]};
                    if (pass == 1) {
                        backtrace(U, V, W, p, dt, h0);
                        q_out[t] = tricubic(Q, p[0] - st[0], p[1] - st[1], p[2] - st[2]);
                    } else if (pass == 2) {
                        backtrace(U, V, W, p, -dt, h0);
                        q_tmp[t] = tricubic(Q, p[0] - st[0], p[1] - st[1], p[2] - st[2]);
                    } else {
                        backtrace(U, V, W, p, dt, h0);
                        const float q_sl = tricubic(Q, p[0] - st[0], p[1] - st[1], p[2] - st[2]);
                        q_out[t] = q_out[t] + 0.5f * (q_in[t] - q_tmp[t]);
                        float mn, mx;
                        sample_minmax(Q, p, h0, st, &mn, &mx);
                        if (!(mn < q_out[t] && q_out[t] < mx)) q_out[t] = q_sl;
                    }
                }
    }
}

/* advect.py:38-56 advect_phi: the backtraced value where it lands inside the
 * array ([0, g) per axis, in grid units), else h0. */
void orc_advect_phi(const float *u, const float *v, const float *w, int nx, int ny, int nz, const float *phi,
                    float *new_phi, const float *st, float dt, float h0)
{
    const Arr U = {u, nx + 1, ny, nz}, V = {v, nx, ny + 1, nz}, W = {w, nx, ny, nz + 1}, Q = {phi, nx, ny, nz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                float p[3] = {(float)i + st[0], (float)j + st[1], (float)k + st[2]};
                backtrace(U, V, W, p, dt, h0);
                const int inb = 0.0f <= p[0] && p[0] < (float)nx && 0.0f <= p[1] && p[1] < (float)ny &&
                                0.0f <= p[2] && p[2] < (float)nz;
                new_phi[AIDX(i, j, k, ny, nz)] = inb ? tricubic(Q, p[0] - st[0], p[1] - st[1], p[2] - st[2]) : h0;
            }
}

/* advect.py:25-35 advect_simple_3d with trilinear sampling (utils/util.py:90-95). */
void orc_advect_simple(const float *u, const float *v, const float *w, int nx, int ny, int nz, const float *q,
                       float *new_q, float dt, float h)
{
    const Arr U = {u, nx + 1, ny, nz}, V = {v, nx, ny + 1, nz}, W = {w, nx, ny, nz + 1}, Q = {q, nx, ny, nz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                float p[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
                backtrace(U, V, W, p, dt, h);
                new_q[AIDX(i, j, k, ny, nz)] = trilerp_(Q, p[0] - 0.5f, p[1] - 0.5f, p[2] - 0.5f);
            }
}
