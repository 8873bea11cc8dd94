// Internal layout of the MGPCG handle, shared by the FP32 solver (mgpcg.hip)
// and the mixed-precision flexible-CG solver (mixed.hip).
#pragma once
#include "common.h"

#define MAXL 12
#define RBS 256                 // reduction block size
#define MAXBLK 2048             // max partials per reduction
#define MAXPART 8192            // partial buffer size (the blocked level-0 sweep writes one per tile)
#define BOTTOM_ONE_WG_MAX 2048  // bottom levels up to this many cells per colour run in one workgroup
#define SOLVE_SAFETY_CAP 20000  // iteration cap applied when max_iters == -1

// Device scalar slots (FP64).
// SC_ZINV0/1: 1/scale of the mixed-precision V-cycle output held in the split z
// buffer of that slot (the buffer holds z times the storage scale).
// SC_APREV: the alpha of a deferred x update (see PxState).
enum { SC_ALPHA = 0, SC_BETA, SC_SUM, SC_OLDZTR, SC_RTR, SC_MEAN, SC_COUNT, SC_RSCALE, SC_AUX0, SC_AUX1,
       SC_AUX2, SC_AUX3, SC_ZINV0, SC_ZINV1, SC_APREV, SC_NUM = 16 };
// k_finish operations.
enum { FIN_SUM = 0, FIN_ALPHA, FIN_BETA, FIN_OLDZTR, FIN_RTR, FIN_MEAN, FIN_OUT };

struct Level {
    int nx, ny, nz;
    uint32_t n;
    uint8_t *tag;   // packed tags
    uint16_t *code; // per-cell stencil code (neighbour classes, divd, FLUID bit)
    float *r, *z;   // FP32 storage
    float *zpre;    // pre-smoothed z of the fast V-cycle (post-smoothing lands in z)
    void *rh, *zh;  // FP16/BF16 storage (mixed precision only)
};

struct MixedState;

// Level 0 in the colour-split layout (split.hip), used inside the FP32 solve.
struct SplitBuf {
    float *r, *zpre;
    // The solution x in FP64 (split layout).  Accumulating the FP32 updates
    // alpha p in FP32, as the reference does, rounds x at every iteration:
    // at 256^3 those roundings leave the FP64 true residual 11x above the
    // recursive one when the 1e-6 test passes.  In FP64 the two agree, and
    // the FP32 x handed back is x rounded once.
    double *x;
    float *z[2];    // two z buffers: the solve preconditions speculatively into the spare one
    uint16_t *code;
    // mixed precision: r and pre-smoothed z of the V-cycle in FP16/BF16, the
    // post-smoothing's FP32 working z
    void *rh, *zh;
    float *zw;
    float *b;       // the right-hand side (split), for the FP64 true residual
    // split stencil codes of the split levels: lcode[0] = code; levels
    // 1 .. nsplit-1 keep their r, zpre and z in the level's own buffers
    uint16_t *lcode[MAXL];
    // levels 1 .. nsplit-1 that sweep in pairs (split_sweep_pair): a third z
    // buffer of the level (nullptr: one launch per half-sweep)
    float *zt[MAXL];
};

// In-solve timing of the level-0 plain half-sweeps (flub_mgpcg_probe): HIP
// events recorded around those launches inside the solves' V-cycles.
#define PROBE_MAX 512
struct Probe {
    int n, max;
    hipEvent_t ev[2 * PROBE_MAX];
};

struct flub_mgpcg {
    int L, prec;
    Level lv[MAXL];
    float *x, *p, *Ap;
    float *p_alt;   // second p buffer: the reference-layout fused p-update + Ap writes out of place
    double *sc;     // device scalars
    bool fast;      // quad kernels usable (tagged smoother, levels >= 2, pre_post >= 1)
    int vc_dot_nblk;  // partial count of the z.r fused into the fast V-cycle (0: none)
    double *part;   // reduction partials (2*MAXPART)
    double *hsc;    // pinned host mirror of sc
    double *hsc_dev;  // device address of hsc (nullptr: not mapped, values are copied)
    // pinned host buffer of 2 * MAXBLK reduction partials, mapped for the
    // device: the r update stores its r.r partials there and the host sums
    // them in the finish kernel's order (nullptr: a finish kernel sums them)
    double *hpart, *hpart_dev;
    int pre_post, bottom;
    hipStream_t s;     // work stream (caller's; may be the null stream)
    hipStream_t cap;   // private stream used only to capture graphs
    bool tags_set;
    bool no_graph;
    // The split V-cycle replays its captured graph (else its kernels are
    // launched one by one).  Off by default: the fused solves queue the cycle
    // while the host is far ahead of the GPU, and at 256^3 the replay left
    // gaps between its kernels that direct launches do not
    // (13.98 ms against 13.83 ms per FP32 projection).
    bool split_graph;
    bool no_tb;        // launch-per-half-sweep smoothing on every level (no temporal blocking)
    bool no_tb_l0;     // no temporal blocking on level 0
    bool tb_g8;        // blocked smoothing may take 8 half-sweeps per launch on small levels
    uint32_t *tb_order;  // update-order tables of the blocked smoother
    hipGraphExec_t graph[5];   // per smoother; [3], [4]: the V-cycle on the split level 0 into sp.z[0], sp.z[1]
    bool graph_ok[5];
    hipEvent_t ev;     // marks the residual norm's copy to the host
    bool split;        // level 0 of the FP32 solve runs in the colour-split layout
    int nsplit;        // levels 0 .. nsplit-1 of the split V-cycle are colour-split (0: none)
    bool x_only;       // the solve's caller reads only x afterwards (the fused projection):
                       // r, z, p and Ap are left in the solver's internal layout
    bool presplit;     // the fused projection wrote r (and zero x) in the split layout already
    bool pair_l0;      // FP32 split level 0: the first pre- and post-smoothing half-sweeps in pairs
    bool mx_unit;      // BF16 split solve: storage scale 1, the r update stores the next cycle's BF16 r
    int rh_ready;      // z slot whose BF16 r the r update stored for the next V-cycle (-1: none)
    bool x64_valid;    // sp.x holds the FP64 x of the last solve (h->x is it rounded to FP32)
    double *x64n;      // the fused projection: the FP64 x in the reference layout, for the gradient
    bool want_x64n;    // the next split solve also writes x64n
    SplitBuf sp;
    MixedState *mx;
    Probe *probe;      // armed in-solve timing (nullptr: off)
};

// Event before / after a probed launch (no-ops unless a probe is armed and has room).
inline int probe_begin(flub_mgpcg *h)
{
    Probe *p = h->probe;
    if (p && p->n < p->max) FLUB_CHECK_HIP(hipEventRecord(p->ev[2 * p->n], h->s));
    return FLUB_OK;
}
inline int probe_end(flub_mgpcg *h)
{
    Probe *p = h->probe;
    if (p && p->n < p->max) {
        FLUB_CHECK_HIP(hipEventRecord(p->ev[2 * p->n + 1], h->s));
        ++p->n;
    }
    return FLUB_OK;
}

// The p step of the split solves over float4 vectors (n4 = n / 4), any
// layout: p_out = beta p_prev + z (mgpcg.py:229-232), z times zinv[0] when
// zinv is set (the z buffer holds z times a power-of-two scale: exact), with
// the pending x updates of earlier iterations (mgpcg.py:207-210) folded in.
// x is FP64: each update adds (double)alpha * (double)p (exact) with alpha the
// FP32 value the r update used, one rounding in FP64 per update.  p_prev is
// the previous iteration's p; the solve keeps p in two buffers alternately
// (p_out != p_prev), so the p before it survives one more step.  XU:
//   XU_NONE  nothing pending;
//   XU_DEFER the update alpha p_prev waits for the next step: alpha goes to
//            SC_APREV (this step moves no x);
//   XU_ONE   x += alpha p_prev;
//   XU_TWO   x += alpha' p_prev2 + alpha p_prev with alpha' from SC_APREV and
//            p_prev2 the p before p_prev, read from p_out before it is
//            overwritten (the same elements, by the same thread).
// Deferring every other update moves x (16 bytes a cell) on every second
// step only.  XSET: x is known to be zero (not read).  DOP false: the x
// update alone (the solve's last).
enum { XU_NONE = 0, XU_DEFER, XU_ONE, XU_TWO };
template <int XU, bool XSET, bool DOP>
__global__ void __launch_bounds__(256) k_update_pp(float *p_out, const float *p_prev, double *__restrict__ x,
                                                   const float *__restrict__ z, double *__restrict__ sc,
                                                   const double *__restrict__ zinv, uint32_t n4)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n4) return;
    if (XU == XU_DEFER && t == 0) sc[SC_APREV] = sc[SC_ALPHA];   // read by a later launch only
    float4 pv = reinterpret_cast<const float4 *>(p_prev)[t];
    if (XU == XU_ONE || XU == XU_TWO) {
        const double alpha = (double)(float)sc[SC_ALPHA];
        double xv[4] = {0.0, 0.0, 0.0, 0.0};
        if (!XSET) {
            const double2 a = reinterpret_cast<const double2 *>(x)[2 * t];
            const double2 b = reinterpret_cast<const double2 *>(x)[2 * t + 1];
            xv[0] = a.x; xv[1] = a.y; xv[2] = b.x; xv[3] = b.y;
        }
        if (XU == XU_TWO) {
            const double a0 = (double)(float)sc[SC_APREV];
            const float4 q = reinterpret_cast<const float4 *>(p_out)[t];
            const float qq[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) xv[e] = fma(a0, (double)qq[e], xv[e]);
        }
        const float pp[4] = {pv.x, pv.y, pv.z, pv.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) xv[e] = fma(alpha, (double)pp[e], xv[e]);
        reinterpret_cast<double2 *>(x)[2 * t] = make_double2(xv[0], xv[1]);
        reinterpret_cast<double2 *>(x)[2 * t + 1] = make_double2(xv[2], xv[3]);
    }
    if (DOP) {
        const float beta = (float)sc[SC_BETA];
        float4 zv = reinterpret_cast<const float4 *>(z)[t];
        if (zinv) {
            const float zi = (float)zinv[0];
            zv.x *= zi; zv.y *= zi; zv.z *= zi; zv.w *= zi;
        }
        pv.x = fmaf(beta, pv.x, zv.x);
        pv.y = fmaf(beta, pv.y, zv.y);
        pv.z = fmaf(beta, pv.z, zv.z);
        pv.w = fmaf(beta, pv.w, zv.w);
        reinterpret_cast<float4 *>(p_out)[t] = pv;
    }
}

// The split solves' p buffers and pending x updates.  P[0], P[1] alternate
// as the output of the p steps; after the r update of an iteration its
// alpha p is pending (held in SC_ALPHA and the p buffer `pa`), possibly after
// an older deferred one (alpha in SC_APREV, p in `pb`).
struct PxState {
    float *P[2];
    int cur = 0;          // index of the latest p in P
    int npend = 0;        // pending x updates (0, 1, 2)
    bool deferred = false;  // the one pending update's alpha moved to SC_APREV (p in P[cur])
    bool x_set = true;    // x holds a value (else it is zero and not yet written)
    double *x = nullptr;
    double *sc = nullptr;
    uint32_t n = 0;
    float *p() const { return P[cur]; }
};

template <int XU, bool XSET, bool DOP>
inline void launch_pp(hipStream_t s, const PxState &st, float *p_out, const float *p_prev, const float *z,
                      const double *zinv)
{
    const uint32_t n4 = st.n / 4;
    k_update_pp<XU, XSET, DOP><<<blocks_for(n4, 256), 256, 0, s>>>(p_out, p_prev, st.x, z, st.sc, zinv, n4);
}

// The p step of an iteration: p into the other buffer, the pending x update
// deferred (one pending) or applied together with the deferred one (two).
inline int px_step(hipStream_t s, PxState &st, const float *z, const double *zinv)
{
    FLUB_REQUIRE(st.n % 4 == 0, FLUB_EINVAL, "vector update: n = %u is not a multiple of 4", st.n);
    float *prev = st.P[st.cur], *out = st.P[st.cur ^ 1];
    if (st.npend == 0) {
        launch_pp<XU_NONE, false, true>(s, st, out, prev, z, zinv);
    } else if (st.npend == 1) {
        launch_pp<XU_DEFER, false, true>(s, st, out, prev, z, zinv);
        st.deferred = true;
    } else {
        if (st.x_set) launch_pp<XU_TWO, false, true>(s, st, out, prev, z, zinv);
        else launch_pp<XU_TWO, true, true>(s, st, out, prev, z, zinv);
        st.x_set = true;
        st.npend = 0;
    }
    FLUB_CHECK_LAUNCH();
    st.cur ^= 1;
    return FLUB_OK;
}

// After the r update of an iteration: its alpha p is pending.
inline void px_after_r_update(PxState &st)
{
    st.npend = st.deferred ? 2 : 1;
    st.deferred = false;
}

// Apply every pending x update (before x is read).
inline int px_flush(hipStream_t s, PxState &st)
{
    if (st.npend == 0) return FLUB_OK;
    float *pa = st.P[st.cur];
    if (st.npend == 1) {
        if (st.x_set) launch_pp<XU_ONE, false, false>(s, st, nullptr, pa, nullptr, nullptr);
        else launch_pp<XU_ONE, true, false>(s, st, nullptr, pa, nullptr, nullptr);
    } else {   // the older p in the other buffer, passed as p_out (read only: DOP false)
        float *pb = st.P[st.cur ^ 1];
        if (st.x_set) launch_pp<XU_TWO, false, false>(s, st, pb, pa, nullptr, nullptr);
        else launch_pp<XU_TWO, true, false>(s, st, pb, pa, nullptr, nullptr);
    }
    FLUB_CHECK_LAUNCH();
    st.x_set = true;
    st.npend = 0;
    return FLUB_OK;
}

int mg_dot(flub_mgpcg *h, const float *a, const float *b, int op, double eps);
int mg_zero_centered(flub_mgpcg *h);
int mg_precondition(flub_mgpcg *h, int smoother, bool use_graph);
// The V-cycle on the colour-split level 0 (sp.r -> sp.z[zslot]; FP32 or the
// handle's mixed precision), replayed from a captured graph.
int mg_precondition_split(flub_mgpcg *h, int zslot);

// After the kernel that wrote sc[SC_RTR] (and hsc[SC_RTR] through hsc_dev
// when mapped): queue the copy to hsc when not mapped, then record h->ev;
// hsc[SC_RTR] is valid once h->ev has completed.
int mg_rtr_sent(flub_mgpcg *h);
// The sum of nblk partials in exactly the order of the finish kernels
// (per-thread strided sums over RBS threads, the wavefront shuffle tree, then
// the wavefronts in order): bitwise their result, computed on the host.
double mg_host_partial_sum(const double *part, int nblk);
// After a solve whose r.r stayed on the host: store it into sc[SC_RTR].
int mg_rtr_to_device(flub_mgpcg *h, double rTr);

// The split solves' FP64 solution (SplitBuf::x).  mg_split_enter: r and x
// into the split layout and b = r - ndt(x) for the true residual, unless the
// fused projection wrote r and b already (x_set false: x is zero and not
// written yet).  mg_split_leave: x rounded to FP32 into the reference-layout
// h->x.  mg_split_true_residual: R = b - ndt(x) (FP64 stencil), *rTrue its
// squared norm (synchronizes).
int mg_split_enter(flub_mgpcg *h, bool &x_set);
int mg_split_leave(flub_mgpcg *h, bool x_set);
int mg_split_true_residual(flub_mgpcg *h, float *R, double *rTrue);

// alpha = old_zTr / (sum of the partials + eps) (mgpcg.py:212-216)
int mg_finish_alpha(flub_mgpcg *h, int nblk, double eps);

// mixed.hip
int mg_mixed_rscale(flub_mgpcg *h);
// Storage scale and r.r of r (split V-cycle); to_host: r.r also goes to
// hsc[SC_RTR] (see mg_rtr_sent).
int mg_mixed_rscale_of(flub_mgpcg *h, const float *r, bool to_host = false);
int mg_mixed_setup(flub_mgpcg *h);
void mg_mixed_teardown(flub_mgpcg *h);
int mg_vcycle_mixed(flub_mgpcg *h, int smoother);
int mg_solve_mixed(flub_mgpcg *h, int max_iters, double eps, double abs_tol, double rel_tol, int smoother,
                   int verbose, int *iters, double *rTr);
