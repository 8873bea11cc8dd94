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
enum { SC_ALPHA = 0, SC_BETA, SC_SUM, SC_OLDZTR, SC_RTR, SC_MEAN, SC_COUNT, SC_RSCALE, SC_AUX0, SC_AUX1,
       SC_AUX2, SC_AUX3, SC_ZINV0, SC_ZINV1, SC_NUM = 16 };
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
    // The restriction into this level stores cs times the sum of the children's
    // residuals (1, set at creation: the reference's un-normalised sum,
    // mgpcg.py:94-101).
    float cs;
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
    // p buffers of the solve's ring beyond p and p_alt (PxState), and the
    // alpha saved per ring slot (device)
    float *ph[32];
    int nph;           // ring size, p and p_alt included
    double *ahist;
    // BF16 solve: level 1's r and its pre-smoothed z (and the pairs' second z
    // buffer) in BF16, split layout
    void *r1h, *zp1h, *zt1h;
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
    int graph_rh[5];           // [3], [4]: the mixed V-cycle's 16-bit r state at capture (mx_state)
    hipEvent_t ev;     // marks the residual norm's copy to the host
    bool split;        // level 0 of the FP32 solve runs in the colour-split layout
    int nsplit;        // levels 0 .. nsplit-1 of the split V-cycle are colour-split (0: none)
    bool x_only;       // the solve's caller reads only x afterwards (the fused projection):
                       // r, z, p and Ap are left in the solver's internal layout
    bool presplit;     // the fused projection wrote r (and zero x) in the split layout already
    bool pair_l0;      // FP32 split level 0: the first pre- and post-smoothing half-sweeps in pairs
    bool mx_unit;      // BF16 split solve: storage scale 1, the r update stores the next cycle's BF16 r
    int rh_ready;      // z slot whose BF16 r the r update stored for the next V-cycle (-1: none)
    bool mx_fuse0;     // BF16 split solve: the r update also runs the next cycle's first half-sweep
    bool mx_l1;        // BF16 split solve: level 1 pre-smooths in BF16 (sp.r1h, zp1h, zt1h)
    int zh0_ready;     // z slot whose first pre-smoothing half-sweep the r update ran (-1: none)
    bool x64_valid;    // sp.x holds the FP64 x of the last solve (h->x is it rounded to FP32)
    double *x64n;      // the fused projection: the FP64 x in the reference layout, for the gradient
    bool want_x64n;    // the next split solve also writes x64n
    float *p_out;      // the fused projection: the split solve also writes its pressure here ...
    float p_h;         // ... from the FP32 x through the scale chain with this cell size
    bool p_written;    // the last split solve wrote p_out
    SplitBuf sp;
    MixedState *mx;
    Probe *probe;      // armed in-solve timing (nullptr: off)
    int sweep_cap;     // > 0: at most this many pre/post iterations on any level (0: the reference's doubling)
    bool no_tail;      // the coarsest levels launch kernel by kernel (no one-workgroup tail, tail.hip)
    int pap_ib;        // split solves' p step: -1 p update and Ap in two passes; 0 one pass, rows
                       // per segment chosen by grid size; > 0 one pass with this many rows
};

// The one-workgroup lower part of the V-cycle (tail.hip): levels of at most
// TAIL_MAX_CELLS cells, at most TAIL_MAXL of them.
constexpr uint32_t TAIL_MAX_CELLS = 4096;
constexpr int TAIL_MAXL = 8;
// First level (>= lowest, >= 1) the tail can take, or -1.
int tail_first_level(const flub_mgpcg *h, int lowest);
// The V-cycle from level `first` (r in its lv[].r) down to the bottom and back:
// z of every level from `first` on into lv[].z, the restricted r into lv[].r.
int tail_cycle(hipStream_t s, flub_mgpcg *h, int first);

// Red-black iterations of the pre- and post-smoothing of level l: the
// reference doubles them per level (pre_and_post_smoothing << l,
// mgpcg.py:301,341); a handle may cap them.
inline int mg_sweeps(const flub_mgpcg *h, int l)
{
    const int s = h->pre_post << l;
    return h->sweep_cap > 0 && s > h->sweep_cap ? h->sweep_cap : s;
}

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
// zinv is set (the z buffer holds z times a power-of-two scale: exact).  The
// solve keeps every p since x was last updated (PxState): with asave set,
// thread 0 also stores alpha (SC_ALPHA, the previous iteration's, which the
// r update used) into *asave for the later x update of p_prev.
static __global__ void __launch_bounds__(256) k_update_pp(float *__restrict__ p_out, const float *__restrict__ p_prev,
                                                   const float *__restrict__ z, const double *__restrict__ sc,
                                                   const double *__restrict__ zinv, double *__restrict__ asave,
                                                   uint32_t n4)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n4) return;
    if (asave && t == 0) asave[0] = sc[SC_ALPHA];   // read by a later launch only
    float4 pv = reinterpret_cast<const float4 *>(p_prev)[t];
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

// The deferred x updates of the split solves (mgpcg.py:207-210): x += alpha_i
// p_i for the p of every iteration since the last flush, oldest first, in
// FP64: each term (double)alpha_i * (double)p_i is exact (alpha_i the FP32
// value the r update used) and each addition rounds once, so x is the FP64
// sum of the updates the reference makes in FP32.  alpha_i is *a[i] (the last
// one may be sc + SC_ALPHA).  XSET: x is zero (not read).
#define MAXPH 32
struct XFlush {
    const float *p[MAXPH];
    const double *a[MAXPH];
    int n;
};
template <bool XSET>
__global__ void __launch_bounds__(256) k_flush_x(double *__restrict__ x, XFlush f, uint32_t n4)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n4) return;
    double xv[4] = {0.0, 0.0, 0.0, 0.0};
    if (!XSET) {
        const double2 a = reinterpret_cast<const double2 *>(x)[2 * t];
        const double2 b = reinterpret_cast<const double2 *>(x)[2 * t + 1];
        xv[0] = a.x; xv[1] = a.y; xv[2] = b.x; xv[3] = b.y;
    }
    for (int i = 0; i < f.n; ++i) {
        const double al = (double)(float)f.a[i][0];
        const float4 q = reinterpret_cast<const float4 *>(f.p[i])[t];
        xv[0] = fma(al, (double)q.x, xv[0]);
        xv[1] = fma(al, (double)q.y, xv[1]);
        xv[2] = fma(al, (double)q.z, xv[2]);
        xv[3] = fma(al, (double)q.w, xv[3]);
    }
    reinterpret_cast<double2 *>(x)[2 * t] = make_double2(xv[0], xv[1]);
    reinterpret_cast<double2 *>(x)[2 * t + 1] = make_double2(xv[2], xv[3]);
}

// The split solves' p buffers and deferred x updates.  Each p step writes the
// new p into a free buffer of the ring (the previous p stays readable), and
// the x update alpha p of every iteration waits until x is needed (the true
// residual, the end) or the ring is full: then one pass adds them all.  x
// (FP64, 16 bytes a cell read and written) then moves once per flush instead
// of once per iteration, and each p once more (4 bytes).
struct PxState {
    float *ring[MAXPH];
    int nring = 0;
    double *ahist = nullptr;   // device: the alpha saved for each ring slot
    int cur = 0;               // slot of the latest p
    int pend[MAXPH];           // slots whose x update is pending, oldest first
    int npend = 0;
    bool newest_saved = false; // the newest pending alpha was copied into ahist
    bool x_set = true;         // x holds a value (else it is zero and not yet written)
    double *x = nullptr;
    double *sc = nullptr;
    uint32_t n = 0;
    float *p() const { return ring[cur]; }
};

// Apply every pending x update (before x is read).
inline int px_flush(hipStream_t s, PxState &st)
{
    if (st.npend == 0) return FLUB_OK;
    FLUB_REQUIRE(st.n % 4 == 0, FLUB_EINVAL, "x update: n = %u is not a multiple of 4", st.n);
    XFlush f{};
    f.n = st.npend;
    for (int i = 0; i < st.npend; ++i) {
        const int sl = st.pend[i];
        f.p[i] = st.ring[sl];
        f.a[i] = (i + 1 < st.npend || st.newest_saved) ? st.ahist + sl : st.sc + SC_ALPHA;
    }
    const uint32_t n4 = st.n / 4;
    if (st.x_set) k_flush_x<false><<<blocks_for(n4, 256), 256, 0, s>>>(st.x, f, n4);
    else k_flush_x<true><<<blocks_for(n4, 256), 256, 0, s>>>(st.x, f, n4);
    FLUB_CHECK_LAUNCH();
    st.x_set = true;
    st.npend = 0;
    st.newest_saved = false;
    return FLUB_OK;
}

// The p step of an iteration: p into a free ring buffer; the pending update of
// the previous p gets its alpha saved.  px_step_slot picks that buffer (*out)
// and says whether the kernel writing it also saves alpha into
// ahist[cur] (*save); px_step_done makes it the current p.
inline int px_step_slot(hipStream_t s, PxState &st, int *out, bool *save)
{
    FLUB_REQUIRE(st.n % 4 == 0, FLUB_EINVAL, "vector update: n = %u is not a multiple of 4", st.n);
    if (st.npend + 1 >= st.nring) {   // keep a buffer free besides the pending ones
        const int rc = px_flush(s, st);
        if (rc) return rc;
    }
    bool used[MAXPH] = {};
    used[st.cur] = true;
    for (int i = 0; i < st.npend; ++i) used[st.pend[i]] = true;
    int o = 0;
    while (used[o]) ++o;
    *out = o;
    *save = st.npend > 0 && st.pend[st.npend - 1] == st.cur && !st.newest_saved;
    return FLUB_OK;
}
inline void px_step_done(PxState &st, int out, bool save)
{
    if (save) st.newest_saved = true;
    st.cur = out;
}
inline int px_step(hipStream_t s, PxState &st, const float *z, const double *zinv)
{
    int out;
    bool save;
    const int rc = px_step_slot(s, st, &out, &save);
    if (rc) return rc;
    const uint32_t n4 = st.n / 4;
    k_update_pp<<<blocks_for(n4, 256), 256, 0, s>>>(st.ring[out], st.ring[st.cur], z, st.sc, zinv,
                                                    save ? st.ahist + st.cur : nullptr, n4);
    FLUB_CHECK_LAUNCH();
    px_step_done(st, out, save);
    return FLUB_OK;
}

// The ring of a split solve: p (zero at the start of a solve, as init and
// the fused projection leave it), p_alt and the extra buffers.
inline void px_init(flub_mgpcg *h, PxState &st, double *x, uint32_t n)
{
    st.nring = h->sp.nph;
    st.ring[0] = h->p;
    st.ring[1] = h->p_alt;
    for (int i = 2; i < st.nring; ++i) st.ring[i] = h->sp.ph[i];
    st.ahist = h->sp.ahist;
    st.x = x;
    st.sc = h->sc;
    st.n = n;
}

// After the r update of an iteration: its alpha p is pending.
inline void px_after_r_update(PxState &st)
{
    st.pend[st.npend++] = st.cur;
    st.newest_saved = false;
}

// The split solves' p step followed by Ap = ndt(p) and the partials of p.Ap
// (*nblk of them): one pass (split.hip, k_sp_pAp) when ib >= 0 and the grid
// fits it, else px_step then split_Ap.  p and Ap are bitwise the same either
// way; the partials sum the same products in another grouping.
int split_pstep_Ap(hipStream_t s, PxState &st, const float *z, const double *zinv, const Level &L0, float *Ap,
                   const uint16_t *code, double *part, int *nblk, int ib);

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
