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
};

struct MixedState;

// Level 0 in the colour-split layout (split.hip), used inside the FP32 solve.
struct SplitBuf {
    float *r, *zpre, *x;
    float *z[2];    // two z buffers: the solve preconditions speculatively into the spare one
    uint16_t *code;
    // mixed precision: r and pre-smoothed z of the V-cycle in FP16/BF16, the
    // post-smoothing's FP32 working z
    void *rh, *zh;
    float *zw;
    float *b;       // mixed precision: the right-hand side (split), for the true residual
};

struct flub_mgpcg {
    int L, prec;
    Level lv[MAXL];
    float *x, *p, *Ap;
    float *p_alt;   // second p buffer: the fused p-update + Ap writes out of place
    double *sc;     // device scalars
    bool fast;      // quad kernels usable (tagged smoother, levels >= 2, pre_post >= 1)
    int vc_dot_nblk;  // partial count of the z.r fused into the fast V-cycle (0: none)
    double *part;   // reduction partials (2*MAXPART)
    double *hsc;    // pinned host mirror of sc
    int pre_post, bottom;
    hipStream_t s;     // work stream (caller's; may be the null stream)
    hipStream_t cap;   // private stream used only to capture graphs
    bool tags_set;
    bool no_graph;
    bool no_tb;        // launch-per-half-sweep smoothing on every level (no temporal blocking)
    bool no_tb_l0;     // no temporal blocking on level 0
    bool tb_g8;        // blocked smoothing may take 8 half-sweeps per launch on small levels
    uint32_t *tb_order;  // update-order tables of the blocked smoother
    hipGraphExec_t graph[5];   // per smoother; [3], [4]: the V-cycle on the split level 0 into sp.z[0], sp.z[1]
    bool graph_ok[5];
    hipEvent_t ev;     // marks the residual norm's copy to the host
    bool split;        // level 0 of the FP32 solve runs in the colour-split layout
    bool fused_pAp;    // split solve: p update fused into the Ap kernel (else an elementwise pass first)
    SplitBuf sp;
    MixedState *mx;
};

int mg_dot(flub_mgpcg *h, const float *a, const float *b, int op, double eps);
int mg_zero_centered(flub_mgpcg *h);
int mg_precondition(flub_mgpcg *h, int smoother, bool use_graph);
// The V-cycle on the colour-split level 0 (sp.r -> sp.z[zslot]; FP32 or the
// handle's mixed precision), replayed from a captured graph.
int mg_precondition_split(flub_mgpcg *h, int zslot);

// alpha = old_zTr / (sum of the partials + eps) (mgpcg.py:212-216)
int mg_finish_alpha(flub_mgpcg *h, int nblk, double eps);

// mixed.hip
int mg_mixed_rscale(flub_mgpcg *h);
int mg_mixed_rscale_of(flub_mgpcg *h, const float *r);
int mg_mixed_setup(flub_mgpcg *h);
void mg_mixed_teardown(flub_mgpcg *h);
int mg_vcycle_mixed(flub_mgpcg *h, int smoother);
int mg_solve_mixed(flub_mgpcg *h, int max_iters, double eps, double abs_tol, double rel_tol, int smoother,
                   int verbose, int *iters, double *rTr);
