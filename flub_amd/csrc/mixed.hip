// Mixed-precision V-cycle and flexible CG (no counterpart in the reference).
// Placeholder until the FP16/BF16 kernels land: handles created with a
// reduced-precision V-cycle report FLUB_ESTATE from solve.
#include "mgpcg_internal.h"

int mg_mixed_setup(flub_mgpcg *h) { (void)h; return FLUB_OK; }
void mg_mixed_teardown(flub_mgpcg *h) { (void)h; }

int mg_vcycle_mixed(flub_mgpcg *h, int smoother)
{
    (void)h; (void)smoother;
    flub_set_error("mixed-precision V-cycle not built yet");
    return FLUB_ESTATE;
}

int mg_solve_mixed(flub_mgpcg *h, int max_iters, double eps, double abs_tol, double rel_tol, int smoother,
                   int verbose, int *iters, double *rTr)
{
    (void)h; (void)max_iters; (void)eps; (void)abs_tol; (void)rel_tol; (void)smoother; (void)verbose;
    (void)iters; (void)rTr;
    flub_set_error("mixed-precision solve not built yet");
    return FLUB_ESTATE;
}
