// xs_intel.h -- building blocks of the intel-family Krylov loops
// (xs_intel.hip): multi-pair dots, one-thread scalar steps, linear
// combinations, all with the scalars in the device state.
#pragma once
#include "xs_internal.h"

// scalar slots of the intel-family loops in xs_state::sc
enum xs_islot {
  IS_TAU = XS_NSCAL - 24, IS_TAUO, IS_ALPHA, IS_NALPHA, IS_BETA, IS_RHO, IS_THETA, IS_FAI,
  IS_OMEGA, IS_NOMEGA, IS_DELTA, IS_NDELTA, IS_SIGMA, IS_PI, IS_MU, IS_ROU, IS_ALPHAO, IS_T1,
  IS_NT2, IS_RESL, IS_ONE, IS_MONE
};

// the scalar step that follows a multi-dot (or stands alone)
enum xs_istep {
  IST_CGL_INIT = 0,  // r.r -> res0, tau
  IST_CGL_ALPHA,     // p.q -> alpha
  IST_CGL_RES,       // r.r -> commit, beta
  IST_ICG_ALPHA,     // p.q, r.q, q.q -> alpha, rho, theta
  IST_ICG_RES,       // r.r -> commit, recurrence tau, beta
  IST_ICGM_INIT,     // r.r, r.z -> res0, tau
  IST_ICGM_MAIN,     // p.Ap, r.q, Ap.z, q.Ap, r.r -> alpha, tau, beta, pending ||r||
  IST_COMMIT,        // iteration count, resvec, convergence on the pending ||r||
  IST_IBI_INIT,      // r.r, r.t -> res0, fai, sigma, start scalars
  IST_IBI_S1,        // (no dots) rho, delta, beta, tau, alpha and the update weights
  IST_IBI_S2,        // six dots -> fai, pi, mu, omega, sigma
  IST_IBI_RES        // r.r -> commit
};

#define XS_MDOT_MAX 6
#define XS_MDOT_STRIDE 8
// partial buffer: XS_MDOT_STRIDE doubles per workgroup of the dot grid
int xs_mdot_blocks(int n);

// tot[p] = x_p . y_p (p < np) reduced into the scalar step `step`
void xs_mdot_step(hipStream_t s, int n, int np, const double *const *x, const double *const *y,
                  double *partial, int step, xs_state *st, double *resvec);
// a scalar step with no dots
void xs_istep(hipStream_t s, int step, xs_state *st, double *resvec);
// y = sum_t sc[slot_t] x_t (t < nt <= 4); y may be one of the x_t.  Skipped
// once the state is done.
void xs_lincomb(hipStream_t s, int n, int nt, const double *const *x, const int *slot, double *y,
                const xs_state *st);
// y += x, whatever the state
void xs_add(hipStream_t s, int n, const double *x, double *y);
