// xs_solvers.cpp -- Krylov loops driven from the host with all scalars on the
// device.  Each iteration is a fixed sequence of kernel launches; the host
// reads the device convergence flag only every `check_every` iterations
// (iterations launched past convergence are no-ops: every kernel checks
// st->done first).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "../../include/xsolver_c.h"
#include "xs_solvers.h"

#define XS_CHK(call)                            \
  do {                                          \
    if ((call) != hipSuccess) return XS_ERR_DEVICE; \
  } while (0)

// Launch `iteration` until the device reports convergence/breakdown or
// maxiter iterations have been launched.  Two pinned state slots are used so
// the next batch is queued before the previous batch's flag is inspected.
template <class F>
static int drive(xs_ctx *c, F iteration) {
  hipStream_t s = c->stream;
  XS_CHK(hipMemcpyAsync(c->h_state, c->st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
  XS_CHK(hipStreamSynchronize(s));
  int launched = 0;
  while (!c->h_state->done && launched < c->maxiter) {
    const int k = std::min(c->check_every, c->maxiter - launched);
    for (int j = 0; j < k; j++) iteration();
    launched += k;
    XS_CHK(hipGetLastError());
    XS_CHK(hipMemcpyAsync(c->h_state, c->st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
    XS_CHK(hipStreamSynchronize(s));
  }
  const xs_state &h = *c->h_state;
  c->iters = h.iter;
  c->res0 = h.sc[S_RES0];
  c->resl = h.sc[S_RESL];
  if (h.done == 2) return XS_ERR_BREAKDOWN;
  if (h.done == 1) return XS_OK;
  return XS_NOT_CONVERGED;
}

// Preconditioned CG, hip/Solver/xsolver_cg.cpp:16-74.
int xs_solve_cg(xs_ctx *c, const double *b, double *x) {
  int rc = xs_ctx_ensure_work(c, XS_NWORK);
  if (rc) return rc;
  const int n = c->A.n;
  hipStream_t s = c->stream;
  double *r = xs_ctx_vec(c, 0), *z = xs_ctx_vec(c, 1), *p = xs_ctx_vec(c, 2);
  double *Ap = xs_ctx_vec(c, 3), *t = xs_ctx_vec(c, 4);
  const double *jac = (c->pc == XS_PC_JACOBI) ? c->jac_dinv : nullptr;
  xs_init_state(s, c->st, c->tol, c->abstol, c->maxiter);
  xs_op_spmv(c, x, r, EPI_BMAX, b, nullptr, nullptr);          // r = b - A x
  xs_op_pc(c, r, z, t, nullptr);                                // z = M^-1 r
  xs_reduce rinit = xs_ctx_reduce(c, F_CG_INIT, 0);
  xs_launch_dot(s, n, r, z, r, r, rinit);                       // tau, res0
  xs_copy(s, n, z, p);                                          // p = z
  const xs_reduce ralpha = xs_ctx_reduce(c, F_CG_ALPHA, 0);
  const xs_reduce rres = xs_ctx_reduce(c, F_CG_RES, 0);
  const xs_reduce rrr = xs_ctx_reduce(c, F_STORE, S_SCRATCH0);
  return drive(c, [&]() {
    xs_op_spmv(c, p, Ap, EPI_DOT1, nullptr, p, &ralpha);        // Ap, sigma -> alpha
    if (jac) {
      xs_cg_xr(s, n, x, r, p, Ap, jac, z, rres);                // x, r, z, tau, ||r||
    } else {
      xs_cg_xr(s, n, x, r, p, Ap, nullptr, nullptr, rrr);       // x, r, ||r||^2
      xs_op_pc(c, r, z, t, c->st);                              // z = M^-1 r
      xs_cg_zr_dot(s, n, z, r, rres);                           // tau, beta, convergence
    }
    xs_cg_p(s, n, p, z, c->st);                                 // p = z + beta p
  });
}

// Right-preconditioned BiCGSTAB, hip/Solver/xsolver_bicgstab.cpp:20-124.
int xs_solve_bicgstab(xs_ctx *c, const double *b, double *x) {
  int rc = xs_ctx_ensure_work(c, XS_NWORK);
  if (rc) return rc;
  const int n = c->A.n;
  hipStream_t s = c->stream;
  double *r = xs_ctx_vec(c, 0), *rh = xs_ctx_vec(c, 1), *p = xs_ctx_vec(c, 2);
  double *u = xs_ctx_vec(c, 3), *sv = xs_ctx_vec(c, 4), *v = xs_ctx_vec(c, 5);
  double *z = xs_ctx_vec(c, 6), *z1 = xs_ctx_vec(c, 7), *t = xs_ctx_vec(c, 8);
  const double *jac = (c->pc == XS_PC_JACOBI) ? c->jac_dinv : nullptr;
  xs_init_state(s, c->st, c->tol, c->abstol, c->maxiter);
  XS_CHK(hipMemsetAsync(p, 0, sizeof(double) * n, s));
  XS_CHK(hipMemsetAsync(u, 0, sizeof(double) * n, s));
  const xs_reduce r0 = xs_ctx_reduce(c, F_RES0, 0);
  xs_op_spmv(c, x, r, EPI_BMAXN, b, nullptr, &r0);              // r = b - A x, res0
  xs_copy(s, n, r, rh);                                         // rhat = r0
  const xs_reduce rg = xs_ctx_reduce(c, F_BICG_GAMA, 0);
  const xs_reduce ra = xs_ctx_reduce(c, F_BICG_ALPHA, 0);
  const xs_reduce rr = xs_ctx_reduce(c, F_BICG_RES, 0);
  return drive(c, [&]() {
    xs_bicg_p(s, n, r, u, p, jac, z, c->st);                    // p = r + w (p - a u)
    if (!jac) xs_op_pc(c, p, z, t, c->st);                      // z = M^-1 p
    xs_op_spmv(c, z, u, EPI_DOT1, nullptr, rh, &rg);            // u = A z, gama
    xs_bicg_s(s, n, r, u, sv, jac, z1, c->st);                  // s = r - gama u
    if (!jac) xs_op_pc(c, sv, z1, t, c->st);                    // z1 = M^-1 s
    xs_op_spmv(c, z1, v, EPI_DOT2, nullptr, sv, &ra);           // v = A z1, alpha
    xs_bicg_xr(s, n, x, r, z, z1, sv, v, rh, rr);               // x, r, ||r||, next beta
  });
}

int xs_solve_gcr(xs_ctx *, const double *, double *) { return XS_ERR_UNSUPPORTED; }
int xs_solve_gmres(xs_ctx *, const double *, double *, bool) { return XS_ERR_UNSUPPORTED; }
int xs_parilut_factor(xs_ctx *) { return XS_ERR_UNSUPPORTED; }
int xs_build_transpose(xs_ctx *) { return XS_ERR_UNSUPPORTED; }
int xs_op_pc_t(xs_ctx *, const double *, double *, double *, hipStream_t) { return XS_ERR_UNSUPPORTED; }
void xs_solvers_free(xs_ctx *c) {
  if (c->basis) (void)hipFree(c->basis);
  c->basis = nullptr;
  if (c->small_d) (void)hipFree(c->small_d);
  c->small_d = nullptr;
}
