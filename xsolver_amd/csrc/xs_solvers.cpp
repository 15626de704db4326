// xs_solvers.cpp -- Krylov loops driven from the host with all scalars on the
// device.  Each iteration is a fixed sequence of kernel launches; the host
// reads the device convergence flag only every `check_every` iterations
// (iterations launched past convergence are no-ops: every kernel checks
// st->done first).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "../../include/xsolver_c.h"
#include "xs_intel.h"
#include "xs_krylov.h"
#include "xs_mixed.h"
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
  xs_graph_cache gc;
  while (!c->h_state->done && launched < c->maxiter) {
    const int k = std::min(c->check_every, c->maxiter - launched);
    XS_CHK(xs_run_batch(c, gc, k, iteration) == 0 ? hipSuccess : hipErrorUnknown);
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
  // With the sweep apply, p's update is formed inside the first L pass; the
  // pass reads the old p at neighbouring rows, so the new p goes to the other
  // of two buffers, alternating.  An even batch of captured iterations ends
  // on the parity it started with, which replays of the batch rely on.
  const bool pfuse = !jac && c->fuse_p && xs_pc_fuse_l(c) && c->check_every % 2 == 0;
  double *p2 = xs_ctx_vec(c, 9);
  bool odd = false;
  return drive(c, [&]() {
    if (pfuse) {
      xs_op_pc_p(c, r, u, odd ? p2 : p, odd ? p : p2, z, t, c->st);   // p update, z = M^-1 p
      odd = !odd;
    } else {
      xs_bicg_p(s, n, r, u, p, jac, z, c->st);                  // p = r + w (p - a u)
      if (!jac) xs_op_pc(c, p, z, t, c->st);                    // z = M^-1 p
    }
    xs_op_spmv(c, z, u, EPI_DOT1, nullptr, rh, &rg);            // u = A z, gama
    if (jac) xs_bicg_s(s, n, r, u, sv, jac, z1, c->st);         // s = r - gama u, z1 = D s
    else xs_op_pc_s(c, r, u, sv, z1, t, c->st);                 // s = r - gama u, z1 = M^-1 s
    xs_op_spmv(c, z1, v, EPI_DOT2, nullptr, sv, &ra);           // v = A z1, alpha
    xs_bicg_xr(s, n, x, r, z, z1, sv, v, rh, rr);               // x, r, ||r||, next beta
  });
}

// ---------------------------------------------------------------------------
// basis workspace
// ---------------------------------------------------------------------------
static int ensure_dev(double **p, size_t *len, size_t need) {
  if (*p && *len >= need) return XS_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *len = 0;
  if (hipMalloc((void **)p, sizeof(double) * (need ? need : 1)) != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return XS_ERR_MEMORY;
  }
  *len = need;
  return XS_OK;
}

// reduction slots start out (and are left) holding the sentinel
static int ensure_ts(xs_ctx *c) {
  const size_t need = (size_t)std::max(xs_ts_blocks(c->A.n), XS_MAX_REDUCE_BLOCKS) * XS_TS_KMAX;
  if (c->ts_partial && c->ts_partial_len >= need) return XS_OK;
  int rc = ensure_dev(&c->ts_partial, &c->ts_partial_len, need);
  if (rc) return rc;
  xs_fill_sentinel(c->stream, c->ts_partial, (int)c->ts_partial_len);
  return XS_OK;
}

static xs_gmres_small carve_small(double *base, int m) {
  xs_gmres_small g;
  double *p = base;
  g.H = p; p += (size_t)(m + 1) * m;
  g.cs = p; p += m + 1;
  g.sn = p; p += m + 1;
  g.s = p; p += m + 1;
  g.h1 = p; p += m + 2;
  g.h2 = p; p += m + 2;
  g.y = p; p += m + 1;
  g.misc = p;
  return g;
}
static size_t small_size(int m) { return (size_t)(m + 1) * m + 6 * (size_t)(m + 2) + 16; }

// GMRES(m) (flexible = false: left preconditioning,
// intel/mpisolver-C-v1.cpp:933-1087) and FGMRES(m) (flexible = true: right
// preconditioning with stored Z_i = M^{-1} V_i, :1089-1247).  Arnoldi with
// CGS2 on the device; Hessenberg, Givens rotations, residual estimate and
// back-substitution in FP64 on the device.  On convergence the update uses
// all Krylov directions of the cycle.
int xs_solve_gmres(xs_ctx *c, const double *b, double *x, bool flexible) {
  int rc = xs_ctx_ensure_work(c, XS_NWORK);
  if (rc) return rc;
  const int n = c->A.n;
  const int m = c->restart > 0 ? c->restart : 30;
  hipStream_t s = c->stream;
  const size_t ld = (size_t)n;
  const size_t nbasis = (size_t)(m + 1) + (flexible ? (size_t)m : 0);
  if ((rc = ensure_dev(&c->basis, &c->basis_len, nbasis * ld))) return rc;
  if ((rc = ensure_dev(&c->small_d, &c->small_len, small_size(m)))) return rc;
  if ((rc = ensure_ts(c))) return rc;
  double *V = c->basis;
  double *Z = flexible ? c->basis + (size_t)(m + 1) * ld : nullptr;
  xs_gmres_small g = carve_small(c->small_d, m);
  double *r = xs_ctx_vec(c, 0), *t = xs_ctx_vec(c, 1), *tmp = xs_ctx_vec(c, 2);
  double *w = xs_ctx_vec(c, 3);
  unsigned *tc = c->counters + 10;
  double *P = c->ts_partial;
  xs_state *st = c->st;
  double *rv = c->resvec_d;

  xs_init_state(s, st, c->tol, c->abstol, c->maxiter);
  auto residual = [&]() {   // r = [M^{-1}] (b - A x)
    xs_op_spmv(c, x, flexible ? r : t, EPI_BMAX, b, nullptr, nullptr);
    if (!flexible) xs_op_pc(c, t, r, tmp, nullptr);
  };
  residual();
  xs_ts_dot<double>(s, n, 1, r, ld, r, P, tc, g.misc + XS_G_NRM2, nullptr);
  xs_gmres_cycle(s, m, 1, g, st, rv);
  xs_ts_scale<double>(s, n, r, V, g.misc + XS_G_SCALE, st);

  int launched = 0, i = 0, cycle_start = 0;
  XS_CHK(hipMemcpyAsync(c->h_state, st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
  XS_CHK(hipStreamSynchronize(s));
  bool done = c->h_state->done != 0;
  while (!done && launched < c->maxiter) {
    double *vi = V + (size_t)i * ld, *vn = V + (size_t)(i + 1) * ld;
    if (flexible) {
      double *zi = Z + (size_t)i * ld;
      xs_op_pc(c, vi, zi, tmp, st);
      xs_op_spmv(c, zi, w, EPI_Y, nullptr, nullptr, nullptr);
    } else {
      xs_op_spmv(c, vi, t, EPI_Y, nullptr, nullptr, nullptr);
      xs_op_pc(c, t, w, tmp, st);
    }
    xs_ts_dot<double>(s, n, i + 1, V, ld, w, P, tc, g.h1, st);
    xs_ts_update<double>(s, n, i + 1, V, nullptr, ld, g.h1, w, w, nullptr, st);
    xs_ts_dot<double>(s, n, i + 1, V, ld, w, P, tc, g.h2, st);
    xs_ts_update<double>(s, n, i + 1, V, nullptr, ld, g.h2, w, vn, nullptr, st);
    xs_ts_dot<double>(s, n, 1, vn, ld, vn, P, tc, g.misc + XS_G_NRM2, st);
    xs_gmres_column(s, i, m, g, st, rv);
    xs_ts_scale<double>(s, n, vn, vn, g.misc + XS_G_SCALE, st);
    i++;
    launched++;
    if (i == m || launched == c->maxiter || launched % c->check_every == 0) {
      XS_CHK(hipGetLastError());
      XS_CHK(hipMemcpyAsync(c->h_state, st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
      XS_CHK(hipStreamSynchronize(s));
      done = c->h_state->done != 0;
      if (done) break;
    }
    if (i == m && launched < c->maxiter) {      // restart (intel :1065-1070)
      xs_gmres_backsub(s, m, m, g);
      xs_ts_axpy<double>(s, n, m, flexible ? Z : V, ld, g.y, x);
      residual();
      xs_gmres_cycle(s, m, 0, g, st, rv);
      xs_ts_scale<double>(s, n, r, V, g.misc + XS_G_SCALE, st);
      i = 0;
      cycle_start = launched;
    }
  }
  const xs_state &h = *c->h_state;
  int k;
  if (h.done == 1) k = h.iter - cycle_start;   // directions of the converged cycle
  else k = i;                                  // maxiter: use what this cycle built
  if (h.done != 2 && k > 0) {
    xs_gmres_backsub(s, k, m, g);
    xs_ts_axpy<double>(s, n, k, flexible ? Z : V, ld, g.y, x);
  }
  XS_CHK(hipGetLastError());
  XS_CHK(hipStreamSynchronize(s));
  c->iters = h.iter;
  c->res0 = h.sc[S_RES0];
  c->resl = h.sc[S_RESL];
  if (h.done == 2) return XS_ERR_BREAKDOWN;
  if (h.done == 1) return XS_OK;
  return XS_NOT_CONVERGED;
}

// Right-preconditioned GCR(restart), hip/Solver/xsolver_igcr.cpp:28-108 (the
// reference's default restart is maxiter, xsolver_igcr.h:36); c = A u is
// orthogonalised against the stored c's with CGS2, the same coefficients
// applied to u.  left = true: the intel left-preconditioned GCR
// (intel/mpisolver-C-v1.cpp:685-810 with classical Gram-Schmidt, one pass;
// mgs: :812-931, modified Gram-Schmidt, one coefficient at a time) on the
// residual r = M^-1 (b - A x): u_k = r, c_k = M^-1 A u_k.
int xs_solve_gcr(xs_ctx *c, const double *b, double *x, bool left, bool mgs) {
  int rc = xs_ctx_ensure_work(c, XS_NWORK);
  if (rc) return rc;
  const int n = c->A.n;
  const int R = c->restart > 0 ? c->restart : c->maxiter;
  hipStream_t s = c->stream;
  const size_t ld = (size_t)n;
  if ((rc = ensure_dev(&c->basis, &c->basis_len, 2 * (size_t)R * ld))) return rc;
  if ((rc = ensure_dev(&c->small_d, &c->small_len, 2 * (size_t)R + 16))) return rc;
  if ((rc = ensure_ts(c))) return rc;
  double *U = c->basis, *C = c->basis + (size_t)R * ld;
  double *h1 = c->small_d, *h2 = c->small_d + R, *misc = c->small_d + 2 * R;
  double *r = xs_ctx_vec(c, 0), *tmp = xs_ctx_vec(c, 1);
  unsigned *tc = c->counters + 10;
  double *P = c->ts_partial;
  xs_state *st = c->st;
  double *rv = c->resvec_d;

  xs_init_state(s, st, c->tol, c->abstol, c->maxiter);
  if (left) {
    double *t = xs_ctx_vec(c, 2);
    xs_op_spmv(c, x, t, EPI_BMAX, b, nullptr, nullptr);
    xs_op_pc(c, t, r, tmp, nullptr);                               // r = M^-1 (b - A x)
  } else {
    xs_op_spmv(c, x, r, EPI_BMAX, b, nullptr, nullptr);
  }
  xs_ts_dot<double>(s, n, 1, r, ld, r, P, tc, misc + XS_G_NRM2, nullptr);
  xs_res0_from(s, misc + XS_G_NRM2, st, rv);
  int k = 0, launched = 0;
  XS_CHK(hipMemcpyAsync(c->h_state, st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
  XS_CHK(hipStreamSynchronize(s));
  bool done = c->h_state->done != 0;
  double *t01 = misc + XS_G_T0;
  while (!done && launched < c->maxiter) {
    double *uk = U + (size_t)k * ld, *ck = C + (size_t)k * ld;
    if (left) {
      double *t = xs_ctx_vec(c, 2);
      xs_copy(s, n, r, uk);                                        // u_k = r
      xs_op_spmv(c, uk, t, EPI_Y, nullptr, nullptr, nullptr);
      xs_op_pc(c, t, ck, tmp, st);                                 // c_k = M^-1 A u_k
    } else {
      xs_op_pc(c, r, uk, tmp, st);                                 // u_k = M^-1 r
      xs_op_spmv(c, uk, ck, EPI_Y, nullptr, nullptr, nullptr);     // c_k = A u_k
    }
    if (k > 0 && left && mgs) {                                    // one coefficient at a time
      for (int i = 0; i < k; i++) {
        xs_ts_dot<double>(s, n, 1, C + (size_t)i * ld, ld, ck, P, tc, h1 + i, st);
        xs_ts_update<double>(s, n, 1, C + (size_t)i * ld, U + (size_t)i * ld, ld, h1 + i, ck,
                             ck, uk, st);
      }
    } else if (k > 0 && left) {                                    // all from the unmodified c_k
      xs_ts_dot<double>(s, n, k, C, ld, ck, P, tc, h1, st);
      xs_ts_update<double>(s, n, k, C, U, ld, h1, ck, ck, uk, st);
    } else if (k > 0) {
      xs_ts_dot<double>(s, n, k, C, ld, ck, P, tc, h1, st);
      xs_ts_update<double>(s, n, k, C, U, ld, h1, ck, ck, uk, st);
      xs_ts_dot<double>(s, n, k, C, ld, ck, P, tc, h2, st);
      xs_ts_update<double>(s, n, k, C, U, ld, h2, ck, ck, uk, st);
    }
    xs_ts_dot<double>(s, n, 1, ck, ld, ck, P, tc, t01, st);                // c.c
    xs_ts_dot<double>(s, n, 1, ck, ld, r, P, tc, t01 + 1, st);             // c.r
    xs_gcr_apply(s, n, uk, ck, x, r, t01, P, tc, st, rv);          // x, r, ||r||
    k = (k + 1 == R) ? 0 : k + 1;
    launched++;
    if (launched == c->maxiter || launched % c->check_every == 0) {
      XS_CHK(hipGetLastError());
      XS_CHK(hipMemcpyAsync(c->h_state, st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
      XS_CHK(hipStreamSynchronize(s));
      done = c->h_state->done != 0;
    }
  }
  const xs_state &h = *c->h_state;
  c->iters = h.iter;
  c->res0 = h.sc[S_RES0];
  c->resl = h.sc[S_RESL];
  if (h.done == 2) return XS_ERR_BREAKDOWN;
  if (h.done == 1) return XS_OK;
  return XS_NOT_CONVERGED;
}
// ---------------------------------------------------------------------------
// The rest of the intel Krylov family (intel/mpisolver-C-v1.cpp), each as a
// fixed launch sequence per iteration over the generic kernels of
// xs_intel.hip (multi-dot + scalar step, linear combinations) and the SpMV /
// preconditioner operators; small = 1e-20 in the denominators as there.
// ---------------------------------------------------------------------------
static int ensure_mdot(xs_ctx *c) {
  return ensure_dev(&c->mdot_partial, &c->mdot_len,
                    (size_t)xs_mdot_blocks(c->A.n) * XS_MDOT_STRIDE);
}

// CG on M^-1 A (cpu_cg_solve, :26-102) and, improved = true, ICG
// (cpu_icg_solve, :104-192: p.q, r.q and q.q in one reduction, the next
// beta from the recurrence tao = taoo - 2 alpha rho + alpha^2 theta).  The
// residual is the preconditioned one, r = M^-1 (b - A x).
int xs_solve_cg_left(xs_ctx *c, const double *b, double *x, bool improved) {
  int rc = xs_ctx_ensure_work(c, XS_NWORK);
  if (rc || (rc = ensure_mdot(c))) return rc;
  const int n = c->A.n;
  hipStream_t s = c->stream;
  double *r = xs_ctx_vec(c, 0), *p = xs_ctx_vec(c, 1), *Ap = xs_ctx_vec(c, 2);
  double *q = xs_ctx_vec(c, 3), *t = xs_ctx_vec(c, 4);
  double *P = c->mdot_partial, *rv = c->resvec_d;
  xs_state *st = c->st;
  xs_init_state(s, st, c->tol, c->abstol, c->maxiter);
  xs_op_spmv(c, x, Ap, EPI_BMAX, b, nullptr, nullptr);
  xs_op_pc(c, Ap, r, t, nullptr);                                  // r = M^-1 (b - A x)
  {
    const double *xs_[1] = {r}, *ys_[1] = {r};
    xs_mdot_step(s, n, 1, xs_, ys_, P, IST_CGL_INIT, st, rv);
  }
  xs_copy(s, n, r, p);
  const double *x_pq[3] = {p, r, q}, *y_pq[3] = {q, q, q};
  const double *x_rr[1] = {r}, *y_rr[1] = {r};
  const double *lx[2] = {x, p}, *lr[2] = {r, q}, *lp[2] = {p, r};
  const int sx[2] = {IS_ONE, IS_ALPHA}, sr[2] = {IS_ONE, IS_NALPHA}, sp[2] = {IS_BETA, IS_ONE};
  return drive(c, [&]() {
    xs_op_spmv(c, p, Ap, EPI_Y, nullptr, nullptr, nullptr);        // Ap
    xs_op_pc(c, Ap, q, t, st);                                     // q = M^-1 A p
    xs_mdot_step(s, n, improved ? 3 : 1, x_pq, y_pq, P,
                 improved ? IST_ICG_ALPHA : IST_CGL_ALPHA, st, rv);
    xs_lincomb(s, n, 2, lx, sx, x, st);                            // x += alpha p
    xs_lincomb(s, n, 2, lr, sr, r, st);                            // r -= alpha q
    xs_mdot_step(s, n, 1, x_rr, y_rr, P, improved ? IST_ICG_RES : IST_CGL_RES, st, rv);
    xs_lincomb(s, n, 2, lp, sp, p, st);                            // p = beta p + r
  });
}

// ICGM (cpu_icgm_solve, :273-352): preconditioned CG with the five inner
// products of an iteration in one reduction; convergence is judged on ||r||
// from before the iteration's update (the x update of that iteration is
// still applied, as in the reference).
int xs_solve_icgm(xs_ctx *c, const double *b, double *x) {
  int rc = xs_ctx_ensure_work(c, XS_NWORK);
  if (rc || (rc = ensure_mdot(c))) return rc;
  const int n = c->A.n;
  hipStream_t s = c->stream;
  double *r = xs_ctx_vec(c, 0), *z = xs_ctx_vec(c, 1), *p = xs_ctx_vec(c, 2);
  double *Ap = xs_ctx_vec(c, 3), *q = xs_ctx_vec(c, 4), *t = xs_ctx_vec(c, 5);
  double *P = c->mdot_partial, *rv = c->resvec_d;
  xs_state *st = c->st;
  xs_init_state(s, st, c->tol, c->abstol, c->maxiter);
  xs_op_spmv(c, x, r, EPI_BMAX, b, nullptr, nullptr);              // r = b - A x
  xs_op_pc(c, r, z, t, nullptr);                                   // z = M^-1 r
  {
    const double *xs_[2] = {r, r}, *ys_[2] = {r, z};
    xs_mdot_step(s, n, 2, xs_, ys_, P, IST_ICGM_INIT, st, rv);
  }
  xs_copy(s, n, z, p);
  const double *x5[5] = {p, r, Ap, q, r}, *y5[5] = {Ap, q, z, Ap, r};
  const double *lx[2] = {x, p}, *lr[2] = {r, Ap}, *lp[2] = {p, z};
  const int sx[2] = {IS_ONE, IS_ALPHA}, sr[2] = {IS_ONE, IS_NALPHA}, sp[2] = {IS_BETA, IS_ONE};
  return drive(c, [&]() {
    xs_op_spmv(c, p, Ap, EPI_Y, nullptr, nullptr, nullptr);
    xs_op_pc(c, Ap, q, t, st);                                     // q = M^-1 A p
    xs_mdot_step(s, n, 5, x5, y5, P, IST_ICGM_MAIN, st, rv);       // the one reduction
    xs_lincomb(s, n, 2, lx, sx, x, st);                            // x += alpha p
    xs_lincomb(s, n, 2, lr, sr, r, st);                            // r -= alpha A p
    xs_op_pc(c, r, z, t, st);                                      // z = M^-1 r
    xs_lincomb(s, n, 2, lp, sp, p, st);                            // p = beta p + z
    xs_istep(s, IST_COMMIT, st, rv);                               // ||r|| of before the update
  });
}

// Improved BiCGSTAB (cpu_ibicgstab_solve, :484-684), right-preconditioned in
// y-space: the correction x - x0 = M^-1 y is built from y = 0 (the
// reference starts from y0 = its initial guess; identical for x0 = 0).  Per
// iteration: two SpMVs, two preconditioner applications, six inner products
// in one reduction, and ||r||.
int xs_solve_ibicgstab(xs_ctx *c, const double *b, double *x) {
  int rc = xs_ctx_ensure_work(c, XS_NWORK);
  if (rc || (rc = ensure_mdot(c))) return rc;
  const int n = c->A.n;
  hipStream_t s = c->stream;
  double *res = xs_ctx_vec(c, 0), *reso = xs_ctx_vec(c, 1), *tk = xs_ctx_vec(c, 2);
  double *pk = xs_ctx_vec(c, 3), *qk = xs_ctx_vec(c, 4), *zk = xs_ctx_vec(c, 5);
  double *vk = xs_ctx_vec(c, 6), *vko = xs_ctx_vec(c, 7), *sk = xs_ctx_vec(c, 8);
  double *tm = xs_ctx_vec(c, 9), *phi = xs_ctx_vec(c, 10), *t = xs_ctx_vec(c, 11);
  double *P = c->mdot_partial, *rv = c->resvec_d;
  xs_state *st = c->st;
  xs_init_state(s, st, c->tol, c->abstol, c->maxiter);
  for (double *v : {pk, qk, zk, vk, phi}) XS_CHK(hipMemsetAsync(v, 0, sizeof(double) * n, s));
  xs_op_spmv(c, x, res, EPI_BMAX, b, nullptr, nullptr);            // res = b - A x0
  xs_copy(s, n, res, reso);
  xs_op_pc(c, res, tm, t, nullptr);
  xs_op_spmv(c, tm, tk, EPI_Y, nullptr, nullptr, nullptr);         // tk = A M^-1 res
  {
    const double *xs_[2] = {res, res}, *ys_[2] = {res, tk};
    xs_mdot_step(s, n, 2, xs_, ys_, P, IST_IBI_INIT, st, rv);
  }
  const double *l_vko[1] = {vk};
  const int s_vko[1] = {IS_ONE};
  const double *l_vk[4] = {tk, pk, vko, qk};
  const int s_vk[4] = {IS_ONE, IS_NOMEGA, IS_BETA, IS_NDELTA};
  const double *l_sk[2] = {res, vk};
  const int s_sk[2] = {IS_ONE, IS_NALPHA};
  const double *l_tk[3] = {tk, pk, qk};
  const int s_tk[3] = {IS_ONE, IS_NOMEGA, IS_NALPHA};
  const double *l_zk[3] = {zk, res, vko};
  const int s_zk[3] = {IS_T1, IS_ALPHA, IS_NT2};
  const double *x6[6] = {reso, reso, sk, tk, reso, reso}, *y6[6] = {sk, qk, tk, tk, tk, pk};
  const double *l_res[2] = {sk, tk};
  const int s_res[2] = {IS_ONE, IS_NOMEGA};
  const double *l_phi[3] = {phi, zk, sk};
  const int s_phi[3] = {IS_ONE, IS_ONE, IS_OMEGA};
  const double *x_rr[1] = {res}, *y_rr[1] = {res};
  rc = drive(c, [&]() {
    xs_istep(s, IST_IBI_S1, st, rv);
    xs_lincomb(s, n, 1, l_vko, s_vko, vko, st);                    // vko = vk
    xs_lincomb(s, n, 4, l_vk, s_vk, vk, st);                       // vk = t - w p + b vko - d q
    xs_op_pc(c, vk, tm, t, st);
    xs_op_spmv(c, tm, qk, EPI_Y, nullptr, nullptr, nullptr);       // qk = A M^-1 vk
    xs_lincomb(s, n, 2, l_sk, s_sk, sk, st);                       // sk = res - alpha vk
    xs_lincomb(s, n, 3, l_tk, s_tk, tk, st);                       // tk -= w pk + alpha qk
    xs_lincomb(s, n, 3, l_zk, s_zk, zk, st);                       // zk = t1 zk + a res - t2 vko
    xs_op_pc(c, tk, tm, t, st);
    xs_op_spmv(c, tm, pk, EPI_Y, nullptr, nullptr, nullptr);       // pk = A M^-1 tk
    xs_mdot_step(s, n, 6, x6, y6, P, IST_IBI_S2, st, rv);          // the six inner products
    xs_lincomb(s, n, 2, l_res, s_res, res, st);                    // res = sk - w tk
    xs_lincomb(s, n, 3, l_phi, s_phi, phi, st);                    // y += zk + w sk
    xs_mdot_step(s, n, 1, x_rr, y_rr, P, IST_IBI_RES, st, rv);
  });
  // x = x0 + M^-1 y (the reference's final preconditioner application, :656)
  xs_op_pc(c, phi, tm, t, nullptr);
  xs_add(s, n, tm, x);
  XS_CHK(hipGetLastError());
  return rc;
}

void xs_solvers_free(xs_ctx *c) {
  xs_multi_free(c);
  if (c->mdot_partial) (void)hipFree(c->mdot_partial);
  c->mdot_partial = nullptr;
  c->mdot_len = 0;
  if (c->basis) (void)hipFree(c->basis);
  c->basis = nullptr;
  c->basis_len = 0;
  if (c->small_d) (void)hipFree(c->small_d);
  c->small_d = nullptr;
  c->small_len = 0;
  if (c->ts_partial) (void)hipFree(c->ts_partial);
  c->ts_partial = nullptr;
  c->ts_partial_len = 0;
}

// ---------------------------------------------------------------------------
// Mixed precision: FP64 iterative refinement around an FP32 inner solve.
//   r = b - A x (FP64 SpMV, FP64 norm: the convergence-deciding quantity)
//   repeat: d ~= A^{-1} (r/||r||) by the FP32 inner Krylov method with FP32
//           SpMV and FP32 triangular solves over FP32/BF16/FP16 factors;
//           x += ||r|| d (FP64); r = b - A x (FP64)
// until ||r||/||r0|| < tol.  Iterations reported = inner iterations in total;
// resvec holds the FP64 true residual after each refinement step.
// ---------------------------------------------------------------------------
template <class T>
static int ensure_buf(T **p, size_t need) {
  if (*p) return XS_OK;
  if (hipMalloc((void **)p, sizeof(T) * (need ? need : 1)) != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return XS_ERR_MEMORY;
  }
  return XS_OK;
}

void xs_mixed_free(xs_ctx *c) {
  auto f = [](auto *&p) { if (p) (void)hipFree((void *)p); p = nullptr; };
  f(c->A32); f(c->ell32); f(c->dinv32); f(c->jac32); f(c->F32); f(c->work32); f(c->st_in);
  f(c->dia32);
  f(c->basis32);
  c->basis32_len = 0;
  c->mixed_mat_ready = false;
  c->work32_n = 0;
  c->F32_bytes = 0;
  c->mixed_ready = false;
}

// FP32 / BF16 / FP16 copies of A, the factors and the sweep copies: made
// once and kept until the matrix values, the preconditioner, the factor
// storage type or the SpMV format change (each clears mixed_ready).
static int mixed_prepare(xs_ctx *c) {
  if (c->mixed_ready) return XS_OK;
  hipStream_t s = c->stream;
  const int n = c->A.n;
  int rc;
  // the FP32 copy of A that op_spmv32 reads (diagonal slots, ELL or CSR);
  // a refactorisation leaves it as it is
  if (!c->mixed_mat_ready) {
    if (c->use_ell && c->ell.v && c->ell.dv) {
      if ((rc = ensure_buf(&c->dia32, (size_t)c->ell.nd * c->ell.ld))) return rc;
      xs_d2f(s, c->ell.nd * c->ell.ld, c->ell.dv, c->dia32);
    } else if (c->use_ell && c->ell.v) {
      if ((rc = ensure_buf(&c->ell32, (size_t)c->ell.W * c->ell.ld))) return rc;
      xs_d2f(s, c->ell.W * c->ell.ld, c->ell.v, c->ell32);
    } else {
      if ((rc = ensure_buf(&c->A32, (size_t)c->A.nnz))) return rc;
      xs_d2f(s, c->A.nnz, c->A.v, c->A32);
    }
    c->mixed_mat_ready = true;
  }
  if (c->pc == XS_PC_JACOBI) {
    if ((rc = ensure_buf(&c->jac32, (size_t)n))) return rc;
    xs_vec_d2f(s, n, c->jac_dinv, c->jac32);
  } else if (xs_pc_factored(c->pc)) {
    // the low-precision buffers grow only: a refactorisation reuses them
    auto grow = [](void *&buf, size_t &have, size_t need) -> int {
      if (buf && have >= need) return XS_OK;
      if (buf) (void)hipFree(buf);
      buf = nullptr;
      have = 0;
      if (hipMalloc(&buf, need ? need : 1) != hipSuccess) { (void)hipGetLastError(); return XS_ERR_MEMORY; }
      have = need;
      return XS_OK;
    };
    const size_t es = c->factor_type == XS_F32 ? 4 : 2;
    // the CSR factor in low precision: read by the exact triangular solves
    // and by sweeps that do not run over the sweep copies (fewer than two
    // sweeps, no copies yet); the sweep counts clear mixed_ready
    if (!(c->tell_ready && c->lsweeps >= 2 && c->usweeps >= 2)) {
      if ((rc = grow(c->F32, c->F32_bytes, es * (size_t)c->ilu.nnz))) return rc;
      xs_cvt_factor(s, c->ilu.nnz, c->ilu.lu, c->F32, c->factor_type);
    }
    if ((rc = ensure_buf(&c->dinv32, (size_t)n))) return rc;
    xs_vec_d2f(s, n, c->ilu.dinv, c->dinv32);
    c->tell32_ready = false;
    if (c->tell_ready) {   // sweep copies in the factor storage type
      for (xs_tell *E : {&c->ilu.tl, &c->ilu.tu}) {
        if (E->W == 0) continue;
        if ((rc = grow(E->vlo, E->vlo_bytes, es * E->cap))) return rc;
        if (E->nd > 0 && E->nd <= XS_DIA_MAX && E->dv) {
          // the sweeps read the diagonal-slot copy only
          if ((rc = grow(E->dvlo, E->dvlo_bytes, es * (size_t)E->nd * E->ld))) return rc;
          xs_cvt_factor(s, E->nd * E->ld, E->dv, E->dvlo, c->factor_type);
        } else {
          xs_cvt_factor(s, E->W * E->ld, E->v, E->vlo, c->factor_type);
        }
      }
      c->tell32_ready = true;
    }
  }
  if (c->work32_n != (size_t)n) {
    if (c->work32) (void)hipFree(c->work32);
    c->work32 = nullptr;
    if ((rc = ensure_buf(&c->work32, (size_t)13 * n))) return rc;
    c->work32_n = n;
  }
  if ((rc = ensure_buf(&c->st_in, 1))) return rc;
  c->mixed_ready = true;
  return XS_OK;
}

static void op_spmv32(xs_ctx *c, const float *x, float *y, int epi, const float *b,
                      const float *w, const xs_reduce *red) {
  xs_prof_begin(c, 1);
  if (c->use_ell && c->ell.dv && c->dia32)
    xs_launch_spmv_dia32(c->stream, c->ell, c->dia32, x, y, epi, b, w, red);
  else if (c->use_ell && c->ell32)
    xs_launch_spmv_ell32(c->stream, c->ell, c->ell32, x, y, epi, b, w, red);
  else
    xs_launch_spmv_csr32(c->stream, c->A, c->A32, c->plan, x, y, epi, b, w, red);
  xs_prof_end(c, 1);
}

// FP32 counterpart of the host's ilu_sweeps (k Jacobi sweeps of one factor);
// k >= 2 over the FP32 ELL copies when they are current (sweeps 1 and 2
// fused).  With ru (L, k >= 2, ELL copies current): b is r, and the first
// pass forms s = r - gama u into sv, which the later passes read as b.
static void ilu_sweeps32(xs_ctx *c, bool upper, int k, const float *b, float *out, float *w,
                         const xs_state *st, const float *ru = nullptr, float *sv = nullptr) {
  const int n = c->A.n, m = k < 1 ? 1 : k;
  if (m >= 2 && c->tell32_ready) {
    const xs_tell &E = upper ? c->ilu.tu : c->ilu.tl;
    float *cur = ((m - 2) % 2 == 0) ? out : w;
    if (ru) {
      xs_tri_sweep_ell32_s(c->stream, n, E, c->factor_type, b, ru, sv, cur, st);
      b = sv;
    } else {
      xs_tri_sweep_ell32(c->stream, upper, true, n, E, c->factor_type, c->dinv32, b, nullptr, cur,
                         st);
    }
    for (int j = 3; j <= m; ++j) {
      float *nxt = (cur == out) ? w : out;
      xs_tri_sweep_ell32(c->stream, upper, false, n, E, c->factor_type, c->dinv32, b, cur, nxt, st);
      cur = nxt;
    }
    return;
  }
  float *cur = ((m - 1) % 2 == 0) ? out : w;
  xs_tri_sweep32(c->stream, upper, n, c->ilu.rp, c->ilu.ci, c->F32, c->factor_type, c->ilu.dpos,
                 c->dinv32, b, nullptr, cur, st);
  for (int j = 2; j <= m; ++j) {
    float *nxt = (cur == out) ? w : out;
    xs_tri_sweep32(c->stream, upper, n, c->ilu.rp, c->ilu.ci, c->F32, c->factor_type,
                   c->ilu.dpos, c->dinv32, b, cur, nxt, st);
    cur = nxt;
  }
}

static void op_pc32(xs_ctx *c, const float *x, float *y, float *tmp, const xs_state *st) {
  const int n = c->A.n;
  if (xs_pc_factored(c->pc)) {
    float *w = c->work32 + 11 * (size_t)n;   // sweep scratch
    xs_prof_begin(c, 2);
    if (c->lsweeps >= 0)
      ilu_sweeps32(c, false, c->lsweeps, x, tmp, w, st);
    else
      xs_trsv32(c->stream, false, n, c->ilu.rp, c->ilu.ci, c->F32, c->factor_type, c->ilu.dpos,
                c->dinv32, x, tmp, c->counters + 2, xs_ilu_tsched(c->ilu, false), st);
    xs_prof_end(c, 2);
    xs_prof_begin(c, 3);
    if (c->usweeps >= 0)
      ilu_sweeps32(c, true, c->usweeps, tmp, y, w, st);
    else
      xs_trsv32(c->stream, true, n, c->ilu.rp, c->ilu.ci, c->F32, c->factor_type, c->ilu.dpos,
                c->dinv32, tmp, y, c->counters + 4, xs_ilu_tsched(c->ilu, true), st);
    xs_prof_end(c, 3);
  } else if (c->pc == XS_PC_JACOBI) {
    xs_scale32(c->stream, n, c->jac32, x, y);
  } else {
    xs_copy32(c->stream, n, x, y);
  }
}

// BiCGSTAB's s = r - gama u, then y = M^{-1} s: s formed inside the first L
// pass when the ILU is applied by two or more L sweeps over the FP32 ELL
// copies, else by its own kernel first
static void op_pc32_s(xs_ctx *c, const float *r, const float *u, float *sv, float *y, float *tmp,
                      xs_state *st) {
  const int n = c->A.n;
  if (!(xs_pc_factored(c->pc) && c->lsweeps >= 2 && c->tell32_ready)) {
    xs_bicg_s32(c->stream, n, r, u, sv, nullptr, nullptr, st);
    op_pc32(c, sv, y, tmp, st);
    return;
  }
  float *w = c->work32 + 11 * (size_t)n;   // sweep scratch
  xs_prof_begin(c, 2);
  ilu_sweeps32(c, false, c->lsweeps, r, tmp, w, st, u, sv);
  xs_prof_end(c, 2);
  xs_prof_begin(c, 3);
  if (c->usweeps >= 0)
    ilu_sweeps32(c, true, c->usweeps, tmp, y, w, st);
  else
    xs_trsv32(c->stream, true, n, c->ilu.rp, c->ilu.ci, c->F32, c->factor_type, c->ilu.dpos,
              c->dinv32, tmp, y, c->counters + 4, xs_ilu_tsched(c->ilu, true), st);
  xs_prof_end(c, 3);
}

// BiCGSTAB's p = r + beta (pin - omega u) formed inside the first L pass over
// the FP32 diagonal-slot copy and written to pout, then y = M^{-1} p
static bool pc32_fuse_p(const xs_ctx *c) {
  return xs_pc_factored(c->pc) && c->fuse_p && c->lsweeps >= 2 && c->tell32_ready &&
         xs_tell32_dia(c->ilu.tl);
}
static void op_pc32_p(xs_ctx *c, const float *r, const float *u, const float *pin, float *pout,
                      float *y, float *tmp, xs_state *st) {
  const int n = c->A.n, m = c->lsweeps;
  float *w = c->work32 + 11 * (size_t)n;   // sweep scratch
  xs_prof_begin(c, 2);
  float *cur = ((m - 2) % 2 == 0) ? tmp : w;
  xs_tri_sweep_ell32_p(c->stream, n, c->ilu.tl, c->factor_type, r, u, pin, pout, cur, st);
  for (int j = 3; j <= m; ++j) {
    float *nxt = (cur == tmp) ? w : tmp;
    xs_tri_sweep_ell32(c->stream, false, false, n, c->ilu.tl, c->factor_type, c->dinv32, pout, cur,
                       nxt, st);
    cur = nxt;
  }
  xs_prof_end(c, 2);
  xs_prof_begin(c, 3);
  if (c->usweeps >= 0)
    ilu_sweeps32(c, true, c->usweeps, tmp, y, w, st);
  else
    xs_trsv32(c->stream, true, n, c->ilu.rp, c->ilu.ci, c->F32, c->factor_type, c->ilu.dpos,
              c->dinv32, tmp, y, c->counters + 4, xs_ilu_tsched(c->ilu, true), st);
  xs_prof_end(c, 3);
}

// Relative reduction asked of an inner FP32 solve.  By default sqrt(rtol):
// two refinement steps reach rtol when each inner solve meets its target,
// and 1e-6 is about the reduction FP32 Krylov iterations still deliver.
static double inner_tol_of(const xs_ctx *c) {
  if (c->inner_tol > 0.0) return c->inner_tol;
  return std::min(0.5, std::max(std::sqrt(c->tol), 1e-6));
}

// FP32 BiCGSTAB on A d = rhs from d = 0 (inner solve); returns iterations.
static int inner_bicgstab32(xs_ctx *c, const float *rhs, float *d, int maxit, int *iters) {
  *iters = 0;
  hipStream_t s = c->stream;
  const int n = c->A.n;
  float *W = c->work32;
  float *r = W, *rh = W + (size_t)n, *p = W + 2 * (size_t)n, *u = W + 3 * (size_t)n;
  float *sv = W + 4 * (size_t)n, *v = W + 5 * (size_t)n, *z = W + 6 * (size_t)n;
  float *z1 = W + 7 * (size_t)n, *t = W + 8 * (size_t)n;
  xs_state *st = c->st_in;
  const float *jac = (c->pc == XS_PC_JACOBI) ? c->jac32 : nullptr;
  xs_init_state(s, st, inner_tol_of(c), 0.0, maxit);
  xs_zero32(s, n, d);
  xs_zero32(s, n, p);
  xs_zero32(s, n, u);
  xs_reduce r0{}; r0.partial = c->partial; r0.counter = c->counters; r0.st = st; r0.op = F_RES0;
  op_spmv32(c, d, r, EPI_BMAXN, rhs, nullptr, &r0);
  xs_copy32(s, n, r, rh);
  xs_reduce rg = r0, ra = r0, rr = r0;
  rg.op = F_BICG_GAMA; ra.op = F_BICG_ALPHA; rr.op = F_BICG_RES;
  int launched = 0;
  xs_state &h = c->h_state[1];   // pinned
  XS_CHK(hipMemcpyAsync(&h, st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
  XS_CHK(hipStreamSynchronize(s));
  auto iteration = [&]() {
    xs_bicg_p32(s, n, r, u, p, jac, z, st);
    if (!jac) op_pc32(c, p, z, t, st);
    op_spmv32(c, z, u, EPI_DOT1, nullptr, rh, &rg);
    if (jac) xs_bicg_s32(s, n, r, u, sv, jac, z1, st);
    else op_pc32_s(c, r, u, sv, z1, t, st);
    op_spmv32(c, z1, v, EPI_DOT2, nullptr, sv, &ra);
    xs_bicg_xr32(s, n, d, r, z, z1, sv, v, rh, rr);
  };
  xs_graph_cache gc;
  while (!h.done && launched < maxit) {
    const int k = std::min(c->check_every, maxit - launched);
    XS_CHK(xs_run_batch(c, gc, k, iteration) == 0 ? hipSuccess : hipErrorUnknown);
    launched += k;
    XS_CHK(hipGetLastError());
    XS_CHK(hipMemcpyAsync(&h, st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
    XS_CHK(hipStreamSynchronize(s));
  }
  *iters = h.iter;
  return h.done == 2 ? XS_ERR_BREAKDOWN : XS_OK;
}

// FP32 flexible GMRES(m) on A d = rhs from d = 0 (inner solve of the
// mixed-precision refinement): FP32 basis, SpMV and triangular solves; FP64
// inner products, Hessenberg, Givens rotations and residual estimate.
static int inner_fgmres32(xs_ctx *c, const float *rhs, float *d, int maxit, int *iters) {
  *iters = 0;
  hipStream_t s = c->stream;
  const int n = c->A.n;
  const int m = c->restart > 0 ? c->restart : 30;
  const size_t ld = (size_t)n;
  const size_t need = (size_t)(2 * m + 1) * ld;
  if (c->basis32 && c->basis32_len < need) { (void)hipFree(c->basis32); c->basis32 = nullptr; }
  if (!c->basis32) {
    if (hipMalloc((void **)&c->basis32, sizeof(float) * need) != hipSuccess) {
      (void)hipGetLastError();
      return XS_ERR_MEMORY;
    }
    c->basis32_len = need;
  }
  int rc;
  if ((rc = ensure_dev(&c->small_d, &c->small_len, small_size(m)))) return rc;
  if ((rc = ensure_ts(c))) return rc;
  float *V = c->basis32, *Z = c->basis32 + (size_t)(m + 1) * ld;
  xs_gmres_small g = carve_small(c->small_d, m);
  float *W = c->work32;
  float *r = W, *tmp = W + (size_t)n, *w = W + 2 * (size_t)n;
  unsigned *tc = c->counters + 10;
  double *P = c->ts_partial;
  xs_state *st = c->st_in;
  xs_init_state(s, st, inner_tol_of(c), 0.0, maxit);
  xs_zero32(s, n, d);
  xs_copy32(s, n, rhs, r);
  xs_ts_dot<float>(s, n, 1, r, ld, r, P, tc, g.misc + XS_G_NRM2, nullptr);
  xs_gmres_cycle(s, m, 1, g, st, nullptr);
  xs_ts_scale<float>(s, n, r, V, g.misc + XS_G_SCALE, st);
  int launched = 0, i = 0, cycle_start = 0;
  xs_state &h = c->h_state[1];   // pinned
  XS_CHK(hipMemcpyAsync(&h, st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
  XS_CHK(hipStreamSynchronize(s));
  bool done = h.done != 0;
  while (!done && launched < maxit) {
    float *vi = V + (size_t)i * ld, *vn = V + (size_t)(i + 1) * ld, *zi = Z + (size_t)i * ld;
    op_pc32(c, vi, zi, tmp, st);
    op_spmv32(c, zi, w, EPI_Y, nullptr, nullptr, nullptr);
    xs_ts_dot<float>(s, n, i + 1, V, ld, w, P, tc, g.h1, st);
    xs_ts_update<float>(s, n, i + 1, V, nullptr, ld, g.h1, w, w, nullptr, st);
    xs_ts_dot<float>(s, n, i + 1, V, ld, w, P, tc, g.h2, st);
    xs_ts_update<float>(s, n, i + 1, V, nullptr, ld, g.h2, w, vn, nullptr, st);
    xs_ts_dot<float>(s, n, 1, vn, ld, vn, P, tc, g.misc + XS_G_NRM2, st);
    xs_gmres_column(s, i, m, g, st, nullptr);
    xs_ts_scale<float>(s, n, vn, vn, g.misc + XS_G_SCALE, st);
    i++;
    launched++;
    if (i == m || launched == maxit || launched % c->check_every == 0) {
      XS_CHK(hipGetLastError());
      XS_CHK(hipMemcpyAsync(&h, st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
      XS_CHK(hipStreamSynchronize(s));
      done = h.done != 0;
      if (done) break;
    }
    if (i == m && launched < maxit) {
      xs_gmres_backsub(s, m, m, g);
      xs_ts_axpy<float>(s, n, m, Z, ld, g.y, d);
      op_spmv32(c, d, r, EPI_BMAX, rhs, nullptr, nullptr);
      xs_gmres_cycle(s, m, 0, g, st, nullptr);
      xs_ts_scale<float>(s, n, r, V, g.misc + XS_G_SCALE, st);
      i = 0;
      cycle_start = launched;
    }
  }
  const int k = (h.done == 1) ? h.iter - cycle_start : i;
  if (h.done != 2 && k > 0) {
    xs_gmres_backsub(s, k, m, g);
    xs_ts_axpy<float>(s, n, k, Z, ld, g.y, d);
  }
  XS_CHK(hipGetLastError());
  *iters = h.iter;
  return h.done == 2 ? XS_ERR_BREAKDOWN : XS_OK;
}

// FP32 BiCGSTAB with reliable updates (Sleijpen & van der Vorst, "Reliable
// updated residuals in hybrid Bi-CG methods", Computing 56, 1996): one Krylov
// recurrence for the whole solve, in FP32 on the system scaled by sigma =
// ||b - A x0||; every check_every iterations the FP32 correction d is added
// into the FP64 iterate (x += sigma d), the true residual b - A x is formed
// in FP64 (the convergence-deciding norm) and replaces the recursive FP32
// residual, and beta' is recomputed from it (F_BICG_REPL).  Unlike the
// restarted refinement, the Krylov space is never discarded.
static int mixed_bicgstab_reliable(xs_ctx *c, const double *b, double *x) {
  hipStream_t s = c->stream;
  const int n = c->A.n;
  double *r64 = xs_ctx_vec(c, 0);
  float *W = c->work32;
  float *r = W, *rh = W + (size_t)n, *p = W + 2 * (size_t)n, *u = W + 3 * (size_t)n;
  float *sv = W + 4 * (size_t)n, *v = W + 5 * (size_t)n, *z = W + 6 * (size_t)n;
  float *z1 = W + 7 * (size_t)n, *t = W + 8 * (size_t)n;
  float *rhs = W + 9 * (size_t)n, *d = W + 10 * (size_t)n;
  xs_state *st = c->st, *si = c->st_in;
  const float *jac = (c->pc == XS_PC_JACOBI) ? c->jac32 : nullptr;
  xs_init_state(s, st, c->tol, c->abstol, c->maxiter);
  const int SIG = S_SCRATCH0 + 1, TRES = S_SCRATCH0 + 2;
  const xs_reduce rs = xs_ctx_reduce(c, F_STORE, SIG);   // sigma^2 = ||b - A x0||^2
  const xs_reduce rt = xs_ctx_reduce(c, F_STORE, TRES);  // ||b - A x||^2
  xs_op_spmv(c, x, r64, EPI_BMAXN, b, nullptr, &rs);
  xs_state &h = c->h_state[0], &hi = c->h_state[1];   // pinned
  XS_CHK(hipMemcpyAsync(&h, st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
  XS_CHK(hipStreamSynchronize(s));
  const double res0 = std::sqrt(h.sc[SIG]);
  std::vector<double> hist(1, res0);
  double res = res0;
  int rc = XS_NOT_CONVERGED, outer = 0;
  c->iters = 0;
  if (!std::isfinite(res0)) rc = XS_ERR_BREAKDOWN;
  else if (res0 == 0.0 || res0 <= c->abstol) rc = XS_OK;
  if (rc == XS_NOT_CONVERGED) {
    // scaled start: r = (b - A x0) / sigma, rhat = r, d = p = u = 0
    xs_d2f_scaled(s, n, r64, rhs, st, SIG);
    xs_init_state(s, si, c->tol, 0.0, c->maxiter);
    xs_zero32(s, n, d);
    xs_zero32(s, n, p);
    xs_zero32(s, n, u);
    xs_reduce r0{}; r0.partial = c->partial; r0.counter = c->counters; r0.st = si; r0.op = F_RES0;
    op_spmv32(c, d, r, EPI_BMAXN, rhs, nullptr, &r0);
    xs_copy32(s, n, r, rh);
    xs_reduce rg = r0, ra = r0, rr = r0, rp = r0;
    rg.op = F_BICG_GAMA; ra.op = F_BICG_ALPHA; rr.op = F_BICG_RES; rp.op = F_BICG_REPL;
    // p formed inside the first L pass, alternating between p and p2 as in
    // the FP64 loop (an even batch ends on the parity it started with)
    const bool pfuse = !jac && pc32_fuse_p(c) && c->check_every % 2 == 0;
    float *p2 = W + 12 * (size_t)n;
    bool odd = false;
    auto iteration = [&]() {
      if (pfuse) {
        op_pc32_p(c, r, u, odd ? p2 : p, odd ? p : p2, z, t, si);
        odd = !odd;
      } else {
        xs_bicg_p32(s, n, r, u, p, jac, z, si);
        if (!jac) op_pc32(c, p, z, t, si);
      }
      op_spmv32(c, z, u, EPI_DOT1, nullptr, rh, &rg);
      if (jac) xs_bicg_s32(s, n, r, u, sv, jac, z1, si);
      else op_pc32_s(c, r, u, sv, z1, t, si);
      op_spmv32(c, z1, v, EPI_DOT2, nullptr, sv, &ra);
      xs_bicg_xr32(s, n, d, r, z, z1, sv, v, rh, rr);
    };
    xs_graph_cache gc;
    int launched = 0;
    for (;;) {
      const int k = std::min(c->check_every, c->maxiter - launched);
      XS_CHK(xs_run_batch(c, gc, k, iteration) == 0 ? hipSuccess : hipErrorUnknown);
      launched += k;
      // reliable update: x += sigma d, true residual in FP64, replacement
      xs_flush_f2d(s, n, d, x, st, SIG);
      xs_op_spmv(c, x, r64, EPI_BMAXN, b, nullptr, &rt);
      xs_replace32(s, n, r64, r, rh, st, SIG, rp);
      outer++;
      XS_CHK(hipGetLastError());
      XS_CHK(hipMemcpyAsync(&h, st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
      XS_CHK(hipMemcpyAsync(&hi, si, sizeof(xs_state), hipMemcpyDeviceToHost, s));
      XS_CHK(hipStreamSynchronize(s));
      res = std::sqrt(h.sc[TRES]);
      hist.push_back(res);
      c->iters = hi.iter;
      if (!std::isfinite(res) || hi.done == 2) { rc = XS_ERR_BREAKDOWN; break; }
      if (res == 0.0 || res / res0 < c->tol || res <= c->abstol) { rc = XS_OK; break; }
      if (hi.iter >= c->maxiter || launched >= c->maxiter) break;
    }
  }
  XS_CHK(hipGetLastError());
  c->outer_iters = outer;
  c->res0 = res0;
  c->resl = res;
  const int m = std::min<int>((int)hist.size(), c->resvec_cap);
  XS_CHK(hipMemcpy(c->resvec_d, hist.data(), sizeof(double) * m, hipMemcpyHostToDevice));
  return rc;
}

int xs_solve_mixed(xs_ctx *c, const double *b, double *x) {
  if (c->ksp != XS_KSP_BICGSTAB && c->ksp != XS_KSP_GMRES && c->ksp != XS_KSP_FGMRES)
    return XS_ERR_UNSUPPORTED;
  int rc = xs_ctx_ensure_work(c, XS_NWORK);
  if (rc) return rc;
  if ((rc = mixed_prepare(c))) return rc;
  if (c->ksp == XS_KSP_BICGSTAB && c->mixed_reliable) return mixed_bicgstab_reliable(c, b, x);
  hipStream_t s = c->stream;
  const int n = c->A.n;
  double *r64 = xs_ctx_vec(c, 0);
  float *rhs32 = c->work32 + 9 * (size_t)n, *d32 = c->work32 + 10 * (size_t)n;
  xs_state *st = c->st;
  xs_init_state(s, st, c->tol, c->abstol, c->maxiter);
  xs_reduce rn = xs_ctx_reduce(c, F_STORE, S_SCRATCH0 + 1);   // ||r||^2 -> sc[S_SCRATCH0+1]
  std::vector<double> hist;
  double res0 = 0.0, res = 0.0;
  int total = 0, outer = 0;
  rc = XS_NOT_CONVERGED;
  for (;;) {
    xs_op_spmv(c, x, r64, EPI_BMAXN, b, nullptr, &rn);           // FP64 true residual
    xs_state &h = c->h_state[0];   // pinned (the inner solves use [1])
    XS_CHK(hipMemcpyAsync(&h, st, sizeof(xs_state), hipMemcpyDeviceToHost, s));
    XS_CHK(hipStreamSynchronize(s));
    res = std::sqrt(h.sc[S_SCRATCH0 + 1]);
    if (outer == 0) res0 = res;
    hist.push_back(res);
    if (!std::isfinite(res)) { rc = XS_ERR_BREAKDOWN; break; }
    if (res == 0.0 || res / res0 < c->tol || res <= c->abstol) { rc = XS_OK; break; }
    if (total >= c->maxiter) break;
    xs_d2f_scaled(s, n, r64, rhs32, st, S_SCRATCH0 + 1);
    int it = 0;
    int irc = (c->ksp == XS_KSP_BICGSTAB)
                  ? inner_bicgstab32(c, rhs32, d32, c->maxiter - total, &it)
                  : inner_fgmres32(c, rhs32, d32, c->maxiter - total, &it);
    total += it;
    // an inner failure other than a breakdown (memory, device) leaves d32
    // undefined: stop before it reaches x; a breakdown before the first
    // inner iteration, or an inner solve that made no iteration, would
    // repeat forever
    if (irc < 0 && irc != XS_ERR_BREAKDOWN) { rc = irc; break; }
    if (it == 0) { rc = irc == XS_ERR_BREAKDOWN ? XS_ERR_BREAKDOWN : XS_NOT_CONVERGED; break; }
    xs_add_f2d(s, n, d32, x, st, S_SCRATCH0 + 1);
    outer++;
  }
  XS_CHK(hipGetLastError());
  c->iters = total;
  c->outer_iters = outer;
  c->res0 = res0;
  c->resl = res;
  // residual history on the host side of the context (resvec_d holds room)
  const int m = std::min<int>((int)hist.size(), c->resvec_cap);
  XS_CHK(hipMemcpy(c->resvec_d, hist.data(), sizeof(double) * m, hipMemcpyHostToDevice));
  return rc;
}
