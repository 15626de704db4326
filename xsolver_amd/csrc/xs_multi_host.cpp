// xs_multi_host.cpp -- host driver of the k right-hand side block solve.
//
// The k columns run the same preconditioned Krylov recurrences as the single
// right-hand side drivers (BiCGSTAB: hip/Solver/xsolver_bicgstab.cpp:20-124,
// CG: hip/Solver/xsolver_cg.cpp:16-74), each with its own scalars and
// convergence flag, over one SpMM and one k-column triangular solve per
// operator application instead of k SpMVs / k solves.  The host reads the
// per-column flags every `check_every` iterations and stops once every column
// is done (or maxiter iterations were launched).
// No reference counterpart (the reference solves one right-hand side).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/xsolver_c.h"
#include "xs_multi.h"
#include "xs_solvers.h"

#define XS_CHK(call)                                \
  do {                                              \
    if ((call) != hipSuccess) return XS_ERR_DEVICE; \
  } while (0)

static xs_mstate *MS(xs_ctx *c) { return static_cast<xs_mstate *>(c->ms); }
static xs_mstate *HMS(xs_ctx *c) { return static_cast<xs_mstate *>(c->h_ms); }

// Work vectors of n*k doubles: the Krylov vectors (9) plus `extra_vecs`.
int xs_multi_reserve(xs_ctx *c, int k, int extra_vecs) {
  const size_t need = (size_t)(9 + extra_vecs) * (size_t)c->A.n * (size_t)k;
  if (c->mwork_len < need) {
    if (c->mwork) (void)hipFree(c->mwork);
    c->mwork = nullptr;
    c->mwork_len = 0;
    XS_CHK(hipMalloc((void **)&c->mwork, sizeof(double) * need));
    c->mwork_len = need;
  }
  if (!c->ms) XS_CHK(hipMalloc(&c->ms, sizeof(xs_mstate)));
  if (!c->h_ms) XS_CHK(hipHostMalloc(&c->h_ms, sizeof(xs_mstate), hipHostMallocDefault));
  if (!c->mpartial) {
    XS_CHK(hipMalloc((void **)&c->mpartial, sizeof(double) * XS_MGRID * 2 * XS_MAX_RHS));
    // reduction slots hold the sentinel between reductions
    xs_fill_sentinel(c->stream, c->mpartial, XS_MGRID * 2 * XS_MAX_RHS);
  }
  const size_t rv = (size_t)(c->maxiter + 1) * (size_t)k;
  if (c->mresvec_len < rv) {
    if (c->mresvec) (void)hipFree(c->mresvec);
    c->mresvec = nullptr;
    c->mresvec_len = 0;
    XS_CHK(hipMalloc((void **)&c->mresvec, sizeof(double) * rv));
    c->mresvec_len = rv;
  }
  return XS_OK;
}

// k-column SpMM over the ELL copy of A when the product uses it (even k,
// aligned blocks), else over the CSR
static void mspmm(xs_ctx *c, hipStream_t s, int k, const double *X, double *Y, int epi,
                  const double *B, const double *W, const xs_mred *red, const xs_mstate *ms) {
  const void *vecs[4] = {X, Y, B, W};
  if (c->use_ell && xs_mspmm_ell_ok(c->ell, k, vecs, 4))
    xs_mspmm_ell(s, c->ell, k, X, Y, epi, B, W, red, ms, c->spmm_cw);
  else
    xs_mspmm(s, c->A, k, X, Y, epi, B, W, red, ms);
}

double *xs_multi_vec(xs_ctx *c, int k, int idx) {
  return c->mwork + (size_t)idx * (size_t)c->A.n * (size_t)k;
}

void xs_multi_free(xs_ctx *c) {
  if (c->mwork) (void)hipFree(c->mwork);
  if (c->ms) (void)hipFree(c->ms);
  if (c->h_ms) (void)hipHostFree(c->h_ms);
  if (c->mpartial) (void)hipFree(c->mpartial);
  if (c->mresvec) (void)hipFree(c->mresvec);
  c->mwork = nullptr;
  c->mwork_len = 0;
  c->ms = c->h_ms = nullptr;
  c->mpartial = c->mresvec = nullptr;
  c->mresvec_len = 0;
}

// Jacobi sweeps of one factor part for k columns: m <= 1 is the first sweep
// from zero (b, or D b), m >= 2 runs over the ELL copy (sweeps 1 and 2 in one
// pass), ping-ponging with w so that the last pass lands in out.
static void msweeps(xs_ctx *c, hipStream_t s, int k, bool upper, int m, const double *b,
                    double *out, double *w, const xs_mstate *ms) {
  const int n = c->A.n;
  if (m <= 1) {
    if (upper) xs_mscale(s, n, k, c->ilu.dinv, b, out);
    else xs_mcopy(s, (size_t)n * k, b, out);
    return;
  }
  if (!c->tell_ready) { xs_raise_launch_guard(); return; }
  const xs_tell &E = upper ? c->ilu.tu : c->ilu.tl;
  double *cur = ((m - 2) % 2 == 0) ? out : w;
  xs_msweep_ell(s, E, n, k, upper, true, c->ilu.dinv, b, nullptr, cur, ms);
  for (int j = 3; j <= m; ++j) {
    double *nxt = (cur == out) ? w : out;
    xs_msweep_ell(s, E, n, k, upper, false, c->ilu.dinv, b, cur, nxt, ms);
    cur = nxt;
  }
}

// k-column ILU apply: Jacobi sweeps per factor when lsolve_sweeps /
// usolve_sweeps >= 0, else the k-column sync-free triangular solves.
void xs_multi_ilu_apply(xs_ctx *c, hipStream_t s, int k, const double *in, double *out,
                        double *tmp, double *w, const xs_mstate *ms) {
  const int n = c->A.n;
  if (c->lsweeps >= 0)
    msweeps(c, s, k, false, c->lsweeps, in, tmp, w, ms);
  else
    xs_mtrsv(s, false, n, k, c->ilu.rp, c->ilu.ci, c->ilu.lu, c->ilu.dpos, nullptr, in, tmp,
             c->counters + 2, xs_ilu_tsched(c->ilu, false), ms);
  if (c->usweeps >= 0)
    msweeps(c, s, k, true, c->usweeps, tmp, out, w, ms);
  else
    xs_mtrsv(s, true, n, k, c->ilu.rp, c->ilu.ci, c->ilu.lu, c->ilu.dpos, c->ilu.dinv, tmp, out,
             c->counters + 4, xs_ilu_tsched(c->ilu, true), ms);
}

namespace {

struct multi_run {
  xs_ctx *c;
  int n, k;
  hipStream_t s;
  xs_mstate *ms;

  xs_mred red(int op) const {
    xs_mred r;
    r.partial = c->mpartial;
    r.counter = c->counters;   // the reduction ticket, self-resetting
    r.ms = ms;
    r.resvec = c->mresvec;
    r.tol = c->tol;
    r.abstol = c->abstol;
    r.op = op;
    return r;
  }
  // out = M^{-1} in for all k columns (tmp: an n*k scratch block)
  void pc(const double *in, double *out, double *tmp) const {
    switch (c->pc) {
      case XS_PC_JACOBI:
        xs_mscale(s, n, k, c->jac_dinv, in, out);
        break;
      case XS_PC_ILU0:
      case XS_PC_ILUT:
      case XS_PC_IC0:
        xs_multi_ilu_apply(c, s, k, in, out, tmp, xs_multi_vec(c, k, 11), ms);
        break;
      default:
        xs_mcopy(s, (size_t)n * k, in, out);
    }
  }
  // launch `iteration` until every column is done or maxiter were launched
  template <class F>
  int drive(F iteration) const {
    xs_mstate *h = HMS(c);
    XS_CHK(hipMemcpyAsync(h, ms, sizeof(xs_mstate), hipMemcpyDeviceToHost, s));
    XS_CHK(hipStreamSynchronize(s));
    int launched = 0;
    xs_graph_cache gc;
    while (!h->all && launched < c->maxiter) {
      const int step = std::min(c->check_every, c->maxiter - launched);
      if (xs_run_batch(c, gc, step, iteration, true)) return XS_ERR_DEVICE;
      launched += step;
      XS_CHK(hipGetLastError());
      XS_CHK(hipMemcpyAsync(h, ms, sizeof(xs_mstate), hipMemcpyDeviceToHost, s));
      XS_CHK(hipStreamSynchronize(s));
    }
    int worst = XS_OK, maxit = 0;
    c->m_iters.assign(h->iter, h->iter + k);
    c->m_res0.resize(k);
    c->m_resl.resize(k);
    for (int j = 0; j < k; ++j) {
      c->m_res0[j] = h->sc[MS_RES0][j];
      c->m_resl[j] = h->sc[MS_RESL][j];
      maxit = std::max(maxit, h->iter[j]);
      if (h->done[j] == 2) worst = XS_ERR_BREAKDOWN;
      else if (h->done[j] == 0 && worst == XS_OK) worst = XS_NOT_CONVERGED;
    }
    c->mk = k;
    c->m_resvec.resize((size_t)(maxit + 1) * k);
    XS_CHK(hipMemcpyAsync(c->m_resvec.data(), c->mresvec, sizeof(double) * c->m_resvec.size(),
                          hipMemcpyDeviceToHost, s));
    XS_CHK(hipStreamSynchronize(s));
    // a summary for the single-solve getters: the slowest column
    c->iters = maxit;
    c->res0 = c->m_res0.empty() ? 0.0 : c->m_res0[0];
    c->resl = c->m_resl.empty() ? 0.0 : *std::max_element(c->m_resl.begin(), c->m_resl.end());
    return worst;
  }
};

}  // namespace

int xs_solve_block(xs_ctx *c, int k, const double *B, double *X) {
  if (k < 1 || k > XS_MAX_RHS) return XS_ERR_ARG;
  if (c->precision != 0) return XS_ERR_UNSUPPORTED;
  if (c->ksp != XS_KSP_BICGSTAB && c->ksp != XS_KSP_CG) return XS_ERR_UNSUPPORTED;
  int rc = xs_multi_reserve(c, k, 3);   // vectors 9-10: column-major copies, 11: sweep scratch
  if (rc) return rc;
  multi_run m{c, c->A.n, k, c->stream, MS(c)};
  const int n = m.n;
  hipStream_t s = m.s;
  const size_t bytes = sizeof(double) * (size_t)n * k;
  const double *jac = (c->pc == XS_PC_JACOBI) ? c->jac_dinv : nullptr;
  xs_mstate_init(s, m.ms, k);
  if (c->ksp == XS_KSP_BICGSTAB) {
    double *R = xs_multi_vec(c, k, 0), *Rh = xs_multi_vec(c, k, 1), *P = xs_multi_vec(c, k, 2);
    double *U = xs_multi_vec(c, k, 3), *Sv = xs_multi_vec(c, k, 4), *V = xs_multi_vec(c, k, 5);
    double *Z = xs_multi_vec(c, k, 6), *Z1 = xs_multi_vec(c, k, 7), *T = xs_multi_vec(c, k, 8);
    XS_CHK(hipMemsetAsync(P, 0, bytes, s));
    XS_CHK(hipMemsetAsync(U, 0, bytes, s));
    const xs_mred r0 = m.red(MF_BICG_RES0), rg = m.red(MF_BICG_GAMA);
    const xs_mred ra = m.red(MF_BICG_ALPHA), rr = m.red(MF_BICG_RES);
    mspmm(c, s, k, X, R, EPI_BMAXN, B, nullptr, &r0, m.ms);      // R = B - A X, res0
    XS_CHK(hipMemcpyAsync(Rh, R, bytes, hipMemcpyDeviceToDevice, s));  // rhat = r0
    return m.drive([&]() {
      xs_mbicg_p(s, n, k, R, U, P, jac, Z, m.ms);                     // p = r + w (p - a u)
      if (!jac) m.pc(P, Z, T);                                         // z = M^-1 p
      mspmm(c, s, k, Z, U, EPI_DOT1, nullptr, Rh, &rg, m.ms);   // u = A z, gama
      xs_mbicg_s(s, n, k, R, U, Sv, jac, Z1, m.ms);                   // s = r - gama u
      if (!jac) m.pc(Sv, Z1, T);                                       // z1 = M^-1 s
      mspmm(c, s, k, Z1, V, EPI_DOT2, nullptr, Sv, &ra, m.ms);  // v = A z1, alpha
      xs_mbicg_xr(s, n, k, X, R, Z, Z1, Sv, V, Rh, rr);               // x, r, ||r||, beta
    });
  }
  // CG
  double *R = xs_multi_vec(c, k, 0), *Z = xs_multi_vec(c, k, 1), *P = xs_multi_vec(c, k, 2);
  double *AP = xs_multi_vec(c, k, 3), *T = xs_multi_vec(c, k, 4);
  const xs_mred rnone = m.red(MF_NONE), rinit = m.red(MF_CG_INIT);
  const xs_mred ralpha = m.red(MF_CG_ALPHA), rres = m.red(MF_CG_RES), rrr = m.red(MF_CG_RR);
  mspmm(c, s, k, X, R, EPI_BMAXN, B, nullptr, &rnone, m.ms);     // R = B - A X
  m.pc(R, Z, T);                                                       // Z = M^-1 R
  xs_mdot(s, n, k, R, Z, R, R, rinit);                                 // tau, res0
  XS_CHK(hipMemcpyAsync(P, Z, bytes, hipMemcpyDeviceToDevice, s));     // P = Z
  return m.drive([&]() {
    mspmm(c, s, k, P, AP, EPI_DOT1, nullptr, P, &ralpha, m.ms);  // AP, alpha
    if (jac) {
      xs_mcg_xr(s, n, k, X, R, P, AP, jac, Z, rres);                   // x, r, z, tau, ||r||
    } else {
      xs_mcg_xr(s, n, k, X, R, P, AP, nullptr, nullptr, rrr);          // x, r, ||r||^2
      m.pc(R, Z, T);                                                   // z = M^-1 r
      xs_mdot(s, n, k, Z, R, nullptr, nullptr, rres);                  // tau, beta, convergence
    }
    xs_mcg_p(s, n, k, P, Z, m.ms);                                     // p = z + beta p
  });
}
