// xs_mcsr.cpp -- the Fortran-style surface of the reference's CUDA backend
// (cuda/HostMatrix.h:162-193, cuda/DeviceMatrix.h:121-169,
// cuda/mpicudasolver-C-v1.cu): the matrix is registered from host arrays in
// MCSR form (diagonal held apart, off-diagonals in CSR, 0-based, as
// intel/HostMatrix.h:83-125 reads it), uploaded by gpu_mat_setup, and each
// gpu_*_solve runs one Krylov method of the intel family with the reference's
// output conventions: usediter = 0-based index of the last iteration,
// resvec[k] = ||r_{k+1}|| / (||r_0|| + 1e-20) of the residual that method
// monitors.  Every argument is passed by address, as from Fortran (the
// reference's C++ references); one matrix at a time, not thread-safe.
// Errors never exit: gpu_last_error() reports the status of the last call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/xsolver_c.h"
#include "xs_solvers.h"

namespace {
struct mcsr_state {
  // registered host arrays (borrowed, as HostMatrix keeps them)
  int n = 0, halo = 0, noff = 0;
  const double *diag = nullptr, *off = nullptr;
  const int *offptr = nullptr, *offcol = nullptr;
  bool registered = false;
  // device solver and the CSR built from the MCSR arrays
  xs_ctx *ctx = nullptr;
  std::vector<int> rp, ci;
  std::vector<int> pos_diag, pos_off;   // CSR position of each MCSR value
  std::vector<double> vals;
  int use_ell = 0;
  int err = XS_OK;
};
mcsr_state g;

int set_err(int rc) {
  g.err = rc;
  return rc;
}

bool same(const char *a, const char *b) { return a && std::strcmp(a, b) == 0; }

// CSR values in the order of the MCSR arrays' positions
void gather_values() {
  g.vals.assign(g.rp.empty() ? 0 : g.rp.back(), 0.0);
  for (int i = 0; i < g.n; i++) g.vals[g.pos_diag[i]] = g.diag[i];
  for (int e = 0; e < g.noff; e++) g.vals[g.pos_off[e]] = g.off[e];
}

// MCSR -> sorted CSR with the diagonal in place; positions kept for updates
int build_csr() {
  const int n = g.n;
  g.rp.assign(n + 1, 0);
  g.pos_diag.assign(n, 0);
  g.pos_off.assign(g.noff, 0);
  if (g.offptr[0] != 0 || g.offptr[n] != g.noff) return XS_ERR_ARG;
  for (int i = 0; i < n; i++) {
    if (g.offptr[i + 1] < g.offptr[i]) return XS_ERR_ARG;
    g.rp[i + 1] = g.rp[i] + (g.offptr[i + 1] - g.offptr[i]) + 1;
  }
  g.ci.assign(g.rp[n], 0);
  std::vector<int> order;
  for (int i = 0; i < n; i++) {
    const int e0 = g.offptr[i], e1 = g.offptr[i + 1];
    order.resize(e1 - e0);
    std::iota(order.begin(), order.end(), e0);
    for (int e = e0; e < e1; e++)
      if (g.offcol[e] < 0 || g.offcol[e] >= n || g.offcol[e] == i) return XS_ERR_ARG;
    std::sort(order.begin(), order.end(), [](int a, int b) { return g.offcol[a] < g.offcol[b]; });
    int at = g.rp[i];
    bool placed = false;
    for (int e : order) {
      if (!placed && g.offcol[e] > i) {
        g.ci[at] = i;
        g.pos_diag[i] = at++;
        placed = true;
      }
      g.ci[at] = g.offcol[e];
      g.pos_off[e] = at++;
    }
    if (!placed) {
      g.ci[at] = i;
      g.pos_diag[i] = at++;
    }
  }
  gather_values();
  return XS_OK;
}

// the library's result in the reference's conventions
void report(int rc, int maxiter, bool gcr_family, int *usediter, double *resvec) {
  xs_ctx *c = g.ctx;
  const int it = c->iters;
  if (usediter) {
    if (rc == XS_OK) *usediter = it > 0 ? it - 1 : 0;
    else *usediter = gcr_family ? maxiter : maxiter - 1;   // the loop index after exhaustion
  }
  if (resvec && !c->resvec.empty()) {
    const double r0 = c->resvec[0];
    for (int k = 0; k + 1 < (int)c->resvec.size(); k++) resvec[k] = c->resvec[k + 1] / (r0 + 1e-20);
  }
}

void solve(const char *ksp, double *rhs, double *sol, const int *restart, const double *tol,
           const int *maxiter, const double *sol_init, int *usediter, double *resvec,
           bool gcr_family) {
  if (!g.ctx || !g.ctx->mat_ready) { set_err(XS_ERR_STATE); return; }
  if (!rhs || !sol || !tol || !maxiter || *maxiter <= 0) { set_err(XS_ERR_ARG); return; }
  int rc = xs_ksp_settype(g.ctx, ksp);
  if (rc == XS_OK) rc = xs_ksp_setoption(g.ctx, *tol, *maxiter, restart ? std::max(*restart, 0) : 0);
  if (rc) { set_err(rc); return; }
  std::vector<double> x(sol_init ? sol_init : sol, (sol_init ? sol_init : sol) + g.n);
  rc = xs_solve(g.ctx, rhs, x.data());
  set_err(rc);
  if (rc >= 0 || rc == XS_ERR_BREAKDOWN) {
    std::copy(x.begin(), x.end(), sol);
    report(rc, *maxiter, gcr_family, usediter, resvec);
  }
}
}  // namespace

extern "C" {

void cpu_mat_setup(int *nInterior, int *nHalo, int *nSizes, double *diag_val, double *offdiag_val,
                   int *offdiag_row_offset, int *offdiag_col_index) {
  g.registered = false;
  if (!nInterior || !nHalo || !nSizes || *nInterior <= 0 || *nSizes < 0 || !diag_val ||
      !offdiag_row_offset || (*nSizes > 0 && (!offdiag_val || !offdiag_col_index))) {
    set_err(XS_ERR_ARG);
    return;
  }
  // one rank: the MPI halo of the reference (cuda/HostMatrix.h:38-45) is dropped
  if (*nHalo != 0) { set_err(XS_ERR_UNSUPPORTED); return; }
  g.n = *nInterior;
  g.halo = *nHalo;
  g.noff = *nSizes;
  g.diag = diag_val;
  g.off = offdiag_val;
  g.offptr = offdiag_row_offset;
  g.offcol = offdiag_col_index;
  g.registered = true;
  set_err(XS_OK);
}

void cpu_mat_update(double *diag_val, double *offdiag_val) {
  if (!g.registered) { set_err(XS_ERR_STATE); return; }
  if (!diag_val || (g.noff > 0 && !offdiag_val)) { set_err(XS_ERR_ARG); return; }
  g.diag = diag_val;
  g.off = offdiag_val;
  set_err(XS_OK);
}

void gpu_mat_setup(void) {
  if (!g.registered) { set_err(XS_ERR_STATE); return; }
  if (!g.ctx && set_err(xs_create(&g.ctx)) != XS_OK) return;
  int rc = build_csr();
  if (rc) { set_err(rc); return; }
  const int pam[3] = {0, g.use_ell, 1};
  set_err(xs_mat_setup(g.ctx, g.n, g.rp.data(), g.ci.data(), g.vals.data(), pam));
}

void gpu_mat_format(char *fmt, int *num_nz) {
  if (!same(fmt, "ELL")) { set_err(XS_OK); return; }
  if (!num_nz || *num_nz <= 0 || *num_nz > 32) { set_err(XS_ERR_ARG); return; }
  // the reference pads to 16 or 32 columns and writes the width back
  // (cuda/DeviceMatrix.h:131-139); here the ELL copy has the true width
  *num_nz = *num_nz <= 16 ? 16 : 32;
  g.use_ell = 1;
  set_err(g.ctx && g.ctx->mat_ready ? xs_set_option(g.ctx, "use_ell", 1) : XS_OK);
}

void gpu_mat_update(void) {
  if (!g.ctx || !g.ctx->mat_ready || !g.registered) { set_err(XS_ERR_STATE); return; }
  gather_values();
  set_err(xs_mat_update_values(g.ctx, g.vals.data(), 0));
}

void gpu_mat_free(void) {
  if (g.ctx) xs_destroy(g.ctx);
  g.ctx = nullptr;
  g.use_ell = 0;
  set_err(XS_OK);
}

// "Jacobi" as in the reference (cuda/DeviceMatrix.h:153-158); the library's
// other preconditioners ("Ilu0", "Ic0", "Ilut", ...) are accepted too; any
// other string leaves the system unpreconditioned.
void gpu_preconditioner_setup(char *fmt) {
  if (!g.ctx || !g.ctx->mat_ready) { set_err(XS_ERR_STATE); return; }
  int rc = xs_pc_settype(g.ctx, fmt ? fmt : "none");
  if (rc == XS_ERR_ARG) rc = xs_pc_settype(g.ctx, "none");
  if (rc == XS_OK) rc = xs_pc_setup(g.ctx);
  set_err(rc);
}

void gpu_preconditioner_update(char *fmt) {
  if (!g.ctx || !g.ctx->mat_ready) { set_err(XS_ERR_STATE); return; }
  (void)fmt;
  int rc = xs_set_option(g.ctx, "refactor", 1);
  if (rc == XS_OK) rc = xs_pc_setup(g.ctx);
  set_err(rc);
}

void gpu_preconditioner_free(char *fmt) {
  (void)fmt;
  if (!g.ctx) { set_err(XS_OK); return; }
  set_err(xs_pc_settype(g.ctx, "none"));
}

void gpu_cg_solve(double *rhs, double *sol, double *tol, int *maxiter, double *sol_init,
                  int *usediter, double *resvec) {
  solve("cg_left", rhs, sol, nullptr, tol, maxiter, sol_init, usediter, resvec, false);
}
void gpu_icg_solve(double *rhs, double *sol, double *tol, int *maxiter, double *sol_init,
                   int *usediter, double *resvec) {
  solve("icg", rhs, sol, nullptr, tol, maxiter, sol_init, usediter, resvec, false);
}
void gpu_cgm_solve(double *rhs, double *sol, double *tol, int *maxiter, double *sol_init,
                   int *usediter, double *resvec) {
  solve("cgm", rhs, sol, nullptr, tol, maxiter, sol_init, usediter, resvec, false);
}
void gpu_icgm_solve(double *rhs, double *sol, double *tol, int *maxiter, double *sol_init,
                    int *usediter, double *resvec) {
  solve("icgm", rhs, sol, nullptr, tol, maxiter, sol_init, usediter, resvec, false);
}
void gpu_bicgstab_solve(double *rhs, double *sol, double *tol, int *maxiter, double *sol_init,
                        int *usediter, double *resvec) {
  solve("bicgstab", rhs, sol, nullptr, tol, maxiter, sol_init, usediter, resvec, false);
}
void gpu_ibicgstab_solve(double *rhs, double *sol, double *tol, int *maxiter, double *sol_init,
                         int *usediter, double *resvec) {
  solve("ibicgstab", rhs, sol, nullptr, tol, maxiter, sol_init, usediter, resvec, false);
}
void gpu_gcr_solve(double *rhs, double *sol, int *restart, double *tol, int *maxiter,
                   double *sol_init, int *usediter, double *resvec) {
  solve("gcr_left", rhs, sol, restart, tol, maxiter, sol_init, usediter, resvec, true);
}
void gpu_igcr_solve(double *rhs, double *sol, int *restart, double *tol, int *maxiter,
                    double *sol_init, int *usediter, double *resvec) {
  solve("igcr_left", rhs, sol, restart, tol, maxiter, sol_init, usediter, resvec, true);
}
void gpu_gmres_solve(double *rhs, double *sol, int *restart, double *tol, int *maxiter,
                     double *sol_init, int *usediter, double *resvec) {
  solve("gmres", rhs, sol, restart, tol, maxiter, sol_init, usediter, resvec, false);
}
void gpu_fgmres_solve(double *rhs, double *sol, int *restart, double *tol, int *maxiter,
                      double *sol_init, int *usediter, double *resvec) {
  solve("fgmres", rhs, sol, restart, tol, maxiter, sol_init, usediter, resvec, false);
}

int gpu_last_error(void) { return g.err; }

}  // extern "C"
