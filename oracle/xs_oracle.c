/*
 * xs_oracle.c -- CPU restatement of the xsolver reference's preconditioned
 * Krylov path.  TEST INFRASTRUCTURE ONLY (the checker, never the product):
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function cites the reference file:line it restates.  How each function
 * is pinned (golden vectors from the compiled reference, closed forms,
 * library routines, brute force) is listed in the test files under tests/.
 */
#include "xs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* BLAS-1 helpers.  Loops over vector entries and matrix rows run on OpenMP  */
/* threads; a dot product sums fixed blocks of OR_BLK entries sequentially   */
/* and then the block sums in block order, so every result is the same for  */
/* any thread count (the serial build gives the same bits).                 */
/* ------------------------------------------------------------------------ */
#define OR_BLK 4096
static double dot(int n, const double *a, const double *b) {
  const int nb = (n + OR_BLK - 1) / OR_BLK;
  double stackp[64];
  double *part = nb <= 64 ? stackp : (double *)malloc(sizeof(double) * (size_t)nb);
#pragma omp parallel for schedule(static) if (nb > 8)
  for (int k = 0; k < nb; k++) {
    const int e = (k + 1) * OR_BLK < n ? (k + 1) * OR_BLK : n;
    double s = 0.0;
    for (int i = k * OR_BLK; i < e; i++) s += a[i] * b[i];
    part[k] = s;
  }
  double s = 0.0;
  for (int k = 0; k < nb; k++) s += part[k];
  if (part != stackp) free(part);
  return s;
}
static void axpy(int n, double a, const double *x, double *y) {
#pragma omp parallel for schedule(static) if (n > 32768)
  for (int i = 0; i < n; i++) y[i] += a * x[i];
}
static void scal(int n, double a, double *x) {
#pragma omp parallel for schedule(static) if (n > 32768)
  for (int i = 0; i < n; i++) x[i] *= a;
}
static void copy(int n, const double *x, double *y) { memcpy(y, x, sizeof(double) * (size_t)n); }

static void apply_M(or_precond_fn M, void *ctx, int n, const double *x, double *y) {
  if (M) M(ctx, x, y);
  else if (x != y) copy(n, x, y);
}

/* ------------------------------------------------------------------------ */
/* Sparse kernels                                                           */
/* ------------------------------------------------------------------------ */

/* y = A x.  intel/HostMatrix.h:83-103: the off-diagonal entries of the row are
 * accumulated in storage order and the diagonal term is added last. */
void or_spmv(int n, const int *rp, const int *ci, const double *v,
             const double *x, double *y) {
#pragma omp parallel for schedule(static) if (n > 16384)
  for (int i = 0; i < n; i++) {
    double r = 0.0, d = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; e++) {
      if (ci[e] == i) d += v[e];
      else r = r + v[e] * x[ci[e]];
    }
    y[i] = r + d * x[i];
  }
}

/* y = b - A x.  intel/HostMatrix.h:105-125 (r = rhs; r -= offdiag; y = r - diag*x). */
void or_bmax(int n, const int *rp, const int *ci, const double *v,
             const double *b, const double *x, double *y) {
#pragma omp parallel for schedule(static) if (n > 16384)
  for (int i = 0; i < n; i++) {
    double r = b[i], d = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; e++) {
      if (ci[e] == i) d += v[e];
      else r = r - v[e] * x[ci[e]];
    }
    y[i] = r - d * x[i];
  }
}

/* y = A^T x (adjoint path; no reference counterpart). */
void or_spmv_t(int n, const int *rp, const int *ci, const double *v,
               const double *x, double *y) {
  memset(y, 0, sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++)
    for (int e = rp[i]; e < rp[i + 1]; e++) y[ci[e]] += v[e] * x[i];
}

/* Forward solve with a strictly-lower CSR factor and an implied unit diagonal:
 * x_i = b_i - sum_j L_ij x_j.  hip/HostMatrix/HostMatrixCSR.cpp:386-394. */
void or_lower_solve(int n, const int *rp, const int *ci, const double *v,
                    const double *b, double *x) {
  for (int i = 0; i < n; i++) {
    double s = b[i];
    for (int e = rp[i]; e < rp[i + 1]; e++) s -= x[ci[e]] * v[e];
    x[i] = s;
  }
}

/* Backward solve with a strictly-upper CSR factor and the inverted diagonal:
 * x_i = dinv_i (b_i - sum_j U_ij x_j).  Same recurrence as
 * hip/HostMatrix/HostMatrixCSR.cpp:398-408 but with an unscaled U. */
void or_upper_solve(int n, const int *rp, const int *ci, const double *v,
                    const double *dinv, const double *b, double *x) {
  for (int i = n - 1; i >= 0; i--) {
    double s = b[i];
    for (int e = rp[i]; e < rp[i + 1]; e++) s -= x[ci[e]] * v[e];
    x[i] = s * dinv[i];
  }
}

/* ------------------------------------------------------------------------ */
/* Preconditioners                                                          */
/* ------------------------------------------------------------------------ */

/* dinv_i = 1/(a_ii + small).  intel/HostMatrix.h:127-137; the HIP framework
 * uses small = 0 (hip/Precond/precon_jacobi.h:18).  Returns the first row
 * with a zero (or missing) pivot + 1, or 0 on success
 * (hip/HostVector/HostVector.cpp:84-95 exits there). */
int or_jacobi_setup(int n, const int *rp, const int *ci, const double *v,
                    double small, double *dinv) {
  int bad = 0;
  for (int i = 0; i < n; i++) {
    double d = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; e++)
      if (ci[e] == i) d = v[e];
    double t = d + small;
    if (t == 0.0 && !bad) bad = i + 1;
    dinv[i] = 1.0 / t;
  }
  return bad;
}

/* y = dinv o x.  intel/HostMatrix.h:139-146. */
typedef struct { int n; const double *dinv; } or_jac_ctx_t;
void or_jacobi_apply(void *ctx, const double *x, double *y) {
  const or_jac_ctx_t *c = (const or_jac_ctx_t *)ctx;
#pragma omp parallel for schedule(static) if (c->n > 32768)
  for (int i = 0; i < c->n; i++) y[i] = c->dinv[i] * x[i];
}

/* In-place textbook IKJ ILU(0) on a copy of A's values; pivots below 1e-20 in
 * magnitude are replaced by 1e-20.  hip/Precond/ilut_diag_in_csr.cpp:347-406.
 * uptr[i] = position of a_ii (or -1 when absent).  Returns 0, or i+1 for the
 * first row without a diagonal entry. */
int or_ilu0(int n, const int *rp, const int *ci, const double *v,
            double *luval, int *uptr) {
  int nnz = rp[n];
  for (int i = 0; i < n; i++) uptr[i] = -1;
  for (int e = 0; e < nnz; e++) luval[e] = v[e];
  for (int i = 0; i < n; i++)
    for (int e = rp[i]; e < rp[i + 1]; e++)
      if (ci[e] == i) { uptr[i] = e; break; }
  for (int i = 0; i < n; i++)
    if (uptr[i] < 0) return i + 1;
  for (int i = 0; i < n; i++) {
    int upart = uptr[i];
    for (int j = rp[i]; j < upart; j++) {
      int k = ci[j];
      if (fabs(luval[uptr[k]]) < 1e-20) luval[uptr[k]] = 1e-20;
      luval[j] = luval[j] / luval[uptr[k]];
      int sub = uptr[k];
      for (int jj = j + 1; jj < rp[i + 1]; jj++) {
        int kk = ci[jj];
        while (ci[sub] < kk && sub < rp[k + 1]) sub++;
        if (sub < rp[k + 1]) {
          if (ci[sub] == kk) luval[jj] = luval[jj] - luval[j] * luval[sub];
        } else {
          break;
        }
      }
    }
    if (fabs(luval[uptr[i]]) < 1e-20) luval[uptr[i]] = 1e-20;
  }
  return 0;
}

/* Split a combined LU (unit-lower strict part | diagonal | upper part) into
 * L (strict lower), dinv = 1/u_ii and U (strict upper, each row scaled by
 * dinv_i).  hip/HostMatrix/HostMatrixCSR.cpp:164-197 (ILU0) and :230-268
 * (ILUT): both use the same split. */
void or_ilu_split(int n, const int *rp, const int *ci, const double *luval,
                  const int *uptr, int *lrp, int *lci, double *lv,
                  int *urp, int *uci, double *uv, double *dinv) {
  int il = 0, iu = 0;
  lrp[0] = 0; urp[0] = 0;
  for (int i = 0; i < n; i++) {
    for (int e = rp[i]; e < uptr[i]; e++) { lv[il] = luval[e]; lci[il++] = ci[e]; }
    dinv[i] = 1.0 / luval[uptr[i]];
    for (int e = uptr[i] + 1; e < rp[i + 1]; e++) {
      uv[iu] = luval[e] * dinv[i];
      uci[iu++] = ci[e];
    }
    lrp[i + 1] = il; urp[i + 1] = iu;
  }
}

/* y = (LU)^{-1} x with the split factors.  hip/HostMatrix/HostMatrixCSR.cpp:385-409:
 * forward with unit L, y *= dinv, backward with the pre-scaled U visiting each
 * row's entries from last to first. */
void or_lusolve(void *ctx, const double *x, double *y) {
  const or_lu_t *f = (const or_lu_t *)ctx;
  int n = f->n;
  double *t = (double *)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) {
    double s = x[i];
    for (int e = f->lrp[i]; e < f->lrp[i + 1]; e++) s -= t[f->lci[e]] * f->lv[e];
    t[i] = s;
  }
  for (int i = 0; i < n; i++) t[i] *= f->dinv[i];
  for (int i = n - 1; i >= 0; i--) {
    double s = t[i];
    for (int e = f->urp[i + 1] - 1; e >= f->urp[i]; e--) s -= t[f->uci[e]] * f->uv[e];
    t[i] = s;
  }
  memcpy(y, t, sizeof(double) * (size_t)n);
  free(t);
}

/* Iterative application of combined ILU factors (the reference's GPU ILU
 * apply, hip/Precond/precon_gpuilu0.cpp:16-19 and precon_gpuilut.cpp:18-21,
 * DeviceMatrixCSR::Lsolve_iter / Usolve_iter,
 * hip/HostMatrix/DeviceMatrix/DeviceMatrixCSR.cpp:281-304, kernels
 * LMultCSR / lsolveiterCSR / UMultCSR / usolveiterCSR :14-58):
 *   L: t_0 = 0, t_{k+1} = x - L_strict t_k           (kl sweeps)
 *   U: y_0 = 0, y_{k+1} = (t - U_strict y_k) / u_ii  (ku sweeps)
 * Off-diagonal sums run over each row's entries in ascending column order.
 * luval/uptr: combined factor on pattern (rp, ci) with the pivot at uptr[i].
 * kl = 0 copies x (Lsolve_iter's maxiter == 0 branch); ku = 0 scales by
 * 1/u_ii (Usolve_iter's jacobiSolve branch). */
void or_ilu_sweeps(int n, const int *rp, const int *ci, const double *luval, const int *uptr,
                   int kl, int ku, const double *x, double *y) {
  double *t = (double *)calloc((size_t)n, sizeof(double));
  double *u = (double *)calloc((size_t)n, sizeof(double));
  double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
  if (kl == 0) memcpy(t, x, sizeof(double) * (size_t)n);
  for (int k = 0; k < kl; k++) {
#pragma omp parallel for schedule(static) if (n > 16384)
    for (int i = 0; i < n; i++) {
      double r = 0.0;
      for (int e = rp[i]; e < uptr[i]; e++) r += luval[e] * t[ci[e]];
      tmp[i] = r;
    }
#pragma omp parallel for schedule(static) if (n > 32768)
    for (int i = 0; i < n; i++) t[i] = x[i] - tmp[i];
  }
  if (ku == 0)
    for (int i = 0; i < n; i++) u[i] = (1.0 / luval[uptr[i]]) * t[i];
  for (int k = 0; k < ku; k++) {
#pragma omp parallel for schedule(static) if (n > 16384)
    for (int i = 0; i < n; i++) {
      double r = 0.0;
      for (int e = uptr[i] + 1; e < rp[i + 1]; e++) r += luval[e] * u[ci[e]];
      tmp[i] = r;
    }
#pragma omp parallel for schedule(static) if (n > 32768)
    for (int i = 0; i < n; i++) u[i] = (1.0 / luval[uptr[i]]) * (t[i] - tmp[i]);
  }
  memcpy(y, u, sizeof(double) * (size_t)n);
  free(t); free(u); free(tmp);
}

typedef struct { int n; const int *rp, *ci, *uptr; const double *luval; int kl, ku; } or_sweep_t;
static void or_sweep_apply(void *ctx, const double *x, double *y) {
  const or_sweep_t *c = (const or_sweep_t *)ctx;
  or_ilu_sweeps(c->n, c->rp, c->ci, c->luval, c->uptr, c->kl, c->ku, x, y);
}

/* --- Incomplete Cholesky (CPU preconditioners of the reference) ---------- */
/* IC(0), hip/Precond/ic0.cpp:5-75 (ic0_csr_A).  initLByA (:5-27): L = lower
 * part of A with the diagonal (last in each row), tnorm_i = droptol ||a_i||_2
 * over the whole row.  ic0_fact (:29-67), row by row:
 *   l_ij = (a_ij - sum_{k<j} l_ik l_jk) / l_jj  (merge of row i's entries
 *          before j with row j's strict entries; subtraction in ascending k),
 *   l_ii = sqrt(a_ii - sum_j l_ij^2),
 * with |l_ii| < 1e-20 replaced by smallnum / droptol tnorm_i + tnorm_i, and
 * 1/l_ii stored in place of l_ii (so "/ l_jj" is a product with that
 * entry).  smallnum = droptol = 1e-4 in the reference (precon_ic0.h:14-15). */
int or_ic0(int n, const int *rp, const int *ci, const double *v, double smallnum,
           double droptol, int *lp, int *lc, double *lv) {
  double *tn = (double *)calloc((size_t)n, sizeof(double));
  int len = 0;
  lp[0] = 0;
  for (int i = 0; i < n; i++) {
    for (int e = rp[i]; e < rp[i + 1]; e++) {
      if (ci[e] <= i) { lv[len] = v[e]; lc[len] = ci[e]; len++; }
      tn[i] += v[e] * v[e];
    }
    lp[i + 1] = len;
    tn[i] = sqrt(tn[i]) * droptol;
    if (lp[i + 1] == lp[i] || lc[lp[i + 1] - 1] != i) { free(tn); return i + 1; }
  }
  lv[0] = sqrt(lv[0]);
  if (fabs(lv[0]) < 1e-20) lv[0] = smallnum / droptol * tn[0] + tn[0];
  lv[0] = 1.0 / lv[0];
  for (int i = 1; i < n; i++) {
    const int dl = lp[i + 1] - 1;
    for (int j = lp[i]; j < dl; j++) {
      const int c = lc[j];
      int from = lp[c];
      const int du = lp[c + 1] - 1;
      for (int k = lp[i]; k < j; k++)
        for (int p = from; p < du; p++) {
          if (lc[p] == lc[k]) {
            lv[j] -= lv[k] * lv[p];
          } else if (lc[p] > lc[k]) {
            from = p;
            break;
          }
        }
      lv[j] *= lv[du];
    }
    for (int j = lp[i]; j < dl; j++) lv[dl] -= lv[j] * lv[j];
    lv[dl] = sqrt(lv[dl]);
    if (fabs(lv[dl]) < 1e-20) lv[dl] = smallnum / droptol * tn[i] + tn[i];
    lv[dl] = 1.0 / lv[dl];
  }
  free(tn);
  return 0;
}

/* Min-heap on |val| over the first m entries (adjustHeap_l,
 * hip/Precond/ict_l.cpp:4-24): sift down from i, the smaller child taken on
 * |left| > |right|, moved up while strictly smaller than the sifted value. */
static void ic_sift(double *val, int *col, int i, int m) {
  const double tv = val[i];
  const int tc = col[i];
  for (int k = 2 * i + 1; k < m; k = 2 * k + 1) {
    if (k + 1 < m && fabs(val[k]) > fabs(val[k + 1])) k++;
    if (fabs(val[k]) < fabs(tv)) {
      val[i] = val[k];
      col[i] = col[k];
      i = k;
    } else {
      break;
    }
  }
  val[i] = tv;
  col[i] = tc;
}

/* the m largest |val| of cnt candidates kept in val[0..m) (heapSort_part_l,
 * ict_l.cpp:25-41): heapify from m/2 down, then a later candidate replaces
 * the root only when strictly larger in magnitude */
static void ic_keep_largest(double *val, int *col, int m, int cnt) {
  for (int i = m / 2; i >= 0; i--) ic_sift(val, col, i, m);
  for (int i = m; i < cnt; i++)
    if (fabs(val[i]) > fabs(val[0])) {
      val[0] = val[i];
      col[0] = col[i];
      ic_sift(val, col, 0, m);
    }
}

/* ICT, hip/Precond/ict_l.cpp:139-232 (ict_csr_A_l / ict_fact_l).
 * tnorm_i = droptol ||a_i||_2 (calTnorm :117-137).  Row i (left-looking):
 * a dense work row w holds A's row i (every entry of the row, the upper part
 * included); w is cleared first over [first column of L's row c0, i], c0 the
 * first column of A's row i -- the one work array serves every row, and
 * positions below that range keep what an earlier row left there, as in the
 * reference.  Then for j = c0 .. i-1 in order:
 *   w_j = (w_j - sum_{p in L_j strict} w[col p] l_p) / l_jj   (sum in p order),
 *   kept when |w_j| > tnorm_i, and then w_i -= w_j^2 (a dropped w_j still
 *   feeds the later columns of the row);
 * |w_i| < 1e-20 -> smallnum / droptol tnorm_i + tnorm_i; 1/sqrt(|w_i|) is
 * stored as the diagonal (last).  More than maxfil kept: the maxfil largest
 * in magnitude (heap semantics above), in ascending column order.  Row 0 is
 * a_00 alone.  The reference passes the macro small_num = 1e-4 as smallnum
 * (HostMatrixCSR.cpp:331, ilut_diag_in_csr.h:7). */
int or_ict(int n, const int *rp, const int *ci, const double *v, double smallnum,
           double droptol, int maxfil, int *lp, int *lc, double *lv) {
  double *tn = (double *)calloc((size_t)n, sizeof(double));
  double *w = (double *)calloc((size_t)n, sizeof(double));
  double *tv = (double *)malloc(sizeof(double) * (size_t)n);
  int *tc = (int *)malloc(sizeof(int) * (size_t)n);
  int rc = 0;
  for (int i = 0; i < n; i++) {
    int has = 0;
    for (int e = rp[i]; e < rp[i + 1]; e++) {
      tn[i] += v[e] * v[e];
      if (ci[e] == i) has = 1;
    }
    if (!has) { rc = i + 1; goto out; }
    tn[i] = sqrt(tn[i]) * droptol;
  }
  lp[0] = 0;
  lv[0] = v[0];
  lc[0] = 0;
  if (fabs(lv[0]) < 1e-20) lv[0] = smallnum * (1.0 / droptol) * tn[0] + tn[0];
  lv[0] = 1.0 / sqrt(fabs(lv[0]));
  lp[1] = 1;
  for (int i = 1; i < n; i++) {
    int cnt = 0;
    const int c0 = ci[rp[i]];
    /* a row without entries left of the diagonal would make the reference
     * read L's row i before it is written (lc[lp[i]]); w_i alone is cleared */
    for (int j = c0 < i ? lc[lp[c0]] : i; j <= i; j++) w[j] = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; e++) w[ci[e]] = v[e];
    for (int j = c0; j < i; j++) {
      const int du = lp[j + 1] - 1;
      double t = 0.0;
      for (int p = lp[j]; p < du; p++) t += w[lc[p]] * lv[p];
      w[j] -= t;
      w[j] *= lv[du];
      if (fabs(w[j]) > tn[i]) {
        tv[cnt] = w[j];
        tc[cnt] = j;
        cnt++;
        w[i] -= w[j] * w[j];
      }
    }
    if (fabs(w[i]) < 1e-20) w[i] = smallnum * (1.0 / droptol) * tn[i] + tn[i];
    w[i] = 1.0 / sqrt(fabs(w[i]));
    int keep = cnt;
    if (cnt > maxfil) {
      ic_keep_largest(tv, tc, maxfil, cnt);
      keep = maxfil;
    }
    /* ascending columns (the columns are distinct, so the order is unique) */
    for (int a = 1; a < keep; a++) {
      const int c = tc[a];
      const double x = tv[a];
      int b = a - 1;
      while (b >= 0 && tc[b] > c) { tc[b + 1] = tc[b]; tv[b + 1] = tv[b]; b--; }
      tc[b + 1] = c;
      tv[b + 1] = x;
    }
    int at = lp[i];
    for (int a = 0; a < keep; a++) { lv[at] = tv[a]; lc[at] = tc[a]; at++; }
    lv[at] = w[i];
    lc[at] = i;
    lp[i + 1] = lp[i] + keep + 1;
  }
out:
  free(tn); free(w); free(tv); free(tc);
  return rc;
}

/* seqic0 / seqict (hip/HostMatrix/HostMatrixCSR.cpp:269-326, :327-384):
 * d_i = stored 1/l_ii; L~_ij = l_ij d_j (strict, CSR in row order); U =
 * L~^T as CSR (row j lists i ascending, filled in row order of i);
 * dinv_i = d_i^2.  LUsolve (:385-409) then applies L~, dinv, U. */
void or_ic_split(int n, const int *lp, const int *lc, const double *lv, int *lrp, int *lci,
                 double *lvs, int *urp, int *uci, double *uvs, double *dinv) {
  for (int i = 0; i <= n; i++) urp[i] = 0;
  for (int i = 0; i < n; i++)
    for (int e = lp[i]; e < lp[i + 1]; e++) {
      if (lc[e] == i) dinv[i] = lv[e];
      else urp[lc[e] + 1]++;
    }
  for (int i = 1; i <= n; i++) urp[i] += urp[i - 1];
  int *fill = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  memcpy(fill, urp, sizeof(int) * (size_t)(n + 1));
  int il = 0;
  lrp[0] = 0;
  for (int i = 0; i < n; i++) {
    for (int e = lp[i]; e < lp[i + 1]; e++) {
      const int j = lc[e];
      if (j == i) continue;
      const double t = lv[e] * dinv[j];
      lvs[il] = t;
      lci[il++] = j;
      uvs[fill[j]] = t;
      uci[fill[j]++] = i;
    }
    lrp[i + 1] = il;
  }
  for (int i = 0; i < n; i++) dinv[i] *= dinv[i];
  free(fill);
}

/* --- ILUT (dual threshold, per-row heaps) -------------------------------- */
/* Min-heap of column ids ordered by |w[id]| ("less" = larger magnitude, so the
 * top is the smallest magnitude).  vh_push restates libstdc++'s push_heap
 * sift-up (strict comparison) as used by hip/Precond/ilut_diag_in_csr.cpp:120,
 * vh_replace_top restates heap_swap (ilut_diag_in_csr.cpp:2-18). */
static int vh_less(const double *w, int a, int b) { return fabs(w[a]) > fabs(w[b]); }
static void vh_push(int *h, int len, const double *w) {
  int hole = len - 1, val = h[hole];
  int parent = (hole - 1) / 2;
  while (hole > 0 && vh_less(w, h[parent], val)) {
    h[hole] = h[parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  h[hole] = val;
}
static void vh_replace_top(int *h, int len, const double *w) {
  int i = 0, j = 1, k = h[0];
  double vk = fabs(w[k]);
  while (j < len) {
    if (j + 1 < len && fabs(w[h[j]]) > fabs(w[h[j + 1]])) j = j + 1;
    if (fabs(w[h[j]]) > vk) break;
    h[i] = h[j];
    i = j;
    j = 2 * j + 1;
  }
  h[i] = k;
}
/* Plain binary min-heap of ints (column order of the elimination). */
static void ih_push(int *h, int len) {
  int hole = len - 1, val = h[hole], parent = (hole - 1) / 2;
  while (hole > 0 && h[parent] > val) { h[hole] = h[parent]; hole = parent; parent = (hole - 1) / 2; }
  h[hole] = val;
}
static int ih_pop(int *h, int len) {
  int top = h[0], last = h[len - 1];
  len -= 1;
  int hole = 0;
  for (;;) {
    int c = 2 * hole + 1;
    if (c >= len) break;
    if (c + 1 < len && h[c + 1] < h[c]) c++;
    if (h[c] >= last) break;
    h[hole] = h[c];
    hole = c;
  }
  if (len > 0) h[hole] = last;
  return top;
}
static void sort_prefix_by_col(int *cols, int len) {
  /* selection sort, ilut_diag_in_csr.cpp:147-158 */
  for (int a = 0; a < len; a++) {
    int im = a;
    for (int b = a + 1; b < len; b++) if (cols[b] < cols[im]) im = b;
    int t = cols[a]; cols[a] = cols[im]; cols[im] = t;
  }
}

/* Dual-threshold ILUT: drop |v| < droptol*||a_i||_2, keep at most lfil entries
 * in each of the L and U parts of a row (the largest in magnitude), substitute
 * (1e-4/droptol + 1)*droptol*||a_i||_2 for a zero pivot.
 * hip/Precond/ilut_diag_in_csr.cpp:20-251 (ilutp_final, permtol/iperm unused).
 * Output: combined rows lurows/lucols/luval with uptr[i] = position of the
 * pivot, diag[i] = 1/pivot.  cap = capacity of lucols/luval.  Returns 0, or -1
 * when cap is exceeded. */
int or_ilut(int n, const int *rp, const int *ci, const double *v, int lfil,
            double droptol, int cap, double *luval, int *lucols, int *lurows,
            int *uptr, double *diag) {
  double *w = (double *)calloc((size_t)n, sizeof(double));
  int *uj = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  int *lj = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  int nnz = 0, rc = 0;
  lurows[0] = 0;
  for (int i = 0; i < n; i++) diag[i] = 0.0;
  for (int i = 0; i < n && rc == 0; i++) {
    int nu = 0, nl = 0;
    double t1 = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; e++) {
      int k = ci[e];
      w[k] = v[e];
      t1 += v[e] * v[e];
      if (k < i) { lj[nl++] = k; ih_push(lj, nl); }
      if (k > i) uj[nu++] = k;
    }
    double tnorm = sqrt(t1) * droptol;
    if (nnz + 2 * lfil + 1 > cap) { rc = -1; break; }
    int *vh = lucols + nnz;
    int nv = 0;
    while (nl > 0) {
      int k = ih_pop(lj, nl);
      nl -= 1;
      double factor = w[k] * diag[k];
      if (fabs(factor) < tnorm) { w[k] = 0.0; continue; }
      for (int e = uptr[k] + 1; e < lurows[k + 1]; e++) {
        int j = lucols[e];
        double twj = w[j];
        if (twj == 0.0) {
          if (j < i) { lj[nl++] = j; ih_push(lj, nl); }
          if (i < j) uj[nu++] = j;
        }
        w[j] = twj - factor * luval[e];
      }
      w[k] = factor;
      if (nv < lfil) {
        vh[nv++] = k;
        vh_push(vh, nv, w);
      } else {
        int top = vh[0];
        if (fabs(factor) > fabs(w[top])) {
          w[top] = 0.0;
          vh[0] = k;
          vh_replace_top(vh, nv, w);
        } else {
          w[k] = 0.0;
        }
      }
    }
    sort_prefix_by_col(vh, nv);
    for (int t = 0; t < nv; t++) { luval[nnz] = w[vh[t]]; nnz++; w[vh[t]] = 0.0; }
    uptr[i] = nnz;
    if (w[i] != 0.0) {
      diag[i] = 1.0 / w[i];
      luval[nnz] = w[i];
      lucols[nnz] = i;
      w[i] = 0.0;
      nnz++;
    } else {
      double sub = 1e-4 / droptol * tnorm + tnorm;
      diag[i] = 1.0 / sub;
      luval[nnz] = sub;
      lucols[nnz] = i;
      nnz++;
    }
    vh = lucols + nnz;
    nv = 0;
    for (int t = 0; t < nu; t++) {
      int ki = uj[t];
      if (fabs(w[ki]) > tnorm) {
        if (nv < lfil) {
          vh[nv++] = ki;
          vh_push(vh, nv, w);
        } else {
          int top = vh[0];
          if (fabs(w[ki]) > fabs(w[top])) {
            w[top] = 0.0;
            vh[0] = ki;
            vh_replace_top(vh, nv, w);
          } else {
            w[ki] = 0.0;
          }
        }
      } else {
        w[ki] = 0.0;
      }
    }
    sort_prefix_by_col(vh, nv);
    for (int t = 0; t < nv; t++) { luval[nnz] = w[vh[t]]; nnz++; w[vh[t]] = 0.0; }
    lurows[i + 1] = nnz;
  }
  free(w); free(uj); free(lj);
  return rc;
}

/* One synchronous ParILU(0) fixed-point sweep over pattern(A).
 * hip/Precond/gpuparilu.cpp:46-102 (ilu0_ij_aw): for (i,j) in pattern(A),
 *   s = a_ij - sum_{k<min(i,j)} l_ik u_kj   (sorted merge of L row i, U col j)
 *   i > j : l_ij = s / u_jj  (|u_jj| < 1e-10 clamped to +-1e-10)
 *   i <= j: u_ij = s         (diagonal |s| < 1e-10 clamped to +-1e-10)
 *   a non-finite result leaves the old value in place.
 * The reference sweeps asynchronously in place; this is the Jacobi
 * (read-old / write-new) form.  L: strict lower CSR (unit diagonal implied).
 * U: upper incl. diagonal in CSC with the diagonal last in each column. */
void or_parilu_sweep(int n, const int *rp, const int *ci, const double *a,
                     const int *lrp, const int *lci, const double *lv_in,
                     const int *ucp, const int *uri, const double *uv_in,
                     double *lv_out, double *uv_out) {
  memcpy(lv_out, lv_in, sizeof(double) * (size_t)lrp[n]);
  memcpy(uv_out, uv_in, sizeof(double) * (size_t)ucp[n]);
  for (int i = 0; i < n; i++) {
    for (int e = rp[i]; e < rp[i + 1]; e++) {
      int j = ci[e];
      int m = i < j ? i : j;
      double s = a[e];
      int il = lrp[i], iu = ucp[j];
      int le = lrp[i + 1], ue = ucp[j + 1];
      while (il < le && iu < ue) {
        int kl = lci[il], ku = uri[iu];
        if (kl >= m || ku >= m) break;
        if (kl == ku) { s -= lv_in[il] * uv_in[iu]; il++; iu++; }
        else if (kl < ku) il++;
        else iu++;
      }
      if (i > j) {
        double d = uv_in[ucp[j + 1] - 1];
        if (fabs(d) < 1e-10) d = d >= 0 ? 1e-10 : -1e-10;
        double t = s / d;
        /* locate l_ij */
        for (int q = lrp[i]; q < le; q++)
          if (lci[q] == j) { if (isfinite(t)) lv_out[q] = t; break; }
      } else {
        if (i == j && fabs(s) < 1e-10) s = s >= 0 ? 1e-10 : -1e-10;
        for (int q = ucp[j]; q < ue; q++)
          if (uri[q] == i) { if (isfinite(s)) uv_out[q] = s; break; }
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Krylov solvers                                                           */
/* ------------------------------------------------------------------------ */

/* Preconditioned CG.  hip/Solver/xsolver_cg.cpp:16-74 (small = 0), which is
 * the same recurrence as intel/mpisolver-C-v1.cpp:193-272 (cpu_cgm_solve).
 * Returns 0 on convergence, 1 when maxiter is reached, 2 on a non-finite
 * residual. */
int or_cg(int n, const int *rp, const int *ci, const double *v,
          const double *b, double *x, double tol, double abstol, int maxiter,
          or_precond_fn M, void *Mctx, double *resvec, int *usediter) {
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
  int rc = 1;
  or_bmax(n, rp, ci, v, b, x, r);
  apply_M(M, Mctx, n, r, z);
  double tau = dot(n, r, z);
  double res0 = sqrt(dot(n, r, r));
  copy(n, z, p);
  if (resvec) resvec[0] = res0;
  *usediter = 0;
  if (res0 == 0.0 || res0 <= abstol) { rc = 0; goto done; }
  for (int it = 0; it < maxiter; it++) {
    double taoo = tau;
    or_spmv(n, rp, ci, v, p, Ap);
    double sigma = dot(n, Ap, p);
    double alpha = tau / sigma;
    axpy(n, alpha, p, x);
    axpy(n, -alpha, Ap, r);
    apply_M(M, Mctx, n, r, z);
    tau = dot(n, z, r);
    double beta = tau / taoo;
    scal(n, beta, p);
    axpy(n, 1.0, z, p);
    double resl = sqrt(dot(n, r, r));
    if (resvec) resvec[it + 1] = resl;
    *usediter = it + 1;
    if (!isfinite(resl)) { rc = 2; goto done; }
    if (resl / res0 < tol || resl <= abstol) { rc = 0; goto done; }
  }
done:
  free(r); free(z); free(p); free(Ap);
  return rc;
}

/* Right-preconditioned BiCGSTAB in x-space.  hip/Solver/xsolver_bicgstab.cpp:20-124
 * with the fused updates of hip/HostVector/DeviceVector/DeviceVector.cpp:55-72
 * (p = r + omega (p - alpha u); s = r - gamma u; x += gamma z + alpha z1),
 * small = 0 except the 1e-40 guard on omega's denominator.  In exact arithmetic
 * the iterates equal intel/mpisolver-C-v1.cpp:365-482 (y-space, x = M^{-1} y). */
int or_bicgstab(int n, const int *rp, const int *ci, const double *v,
                const double *b, double *x, double tol, double abstol,
                int maxiter, or_precond_fn M, void *Mctx, double *resvec,
                int *usediter) {
  double *r = calloc(n, sizeof(double)), *rh = calloc(n, sizeof(double));
  double *p = calloc(n, sizeof(double)), *u = calloc(n, sizeof(double));
  double *s = calloc(n, sizeof(double)), *vv = calloc(n, sizeof(double));
  double *z = calloc(n, sizeof(double)), *z1 = calloc(n, sizeof(double));
  int rc = 1;
  or_bmax(n, rp, ci, v, b, x, r);
  double res0 = sqrt(dot(n, r, r));
  copy(n, r, rh);
  if (resvec) resvec[0] = res0;
  *usediter = 0;
  if (res0 == 0.0 || res0 <= abstol) { rc = 0; goto done; }
  double alpha = 1.0, beto = 1.0, gama = 1.0;
  for (int it = 0; it < maxiter; it++) {
    double beta = dot(n, rh, r);
    double omega = (beta * gama) / (alpha * beto + 1e-40);
    beto = beta;
#pragma omp parallel for schedule(static) if (n > 32768)
    for (int i = 0; i < n; i++) p[i] = r[i] + omega * (p[i] - alpha * u[i]);
    apply_M(M, Mctx, n, p, z);
    or_spmv(n, rp, ci, v, z, u);
    double tmp = dot(n, rh, u);
    gama = beta / tmp;
#pragma omp parallel for schedule(static) if (n > 32768)
    for (int i = 0; i < n; i++) s[i] = r[i] - gama * u[i];
    apply_M(M, Mctx, n, s, z1);
    or_spmv(n, rp, ci, v, z1, vv);
    double t0 = dot(n, s, vv), t1 = dot(n, vv, vv);
    /* the reference divides unguarded (xsolver_bicgstab.cpp:95); when the
       half step already solved the system (s = 0, so v = 0) that is 0/0 and
       every later value is NaN -- take alpha = 0, keeping x + gama z */
    alpha = t1 != 0.0 ? t0 / t1 : 0.0;
#pragma omp parallel for schedule(static) if (n > 32768)
    for (int i = 0; i < n; i++) x[i] = x[i] + gama * z[i] + alpha * z1[i];
    copy(n, s, r);
    axpy(n, -alpha, vv, r);
    double resl = sqrt(dot(n, r, r));
    if (resvec) resvec[it + 1] = resl;
    *usediter = it + 1;
    if (!isfinite(resl)) { rc = 2; goto done; }
    if (resl / res0 < tol || resl <= abstol) { rc = 0; goto done; }
  }
done:
  free(r); free(rh); free(p); free(u); free(s); free(vv); free(z); free(z1);
  return rc;
}

/* Right-preconditioned GCR(restart) with modified Gram-Schmidt on c = A u.
 * hip/Solver/xsolver_igcr.cpp:28-108 (small = 0). */
int or_gcr(int n, const int *rp, const int *ci, const double *v,
           const double *b, double *x, double tol, double abstol, int maxiter,
           int restart, or_precond_fn M, void *Mctx, double *resvec,
           int *usediter) {
  if (restart <= 0) restart = maxiter;
  double *r = malloc(sizeof(double) * n);
  double *U = malloc(sizeof(double) * (size_t)n * restart);
  double *C = malloc(sizeof(double) * (size_t)n * restart);
  int rc = 1;
  or_bmax(n, rp, ci, v, b, x, r);
  double res0 = sqrt(dot(n, r, r));
  if (resvec) resvec[0] = res0;
  *usediter = 0;
  if (res0 == 0.0 || res0 <= abstol) { rc = 0; goto done; }
  int iiter = 0;
  while (iiter < maxiter) {
    int k = 0;
    while (k < restart && iiter < maxiter) {
      double *uk = U + (size_t)k * n, *ck = C + (size_t)k * n;
      apply_M(M, Mctx, n, r, uk);
      or_spmv(n, rp, ci, v, uk, ck);
      for (int i = 0; i < k; i++) {
        double a = dot(n, C + (size_t)i * n, ck);
        axpy(n, -a, C + (size_t)i * n, ck);
        axpy(n, -a, U + (size_t)i * n, uk);
      }
      double t0 = sqrt(dot(n, ck, ck)), t1 = dot(n, ck, r);
      double inv = 1.0 / t0;
      scal(n, inv, uk);
      scal(n, inv, ck);
      double tmp = t1 / t0;
      axpy(n, tmp, uk, x);
      axpy(n, -tmp, ck, r);
      double rn = sqrt(dot(n, r, r));
      if (resvec) resvec[iiter + 1] = rn;
      *usediter = iiter + 1;
      if (!isfinite(rn)) { rc = 2; goto done; }
      if (rn / res0 < tol || rn <= abstol) { rc = 0; goto done; }
      k++; iiter++;
    }
  }
done:
  free(r); free(U); free(C);
  return rc;
}

/* Givens rotation helpers.  intel/mpisolver-C-v1.cpp:1249-1279. */
static void rot_apply(double *dx, double *dy, double cs, double sn) {
  double t = cs * *dx + sn * *dy;
  *dy = -sn * *dx + cs * *dy;
  *dx = t;
}
static void rot_gen(double dx, double dy, double *cs, double *sn) {
  if (dy == 0.0) { *cs = 1.0; *sn = 0.0; }
  else if (fabs(dy) > fabs(dx)) {
    double t = dx / dy;
    *sn = 1.0 / sqrt(1.0 + t * t);
    *cs = t * *sn;
  } else {
    double t = dy / dx;
    *cs = 1.0 / sqrt(1.0 + t * t);
    *sn = t * *cs;
  }
}
/* y = H(0:k,0:k)^{-1} s (H_ii + 1e-20), x += sum_i y_i B_i.
 * intel/mpisolver-C-v1.cpp:1281-1301 (updatec). */
static void gm_update(int n, double *x, int k, const double *H, int ldh,
                      const double *s, const double *B) {
  double *y = malloc(sizeof(double) * (size_t)(k + 1));
  for (int i = 0; i < k; i++) y[i] = s[i];
  for (int i = k - 1; i >= 0; i--) {
    y[i] = y[i] / (H[i * ldh + i] + 1e-20);
    for (int j = i - 1; j >= 0; j--) y[j] = y[j] - H[j * ldh + i] * y[i];
  }
  for (int i = 0; i < k; i++) axpy(n, y[i], B + (size_t)i * n, x);
  free(y);
}

/* GMRES(restart) with MGS Arnoldi and Givens rotations.
 * flexible = 0: left preconditioning, intel/mpisolver-C-v1.cpp:933-1087
 *   (residuals are norms of M^{-1}(b - A x)).
 * flexible = 1: right preconditioning with stored Z_i = M V_i,
 *   intel/mpisolver-C-v1.cpp:1089-1247 (FGMRES; residuals unpreconditioned).
 * Restart keeps the reference's rule beta = |s[m]| (:1069).  Unlike the
 * reference (which calls updatec(phi, i, ...) on convergence and so drops the
 * newest direction), the converged update uses all i+1 directions.
 * resvec[0] = res0, resvec[j+1] = |s_{i+1}| after global step j. */
int or_gmres(int n, const int *rp, const int *ci, const double *v,
             const double *b, double *x, double tol, int maxiter, int restart,
             int flexible, or_precond_fn M, void *Mctx, double *resvec,
             int *usediter) {
  int m = restart;
  double *V = calloc((size_t)n * (m + 1), sizeof(double));
  double *Z = flexible ? calloc((size_t)n * (m + 1), sizeof(double)) : NULL;
  double *H = calloc((size_t)(m + 1) * m, sizeof(double));
  double *s = calloc(m + 1, sizeof(double)), *cs = calloc(m + 1, sizeof(double));
  double *sn = calloc(m + 1, sizeof(double));
  double *r = malloc(sizeof(double) * n), *w = malloc(sizeof(double) * n);
  double *t = malloc(sizeof(double) * n);
  int rc = 1;
  or_bmax(n, rp, ci, v, b, x, r);
  if (!flexible) { apply_M(M, Mctx, n, r, t); copy(n, t, r); }
  double res0 = sqrt(dot(n, r, r));
  double beta = res0;
  if (resvec) resvec[0] = res0;
  *usediter = 0;
  if (res0 == 0.0) { rc = 0; goto done; }
  int j = 0;
  while (j < maxiter) {
    memset(H, 0, sizeof(double) * (size_t)(m + 1) * m);
    memset(s, 0, sizeof(double) * (m + 1));
    memset(cs, 0, sizeof(double) * (m + 1));
    memset(sn, 0, sizeof(double) * (m + 1));
    scal(n, 1.0 / (beta + 1e-20), r);
    copy(n, r, V);
    s[0] = beta;
    int i = 0;
    while (i < m && j < maxiter) {
      double *vi = V + (size_t)i * n;
      if (flexible) {
        double *zi = Z + (size_t)i * n;
        apply_M(M, Mctx, n, vi, zi);
        or_spmv(n, rp, ci, v, zi, w);
      } else {
        or_spmv(n, rp, ci, v, vi, t);
        apply_M(M, Mctx, n, t, w);
      }
      for (int k = 0; k <= i; k++) {
        double h = dot(n, w, V + (size_t)k * n);
        H[k * m + i] = h;
        axpy(n, -h, V + (size_t)k * n, w);
      }
      double h1 = sqrt(dot(n, w, w));
      H[(i + 1) * m + i] = h1;
      scal(n, 1.0 / (h1 + 1e-20), w);
      copy(n, w, V + (size_t)(i + 1) * n);
      for (int k = 0; k < i; k++) rot_apply(&H[k * m + i], &H[(k + 1) * m + i], cs[k], sn[k]);
      rot_gen(H[i * m + i], H[(i + 1) * m + i], &cs[i], &sn[i]);
      rot_apply(&H[i * m + i], &H[(i + 1) * m + i], cs[i], sn[i]);
      rot_apply(&s[i], &s[i + 1], cs[i], sn[i]);
      double err = fabs(s[i + 1]);
      if (resvec) resvec[j + 1] = err;
      *usediter = j + 1;
      if (!isfinite(err)) { rc = 2; goto done; }
      if (err / (res0 + 1e-20) < tol) {
        gm_update(n, x, i + 1, H, m, s, flexible ? Z : V);
        rc = 0;
        goto done;
      }
      i++; j++;
    }
    gm_update(n, x, m, H, m, s, flexible ? Z : V);
    or_bmax(n, rp, ci, v, b, x, r);
    if (!flexible) { apply_M(M, Mctx, n, r, t); copy(n, t, r); }
    beta = fabs(s[m]);
  }
done:
  free(V); free(Z); free(H); free(s); free(cs); free(sn); free(r); free(w); free(t);
  return rc;
}

/* Independent systems (the diagonal blocks sys_ptr[s]..sys_ptr[s+1] of a
 * block-diagonal CSR matrix), each solved by or_cg with Jacobi
 * (intel/mpisolver-C-v1.cpp:193-272 per system; the batch itself has no
 * reference counterpart).  Systems run on OpenMP threads, one system per
 * thread at a time (the solves inside are serial), so results do not depend
 * on the thread count.  iters[s] = iterations, relres[s] = ||b - A x|| / ||b||
 * recomputed at the end; returns the number of systems that did not
 * converge. */
int or_batch_cg(int nsys, const int *sys_ptr, const int *rp, const int *ci, const double *v,
                const double *b, double *x, double tol, int maxiter, int *iters,
                double *relres) {
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
  for (int s = 0; s < nsys; s++) {
    const int r0 = sys_ptr[s], m = sys_ptr[s + 1] - r0;
    const int e0 = rp[r0], ne = rp[r0 + m] - e0;
    int *lrp = malloc(sizeof(int) * (size_t)(m + 1));
    int *lci = malloc(sizeof(int) * (size_t)(ne + 1));
    double *dinv = malloc(sizeof(double) * (size_t)(m + 1));
    double *res = malloc(sizeof(double) * (size_t)(m + 1));
    for (int i = 0; i <= m; i++) lrp[i] = rp[r0 + i] - e0;
    for (int e = 0; e < ne; e++) lci[e] = ci[e0 + e] - r0;
    const double *lv = v + e0;
    or_jac_ctx_t jac = {m, dinv};
    int it = 0, rc;
    if (or_jacobi_setup(m, lrp, lci, lv, 0.0, dinv)) rc = -1;
    else rc = or_cg(m, lrp, lci, lv, b + r0, x + r0, tol, 0.0, maxiter, or_jacobi_apply, &jac,
                    NULL, &it);
    or_bmax(m, lrp, lci, lv, b + r0, x + r0, res);
    double rn = 0.0, bn = 0.0;
    for (int i = 0; i < m; i++) { rn += res[i] * res[i]; bn += b[r0 + i] * b[r0 + i]; }
    iters[s] = it;
    relres[s] = bn > 0.0 ? sqrt(rn / bn) : 0.0;
    bad += rc != 0;
    free(lrp); free(lci); free(dinv); free(res);
  }
  return bad;
}

/* ------------------------------------------------------------------------ */
/* Driver: build the preconditioner named by pc and run the solver named by  */
/* ksp, the way hip/xsolver_c.cpp:150-292 wires them (Precond_jacobi,        */
/* Precond_ilu0 = seqilu0 + LUsolve, Precond_ilup = seqilut + LUsolve).      */
/* Returns the solver status, or -1 for an unknown name / setup failure.     */
/* ------------------------------------------------------------------------ */
/* ------------------------------------------------------------------------ */
/* The rest of the intel Krylov family (intel/mpisolver-C-v1.cpp), each step  */
/* for step, small = 1e-20 in every denominator as there.  Reported in this   */
/* oracle's conventions: resvec[0] = ||r0||, resvec[k] = ||r_k|| (absolute    */
/* norms of the residual each method monitors: M^-1 (b - A x) for the left-   */
/* preconditioned ones), *usediter = iterations performed (the reference      */
/* stores the 0-based index of the last one and relative norms).              */
/* ------------------------------------------------------------------------ */
#define OR_SMALL 1e-20

/* the convergence test of every intel loop: ||r|| / (||r0|| + small) < tol */
static int or_intel_check(double resl, double res0, double tol, double abstol, int it,
                          double *resvec, int *usediter, int *rc) {
  if (resvec) resvec[it] = resl;
  *usediter = it;
  if (!isfinite(resl)) { *rc = 2; return 1; }
  if (resl / (res0 + OR_SMALL) < tol || resl <= abstol) { *rc = 0; return 1; }
  return 0;
}

/* CG on M^-1 A (cpu_cg_solve, :26-102): r = M^-1 (b - A x); q = M^-1 A p;
 * sigma = p.q; tao = r.r */
int or_cg_left(int n, const int *rp, const int *ci, const double *v, const double *b,
               double *x, double tol, double abstol, int maxiter, or_precond_fn M, void *Mctx,
               double *resvec, int *usediter) {
  double *r = malloc(sizeof(double) * n), *t = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
  double *q = malloc(sizeof(double) * n);
  int rc = 1;
  or_bmax(n, rp, ci, v, b, x, t);
  apply_M(M, Mctx, n, t, r);
  double taoo = dot(n, r, r);
  double res0 = sqrt(taoo);
  copy(n, r, p);
  if (resvec) resvec[0] = res0;
  *usediter = 0;
  if (res0 == 0.0 || res0 <= abstol) { rc = 0; goto done; }
  for (int it = 0; it < maxiter; it++) {
    or_spmv(n, rp, ci, v, p, Ap);
    apply_M(M, Mctx, n, Ap, q);
    double sigma = dot(n, p, q);
    double alpha = taoo / (sigma + OR_SMALL);
    axpy(n, alpha, p, x);
    axpy(n, -alpha, q, r);
    double tao = dot(n, r, r);
    if (or_intel_check(sqrt(tao), res0, tol, abstol, it + 1, resvec, usediter, &rc)) goto done;
    double beta = tao / (taoo + OR_SMALL);
    taoo = tao;
    scal(n, beta, p);
    axpy(n, 1.0, r, p);
  }
done:
  free(r); free(t); free(p); free(Ap); free(q);
  return rc;
}

/* ICG (cpu_icg_solve, :104-192): CG on M^-1 A with sigma = p.q, rho = r.q,
 * theta = q.q in one reduction and the recurrence
 * tao = taoo - 2 alpha rho + alpha^2 theta for the next beta (the
 * convergence test still uses the computed r.r) */
int or_icg(int n, const int *rp, const int *ci, const double *v, const double *b, double *x,
           double tol, double abstol, int maxiter, or_precond_fn M, void *Mctx, double *resvec,
           int *usediter) {
  double *r = malloc(sizeof(double) * n), *t = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
  double *q = malloc(sizeof(double) * n);
  int rc = 1;
  or_bmax(n, rp, ci, v, b, x, t);
  apply_M(M, Mctx, n, t, r);
  double taoo = dot(n, r, r);
  double res0 = sqrt(taoo);
  copy(n, r, p);
  if (resvec) resvec[0] = res0;
  *usediter = 0;
  if (res0 == 0.0 || res0 <= abstol) { rc = 0; goto done; }
  for (int it = 0; it < maxiter; it++) {
    or_spmv(n, rp, ci, v, p, Ap);
    apply_M(M, Mctx, n, Ap, q);
    double sigma = dot(n, p, q), rho = dot(n, r, q), theta = dot(n, q, q);
    double alpha = taoo / (sigma + OR_SMALL);
    axpy(n, alpha, p, x);
    axpy(n, -alpha, q, r);
    double resl = dot(n, r, r);
    if (or_intel_check(sqrt(resl), res0, tol, abstol, it + 1, resvec, usediter, &rc)) goto done;
    double tao = taoo - 2.0 * alpha * rho + alpha * alpha * theta;
    double beta = tao / (taoo + OR_SMALL);
    taoo = tao;
    scal(n, beta, p);
    axpy(n, 1.0, r, p);
  }
done:
  free(r); free(t); free(p); free(Ap); free(q);
  return rc;
}

/* ICGM (cpu_icgm_solve, :273-352): preconditioned CG with one reduction per
 * iteration: sigma = p.Ap, rho = r.q, fai = Ap.z, theta = q.Ap, rr = r.r
 * (q = M^-1 A p, z = M^-1 r), tau' = tau - alpha fai - alpha rho +
 * alpha^2 theta; the test uses ||r|| from before the update (so x has moved
 * one step further than the residual it is judged by) and ||r0|| of the
 * unpreconditioned residual */
int or_icgm(int n, const int *rp, const int *ci, const double *v, const double *b, double *x,
            double tol, double abstol, int maxiter, or_precond_fn M, void *Mctx, double *resvec,
            int *usediter) {
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
  double *q = malloc(sizeof(double) * n);
  int rc = 1;
  or_bmax(n, rp, ci, v, b, x, r);
  apply_M(M, Mctx, n, r, z);
  double res0 = sqrt(dot(n, r, r));
  double tau = dot(n, r, z);
  copy(n, z, p);
  if (resvec) resvec[0] = res0;
  *usediter = 0;
  if (res0 == 0.0 || res0 <= abstol) { rc = 0; goto done; }
  for (int it = 0; it < maxiter; it++) {
    or_spmv(n, rp, ci, v, p, Ap);
    apply_M(M, Mctx, n, Ap, q);
    double sigma = dot(n, p, Ap), rho = dot(n, r, q), fai = dot(n, Ap, z);
    double theta = dot(n, q, Ap), rr = dot(n, r, r);
    double alpha = tau / (sigma + OR_SMALL);
    axpy(n, alpha, p, x);
    axpy(n, -alpha, Ap, r);
    apply_M(M, Mctx, n, r, z);
    double tauo = tau;
    tau = tau - alpha * fai - alpha * rho + alpha * alpha * theta;
    double beta = tau / (tauo + OR_SMALL);
    scal(n, beta, p);
    axpy(n, 1.0, z, p);
    if (or_intel_check(sqrt(rr), res0, tol, abstol, it + 1, resvec, usediter, &rc)) goto done;
  }
done:
  free(r); free(z); free(p); free(Ap); free(q);
  return rc;
}

/* Improved BiCGSTAB (cpu_ibicgstab_solve, :484-684), right-preconditioned in
 * y-space (x = x0 + M^-1 y), started from y = 0: t = A M^-1 r, p = 0
 * (:539-545).  The reference takes a nonzero initial guess as y0 (:512) and
 * its residual as b - A y0; here the guess is x0 and the correction is
 * solved from y = 0 -- identical when x0 = 0.
 * Per iteration two SpMVs, two preconditioner applications and six dots in
 * one reduction (:588-600) plus ||r||. */
int or_ibicgstab(int n, const int *rp, const int *ci, const double *v, const double *b,
                 double *x, double tol, double abstol, int maxiter, or_precond_fn M, void *Mctx,
                 double *resvec, int *usediter) {
  double *buf = calloc((size_t)n * 12, sizeof(double));
  double *res = buf, *reso = res + n, *tk = reso + n, *pk = tk + n, *qk = pk + n, *zk = qk + n;
  double *vk = zk + n, *vko = vk + n, *sk = vko + n, *tmp = sk + n, *phi = tmp + n;
  double *t2 = phi + n;
  int rc = 1;
  or_bmax(n, rp, ci, v, b, x, res);   /* x = y0 = 0: res = b */
  double res0 = sqrt(dot(n, res, res));
  copy(n, res, reso);
  double alpha = 1.0, alphao = 1.0, rou = 1.0, rouo = 1.0, omega = 1.0;
  double pi = 0.0, tau = 0.0, mu = 0.0;
  apply_M(M, Mctx, n, res, tmp);
  or_spmv(n, rp, ci, v, tmp, tk);
  double fai = dot(n, res, res), sigma = dot(n, res, tk);
  if (resvec) resvec[0] = res0;
  *usediter = 0;
  if (res0 == 0.0 || res0 <= abstol) { rc = 0; goto done; }
  for (int it = 0; it < maxiter; it++) {
    rouo = rou;
    rou = fai - omega * mu;
    double delta = (rou * alpha) / (rouo + OR_SMALL);
    double beta = delta / (omega + OR_SMALL);
    tau = sigma + beta * tau - delta * pi;
    alphao = alpha;
    alpha = rou / (tau + OR_SMALL);
    copy(n, vk, vko);
    copy(n, tk, vk);
    axpy(n, -omega, pk, vk);
    axpy(n, beta, vko, vk);
    axpy(n, -delta, qk, vk);
    apply_M(M, Mctx, n, vk, tmp);
    or_spmv(n, rp, ci, v, tmp, qk);
    copy(n, res, sk);
    axpy(n, -alpha, vk, sk);
    axpy(n, -omega, pk, tk);
    axpy(n, -alpha, qk, tk);
    double tmp1 = (beta * alpha) / (alphao + OR_SMALL), tmp2 = alpha * delta;
    scal(n, tmp1, zk);
    axpy(n, alpha, res, zk);
    axpy(n, -tmp2, vko, zk);
    apply_M(M, Mctx, n, tk, tmp);
    or_spmv(n, rp, ci, v, tmp, pk);
    fai = dot(n, reso, sk);
    pi = dot(n, reso, qk);
    double theta = dot(n, sk, tk), kappa = dot(n, tk, tk);
    mu = dot(n, reso, tk);
    double nu = dot(n, reso, pk);
    omega = theta / (kappa + OR_SMALL);
    sigma = mu - omega * nu;
    copy(n, sk, res);
    axpy(n, -omega, tk, res);
    axpy(n, 1.0, zk, phi);
    axpy(n, omega, sk, phi);
    double resl = sqrt(dot(n, res, res));
    if (or_intel_check(resl, res0, tol, abstol, it + 1, resvec, usediter, &rc)) break;
  }
  apply_M(M, Mctx, n, phi, t2);   /* sol = M^-1 y (:656, :667) */
  axpy(n, 1.0, t2, x);
done:
  free(buf);
  return rc;
}

/* Left-preconditioned GCR(restart) (cpu_gcr_solve, :685-810, classical
 * Gram-Schmidt: every coefficient from the unmodified c_k) and
 * (cpu_igcr_solve, :812-931, modified Gram-Schmidt): r = M^-1 (b - A x),
 * u_k = r, c_k = M^-1 A u_k orthogonalised against c_0..c_{k-1} (the same
 * coefficients applied to u_k), both scaled by 1/(||c_k|| + small), then
 * x += (c_k.r / (||c_k|| + small)) u_k, r -= (same) c_k.  After `restart`
 * directions k starts again at 0 (the stored directions are overwritten). */
int or_gcr_left(int n, const int *rp, const int *ci, const double *v, const double *b,
                double *x, double tol, double abstol, int maxiter, int restart, int mgs,
                or_precond_fn M, void *Mctx, double *resvec, int *usediter) {
  if (restart <= 0) restart = maxiter > 0 ? maxiter : 1;
  double *U = calloc((size_t)restart * n, sizeof(double));
  double *C = calloc((size_t)restart * n, sizeof(double));
  double *r = malloc(sizeof(double) * n), *t = malloc(sizeof(double) * n);
  double *al = calloc((size_t)restart, sizeof(double));
  int rc = 1, it = 0;
  or_bmax(n, rp, ci, v, b, x, t);
  apply_M(M, Mctx, n, t, r);
  double res0 = sqrt(dot(n, r, r));
  if (resvec) resvec[0] = res0;
  *usediter = 0;
  if (res0 == 0.0 || res0 <= abstol) { rc = 0; goto done; }
  while (it < maxiter) {
    for (int k = 0; k < restart && it < maxiter; k++) {
      double *uk = U + (size_t)k * n, *ck = C + (size_t)k * n;
      copy(n, r, uk);
      or_spmv(n, rp, ci, v, uk, t);
      apply_M(M, Mctx, n, t, ck);
      if (mgs) {
        for (int i = 0; i < k; i++) {
          double a = dot(n, C + (size_t)i * n, ck);
          axpy(n, -a, C + (size_t)i * n, ck);
          axpy(n, -a, U + (size_t)i * n, uk);
        }
      } else {
        for (int i = 0; i < k; i++) al[i] = dot(n, C + (size_t)i * n, ck);
        for (int i = 0; i < k; i++) {
          axpy(n, -al[i], C + (size_t)i * n, ck);
          axpy(n, -al[i], U + (size_t)i * n, uk);
        }
      }
      double cc = sqrt(dot(n, ck, ck)), cr = dot(n, ck, r);
      double sc = 1.0 / (cc + OR_SMALL);
      scal(n, sc, uk);
      scal(n, sc, ck);
      double step = cr / (cc + OR_SMALL);
      axpy(n, step, uk, x);
      axpy(n, -step, ck, r);
      it++;
      if (or_intel_check(sqrt(dot(n, r, r)), res0, tol, abstol, it, resvec, usediter, &rc))
        goto done;
    }
  }
done:
  free(U); free(C); free(r); free(t); free(al);
  return rc;
}

static int streq(const char *a, const char *b) { return strcmp(a, b) == 0; }
int or_parilut(int n, const int *arp, const int *aci, const double *av, int sweeps,
               double fill, double droptol, int cap, int *lrp, int *lci, double *lv,
               int *urp, int *uci, double *uv, int *trace);

int or_solve(const char *ksp, const char *pc, int n, const int *rp,
             const int *ci, const double *v, const double *b, double *x,
             double tol, double abstol, int maxiter, int restart,
             double ilut_tol, int ilut_fill, int lsweeps, int usweeps,
             double *resvec, int *usediter) {
  or_precond_fn M = NULL;
  void *Mctx = NULL;
  or_jac_ctx_t jac;
  or_lu_t lu;
  or_sweep_t sw;
  double *dinv = NULL, *luval = NULL, *lv = NULL, *uv = NULL, *diag = NULL;
  int *uptr = NULL, *lrp = NULL, *lci = NULL, *urp = NULL, *uci = NULL;
  int *lucols = NULL, *lurows = NULL;
  int rc = -1;
  if (streq(pc, "none")) {
  } else if (streq(pc, "jacobi")) {
    dinv = malloc(sizeof(double) * n);
    if (or_jacobi_setup(n, rp, ci, v, 0.0, dinv)) goto out;
    jac.n = n; jac.dinv = dinv;
    M = or_jacobi_apply; Mctx = &jac;
  } else if (streq(pc, "ilu0") || streq(pc, "ilut")) {
    int cap, nnz = rp[n];
    const int *frp, *fci;
    uptr = malloc(sizeof(int) * (n + 1));
    dinv = malloc(sizeof(double) * n);
    if (streq(pc, "ilu0")) {
      cap = nnz;
      luval = malloc(sizeof(double) * nnz);
      if (or_ilu0(n, rp, ci, v, luval, uptr)) goto out;
      frp = rp; fci = ci;
    } else {
      cap = (2 * ilut_fill + 2) * n;
      luval = malloc(sizeof(double) * (size_t)cap);
      lucols = malloc(sizeof(int) * (size_t)cap);
      lurows = malloc(sizeof(int) * (n + 1));
      diag = malloc(sizeof(double) * n);
      if (or_ilut(n, rp, ci, v, ilut_fill, ilut_tol, cap, luval, lucols, lurows, uptr, diag)) goto out;
      frp = lurows; fci = lucols;
    }
    if (lsweeps > 0 || usweeps > 0) {
      /* iterative application (the reference's GPU apply) */
      sw.n = n; sw.rp = frp; sw.ci = fci; sw.uptr = uptr; sw.luval = luval;
      sw.kl = lsweeps; sw.ku = usweeps;
      M = or_sweep_apply; Mctx = &sw;
    } else {
      lrp = malloc(sizeof(int) * (n + 1)); urp = malloc(sizeof(int) * (n + 1));
      lci = malloc(sizeof(int) * (size_t)cap); uci = malloc(sizeof(int) * (size_t)cap);
      lv = malloc(sizeof(double) * (size_t)cap); uv = malloc(sizeof(double) * (size_t)cap);
      or_ilu_split(n, frp, fci, luval, uptr, lrp, lci, lv, urp, uci, uv, dinv);
      lu.n = n; lu.lrp = lrp; lu.lci = lci; lu.lv = lv;
      lu.urp = urp; lu.uci = uci; lu.uv = uv; lu.dinv = dinv;
      M = or_lusolve; Mctx = &lu;
    }
  } else if (streq(pc, "ic0") || streq(pc, "ict")) {
    /* IC0: smallnum = droptol = 1e-4 (precon_ic0.h:14-15); ICT: droptol =
     * ilut_tol (xsolver_pc_setict, precon_ict.h:27-30), maxfil = ilut_fill,
     * capped at n - 1 (precon_ict.cpp:3-4), smallnum = 1e-4 (the macro the
     * reference passes, HostMatrixCSR.cpp:331) */
    int nnz = rp[n];
    int maxfil = ilut_fill > n - 1 ? n - 1 : ilut_fill;
    int cap = streq(pc, "ic0") ? nnz : (maxfil + 1) * n;
    int *lp = malloc(sizeof(int) * (n + 1));
    int *lc = malloc(sizeof(int) * (size_t)cap);
    double *lvv = malloc(sizeof(double) * (size_t)cap);
    int frc = streq(pc, "ic0") ? or_ic0(n, rp, ci, v, 1e-4, 1e-4, lp, lc, lvv)
                               : or_ict(n, rp, ci, v, 1e-4, ilut_tol, maxfil, lp, lc, lvv);
    if (frc) { free(lp); free(lc); free(lvv); goto out; }
    lrp = malloc(sizeof(int) * (n + 1)); urp = malloc(sizeof(int) * (n + 1));
    lci = malloc(sizeof(int) * (size_t)cap); uci = malloc(sizeof(int) * (size_t)cap);
    lv = malloc(sizeof(double) * (size_t)cap); uv = malloc(sizeof(double) * (size_t)cap);
    dinv = malloc(sizeof(double) * n);
    or_ic_split(n, lp, lc, lvv, lrp, lci, lv, urp, uci, uv, dinv);
    free(lp); free(lc); free(lvv);
    lu.n = n; lu.lrp = lrp; lu.lci = lci; lu.lv = lv;
    lu.urp = urp; lu.uci = uci; lu.uv = uv; lu.dinv = dinv;
    M = or_lusolve; Mctx = &lu;
  } else if (streq(pc, "parilut")) {
    /* ParILUT: 10 outer steps (hip/Precond/precon_gpuilut.cpp:5), fill factor
     * ilut_fill, relative drop tolerance ilut_tol; split as seqilut does. */
    int cap = (int)((ilut_fill + 3.0) * rp[n]) + 2 * n + 16;
    lrp = malloc(sizeof(int) * (n + 1)); urp = malloc(sizeof(int) * (n + 1));
    lci = malloc(sizeof(int) * (size_t)cap); uci = malloc(sizeof(int) * (size_t)cap);
    lv = malloc(sizeof(double) * (size_t)cap); uv = malloc(sizeof(double) * (size_t)cap);
    dinv = malloc(sizeof(double) * n);
    if (or_parilut(n, rp, ci, v, 10, (double)ilut_fill, ilut_tol, cap, lrp, lci, lv, urp, uci, uv, NULL))
      goto out;
    /* U rows: diagonal first; strict part scaled by 1/u_ii, stored from the
     * second entry on (the split of hip/HostMatrix/HostMatrixCSR.cpp:230-268) */
    for (int i = 0; i < n; i++) {
      dinv[i] = 1.0 / uv[urp[i]];
      for (int e = urp[i] + 1; e < urp[i + 1]; e++) uv[e] *= dinv[i];
    }
    /* drop the diagonal from U's index range: shift each row's pointer */
    int *urp2 = malloc(sizeof(int) * (n + 1));
    int w = 0;
    urp2[0] = 0;
    for (int i = 0; i < n; i++) {
      for (int e = urp[i] + 1; e < urp[i + 1]; e++) { uci[w] = uci[e]; uv[w++] = uv[e]; }
      urp2[i + 1] = w;
    }
    memcpy(urp, urp2, sizeof(int) * (n + 1));
    free(urp2);
    lu.n = n; lu.lrp = lrp; lu.lci = lci; lu.lv = lv;
    lu.urp = urp; lu.uci = uci; lu.uv = uv; lu.dinv = dinv;
    M = or_lusolve; Mctx = &lu;
  } else {
    goto out;
  }
  if (streq(ksp, "cg") || streq(ksp, "cgm")) rc = or_cg(n, rp, ci, v, b, x, tol, abstol, maxiter, M, Mctx, resvec, usediter);
  else if (streq(ksp, "cg_left")) rc = or_cg_left(n, rp, ci, v, b, x, tol, abstol, maxiter, M, Mctx, resvec, usediter);
  else if (streq(ksp, "icg")) rc = or_icg(n, rp, ci, v, b, x, tol, abstol, maxiter, M, Mctx, resvec, usediter);
  else if (streq(ksp, "icgm")) rc = or_icgm(n, rp, ci, v, b, x, tol, abstol, maxiter, M, Mctx, resvec, usediter);
  else if (streq(ksp, "ibicgstab")) rc = or_ibicgstab(n, rp, ci, v, b, x, tol, abstol, maxiter, M, Mctx, resvec, usediter);
  else if (streq(ksp, "gcr_left")) rc = or_gcr_left(n, rp, ci, v, b, x, tol, abstol, maxiter, restart, 0, M, Mctx, resvec, usediter);
  else if (streq(ksp, "igcr_left")) rc = or_gcr_left(n, rp, ci, v, b, x, tol, abstol, maxiter, restart, 1, M, Mctx, resvec, usediter);
  else if (streq(ksp, "bicgstab")) rc = or_bicgstab(n, rp, ci, v, b, x, tol, abstol, maxiter, M, Mctx, resvec, usediter);
  else if (streq(ksp, "gcr")) rc = or_gcr(n, rp, ci, v, b, x, tol, abstol, maxiter, restart, M, Mctx, resvec, usediter);
  else if (streq(ksp, "gmres")) rc = or_gmres(n, rp, ci, v, b, x, tol, maxiter, restart > 0 ? restart : 30, 0, M, Mctx, resvec, usediter);
  else if (streq(ksp, "fgmres")) rc = or_gmres(n, rp, ci, v, b, x, tol, maxiter, restart > 0 ? restart : 30, 1, M, Mctx, resvec, usediter);
out:
  free(dinv); free(luval); free(lv); free(uv); free(diag); free(uptr);
  free(lrp); free(lci); free(urp); free(uci); free(lucols); free(lurows);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* ParILUT (deterministic, synchronous form)                                 */
/* ------------------------------------------------------------------------ */
/* Restates the structure of hip/Precond/parilut.cpp:621-1206 (parilut_clean):
 * per outer step s (of `sweeps`):
 *   1. candidates C_i = pattern(A_i) u L_i u U_i u (union over k in L_i of U_k)
 *      (the pattern of A - L U with L's unit diagonal, :760-767);
 *   2. new candidate entries are initialised from R = A - L U:
 *      l_ij = r_ij / u_jj, u_ij = r_ij; existing entries keep their values
 *      (embedA_intoB_L2/U2, :178-233);
 *   3. one fixed-point sweep over the candidates (LU_sweep, :915-919);
 *   4. selection (:921-1085), on the magnitudes of the unscaled factors:
 *      - nothing is dropped by rank unless the candidate count exceeds
 *        round(rate * nnz(A)) (:926-931), rate = 1 + (fill - 1)(s + 1)/sweeps
 *        (:692);
 *      - L's set holds its strict entries and the n unit diagonals (value 1,
 *        last in each L row, :91-123); U's set its entries with the pivots.
 *        keep = (int)(rate * base) with base = that factor's count in A,
 *        diagonal included (l0nnz / u0nnz, :681-682, :1008-1013); the
 *        threshold is the value of rank (count - keep) in ascending |v|
 *        (find_kth_element, select_kth.h:501-634);
 *      - an entry survives when |v| >= threshold; the diagonals always do
 *        (thresold_remove_nnz2 / thresold_remove2, :303-384);
 *      - additionally |v| < droptol * ||a_i||_2 is dropped, the dual-threshold
 *        rule of ilutp_final (ilut_diag_in_csr.cpp:67-80; the GPU ParILUT of
 *        the reference has no drop tolerance, droptol = 0 reproduces it);
 *   5. one fixed-point sweep over the kept pattern (L_sweep/U_sweep, :1155-1169).
 * Sweeps are Jacobi (old values in, new values out); the reference sweeps
 * asynchronously in place.  The sum over k in each sweep runs in ascending k.
 * Factors: L strict lower CSR (unit diagonal implied), U upper CSR including
 * the diagonal (first in each row).  Returns 0, -1 on capacity overflow.
 * The sweeps run on the symmetrically scaled matrix S A S,
 * S = |diag A|^{-1/2} (the reference sweeps the unscaled values; from that
 * start the synchronous sweep diverges on the 27-point lognormal systems);
 * ILU is equivariant under diagonal scaling (L = S^-1 L~ S, U = S^-1 U~ S^-1),
 * so the fixed points are unchanged, and the selection looks at the unscaled
 * values l_ij = l~_ij (s_j / s_i), u_ij = u~_ij / (s_i s_j), as the
 * reference's does.
 * trace (optional, 4 ints per step): strict-L candidates, U candidates
 * (pivots included), kept strict-L entries, kept U entries (pivots included). */
typedef struct { int n; int *rp, *ci; double *v; int cap; } or_fac_t;

static int fac_find(const int *ci, int a, int b, int col) {
  while (a < b) {
    int m = (a + b) >> 1;
    if (ci[m] < col) a = m + 1; else b = m;
  }
  return a;
}
static double fac_get(const or_fac_t *F, int row, int col) {
  int a = F->rp[row], b = F->rp[row + 1];
  int p = fac_find(F->ci, a, b, col);
  return (p < b && F->ci[p] == col) ? F->v[p] : 0.0;
}
/* S_ij = sum_{k in L_i, k < min(i,j)} l_ik u_kj */
static double lu_partial(const or_fac_t *L, const or_fac_t *U, int i, int j) {
  int m = i < j ? i : j;
  double s = 0.0;
  for (int e = L->rp[i]; e < L->rp[i + 1]; e++) {
    int k = L->ci[e];
    if (k >= m) break;
    s += L->v[e] * fac_get(U, k, j);
  }
  return s;
}
static double clamp_piv(double d) {
  if (fabs(d) < 1e-10) d = d >= 0 ? 1e-10 : -1e-10;
  return d;
}
/* one Jacobi sweep over the current patterns of L and U (in place via copies) */
static void fac_sweep(int n, const int *arp, const int *aci, const double *av,
                      or_fac_t *L, or_fac_t *U) {
  double *lv = malloc(sizeof(double) * (size_t)(L->rp[n] + 1));
  double *uv = malloc(sizeof(double) * (size_t)(U->rp[n] + 1));
  for (int i = 0; i < n; i++) {
    for (int e = L->rp[i]; e < L->rp[i + 1]; e++) {
      int j = L->ci[e];
      double a = 0.0;
      int p = fac_find(aci, arp[i], arp[i + 1], j);
      if (p < arp[i + 1] && aci[p] == j) a = av[p];
      double t = (a - lu_partial(L, U, i, j)) / clamp_piv(U->v[U->rp[j]]);
      lv[e] = isfinite(t) ? t : L->v[e];
    }
    for (int e = U->rp[i]; e < U->rp[i + 1]; e++) {
      int j = U->ci[e];
      double a = 0.0;
      int p = fac_find(aci, arp[i], arp[i + 1], j);
      if (p < arp[i + 1] && aci[p] == j) a = av[p];
      double t = a - lu_partial(L, U, i, j);
      if (i == j) t = clamp_piv(t);
      uv[e] = isfinite(t) ? t : U->v[e];
    }
  }
  memcpy(L->v, lv, sizeof(double) * (size_t)L->rp[n]);
  memcpy(U->v, uv, sizeof(double) * (size_t)U->rp[n]);
  free(lv); free(uv);
}
static int cmp_dbl(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}
/* value of rank k (0-based) in ascending |v| over the entries selected by
 * `mask` (NULL = all) */
static double kth_abs(const double *v, int m, int k) {
  double *t = malloc(sizeof(double) * (size_t)(m + 1));
  for (int i = 0; i < m; i++) t[i] = fabs(v[i]);
  qsort(t, (size_t)m, sizeof(double), cmp_dbl);
  double r = t[k];
  free(t);
  return r;
}

int or_parilut(int n, const int *arp, const int *aci, const double *av_in, int sweeps,
               double fill, double droptol, int cap, int *lrp, int *lci, double *lv,
               int *urp, int *uci, double *uv, int *trace) {
  or_fac_t L = {n, lrp, lci, lv, cap}, U = {n, urp, uci, uv, cap};
  double *sc = malloc(sizeof(double) * n);
  double *av = malloc(sizeof(double) * (size_t)(arp[n] + 1));
  for (int i = 0; i < n; i++) {
    double d = 0.0;
    for (int e = arp[i]; e < arp[i + 1]; e++) if (aci[e] == i) d = av_in[e];
    sc[i] = d != 0.0 ? 1.0 / sqrt(fabs(d)) : 1.0;
  }
  for (int i = 0; i < n; i++)
    for (int e = arp[i]; e < arp[i + 1]; e++) av[e] = av_in[e] * sc[i] * sc[aci[e]];
  int nl0 = 0, nu0 = 0;
  lrp[0] = urp[0] = 0;
  for (int i = 0; i < n; i++) {
    for (int e = arp[i]; e < arp[i + 1]; e++) {
      if (aci[e] < i) { lci[nl0] = aci[e]; lv[nl0++] = av[e]; }
      else { uci[nu0] = aci[e]; uv[nu0++] = av[e]; }
    }
    lrp[i + 1] = nl0; urp[i + 1] = nu0;
  }
  /* keep bases with the diagonals: nnz of L with its unit diagonal and of U
     (parilut.cpp:681-682) */
  const long baseL = (long)nl0 + n, baseU = (long)nu0;
  double *anorm = malloc(sizeof(double) * n);   /* ||a_i||_2 of the unscaled rows */
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int e = arp[i]; e < arp[i + 1]; e++) s += av_in[e] * av_in[e];
    anorm[i] = sqrt(s);
  }
  int *mark = malloc(sizeof(int) * n);
  int *list = malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) mark[i] = -1;
  int rc = 0;
  for (int s = 0; s < sweeps && rc == 0; s++) {
    /* 1-2: candidate pattern and embedded values (new arrays) */
    int *nlrp = malloc(sizeof(int) * (n + 1)), *nurp = malloc(sizeof(int) * (n + 1));
    int *nlci = malloc(sizeof(int) * (size_t)cap), *nuci = malloc(sizeof(int) * (size_t)cap);
    double *nlv = malloc(sizeof(double) * (size_t)cap), *nuv = malloc(sizeof(double) * (size_t)cap);
    int nl = 0, nu = 0;
    nlrp[0] = nurp[0] = 0;
    for (int i = 0; i < n && rc == 0; i++) {
      int cnt = 0;
#define ADDC(c) do { int c_ = (c); if (mark[c_] != i) { mark[c_] = i; list[cnt++] = c_; } } while (0)
      for (int e = arp[i]; e < arp[i + 1]; e++) ADDC(aci[e]);
      for (int e = L.rp[i]; e < L.rp[i + 1]; e++) {
        ADDC(L.ci[e]);
        int k = L.ci[e];
        for (int q = U.rp[k]; q < U.rp[k + 1]; q++) ADDC(U.ci[q]);
      }
      for (int e = U.rp[i]; e < U.rp[i + 1]; e++) ADDC(U.ci[e]);
#undef ADDC
      /* sort the row's candidate columns */
      for (int a = 1; a < cnt; a++) {
        int key = list[a], b = a - 1;
        while (b >= 0 && list[b] > key) { list[b + 1] = list[b]; b--; }
        list[b + 1] = key;
      }
      if (nl + cnt > cap || nu + cnt > cap) { rc = -1; break; }
      for (int t = 0; t < cnt; t++) {
        int j = list[t];
        double a = 0.0;
        int p = fac_find(aci, arp[i], arp[i + 1], j);
        if (p < arp[i + 1] && aci[p] == j) a = av[p];
        if (j < i) {
          int q = fac_find(L.ci, L.rp[i], L.rp[i + 1], j);
          double val;
          if (q < L.rp[i + 1] && L.ci[q] == j) val = L.v[q];
          else val = (a - lu_partial(&L, &U, i, j)) / clamp_piv(U.v[U.rp[j]]);
          nlci[nl] = j; nlv[nl++] = isfinite(val) ? val : 0.0;
        } else {
          int q = fac_find(U.ci, U.rp[i], U.rp[i + 1], j);
          double val;
          if (q < U.rp[i + 1] && U.ci[q] == j) val = U.v[q];
          else val = a - lu_partial(&L, &U, i, j);
          nuci[nu] = j; nuv[nu++] = isfinite(val) ? val : 0.0;
        }
      }
      nlrp[i + 1] = nl; nurp[i + 1] = nu;
    }
    if (rc) { free(nlrp); free(nurp); free(nlci); free(nuci); free(nlv); free(nuv); break; }
    or_fac_t CL = {n, nlrp, nlci, nlv, cap}, CU = {n, nurp, nuci, nuv, cap};
    /* 3: sweep over the candidates */
    fac_sweep(n, arp, aci, av, &CL, &CU);
    /* 4: thresholds on the unscaled magnitudes */
    double rate = 1.0 + (fill - 1.0) * (double)(s + 1) / (double)sweeps;
    long keepL = (long)(rate * (double)baseL), keepU = (long)(rate * (double)baseU);
    long keep_all = lround(rate * (double)arp[n]);
    double tL = 0.0, tU = 0.0;
    /* |l_ij| and |u_ij| of the unscaled factors; L's n unit diagonals after them */
    double *ml = malloc(sizeof(double) * (size_t)(nl + n + 1));
    double *mu = malloc(sizeof(double) * (size_t)(nu + 1));
    for (int i = 0; i < n; i++) {
      for (int e = nlrp[i]; e < nlrp[i + 1]; e++) ml[e] = fabs(nlv[e] * (sc[nlci[e]] / sc[i]));
      for (int e = nurp[i]; e < nurp[i + 1]; e++) mu[e] = fabs(nuv[e] / (sc[i] * sc[nuci[e]]));
      ml[nl + i] = 1.0;
    }
    if ((long)nl + nu - keep_all > 0) {
      if ((long)nl + n > keepL) tL = kth_abs(ml, nl + n, (int)(nl + n - keepL));
      if ((long)nu > keepU) tU = kth_abs(mu, nu, (int)(nu - keepU));
    }
    int ol = 0, ou = 0;
    for (int i = 0; i < n; i++) {
      double dt = droptol * anorm[i];
      for (int e = nlrp[i]; e < nlrp[i + 1]; e++)
        if (ml[e] >= tL && ml[e] >= dt) { L.ci[ol] = nlci[e]; L.v[ol++] = nlv[e]; }
      L.rp[i + 1] = ol;
      for (int e = nurp[i]; e < nurp[i + 1]; e++)
        if (nuci[e] == i || (mu[e] >= tU && mu[e] >= dt)) { U.ci[ou] = nuci[e]; U.v[ou++] = nuv[e]; }
      U.rp[i + 1] = ou;
    }
    free(ml); free(mu);
    if (trace) { trace[4 * s] = nl; trace[4 * s + 1] = nu; trace[4 * s + 2] = ol; trace[4 * s + 3] = ou; }
    free(nlrp); free(nurp); free(nlci); free(nuci); free(nlv); free(nuv);
    /* 5: sweep over the kept pattern */
    fac_sweep(n, arp, aci, av, &L, &U);
  }
  for (int i = 0; i < n; i++) {
    for (int e = L.rp[i]; e < L.rp[i + 1]; e++) L.v[e] *= sc[L.ci[e]] / sc[i];
    for (int e = U.rp[i]; e < U.rp[i + 1]; e++) U.v[e] /= sc[i] * sc[U.ci[e]];
  }
  free(anorm); free(mark); free(list); free(sc); free(av);
  return rc;
}
