// xs_ic.hip -- ParIC(0): incomplete Cholesky by fixed-point sweeps on the
// device (no reference counterpart on the GPU: the reference factors IC(0)
// only on the host, hip/Precond/ic0.cpp:29-67, and stops with "IC0/ICT
// precond can not run on GPU", hip/xsolver_c.cpp:174-177).
//
// The factor lives on the lower triangle of A's pattern (diagonal included),
// in the sweep-order buffers of the ParILU(0) machinery (xs_ilu::lc / as /
// cmap / app / apd / pl / pu).  Each sweep updates every lower entry from
// the previous iterate (the Chow--Patel fixed point of IC(0)):
//   i > j : l_ij = (a_ij - sum_{k<j} l_ik l_jk) / l_jj
//   i = j : l_ii = sqrt(a_ii - sum_{k<i} l_ik^2)
// with the sums in ascending k over the pattern, which is the IC(0) of
// ic0_fact at its fixed point.  A non-finite or non-positive pivot keeps the
// previous value.  The converged factor is written as combined ILU factors
// on A's full pattern -- L~ = L diag(l)^{-1} below the diagonal (unit
// diagonal implied), U = diag(l)^2 L~^T on and above it -- so L~ D L~^T =
// L L^T is applied by the ILU solves (exact sync-free SpTRSV or Jacobi
// sweeps) unchanged.  This needs a structurally symmetric pattern: an entry
// above the diagonal takes its value from its mirror below it.
#include "xs_device.h"
#include "xs_ic.h"

// mpos[k] = position of the mirror (ci[k], rowidx[k]) of entry k, or -1;
// *asym = 1 when some entry has no mirror.  Binary search in the sorted row.
__global__ void k_ic_mirror(int nnz, const int *__restrict__ rowidx, const int *__restrict__ rp,
                            const int *__restrict__ ci, int *__restrict__ mpos, int *asym) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int i = rowidx[k], j = ci[k];
  int lo = rp[j], hi = rp[j + 1];
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (ci[m] < i) lo = m + 1;
    else hi = m;
  }
  const int p = (lo < rp[j + 1] && ci[lo] == i) ? lo : -1;
  mpos[k] = p;
  if (p < 0) atomicOr(asym, 1);
}

void xs_ic_mirror(hipStream_t s, const xs_csr &A, const int *rowidx, int *mpos, int *asym) {
  XS_GUARD(A.rp && A.ci && rowidx && mpos && asym);
  hipLaunchKernelGGL(k_ic_mirror, dim3((A.nnz + 255) / 256), dim3(256), 0, s, A.nnz, rowidx,
                     A.rp, A.ci, mpos, asym);
}

// Update lists of the sweep: entry (i, j), j <= i, subtracts l_ik l_jk for
// the common columns k < j of L's rows i and j (positions il, jl in A's
// pattern, ascending k); pd = position of l_jj for j < i, -2 for the
// diagonal, -1 above it (no update).
template <bool FILL>
__global__ __launch_bounds__(XS_BLOCK) void k_ic_pairs(
    int nnz, const int *__restrict__ rowidx, const int *__restrict__ rp,
    const int *__restrict__ ci, const int *__restrict__ dpos, int *__restrict__ cnt,
    const int *__restrict__ pp, int *__restrict__ pl, int *__restrict__ pu,
    int *__restrict__ pd) {
  const int k = blockIdx.x * XS_BLOCK + threadIdx.x;
  if (k >= nnz) return;
  const int i = rowidx[k], j = ci[k];
  int c = 0;
  if (j <= i) {
    int il = rp[i], jl = rp[j];
    const int ie = dpos[i], je = dpos[j];
    const int base = FILL ? pp[k] : 0;
    while (il < ie && jl < je) {
      const int ki = ci[il], kj = ci[jl];
      if (ki >= j || kj >= j) break;
      if (ki == kj) {
        if (FILL) { pl[base + c] = il; pu[base + c] = jl; }
        ++c; ++il; ++jl;
      } else if (ki < kj) {
        ++il;
      } else {
        ++jl;
      }
    }
  }
  if (!FILL) {
    cnt[k] = c;
    pd[k] = j < i ? dpos[j] : (j == i ? -2 : -1);
  }
}

void xs_ic_pairs(hipStream_t s, const xs_csr &A, const xs_ilu &F, int *cnt, const int *pp,
                 int *pl, int *pu, int *pd) {
  XS_GUARD(A.rp && A.ci && F.rowidx && F.dpos);
  const dim3 grid((A.nnz + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
  if (!pl) {
    XS_GUARD(cnt && pd);
    hipLaunchKernelGGL(k_ic_pairs<false>, grid, blk, 0, s, A.nnz, F.rowidx, A.rp, A.ci, F.dpos,
                       cnt, nullptr, nullptr, nullptr, pd);
  } else {
    XS_GUARD(pp && pu);
    hipLaunchKernelGGL(k_ic_pairs<true>, grid, blk, 0, s, A.nnz, F.rowidx, A.rp, A.ci, F.dpos,
                       nullptr, pp, pl, pu, nullptr);
  }
}

// Start values in sweep order: as = a_ij s_i s_j (the sweep's target), the
// iterate lc (and lc2) = the same below the diagonal and sqrt|a_ii s_i^2|
// on it (with the symmetric scaling that is 1).
__global__ void k_ic_start(int nnz, const int *__restrict__ rowidx, const int *__restrict__ ci,
                           const double *__restrict__ a, const double *__restrict__ sc,
                           const int *__restrict__ cmap, double *__restrict__ as,
                           double *__restrict__ lc, double *__restrict__ lc2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int i = rowidx[k], j = ci[k];
  const double v = sc ? a[k] * sc[i] * sc[j] : a[k];
  const double l = i == j ? sqrt(fabs(v)) : v;
  const int p = cmap[k];
  as[p] = v;
  lc[p] = l;
  if (lc2) lc2[p] = l;
}

void xs_ic_start(hipStream_t s, const xs_csr &A, const xs_ilu &F, const double *sc, double *lc2) {
  XS_GUARD(A.ci && A.v && F.rowidx && F.cmap && F.as && F.lc);
  hipLaunchKernelGGL(k_ic_start, dim3((A.nnz + 255) / 256), dim3(256), 0, s, A.nnz, F.rowidx,
                     A.ci, A.v, sc, F.cmap, F.as, F.lc, lc2);
}

// One sweep over the active (lower) entries in sweep order; lc_in == lc_out
// gives the in-place asynchronous form.
__global__ __launch_bounds__(XS_BLOCK) void k_ic_sweep_pairs(
    int nact, const double *__restrict__ as, const int *__restrict__ app,
    const int *__restrict__ pl, const int *__restrict__ pu, const int *__restrict__ apd,
    const double *lc_in, double *lc_out) {
  const int t = xcd_chunk((int)blockIdx.x, (int)gridDim.x) * XS_BLOCK + (int)threadIdx.x;
  if (t >= nact) return;
  const int e0 = app[t], e1 = app[t + 1], d = apd[t];
  double s = as[t];
  for (int e = e0; e < e1; ++e) s -= lc_in[pl[e]] * lc_in[pu[e]];
  double nv;
  if (d >= 0) {
    nv = s / lc_in[d];
  } else {
    nv = s > 0.0 ? sqrt(s) : __builtin_nan("");
  }
  lc_out[t] = isfinite(nv) ? nv : lc_in[t];
}

void xs_ic_sweep_pairs(hipStream_t s, const xs_ilu &F, const double *lc_in, double *lc_out) {
  XS_GUARD(F.app && F.apd && F.pl && F.pu && F.as && lc_in && lc_out);
  if (F.nact == 0) return;
  hipLaunchKernelGGL(k_ic_sweep_pairs, dim3((F.nact + XS_BLOCK - 1) / XS_BLOCK), dim3(XS_BLOCK),
                     0, s, F.nact, F.as, F.app, F.pl, F.pu, F.apd, lc_in, lc_out);
}

// Combined factors on A's pattern from the (scaled) Cholesky iterate: the
// sweeps factor S A S = L' L'^T, so A = L L^T with l_ij = l'_ij / s_i, and
//   j < i : L~_ij = l_ij / l_jj
//   j = i : u_ii  = l_ii^2
//   j > i : u_ij  = l_ii l_ji            (the mirror below the diagonal)
__global__ void k_ic_finish(int nnz, const int *__restrict__ rowidx, const int *__restrict__ ci,
                            const int *__restrict__ dpos, const int *__restrict__ mpos,
                            const double *__restrict__ sc, const int *__restrict__ cmap,
                            const double *__restrict__ lc, double *__restrict__ lu) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int i = rowidx[k], j = ci[k];
  const double si = sc ? sc[i] : 1.0;
  const double lii = lc[cmap[dpos[i]]] / si;
  double v;
  if (j < i) {
    const double sj = sc ? sc[j] : 1.0;
    v = (lc[cmap[k]] / si) / (lc[cmap[dpos[j]]] / sj);
  } else if (j == i) {
    v = lii * lii;
  } else {
    const double sj = sc ? sc[j] : 1.0;
    v = lii * (lc[cmap[mpos[k]]] / sj);
  }
  lu[k] = v;
}

void xs_ic_finish(hipStream_t s, const xs_csr &A, const xs_ilu &F, const int *mpos,
                  const double *sc, const double *lc) {
  XS_GUARD(A.ci && F.rowidx && F.dpos && F.cmap && F.lu && mpos && lc);
  hipLaunchKernelGGL(k_ic_finish, dim3((A.nnz + 255) / 256), dim3(256), 0, s, A.nnz, F.rowidx,
                     A.ci, F.dpos, mpos, sc, F.cmap, lc, F.lu);
}
