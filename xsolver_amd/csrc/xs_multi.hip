// xs_multi.hip -- k right-hand sides at once.
//
// Vectors are n x k row-major blocks (X[i*k + j]): the k values that one
// matrix entry multiplies are contiguous, so the SpMM and the triangular
// solves read the matrix once for all k columns (SpMM bytes:
// 12 nnz + 4(n+1) + 16 n k).  A workgroup of XS_BLOCK threads covers
// XS_BLOCK / k rows x k columns (thread t: column t % k, row t / k), so a
// row block of the vectors is one contiguous, coalesced segment.
// Every column keeps its own Krylov scalars and convergence flag; a column
// that has converged stops updating while the others go on, and once every
// column is done all kernels exit at entry.
// No reference counterpart (the reference solves one right-hand side).
#include <algorithm>
#include <climits>
#include <cstdint>
#include <mutex>
#include <unordered_map>

#include "xs_device.h"
#include "xs_multi.h"

__device__ __forceinline__ double &MSC(xs_mstate *ms, int slot, int j) { return ms->sc[slot][j]; }

__device__ inline void m_conv(const xs_mred &R, int j, double resl) {
  xs_mstate *ms = R.ms;
  MSC(ms, MS_RESL, j) = resl;
  if (!isfinite(resl)) ms->done[j] = 2;
  else if (resl / MSC(ms, MS_RES0, j) < R.tol || resl <= R.abstol) ms->done[j] = 1;
}

__device__ inline void m_res0(const xs_mred &R, int j, double r0) {
  xs_mstate *ms = R.ms;
  MSC(ms, MS_RES0, j) = r0;
  MSC(ms, MS_RESL, j) = r0;
  if (R.resvec) R.resvec[j] = r0;
  if (!isfinite(r0)) ms->done[j] = 2;
  else if (r0 == 0.0 || r0 <= R.abstol) ms->done[j] = 1;
}

// Per-column scalar updates; the same formulas as the single right-hand
// side finishing ops (hip/Solver/xsolver_bicgstab.cpp:36-103,
// hip/Solver/xsolver_cg.cpp:28-58).
__device__ inline void m_finish(const xs_mred &R, int j, int k, const double *tot, int NV) {
  xs_mstate *ms = R.ms;
  if (ms->done[j]) return;
  switch (R.op) {
    case MF_BICG_RES0:
      m_res0(R, j, sqrt(tot[0]));
      MSC(ms, MS_ALPHA, j) = 1.0;
      MSC(ms, MS_GAMA, j) = 1.0;
      MSC(ms, MS_BETA, j) = tot[0];
      MSC(ms, MS_OMEGA, j) = (tot[0] * 1.0) / (1.0 * 1.0 + 1e-40);
      MSC(ms, MS_BETO, j) = tot[0];
      break;
    case MF_BICG_GAMA:
      MSC(ms, MS_GAMA, j) = MSC(ms, MS_BETA, j) / tot[0];
      break;
    case MF_BICG_ALPHA:
      MSC(ms, MS_ALPHA, j) = tot[1] != 0.0 ? tot[0] / tot[1] : 0.0;   // v = 0: s solved exactly
      break;
    case MF_BICG_RES: {
      const double resl = sqrt(tot[0]);
      const int it = ms->iter[j] + 1;
      ms->iter[j] = it;
      if (R.resvec) R.resvec[(size_t)it * k + j] = resl;
      m_conv(R, j, resl);
      const double beta = tot[1];
      MSC(ms, MS_OMEGA, j) =
          (beta * MSC(ms, MS_GAMA, j)) / (MSC(ms, MS_ALPHA, j) * MSC(ms, MS_BETO, j) + 1e-40);
      MSC(ms, MS_BETO, j) = beta;
      MSC(ms, MS_BETA, j) = beta;
      break;
    }
    case MF_CG_INIT:
      MSC(ms, MS_TAU, j) = tot[0];
      m_res0(R, j, sqrt(tot[1]));
      break;
    case MF_CG_ALPHA:
      MSC(ms, MS_CGALPHA, j) = MSC(ms, MS_TAU, j) / tot[0];
      break;
    case MF_CG_RR:
      MSC(ms, MS_RR, j) = tot[0];
      break;
    case MF_CG_RES: {
      const double taoo = MSC(ms, MS_TAU, j);
      MSC(ms, MS_TAU, j) = tot[0];
      MSC(ms, MS_CGBETA, j) = tot[0] / taoo;
      const double rr = NV > 1 ? tot[1] : MSC(ms, MS_RR, j);
      const double resl = sqrt(rr);
      const int it = ms->iter[j] + 1;
      ms->iter[j] = it;
      if (R.resvec) R.resvec[(size_t)it * k + j] = resl;
      m_conv(R, j, resl);
      break;
    }
    default:
      break;
  }
}

// Column sums of the workgroup's (row, column) values, published per
// workgroup with agent-scope stores into slots that hold XS_SENTINEL between
// reductions; the workgroup with the highest index waits for every slot,
// combines the partials in a fixed order (deterministic), writes the sentinel
// back and runs the finishing op for every column (the grid-reduction scheme
// of reduce_finish, no fence and no counter).
// Cross-workgroup part: threads t < k hold the workgroup's column sums bs.
template <int NV>
__device__ __forceinline__ void mcol_cross(int k, const double (&bs)[NV], const xs_mred &R) {
  __shared__ double red[NV][XS_BLOCK];
  const int per = XS_BLOCK / k;
  if ((int)threadIdx.x < k)
    for (int q = 0; q < NV; ++q)
      st_ready(&R.partial[((size_t)blockIdx.x * NV + q) * k + threadIdx.x], bs[q]);
  if (blockIdx.x != gridDim.x - 1) return;
  const int j = threadIdx.x % k, g = threadIdx.x / k;
  double s[NV];
  for (int q = 0; q < NV; ++q) s[q] = 0.0;
  if (g < per) {
    for (unsigned b = g; b < gridDim.x; b += per) {
      unsigned long long u[NV];
      for (int q = 0; q < NV; ++q) u[q] = ld_relaxed(&R.partial[((size_t)b * NV + q) * k + j]);
      for (int q = 0; q < NV; ++q) {
        double *slot = &R.partial[((size_t)b * NV + q) * k + j];
        s[q] += u[q] == XS_SENTINEL ? ld_ready(slot, R.counter) : __longlong_as_double((long long)u[q]);
        __hip_atomic_store(reinterpret_cast<unsigned long long *>(slot), XS_SENTINEL,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
  __syncthreads();
  for (int q = 0; q < NV; ++q) red[q][threadIdx.x] = s[q];
  __syncthreads();
  if ((int)threadIdx.x < k) {
    double tot[2] = {0.0, 0.0};
    for (int q = 0; q < NV; ++q)
      for (int r = 0; r < per; ++r) tot[q] += red[q][r * k + threadIdx.x];
    m_finish(R, threadIdx.x, k, tot, NV);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int all = 1;
    for (int c = 0; c < k; ++c) all &= (R.ms->done[c] != 0);
    R.ms->all = all;
  }
}

// Thread t holds the values of column t % k (rows t / k).
template <int NV>
__device__ __forceinline__ void mcol_reduce(int k, const double (&v)[NV], const xs_mred &R) {
  __shared__ double rowv[NV][XS_BLOCK];
  const int per = XS_BLOCK / k;
  for (int q = 0; q < NV; ++q) rowv[q][threadIdx.x] = v[q];
  __syncthreads();
  double bs[NV];
  for (int q = 0; q < NV; ++q) {
    bs[q] = 0.0;
    if ((int)threadIdx.x < k)
      for (int r = 0; r < per; ++r) bs[q] += rowv[q][r * k + threadIdx.x];
  }
  mcol_cross<NV>(k, bs, R);
}

// ---------------------------------------------------------------------------
// SpMM with epilogues: Y = A X (EPI_Y), + w.y (EPI_DOT1), + w.y and y.y
// (EPI_DOT2), Y = B - A X with y.y (EPI_BMAXN).  Columns already done are
// skipped.
// ---------------------------------------------------------------------------
template <int EPI>
__global__ __launch_bounds__(XS_BLOCK) void k_mspmm(int n, int k, const int *__restrict__ rp,
                                                    const int *__restrict__ ci,
                                                    const double *__restrict__ v,
                                                    const double *__restrict__ X,
                                                    double *__restrict__ Y,
                                                    const double *__restrict__ B,
                                                    const double *__restrict__ W, xs_mred R,
                                                    const xs_mstate *ms) {
  if (ms && ms->all) return;
  constexpr int NV = (EPI == EPI_DOT2) ? 2 : 1;
  const int per = XS_BLOCK / k;
  const int j = threadIdx.x % k, rl = threadIdx.x / k;
  double acc[NV];
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  if (rl < per && !(ms && ms->done[j])) {
    for (int row = blockIdx.x * per + rl; row < n; row += gridDim.x * per) {
      double s = 0.0;
      const int e1 = rp[row + 1];
      int e = rp[row];
      // four independent gathers in flight per thread, summed in entry order
      for (; e + 4 <= e1; e += 4) {
        const int c0 = ci[e], c1 = ci[e + 1], c2 = ci[e + 2], c3 = ci[e + 3];
        const double a0 = v[e], a1 = v[e + 1], a2 = v[e + 2], a3 = v[e + 3];
        const double x0 = X[(size_t)c0 * k + j], x1 = X[(size_t)c1 * k + j];
        const double x2 = X[(size_t)c2 * k + j], x3 = X[(size_t)c3 * k + j];
        s += a0 * x0;
        s += a1 * x1;
        s += a2 * x2;
        s += a3 * x3;
      }
      for (; e < e1; ++e) s += v[e] * X[(size_t)ci[e] * k + j];
      const size_t p = (size_t)row * k + j;
      if (EPI == EPI_BMAXN) {
        const double t = B[p] - s;
        Y[p] = t;
        acc[0] += t * t;
      } else {
        Y[p] = s;
        if (EPI == EPI_DOT1 || EPI == EPI_DOT2) acc[0] += W[p] * s;
        if (EPI == EPI_DOT2) acc[NV - 1] += s * s;
      }
    }
  }
  if (EPI == EPI_Y) return;
  mcol_reduce<NV>(k, acc, R);
}

// Slab entries k0 .. k0 + 3 of row i of an ELL copy: columns (-1 for
// padding) and values.
template <bool O16>
__device__ __forceinline__ void ell_batch(int i, int k0, int ld, int W, const int *__restrict__ eci,
                                          const short *__restrict__ eco,
                                          const double *__restrict__ ev, int (&col)[4],
                                          double (&a)[4]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (O16) {
      const int o = k0 + q < W ? (int)eco[(size_t)(k0 + q) * ld + i] : XS_OFF16_PAD;
      col[q] = o == XS_OFF16_PAD ? -1 : i + o;
    } else {
      col[q] = k0 + q < W ? eci[(size_t)(k0 + q) * ld + i] : -1;
    }
    a[q] = col[q] >= 0 ? ev[(size_t)(k0 + q) * ld + i] : 0.0;
  }
}

// SpMM over the ELL copy of A, CW adjacent columns per thread (k / CW
// threads per row): the row's slab entries are loaded once per CW columns
// and each gather moves CW contiguous values.  Per column the entry-order
// sum of k_mspmm.  Same epilogues; the workgroup's column sums go through
// LDS (rows in order) to the shared cross-workgroup combine.
template <int EPI, bool O16, int CW>
__global__ __launch_bounds__(XS_BLOCK) void k_mspmm_ell(
    int n, int k, int ld, int W, const int *__restrict__ eci, const short *__restrict__ eco,
    const double *__restrict__ ev, const double *__restrict__ X, double *__restrict__ Y,
    const double *__restrict__ B, const double *__restrict__ Wv, xs_mred R, const xs_mstate *ms) {
  if (ms && ms->all) return;
  constexpr int NV = (EPI == EPI_DOT2) ? 2 : 1;
  const int L = k / CW, per = XS_BLOCK / L;
  const int g = threadIdx.x % L, rl = threadIdx.x / L, j0 = g * CW;
  bool live[CW];
  bool any = false;
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    live[c] = rl < per && !(ms && ms->done[j0 + c]);
    any |= live[c];
  }
  double acc[NV][CW];
#pragma unroll
  for (int c = 0; c < CW; ++c)
    for (int q = 0; q < NV; ++q) acc[q][c] = 0.0;
  if (any) {
    // the first slab batch of the thread's next row is loaded while the
    // current row's gathers are in flight
    const int stride = (int)gridDim.x * per;
    int i = (int)blockIdx.x * per + rl;
    int ncol[4];
    double na[4];
    if (i < n) ell_batch<O16>(i, 0, ld, W, eci, eco, ev, ncol, na);
    for (; i < n; i += stride) {
      // the epilogue's operand row (B or W), loaded before the gathers
      double ep[CW];
#pragma unroll
      for (int c = 0; c < CW; ++c) ep[c] = 0.0;
      if (EPI != EPI_Y) {
        const double2 *src = reinterpret_cast<const double2 *>(
            (EPI == EPI_BMAXN ? B : Wv) + (size_t)i * k + j0);
#pragma unroll
        for (int h = 0; h < CW / 2; ++h) {
          const double2 t = src[h];
          ep[2 * h] = t.x;
          ep[2 * h + 1] = t.y;
        }
      }
      double s[CW];
#pragma unroll
      for (int c = 0; c < CW; ++c) s[c] = 0.0;
      for (int k0 = 0; k0 < W; k0 += 4) {
        int col[4];
        double a[4];
        if (k0 == 0) {
#pragma unroll
          for (int q = 0; q < 4; ++q) { col[q] = ncol[q]; a[q] = na[q]; }
          if (i + stride < n) ell_batch<O16>(i + stride, 0, ld, W, eci, eco, ev, ncol, na);
        } else {
          ell_batch<O16>(i, k0, ld, W, eci, eco, ev, col, a);
        }
        double xv[4][CW];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (col[q] >= 0) {
            const double2 *src = reinterpret_cast<const double2 *>(X + (size_t)col[q] * k + j0);
#pragma unroll
            for (int h = 0; h < CW / 2; ++h) {
              const double2 t = src[h];
              xv[q][2 * h] = t.x;
              xv[q][2 * h + 1] = t.y;
            }
          } else {
#pragma unroll
            for (int c = 0; c < CW; ++c) xv[q][c] = 0.0;
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (col[q] >= 0)
#pragma unroll
            for (int c = 0; c < CW; ++c) s[c] += a[q] * xv[q][c];
      }
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        if (!live[c]) continue;
        const size_t p = (size_t)i * k + j0 + c;
        if (EPI == EPI_BMAXN) {
          const double t = ep[c] - s[c];
          Y[p] = t;
          acc[0][c] += t * t;
        } else {
          Y[p] = s[c];
          if (EPI == EPI_DOT1 || EPI == EPI_DOT2) acc[0][c] += ep[c] * s[c];
          if (EPI == EPI_DOT2) acc[NV - 1][c] += s[c] * s[c];
        }
      }
    }
  }
  if (EPI == EPI_Y) return;
  // workgroup column sums: rows in increasing order (fixed)
  __shared__ double cs[NV][XS_BLOCK * CW];
  if (rl < per)
#pragma unroll
    for (int c = 0; c < CW; ++c)
      for (int q = 0; q < NV; ++q) cs[q][rl * k + j0 + c] = acc[q][c];
  __syncthreads();
  double bs[NV];
  for (int q = 0; q < NV; ++q) {
    bs[q] = 0.0;
    if ((int)threadIdx.x < k)
      for (int r = 0; r < per; ++r) bs[q] += cs[q][r * k + threadIdx.x];
  }
  mcol_cross<NV>(k, bs, R);
}

// a.b (and c.d) per column
template <int NV>
__global__ __launch_bounds__(XS_BLOCK) void k_mdot(int n, int k, const double *__restrict__ a,
                                                   const double *__restrict__ b,
                                                   const double *__restrict__ c,
                                                   const double *__restrict__ d, xs_mred R) {
  if (R.ms->all) return;
  const int per = XS_BLOCK / k;
  const int j = threadIdx.x % k, rl = threadIdx.x / k;
  double acc[NV];
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  if (rl < per && !R.ms->done[j]) {
    for (int row = blockIdx.x * per + rl; row < n; row += gridDim.x * per) {
      const size_t p = (size_t)row * k + j;
      acc[0] += a[p] * b[p];
      if (NV == 2) acc[NV - 1] += c[p] * d[p];
    }
  }
  mcol_reduce<NV>(k, acc, R);
}

// P = R + omega (P - alpha U); Z = D^{-1} P with Jacobi
__global__ __launch_bounds__(XS_BLOCK) void k_mbicg_p(int n, int k, const double *__restrict__ R,
                                                      const double *__restrict__ U,
                                                      double *__restrict__ P,
                                                      const double *__restrict__ dinv,
                                                      double *__restrict__ Z,
                                                      const xs_mstate *ms) {
  if (ms->all) return;
  const int per = XS_BLOCK / k;
  const int j = threadIdx.x % k, rl = threadIdx.x / k;
  if (rl >= per || ms->done[j]) return;
  const double om = ms->sc[MS_OMEGA][j], al = ms->sc[MS_ALPHA][j];
  for (int row = blockIdx.x * per + rl; row < n; row += gridDim.x * per) {
    const size_t p = (size_t)row * k + j;
    const double t = R[p] + om * (P[p] - al * U[p]);
    P[p] = t;
    if (dinv) Z[p] = dinv[row] * t;
  }
}

// S = R - gama U; Z1 = D^{-1} S with Jacobi
__global__ __launch_bounds__(XS_BLOCK) void k_mbicg_s(int n, int k, const double *__restrict__ R,
                                                      const double *__restrict__ U,
                                                      double *__restrict__ Sv,
                                                      const double *__restrict__ dinv,
                                                      double *__restrict__ Z1,
                                                      const xs_mstate *ms) {
  if (ms->all) return;
  const int per = XS_BLOCK / k;
  const int j = threadIdx.x % k, rl = threadIdx.x / k;
  if (rl >= per || ms->done[j]) return;
  const double ga = ms->sc[MS_GAMA][j];
  for (int row = blockIdx.x * per + rl; row < n; row += gridDim.x * per) {
    const size_t p = (size_t)row * k + j;
    const double t = R[p] - ga * U[p];
    Sv[p] = t;
    if (dinv) Z1[p] = dinv[row] * t;
  }
}

// X += gama Z + alpha Z1; R = S - alpha V; per column ||r|| and rhat.r
__global__ __launch_bounds__(XS_BLOCK) void k_mbicg_xr(int n, int k, double *__restrict__ X,
                                                       double *__restrict__ Rv,
                                                       const double *__restrict__ Z,
                                                       const double *__restrict__ Z1,
                                                       const double *__restrict__ Sv,
                                                       const double *__restrict__ V,
                                                       const double *__restrict__ Rh, xs_mred R) {
  xs_mstate *ms = R.ms;
  if (ms->all) return;
  const int per = XS_BLOCK / k;
  const int j = threadIdx.x % k, rl = threadIdx.x / k;
  double acc[2] = {0.0, 0.0};
  if (rl < per && !ms->done[j]) {
    const double ga = ms->sc[MS_GAMA][j], al = ms->sc[MS_ALPHA][j];
    for (int row = blockIdx.x * per + rl; row < n; row += gridDim.x * per) {
      const size_t p = (size_t)row * k + j;
      X[p] = X[p] + ga * Z[p] + al * Z1[p];
      const double ri = Sv[p] - al * V[p];
      Rv[p] = ri;
      acc[0] += ri * ri;
      acc[1] += Rh[p] * ri;
    }
  }
  mcol_reduce<2>(k, acc, R);
}

// X += alpha P; R -= alpha AP; with Jacobi also Z = D^{-1} R and (z.r, r.r),
// otherwise r.r only
template <bool JAC>
__global__ __launch_bounds__(XS_BLOCK) void k_mcg_xr(int n, int k, double *__restrict__ X,
                                                     double *__restrict__ Rv,
                                                     const double *__restrict__ P,
                                                     const double *__restrict__ AP,
                                                     const double *__restrict__ dinv,
                                                     double *__restrict__ Z, xs_mred R) {
  constexpr int NV = JAC ? 2 : 1;
  xs_mstate *ms = R.ms;
  if (ms->all) return;
  const int per = XS_BLOCK / k;
  const int j = threadIdx.x % k, rl = threadIdx.x / k;
  double acc[NV];
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  if (rl < per && !ms->done[j]) {
    const double al = ms->sc[MS_CGALPHA][j];
    for (int row = blockIdx.x * per + rl; row < n; row += gridDim.x * per) {
      const size_t p = (size_t)row * k + j;
      X[p] += al * P[p];
      const double ri = Rv[p] - al * AP[p];
      Rv[p] = ri;
      if (JAC) {
        const double zi = dinv[row] * ri;
        Z[p] = zi;
        acc[0] += zi * ri;
        acc[NV - 1] += ri * ri;
      } else {
        acc[0] += ri * ri;
      }
    }
  }
  mcol_reduce<NV>(k, acc, R);
}

// P = Z + beta P
__global__ __launch_bounds__(XS_BLOCK) void k_mcg_p(int n, int k, double *__restrict__ P,
                                                    const double *__restrict__ Z,
                                                    const xs_mstate *ms) {
  if (ms->all) return;
  const int per = XS_BLOCK / k;
  const int j = threadIdx.x % k, rl = threadIdx.x / k;
  if (rl >= per || ms->done[j]) return;
  const double be = ms->sc[MS_CGBETA][j];
  for (int row = blockIdx.x * per + rl; row < n; row += gridDim.x * per) {
    const size_t p = (size_t)row * k + j;
    P[p] = Z[p] + be * P[p];
  }
}

// Y = D^{-1} X (all columns)
__global__ __launch_bounds__(XS_BLOCK) void k_mscale(int n, int k, const double *__restrict__ d,
                                                     const double *__restrict__ X,
                                                     double *__restrict__ Y) {
  const int per = XS_BLOCK / k;
  const int j = threadIdx.x % k, rl = threadIdx.x / k;
  if (rl >= per) return;
  for (int row = blockIdx.x * per + rl; row < n; row += gridDim.x * per) {
    const size_t p = (size_t)row * k + j;
    Y[p] = d[row] * X[p];
  }
}

__global__ void k_mfill(size_t m, unsigned long long *__restrict__ p, unsigned long long val) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < m;
       q += (size_t)gridDim.x * blockDim.x)
    p[q] = val;
}

__global__ void k_mcopy(size_t m, const double *__restrict__ in, double *__restrict__ out) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < m;
       q += (size_t)gridDim.x * blockDim.x)
    out[q] = in[q];
}

// column-major (in[j*n + i]) <-> interleaved (out[i*k + j])
__global__ void k_mlayout(int n, int k, const double *__restrict__ in, double *__restrict__ out,
                          int to_interleaved) {
  const size_t m = (size_t)n * k;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < m;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t i = q / k, j = q % k;
    if (to_interleaved) out[q] = in[j * n + i];
    else out[j * n + i] = in[q];
  }
}

__global__ void k_mstate_init(xs_mstate *ms, int k) {
  const int t = threadIdx.x;
  for (int q = t; q < MS_NSLOT * XS_MAX_RHS; q += blockDim.x) (&ms->sc[0][0])[q] = 0.0;
  for (int q = t; q < XS_MAX_RHS; q += blockDim.x) {
    ms->done[q] = 0;
    ms->iter[q] = 0;
  }
  if (t == 0) {
    ms->all = 0;
    ms->k = k;
  }
}

// ---------------------------------------------------------------------------
// Sync-free triangular solve with KC right-hand sides: the ticket / sentinel
// scheme of the single right-hand side solve (each wave owns 64 consecutive
// rows, taken in ticket order; a row's value slot holds the sentinel until
// published), with each lane carrying its row for KC columns.  X must hold
// the sentinel in columns [j0, j0 + KC) when the kernel starts.
// ---------------------------------------------------------------------------
template <bool UPPER, int KC>
__global__ __launch_bounds__(XS_BLOCK) void k_mtrsv(int n, int k, int j0,
                                                    const int *__restrict__ rp,
                                                    const int *__restrict__ ci,
                                                    const double *__restrict__ lu,
                                                    const int *__restrict__ dpos,
                                                    const double *__restrict__ dinv,
                                                    const double *__restrict__ B, double *X,
                                                    unsigned *tickets,
                                                    const xs_tsched S,
                                                    const xs_mstate *ms) {
  if (ms && ms->all) return;
  unsigned *err = tickets + 4;
  const int lane = threadIdx.x & (XS_WAVE - 1);
  const unsigned nw = trsv_ngroups(n, S);
  for (;;) {
  unsigned base;
  const unsigned t = trsv_next(tickets, base);
  if (base >= nw) break;
  if (t < nw) {
    const unsigned w = S.order ? (unsigned)S.order[t] : t;
    int r0, r1;
    trsv_rows<UPPER>(n, S, w, r0, r1);
    const int nrows = r1 - r0;
    const bool act = lane < nrows;
    const int r = UPPER ? r1 - 1 - lane : r0 + lane;  // lanes along the solve order
    double acc[KC];
    double di = 1.0;
    int eb = 0, ee = 0;  // in-wave entries [eb, ee)
    int iw_c[XS_INW];    // first in-wave entries in solve order, loaded before any wait
    double iw_v[XS_INW];
#pragma unroll
    for (int q = 0; q < XS_INW; ++q) { iw_c[q] = -1; iw_v[q] = 0.0; }
#pragma unroll
    for (int q = 0; q < KC; ++q) acc[q] = 0.0;
    if (act) {
      const double *br = B + (size_t)r * k + j0;
#pragma unroll
      for (int q = 0; q < KC; ++q) acc[q] = br[q];
      if (dinv) di = dinv[r];
      int xb, xe;  // entries outside the wave [xb, xe)
      if (!UPPER) {
        xb = rp[r];
        const int end = dpos[r];
        xe = xb;
        while (xe < end && ci[xe] < r0) ++xe;
        eb = xe; ee = end;
      } else {
        const int beg = dpos[r] + 1;
        xe = rp[r + 1];
        xb = xe;
        while (xb > beg && ci[xb - 1] >= r1) --xb;
        eb = beg; ee = xb;
      }
#pragma unroll
      for (int q = 0; q < XS_INW; ++q) {
        const int e = UPPER ? ee - 1 - q : eb + q;
        if (q < ee - eb) { iw_c[q] = ci[e]; iw_v[q] = lu[e]; }
      }
      for (int e = xb; e < xe; ++e) {
        const double a = lu[e];
        const double *xc = X + (size_t)ci[e] * k + j0;
        // issue all KC loads, then re-poll the pending ones together
        unsigned long long u[KC];
#pragma unroll
        for (int q = 0; q < KC; ++q) u[q] = ld_relaxed(xc + q);
        unsigned polls = 0;
        for (;;) {
          bool pend = false;
#pragma unroll
          for (int q = 0; q < KC; ++q) pend |= u[q] == XS_SENTINEL;
          if (!pend) break;
          if (++polls > XS_SPIN_LIMIT) {
            atomicOr(err, 1u);
#pragma unroll
            for (int q = 0; q < KC; ++q)
              if (u[q] == XS_SENTINEL) u[q] = 0x7FF8000000000000ull;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
#pragma unroll
          for (int q = 0; q < KC; ++q)
            if (u[q] == XS_SENTINEL) u[q] = ld_relaxed(xc + q);
        }
#pragma unroll
        for (int q = 0; q < KC; ++q) acc[q] -= a * __longlong_as_double((long long)u[q]);
      }
    }
    const int nint = ee - eb;
    const int nb = UPPER ? r + 1 : r - 1;
    const bool chain = !act || nint == 0 || (nint == 1 && iw_c[0] == nb);
    if (__all(chain)) {
      // x_r = c + m x_nb per column, composed along the lanes
      double m = (act && nint == 1) ? -di * iw_v[0] : 0.0;
#pragma unroll
      for (int q = 0; q < KC; ++q) acc[q] = act ? di * acc[q] : 0.0;
      wave_affine_scan<KC>(acc, m);
    } else {
      // lane kk's rows are final once the lanes below are: broadcast them;
      // in-wave entries consumed from registers in solve order (as k_trsv)
      int left = act ? nint - XS_INW : 0;
      int enext = UPPER ? ee - 1 - XS_INW : eb + XS_INW;
      for (int kk = 0; kk < nrows; ++kk) {
        const int rk = UPPER ? r1 - 1 - kk : r0 + kk;
        const bool hit = iw_c[0] == rk;
        const double a = hit ? iw_v[0] : 0.0;
#pragma unroll
        for (int q = 0; q < KC; ++q) {
          const double xk = __shfl(acc[q] * di, kk, XS_WAVE);
          if (hit) acc[q] -= a * xk;
        }
        if (hit) {
#pragma unroll
          for (int q = 0; q + 1 < XS_INW; ++q) { iw_c[q] = iw_c[q + 1]; iw_v[q] = iw_v[q + 1]; }
          if (left > 0) {
            iw_c[XS_INW - 1] = ci[enext];
            iw_v[XS_INW - 1] = lu[enext];
            enext += UPPER ? -1 : 1;
            --left;
          } else {
            iw_c[XS_INW - 1] = -1;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) acc[q] *= di;
    }
    if (act) {
      double *xr = X + (size_t)r * k + j0;
#pragma unroll
      for (int q = 0; q < KC; ++q) st_ready(xr + q, acc[q]);
    }
  }
  }
  trsv_retire(tickets);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int mgrid_rows(int n, int k) {
  const int per = XS_BLOCK / k;
  long g = ((long)n + per - 1) / per;
  if (g > XS_MGRID) g = XS_MGRID;
  return g < 1 ? 1 : (int)g;
}
// kernels without a grid reduction: enough workgroups that the dependent
// loads of a row (slabs or row pointers, then gathers) overlap across waves
// instead of running back to back in a grid-stride loop
static inline int mgrid_rows_wide(int n, int k) {
  const int per = XS_BLOCK / k;
  long g = ((long)n + per - 1) / per;
  if (g > 65536) g = 65536;
  return g < 1 ? 1 : (int)g;
}
static inline int mgrid_flat(size_t m) {
  size_t g = (m + XS_BLOCK * 4 - 1) / (XS_BLOCK * 4);
  if (g > 8192) g = 8192;
  return g < 1 ? 1 : (int)g;
}
// k-column Jacobi sweep of a factor over its ELL copy (one pass, columns
// interleaved): CW adjacent columns per thread, k / CW threads per row; the
// row's slab entries are loaded once per CW columns and each gather moves CW
// contiguous values.  Per column the arithmetic of k_tri_sweep_ell (FUSED:
// sweeps 1 and 2 from zero), so every column equals the single right-hand
// side sweep.
template <bool UPPER, bool FUSED, bool O16, int CW>
__global__ __launch_bounds__(XS_BLOCK) void k_msweep_ell(int n, int k, int ld, int W,
                                                         const int *__restrict__ eci,
                                                         const short *__restrict__ eco,
                                                         const double *__restrict__ ev,
                                                         const int *__restrict__ ww,
                                                         const double *__restrict__ dinv,
                                                         const double *__restrict__ B,
                                                         const double *__restrict__ Yo,
                                                         double *__restrict__ Yn,
                                                         const xs_mstate *ms) {
  if (ms && ms->all) return;
  const int L = k / CW, per = XS_BLOCK / L;
  const int g = threadIdx.x % L, rl = threadIdx.x / L, j0 = g * CW;
  if (rl >= per) return;
  bool live[CW];
  bool any = false;
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    live[c] = !(ms && ms->done[j0 + c]);
    any |= live[c];
  }
  if (!any) return;
  // CW contiguous values of row r of an interleaved block
  auto load = [&](const double *base, int r, double (&out)[CW]) {
    const double *src = base + (size_t)r * k + j0;
    if (CW == 1) {
      out[0] = src[0];
    } else {
#pragma unroll
      for (int h = 0; h < CW / 2; ++h) {
        const double2 t = reinterpret_cast<const double2 *>(src)[h];
        out[2 * h] = t.x;
        out[2 * h + 1] = t.y;
      }
    }
  };
  // the first slab batch of the thread's next row is loaded while the
  // current row's gathers are in flight
  const int stride = (int)gridDim.x * per;
  int i = (int)blockIdx.x * per + rl;
  int ncol[4];
  double na[4];
  if (i < n) ell_batch<O16>(i, 0, ld, W, eci, eco, ev, ncol, na);
  for (; i < n; i += stride) {
    // this row's first batch; the next row's is loaded now, whether or not
    // this row walks its slabs
    int col0[4];
    double a0[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) { col0[q] = ncol[q]; a0[q] = na[q]; }
    if (i + stride < n) ell_batch<O16>(i + stride, 0, ld, W, eci, eco, ev, ncol, na);
    double bi[CW];
    load(B, i, bi);
    const double di = UPPER ? dinv[i] : 1.0;
    double r[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c) r[c] = 0.0;
    // slabs past the longest row of the row's 64-row group are padding
    const int Wi = ww ? ww[i / XS_WAVE] : W;
    for (int k0 = 0; k0 < Wi; k0 += 4) {
      int col[4];
      double a[4];
      if (k0 == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) { col[q] = col0[q]; a[q] = a0[q]; }
      } else {
        ell_batch<O16>(i, k0, ld, W, eci, eco, ev, col, a);
      }
      double y[4][CW];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (col[q] >= 0) {
          load(FUSED ? B : Yo, col[q], y[q]);
          if (FUSED) {   // U: the slab values are u_ij / u_jj (E.vs)
#pragma unroll
            for (int c = 0; c < CW; ++c) y[q][c] = y[q][c] - 0.0;
          }
        } else {
#pragma unroll
          for (int c = 0; c < CW; ++c) y[q][c] = 0.0;
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (col[q] >= 0)
#pragma unroll
          for (int c = 0; c < CW; ++c) r[c] += a[q] * y[q][c];
    }
#pragma unroll
    for (int c = 0; c < CW; ++c)
      if (live[c]) Yn[(size_t)i * k + j0 + c] = UPPER ? di * (bi[c] - r[c]) : bi[c] - r[c];
  }
}

static inline bool k_ok(int k) { return k >= 1 && k <= XS_MAX_RHS; }

void xs_mspmm(hipStream_t s, const xs_csr &A, int k, const double *X, double *Y, int epi,
              const double *B, const double *W, const xs_mred *red, const xs_mstate *ms) {
  XS_GUARD(A.rp && A.ci && A.v && X && Y && k_ok(k));
  XS_GUARD(epi == EPI_Y || (red && red->partial && red->counter && red->ms));
  XS_GUARD(epi != EPI_BMAXN || B);
  XS_GUARD(!(epi == EPI_DOT1 || epi == EPI_DOT2) || W);
  const dim3 grid(mgrid_rows(A.n, k)), blk(XS_BLOCK);
  const xs_mred R = red ? *red : xs_mred{};
  switch (epi) {
    case EPI_Y:
      hipLaunchKernelGGL(k_mspmm<EPI_Y>, grid, blk, 0, s, A.n, k, A.rp, A.ci, A.v, X, Y, B, W, R, ms);
      break;
    case EPI_DOT1:
      hipLaunchKernelGGL(k_mspmm<EPI_DOT1>, grid, blk, 0, s, A.n, k, A.rp, A.ci, A.v, X, Y, B, W, R, ms);
      break;
    case EPI_DOT2:
      hipLaunchKernelGGL(k_mspmm<EPI_DOT2>, grid, blk, 0, s, A.n, k, A.rp, A.ci, A.v, X, Y, B, W, R, ms);
      break;
    case EPI_BMAXN:
      hipLaunchKernelGGL(k_mspmm<EPI_BMAXN>, grid, blk, 0, s, A.n, k, A.rp, A.ci, A.v, X, Y, B, W, R, ms);
      break;
    default:
      XS_GUARD(false);
  }
}

// Workgroups of a grid-stride kernel that are resident at once (CUs x
// workgroups per CU at the kernel's register and LDS use), per kernel:
// launching no more than that leaves no partial second round of workgroups.
static int resident_blocks(const void *kernel) {
  static std::mutex mu;
  static std::unordered_map<unsigned long long, int> cache;   // (device, kernel)
  int dev = 0;
  (void)hipGetDevice(&dev);
  const unsigned long long key =
      (reinterpret_cast<uintptr_t>(kernel) << 8) ^ static_cast<unsigned long long>(dev & 0xFF);
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int cus = 0, per_cu = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, XS_BLOCK, 0);
  const int r = cus > 0 && per_cu > 0 ? cus * per_cu : XS_MGRID;
  cache.emplace(key, r);
  return r;
}

static dim3 ell_grid(const void *kernel, int n, int per) {
  long g = ((long)n + per - 1) / per;
  const long cap = std::min(XS_MGRID, resident_blocks(kernel));
  if (g > cap) g = cap;
  return dim3(g < 1 ? 1 : (unsigned)g);
}

// ELL SpMM: k even, every vector 16-byte aligned (checked by the caller)
template <bool O16, int CW>
static void mspmm_ell_launch(hipStream_t s, const xs_ell &E, int k, const double *X, double *Y,
                             int epi, const double *B, const double *W, const xs_mred &R,
                             const xs_mstate *ms) {
  const int per = XS_BLOCK / (k / CW);
  const dim3 blk(XS_BLOCK);
#define XS_MSPMM_ELL(EP)                                                                        \
  hipLaunchKernelGGL((k_mspmm_ell<EP, O16, CW>),                                               \
                     ell_grid(reinterpret_cast<const void *>(&k_mspmm_ell<EP, O16, CW>), E.n, per), \
                     blk, 0, s, E.n, k, E.ld, E.W, E.ci, E.co, E.v, X, Y, B, W, R, ms)
  switch (epi) {
    case EPI_Y: XS_MSPMM_ELL(EPI_Y); break;
    case EPI_DOT1: XS_MSPMM_ELL(EPI_DOT1); break;
    case EPI_DOT2: XS_MSPMM_ELL(EPI_DOT2); break;
    case EPI_BMAXN: XS_MSPMM_ELL(EPI_BMAXN); break;
    default: XS_GUARD(false);
  }
#undef XS_MSPMM_ELL
}

bool xs_mspmm_ell_ok(const xs_ell &E, int k, const void *const *vecs, int nvec) {
  if (!E.v || !(E.ci || E.co) || k < 2 || (k & 1) || k > XS_MAX_RHS) return false;
  for (int q = 0; q < nvec; ++q)
    if (vecs[q] && (reinterpret_cast<uintptr_t>(vecs[q]) & 15)) return false;
  return true;
}

void xs_mspmm_ell(hipStream_t s, const xs_ell &E, int k, const double *X, double *Y, int epi,
                  const double *B, const double *W, const xs_mred *red, const xs_mstate *ms,
                  int cw) {
  const void *vecs[4] = {X, Y, B, W};
  XS_GUARD(X && Y && xs_mspmm_ell_ok(E, k, vecs, 4));
  XS_GUARD(epi == EPI_Y || (red && red->partial && red->counter && red->ms));
  XS_GUARD(epi != EPI_BMAXN || B);
  XS_GUARD(!(epi == EPI_DOT1 || epi == EPI_DOT2) || W);
  const xs_mred R = red ? *red : xs_mred{};
  const bool o16 = E.co != nullptr;
  if (k % 4 == 0 && cw != 2) {
    if (o16) mspmm_ell_launch<true, 4>(s, E, k, X, Y, epi, B, W, R, ms);
    else mspmm_ell_launch<false, 4>(s, E, k, X, Y, epi, B, W, R, ms);
  } else {
    if (o16) mspmm_ell_launch<true, 2>(s, E, k, X, Y, epi, B, W, R, ms);
    else mspmm_ell_launch<false, 2>(s, E, k, X, Y, epi, B, W, R, ms);
  }
}

void xs_mdot(hipStream_t s, int n, int k, const double *a, const double *b, const double *c,
             const double *d, const xs_mred &red) {
  XS_GUARD(a && b && k_ok(k) && red.partial && red.counter && red.ms && (!c || d));
  const dim3 grid(mgrid_rows(n, k)), blk(XS_BLOCK);
  if (c) hipLaunchKernelGGL(k_mdot<2>, grid, blk, 0, s, n, k, a, b, c, d, red);
  else hipLaunchKernelGGL(k_mdot<1>, grid, blk, 0, s, n, k, a, b, c, d, red);
}

void xs_mbicg_p(hipStream_t s, int n, int k, const double *R, const double *U, double *P,
                const double *dinv, double *Z, const xs_mstate *ms) {
  XS_GUARD(R && U && P && ms && k_ok(k) && (!dinv || Z));
  hipLaunchKernelGGL(k_mbicg_p, dim3(mgrid_rows_wide(n, k)), dim3(XS_BLOCK), 0, s, n, k, R, U, P, dinv, Z, ms);
}
void xs_mbicg_s(hipStream_t s, int n, int k, const double *R, const double *U, double *Sv,
                const double *dinv, double *Z1, const xs_mstate *ms) {
  XS_GUARD(R && U && Sv && ms && k_ok(k) && (!dinv || Z1));
  hipLaunchKernelGGL(k_mbicg_s, dim3(mgrid_rows_wide(n, k)), dim3(XS_BLOCK), 0, s, n, k, R, U, Sv, dinv, Z1, ms);
}
void xs_mbicg_xr(hipStream_t s, int n, int k, double *X, double *R, const double *Z,
                 const double *Z1, const double *Sv, const double *V, const double *Rh,
                 const xs_mred &red) {
  XS_GUARD(X && R && Z && Z1 && Sv && V && Rh && k_ok(k) && red.partial && red.counter && red.ms);
  hipLaunchKernelGGL(k_mbicg_xr, dim3(mgrid_rows(n, k)), dim3(XS_BLOCK), 0, s, n, k, X, R, Z, Z1,
                     Sv, V, Rh, red);
}
void xs_mcg_xr(hipStream_t s, int n, int k, double *X, double *R, const double *P,
               const double *AP, const double *dinv, double *Z, const xs_mred &red) {
  XS_GUARD(X && R && P && AP && k_ok(k) && (!dinv || Z) && red.partial && red.counter && red.ms);
  const dim3 grid(mgrid_rows(n, k)), blk(XS_BLOCK);
  if (dinv) hipLaunchKernelGGL(k_mcg_xr<true>, grid, blk, 0, s, n, k, X, R, P, AP, dinv, Z, red);
  else hipLaunchKernelGGL(k_mcg_xr<false>, grid, blk, 0, s, n, k, X, R, P, AP, dinv, Z, red);
}
void xs_mcg_p(hipStream_t s, int n, int k, double *P, const double *Z, const xs_mstate *ms) {
  XS_GUARD(P && Z && ms && k_ok(k));
  hipLaunchKernelGGL(k_mcg_p, dim3(mgrid_rows_wide(n, k)), dim3(XS_BLOCK), 0, s, n, k, P, Z, ms);
}
void xs_mscale(hipStream_t s, int n, int k, const double *d, const double *X, double *Y) {
  XS_GUARD(d && X && Y && k_ok(k));
  hipLaunchKernelGGL(k_mscale, dim3(mgrid_rows_wide(n, k)), dim3(XS_BLOCK), 0, s, n, k, d, X, Y);
}
void xs_mlayout(hipStream_t s, int n, int k, const double *in, double *out, bool to_interleaved) {
  XS_GUARD(in && out && k_ok(k));
  hipLaunchKernelGGL(k_mlayout, dim3(mgrid_flat((size_t)n * k)), dim3(XS_BLOCK), 0, s, n, k, in,
                     out, to_interleaved ? 1 : 0);
}
void xs_msweep_ell(hipStream_t s, const xs_tell &E, int n, int k, bool upper, bool fused,
                   const double *dinv, const double *B, const double *Yo, double *Yn,
                   const xs_mstate *ms) {
  XS_GUARD((E.W == 0 || (E.ci && E.v)) && E.ld >= n && k_ok(k) && B && Yn && (!upper || dinv) &&
           (fused || Yo));
  XS_GUARD(!(upper && fused) || E.W == 0 || E.vs);
  const double *ev = (upper && fused) ? E.vs : E.v;
  // columns per thread: 4 or 2 when k allows and the blocks are 16-byte aligned
  const bool al = ((reinterpret_cast<uintptr_t>(B) | reinterpret_cast<uintptr_t>(Yo) |
                    reinterpret_cast<uintptr_t>(Yn)) & 15) == 0;
  const int cw = !al ? 1 : (k % 4 == 0 ? 4 : (k % 2 == 0 ? 2 : 1));
  const int per = XS_BLOCK / (k / cw);
  const dim3 blk(XS_BLOCK);
#define XS_MSK(U, FU, O, C)                                                                    \
  hipLaunchKernelGGL((k_msweep_ell<U, FU, O, C>),                                             \
                     ell_grid(reinterpret_cast<const void *>(&k_msweep_ell<U, FU, O, C>), n, per), \
                     blk, 0, s, n, k, E.ld, E.W, E.ci, E.co, ev, E.ww, dinv, B, Yo, Yn, ms)
#define XS_MSC(U, FU, O) do { if (cw == 4) XS_MSK(U, FU, O, 4); else if (cw == 2) XS_MSK(U, FU, O, 2); else XS_MSK(U, FU, O, 1); } while (0)
#define XS_MSE(U, FU) do { if (E.co) XS_MSC(U, FU, true); else XS_MSC(U, FU, false); } while (0)
  if (upper) { if (fused) XS_MSE(true, true); else XS_MSE(true, false); }
  else { if (fused) XS_MSE(false, true); else XS_MSE(false, false); }
#undef XS_MSE
#undef XS_MSC
#undef XS_MSK
}

void xs_mcopy(hipStream_t s, size_t m, const double *in, double *out) {
  XS_GUARD(in && out);
  hipLaunchKernelGGL(k_mcopy, dim3(mgrid_flat(m)), dim3(XS_BLOCK), 0, s, m, in, out);
}
void xs_mstate_init(hipStream_t s, xs_mstate *ms, int k) {
  XS_GUARD(ms && k_ok(k));
  hipLaunchKernelGGL(k_mstate_init, dim3(1), dim3(256), 0, s, ms, k);
}

template <bool UPPER>
static void mtrsv_chunk(hipStream_t s, dim3 grid, int kc, int n, int k, int j0, const int *rp,
                        const int *ci, const double *lu, const int *dpos, const double *dinv,
                        const double *B, double *X, unsigned *tickets, const xs_tsched &S,
                        const xs_mstate *ms) {
  const dim3 blk(XS_BLOCK);
#define XS_MT(KC) \
  hipLaunchKernelGGL((k_mtrsv<UPPER, KC>), grid, blk, 0, s, n, k, j0, rp, ci, lu, dpos, dinv, B, X, tickets, S, ms)
  switch (kc) {
    case 16: XS_MT(16); break;
    case 8: XS_MT(8); break;
    case 4: XS_MT(4); break;
    case 2: XS_MT(2); break;
    default: XS_MT(1); break;
  }
#undef XS_MT
}

// X = T^{-1} B for all k columns; the columns are carried in chunks of
// 16/8/4/2/1 per launch.  X is sentinel-filled here.
void xs_mtrsv(hipStream_t s, bool upper, int n, int k, const int *rp, const int *ci,
              const double *lu, const int *dpos, const double *dinv, const double *B, double *X,
              unsigned *tickets, const xs_tsched &S, const xs_mstate *ms) {
  XS_GUARD(rp && ci && lu && dpos && B && X && tickets && k_ok(k) && S.max_wg > 0 &&
           (!S.gb || S.ngroups > 0));
  const size_t m = (size_t)n * k;
  hipLaunchKernelGGL(k_mfill, dim3(mgrid_flat(m)), dim3(XS_BLOCK), 0, s, m,
                     reinterpret_cast<unsigned long long *>(X), XS_SENTINEL);
  const int nw = S.gb ? S.ngroups : (n + XS_WAVE - 1) / XS_WAVE;
  const int tk = (nw + (XS_BLOCK / XS_WAVE) - 1) / (XS_BLOCK / XS_WAVE);
  const dim3 grid(tk < S.max_wg ? tk : S.max_wg);
  for (int j0 = 0; j0 < k;) {
    const int left = k - j0;
    const int kc = left >= 16 ? 16 : left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
    if (upper) mtrsv_chunk<true>(s, grid, kc, n, k, j0, rp, ci, lu, dpos, dinv, B, X, tickets, S, ms);
    else mtrsv_chunk<false>(s, grid, kc, n, k, j0, rp, ci, lu, dpos, dinv, B, X, tickets, S, ms);
    j0 += kc;
  }
}
