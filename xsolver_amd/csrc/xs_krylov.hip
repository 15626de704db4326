// xs_krylov.hip -- tall-skinny kernels for the Krylov methods that keep a
// basis (GMRES/FGMRES Arnoldi, GCR), plus the FP64 Hessenberg / Givens work
// that stays on the device.
//
// The basis is stored as k contiguous vectors of length n (stride ld).  The
// orthogonalisation is classical Gram-Schmidt applied twice (CGS2): one pass
// computes all k inner products V^T w at once (w read once, each basis vector
// read once), one pass subtracts V h.  The reference orthogonalises with
// modified Gram-Schmidt, one dot/axpy pair per basis vector
// (intel/mpisolver-C-v1.cpp:1013-1020, hip/Solver/xsolver_igcr.cpp:55-61);
// CGS2 spans the same Krylov space with MGS-level orthogonality and needs
// O(1) instead of O(k) global reductions per step.
#include <cmath>

#include "xs_device.h"
#include "xs_krylov.h"

#define TS_ROWS 4  // rows per thread in the tall-skinny kernels

__device__ __forceinline__ double ts_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, XS_WAVE);
  return v;
}

// Publish this block's k partial sums (red[j][wave]) and let the block with
// the highest index reduce them in block order into out[0..k).  Partial slots
// hold XS_SENTINEL between reductions (the grid-reduction scheme of
// reduce_finish in xs_device.h: agent-scope stores, no fence, no counter).
__device__ __forceinline__ void ts_finish(int k, double (*red)[XS_BLOCK / XS_WAVE],
                                          double *partial, unsigned *err, double *out) {
  for (int j = threadIdx.x; j < k; j += XS_BLOCK) {
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < XS_BLOCK / XS_WAVE; ++q) a += red[j][q];
    st_ready(&partial[(size_t)blockIdx.x * k + j], a);
  }
  if (blockIdx.x != gridDim.x - 1) return;
  const int lane = threadIdx.x & (XS_WAVE - 1), wv = threadIdx.x / XS_WAVE;
  for (int j = wv; j < k; j += XS_BLOCK / XS_WAVE) {
    double a = 0.0;
    for (unsigned b0 = lane; b0 < gridDim.x; b0 += 4 * XS_WAVE) {
      unsigned long long u[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const unsigned b = b0 + q * XS_WAVE;
        u[q] = b < gridDim.x ? ld_relaxed(&partial[(size_t)b * k + j]) : 0ull;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const unsigned b = b0 + q * XS_WAVE;
        if (b < gridDim.x) {
          double *slot = &partial[(size_t)b * k + j];
          a += u[q] == XS_SENTINEL ? ld_ready(slot, err) : __longlong_as_double((long long)u[q]);
          __hip_atomic_store(reinterpret_cast<unsigned long long *>(slot), XS_SENTINEL,
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    a = ts_wave_sum(a);
    if (lane == 0) out[j] = a;
  }
}

// out[j] = V_j . w  for j < k (k <= XS_TS_KMAX); T = double or float
// vectors, products accumulated in double.
template <class T>
__global__ __launch_bounds__(XS_BLOCK) void k_ts_dot(int n, int k, const T *__restrict__ V,
                                                     size_t ld, const T *__restrict__ w,
                                                     double *partial, unsigned *counter,
                                                     double *out, const xs_state *st) {
  if (st && st->done) return;
  __shared__ double red[XS_TS_KMAX][XS_BLOCK / XS_WAVE];
  const int lane = threadIdx.x & (XS_WAVE - 1), wv = threadIdx.x / XS_WAVE;
  const int base = blockIdx.x * XS_BLOCK * TS_ROWS + threadIdx.x;
  double wr[TS_ROWS];
#pragma unroll
  for (int q = 0; q < TS_ROWS; ++q) {
    const int row = base + q * XS_BLOCK;
    wr[q] = row < n ? (double)w[row] : 0.0;
  }
  for (int j = 0; j < k; ++j) {
    const T *vj = V + (size_t)j * ld;
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < TS_ROWS; ++q) {
      const int row = base + q * XS_BLOCK;
      if (row < n) a += (double)vj[row] * wr[q];
    }
    a = ts_wave_sum(a);
    if (lane == 0) red[j][wv] = a;
  }
  __syncthreads();
  ts_finish(k, red, partial, counter, out);
}

// w -= sum_j h[j] V_j  (and u -= sum_j h[j] U_j when U is given), k <= XS_TS_KMAX;
// the result may be written to wout instead of w.
template <class T>
__global__ __launch_bounds__(XS_BLOCK) void k_ts_update(int n, int k, const T *__restrict__ V,
                                                        const T *__restrict__ U, size_t ld,
                                                        const double *h, T *w, T *wout,
                                                        T *u, double sign, const xs_state *st) {
  if (st && st->done) return;
  __shared__ double hs[XS_TS_KMAX];
  for (int j = threadIdx.x; j < k; j += XS_BLOCK) hs[j] = sign * h[j];
  __syncthreads();
  const int base = blockIdx.x * XS_BLOCK * TS_ROWS + threadIdx.x;
#pragma unroll
  for (int q = 0; q < TS_ROWS; ++q) {
    const int row = base + q * XS_BLOCK;
    if (row >= n) continue;
    double a = (double)w[row];
    for (int j = 0; j < k; ++j) a -= hs[j] * (double)V[(size_t)j * ld + row];
    wout[row] = (T)a;
    if (U) {
      double b = (double)u[row];
      for (int j = 0; j < k; ++j) b -= hs[j] * (double)U[(size_t)j * ld + row];
      u[row] = (T)b;
    }
  }
}

// dst = src * (*scale)
template <class T>
__global__ void k_ts_scale(int n, const T *__restrict__ src, T *__restrict__ dst,
                           const double *scale, const xs_state *st) {
  if (st && st->done) return;
  const double a = *scale;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = (T)((double)src[i] * a);
}

// ---------------------------------------------------------------------------
// Givens rotations (intel/mpisolver-C-v1.cpp:1249-1279)
// ---------------------------------------------------------------------------
__device__ inline void rot_apply(double &dx, double &dy, double cs, double sn) {
  const double t = cs * dx + sn * dy;
  dy = -sn * dx + cs * dy;
  dx = t;
}
__device__ inline void rot_gen(double dx, double dy, double &cs, double &sn) {
  if (dy == 0.0) {
    cs = 1.0; sn = 0.0;
  } else if (fabs(dy) > fabs(dx)) {
    const double t = dx / dy;
    sn = 1.0 / sqrt(1.0 + t * t);
    cs = t * sn;
  } else {
    const double t = dy / dx;
    cs = 1.0 / sqrt(1.0 + t * t);
    sn = t * cs;
  }
}

// Column i of the Hessenberg matrix from the two CGS passes and ||w||, the
// rotations, the residual estimate |s_{i+1}| and the convergence test
// (intel/mpisolver-C-v1.cpp:1013-1042).  One thread.
__global__ void k_gmres_column(int i, int m, xs_gmres_small g, xs_state *st, double *resvec) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (st->done) return;
  double *H = g.H;
  for (int r = 0; r <= i; ++r) H[r * m + i] = g.h1[r] + g.h2[r];
  const double hn = sqrt(g.misc[XS_G_NRM2]);
  H[(i + 1) * m + i] = hn;
  g.misc[XS_G_SCALE] = 1.0 / (hn + 1e-20);
  for (int r = 0; r < i; ++r) rot_apply(H[r * m + i], H[(r + 1) * m + i], g.cs[r], g.sn[r]);
  rot_gen(H[i * m + i], H[(i + 1) * m + i], g.cs[i], g.sn[i]);
  rot_apply(H[i * m + i], H[(i + 1) * m + i], g.cs[i], g.sn[i]);
  rot_apply(g.s[i], g.s[i + 1], g.cs[i], g.sn[i]);
  const double err = fabs(g.s[i + 1]);
  const int it = st->iter + 1;
  st->iter = it;
  if (resvec) resvec[it] = err;
  st->sc[S_RESL] = err;
  if (!isfinite(err)) st->done = 2;
  else if (err / (st->sc[S_RES0] + 1e-20) < st->sc[S_TOL] || err <= st->sc[S_ABSTOL]) st->done = 1;
}

// Start of a restart cycle: beta = sqrt(nrm2) on the first cycle (also sets
// res0 and resvec[0]), |s_m| afterwards (intel/mpisolver-C-v1.cpp:1069);
// clears H, s, cs, sn; s_0 = beta; scale = 1/(beta + 1e-20).  One thread.
__global__ void k_gmres_cycle(int m, int first, xs_gmres_small g, xs_state *st, double *resvec) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (st->done) return;
  double beta;
  if (first) {
    beta = sqrt(g.misc[XS_G_NRM2]);
    st->sc[S_RES0] = beta;
    st->sc[S_RESL] = beta;
    if (resvec) resvec[0] = beta;
    if (!isfinite(beta)) { st->done = 2; return; }
    if (beta == 0.0 || beta <= st->sc[S_ABSTOL]) { st->done = 1; return; }
  } else {
    beta = fabs(g.s[m]);
  }
  for (int q = 0; q < (m + 1) * m; ++q) g.H[q] = 0.0;
  for (int q = 0; q <= m; ++q) { g.s[q] = 0.0; g.cs[q] = 0.0; g.sn[q] = 0.0; }
  g.s[0] = beta;
  g.misc[XS_G_SCALE] = 1.0 / (beta + 1e-20);
}

// y = H(0:k,0:k)^{-1} s with H_ii + 1e-20 (intel/mpisolver-C-v1.cpp:1281-1297).
// Runs even after convergence (it computes the final update).
__global__ void k_gmres_backsub(int k, int m, xs_gmres_small g) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double *y = g.y;
  for (int i = 0; i < k; ++i) y[i] = g.s[i];
  for (int i = k - 1; i >= 0; --i) {
    y[i] = y[i] / (g.H[i * m + i] + 1e-20);
    for (int j = i - 1; j >= 0; --j) y[j] = y[j] - g.H[j * m + i] * y[i];
  }
}

// GCR step (hip/Solver/xsolver_igcr.cpp:62-80): t0 = ||c||, t1 = c.r;
// u /= t0, c /= t0, x += (t1/t0) u, r -= (t1/t0) c; partial ||r||^2.
__global__ __launch_bounds__(XS_BLOCK) void k_gcr_apply(int n, double *__restrict__ u,
                                                        double *__restrict__ c,
                                                        double *__restrict__ x,
                                                        double *__restrict__ r, double *t01,
                                                        double *partial, unsigned *counter,
                                                        xs_state *st, double *resvec) {
  if (st->done) return;
  const double t0 = sqrt(t01[0]);
  const double inv = 1.0 / t0, step = t01[1] / t0;
  double acc = 0.0;
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    const double us = u[i] * inv, cs = c[i] * inv;
    u[i] = us;
    c[i] = cs;
    x[i] += step * us;
    const double ri = r[i] - step * cs;
    r[i] = ri;
    acc += ri * ri;
  }
  __shared__ double red[1][XS_BLOCK / XS_WAVE];
  const int lane = threadIdx.x & (XS_WAVE - 1), wv = threadIdx.x / XS_WAVE;
  acc = ts_wave_sum(acc);
  if (lane == 0) red[0][wv] = acc;
  __syncthreads();
  ts_finish(1, red, partial, counter, t01 + 2);   // ||r||^2 -> t01[2]
}

// Convergence bookkeeping after k_gcr_apply (one thread).
__global__ void k_gcr_check(const double *t01, xs_state *st, double *resvec) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (st->done) return;
  const double resl = sqrt(t01[2]);
  const int it = st->iter + 1;
  st->iter = it;
  if (resvec) resvec[it] = resl;
  st->sc[S_RESL] = resl;
  if (!isfinite(resl)) st->done = 2;
  else if (resl / st->sc[S_RES0] < st->sc[S_TOL] || resl <= st->sc[S_ABSTOL]) st->done = 1;
}

// res0 for GCR from misc[NRM2]: ||r0|| (hip/Solver/xsolver_igcr.cpp:36-41)
__global__ void k_res0_from(const double *nrm2, xs_state *st, double *resvec) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double r0 = sqrt(*nrm2);
  st->sc[S_RES0] = r0;
  st->sc[S_RESL] = r0;
  if (resvec) resvec[0] = r0;
  if (!isfinite(r0)) st->done = 2;
  else if (r0 == 0.0 || r0 <= st->sc[S_ABSTOL]) st->done = 1;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int ts_grid(int n) { return (n + XS_BLOCK * TS_ROWS - 1) / (XS_BLOCK * TS_ROWS); }
int xs_ts_blocks(int n) { return ts_grid(n); }

template <class T>
void xs_ts_dot(hipStream_t s, int n, int k, const T *V, size_t ld, const T *w,
               double *partial, unsigned *counter, double *out, const xs_state *st) {
  for (int j0 = 0; j0 < k; j0 += XS_TS_KMAX) {
    const int kk = k - j0 < XS_TS_KMAX ? k - j0 : XS_TS_KMAX;
    hipLaunchKernelGGL(k_ts_dot<T>, dim3(ts_grid(n)), dim3(XS_BLOCK), 0, s, n, kk,
                       V + (size_t)j0 * ld, ld, w, partial, counter, out + j0, st);
  }
}

template <class T>
void xs_ts_update(hipStream_t s, int n, int k, const T *V, const T *U, size_t ld,
                  const double *h, T *w, T *wout, T *u, const xs_state *st) {
  const T *src = w;
  for (int j0 = 0; j0 < k; j0 += XS_TS_KMAX) {
    const int kk = k - j0 < XS_TS_KMAX ? k - j0 : XS_TS_KMAX;
    T *dst = (j0 + kk >= k) ? wout : w;
    hipLaunchKernelGGL(k_ts_update<T>, dim3(ts_grid(n)), dim3(XS_BLOCK), 0, s, n, kk,
                       V + (size_t)j0 * ld, U ? U + (size_t)j0 * ld : (const T *)nullptr, ld,
                       h + j0, const_cast<T *>(src), dst, u, 1.0, st);
    src = dst;
  }
}

template <class T>
void xs_ts_axpy(hipStream_t s, int n, int k, const T *V, size_t ld, const double *y, T *x) {
  // x += sum_j y_j V_j  ==  x -= sum_j (-y_j) V_j
  for (int j0 = 0; j0 < k; j0 += XS_TS_KMAX) {
    const int kk = k - j0 < XS_TS_KMAX ? k - j0 : XS_TS_KMAX;
    hipLaunchKernelGGL(k_ts_update<T>, dim3(ts_grid(n)), dim3(XS_BLOCK), 0, s, n, kk,
                       V + (size_t)j0 * ld, (const T *)nullptr, ld, y + j0, x, x, (T *)nullptr,
                       -1.0, (const xs_state *)nullptr);
  }
}

template <class T>
void xs_ts_scale(hipStream_t s, int n, const T *src, T *dst, const double *scale,
                 const xs_state *st) {
  int g = (n + XS_BLOCK * 4 - 1) / (XS_BLOCK * 4);
  if (g > 2048) g = 2048;
  hipLaunchKernelGGL(k_ts_scale<T>, dim3(g), dim3(XS_BLOCK), 0, s, n, src, dst, scale, st);
}

#define XS_TS_INST(T)                                                                       \
  template void xs_ts_dot<T>(hipStream_t, int, int, const T *, size_t, const T *, double *, \
                             unsigned *, double *, const xs_state *);                      \
  template void xs_ts_update<T>(hipStream_t, int, int, const T *, const T *, size_t,        \
                                const double *, T *, T *, T *, const xs_state *);          \
  template void xs_ts_axpy<T>(hipStream_t, int, int, const T *, size_t, const double *, T *); \
  template void xs_ts_scale<T>(hipStream_t, int, const T *, T *, const double *, const xs_state *);
XS_TS_INST(double)
XS_TS_INST(float)

void xs_gmres_column(hipStream_t s, int i, int m, const xs_gmres_small &g, xs_state *st,
                     double *resvec) {
  hipLaunchKernelGGL(k_gmres_column, dim3(1), dim3(64), 0, s, i, m, g, st, resvec);
}
void xs_gmres_cycle(hipStream_t s, int m, int first, const xs_gmres_small &g, xs_state *st,
                    double *resvec) {
  hipLaunchKernelGGL(k_gmres_cycle, dim3(1), dim3(64), 0, s, m, first, g, st, resvec);
}
void xs_gmres_backsub(hipStream_t s, int k, int m, const xs_gmres_small &g) {
  hipLaunchKernelGGL(k_gmres_backsub, dim3(1), dim3(64), 0, s, k, m, g);
}
void xs_gcr_apply(hipStream_t s, int n, double *u, double *c, double *x, double *r,
                  double *t01, double *partial, unsigned *counter, xs_state *st,
                  double *resvec) {
  int g = (n + XS_BLOCK * 4 - 1) / (XS_BLOCK * 4);
  if (g > 2048) g = 2048;
  hipLaunchKernelGGL(k_gcr_apply, dim3(g), dim3(XS_BLOCK), 0, s, n, u, c, x, r, t01, partial,
                     counter, st, resvec);
  hipLaunchKernelGGL(k_gcr_check, dim3(1), dim3(64), 0, s, t01, st, resvec);
}
void xs_res0_from(hipStream_t s, const double *nrm2, xs_state *st, double *resvec) {
  hipLaunchKernelGGL(k_res0_from, dim3(1), dim3(64), 0, s, nrm2, st, resvec);
}
