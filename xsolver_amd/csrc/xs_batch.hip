// xs_batch.hip -- many independent systems in one persistent launch.
//
// The systems are the diagonal blocks of one block-diagonal CSR matrix
// (sys_ptr[s]..sys_ptr[s+1] are the rows of system s, with columns in the
// same range).  Each workgroup takes systems from an atomic queue (largest
// first, a longest-processing-time order) and runs preconditioned CG with
// Jacobi (hip/Solver/xsolver_cg.cpp:16-74) on it to convergence with
// workgroup barriers only: every system stops at its own iteration count
// (per-system convergence), and no kernel is launched per iteration.
// Precision 1 runs one FP32 CG recurrence (FP32 values, vectors, SpMV) with
// FP64 reliable updates: the FP64 residual b - A x replaces the FP32 one at
// every 1e-2 drop and decides convergence.
// No reference counterpart (the reference solves one system per call).
#include "xs_device.h"
#include "xs_batch.h"

#define BT_THREADS 1024
#define BT_WAVES (BT_THREADS / XS_WAVE)

template <class T>
__device__ __forceinline__ T bt_wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, XS_WAVE);
  return v;
}

// workgroup sum of two values (double accumulation)
__device__ __forceinline__ void bt_sum2(double &a, double &b, double (*red)[2]) {
  const int lane = threadIdx.x & (XS_WAVE - 1), w = threadIdx.x / XS_WAVE;
  a = bt_wave_sum(a);
  b = bt_wave_sum(b);
  __syncthreads();
  if (lane == 0) { red[w][0] = a; red[w][1] = b; }
  __syncthreads();
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int q = 0; q < BT_WAVES; ++q) { s0 += red[q][0]; s1 += red[q][1]; }
  a = s0;
  b = s1;
}

// Rows of a thread: r0 + threadIdx.x, then every BT_THREADS.  Each thread's
// loads are serial chains (row pointer -> column -> gather; or the operands
// of an update, behind the previous element's stores, which may alias them
// as far as the compiler knows), so the loops below take two rows (four
// elements) per step with every load of the step issued before any store:
// the same arithmetic in the same order, half (a quarter of) the chains.

// A row's first W entries, loaded branch-free and masked (a masked slot
// reads column r, value ignored).
template <int W, class TV, class TX>
struct bt_rowbuf {
  int e0, e1;
  int c[W];
  TV a[W];
  TX g[W];
  __device__ __forceinline__ void load(int r, const int *__restrict__ rp, const int *__restrict__ ci,
                                       const TV *__restrict__ v) {
    e0 = rp[r];
    e1 = rp[r + 1];
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const int e = min(e0 + k, max(e1 - 1, 0));
      const int cc = ci[e];
      c[k] = e0 + k < e1 ? cc : r;
      a[k] = v[e];
    }
  }
  __device__ __forceinline__ void gather(const TX *x) {
#pragma unroll
    for (int k = 0; k < W; ++k) g[k] = x[c[k]];
  }
  // sum_e v[e] x[ci[e]] in CSR order
  __device__ __forceinline__ TX sum() const {
    TX s = 0;
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const TX t = s + (TX)a[k] * g[k];
      s = e0 + k < e1 ? t : s;
    }
    return s;
  }
};

// f(row, (A x)_row) for the thread's rows of [r0, r1), in order, two rows per
// step: each row's first W entries in the window, any further ones after
// them in a loop (rows longer than W are rare), so the sum keeps CSR order.
template <int W, class TV, class TX, class F>
__device__ __forceinline__ void bt_rows(int r0, int r1, const int *__restrict__ rp,
                                        const int *__restrict__ ci, const TV *__restrict__ v,
                                        const TX *x, F &&f) {
  auto tail = [&](const bt_rowbuf<W, TV, TX> &R, TX s) {
    for (int e = R.e0 + W; e < R.e1; ++e) s += (TX)v[e] * x[ci[e]];
    return s;
  };
  for (int r = r0 + threadIdx.x; r < r1; r += 2 * BT_THREADS) {
    const int rb = r + BT_THREADS;
    const bool vb = rb < r1;
    bt_rowbuf<W, TV, TX> A, B;
    A.load(r, rp, ci, v);
    B.load(vb ? rb : r, rp, ci, v);
    A.gather(x);
    B.gather(x);
    f(r, tail(A, A.sum()));
    if (vb) f(rb, tail(B, B.sum()));
  }
}

// y = A p over rows [r0, r1) and sum p.y
template <int W, class TV, class TX>
__device__ __forceinline__ double bt_spmv_dot(int r0, int r1, const int *__restrict__ rp,
                                              const int *__restrict__ ci,
                                              const TV *__restrict__ v, const TX *p, TX *y) {
  double acc = 0.0;
  bt_rows<W, TV, TX>(r0, r1, rp, ci, v, p, [&](int r, TX s) {
    y[r] = s;
    acc += (double)s * (double)p[r];
  });
  return acc;
}

// CG + Jacobi on rows [r0, r1) for A d = rhs from d = 0 (FP32 or FP64);
// stops when ||r|| <= tol * ||rhs||.  Returns iterations (barrier-uniform).
template <int W, class TV, class TX>
__device__ int bt_cg(int r0, int r1, const int *__restrict__ rp, const int *__restrict__ ci,
                     const TV *__restrict__ v, const TX *__restrict__ dinv, const TX *rhs, TX *d,
                     TX *r, TX *z, TX *p, TX *q, double tol, int maxit, double (*red)[2]) {
  constexpr int U = 4;   // elements per step of the update loops
  double tau = 0.0, rr = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += BT_THREADS) {
    const TX ri = rhs[i];
    const TX zi = dinv[i] * ri;
    d[i] = 0;
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    tau += (double)ri * (double)zi;
    rr += (double)ri * (double)ri;
  }
  bt_sum2(tau, rr, red);
  const double res0 = sqrt(rr);
  if (res0 == 0.0) return 0;
  int it = 0;
  while (it < maxit) {
    __syncthreads();
    double sigma = bt_spmv_dot<W, TV, TX>(r0, r1, rp, ci, v, p, q), dummy = 0.0;
    bt_sum2(sigma, dummy, red);
    const TX alpha = (TX)(tau / sigma);
    double tn = 0.0, rn = 0.0;
    for (int i0 = r0 + threadIdx.x; i0 < r1; i0 += U * BT_THREADS) {
      TX pv[U], qv[U], rv[U], dv[U], di[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int i = min(i0 + k * BT_THREADS, r1 - 1);
        pv[k] = p[i];
        qv[k] = q[i];
        rv[k] = r[i];
        dv[k] = d[i];
        di[k] = dinv[i];
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int i = i0 + k * BT_THREADS;
        if (i < r1) {
          d[i] = dv[k] + alpha * pv[k];
          const TX ri = rv[k] - alpha * qv[k];
          r[i] = ri;
          const TX zi = di[k] * ri;
          z[i] = zi;
          tn += (double)zi * (double)ri;
          rn += (double)ri * (double)ri;
        }
      }
    }
    bt_sum2(tn, rn, red);
    const TX beta = (TX)(tn / tau);
    tau = tn;
    ++it;
    if (!(sqrt(rn) > tol * res0)) break;   // also stops on NaN
    for (int i0 = r0 + threadIdx.x; i0 < r1; i0 += U * BT_THREADS) {
      TX zv[U], pv[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int i = min(i0 + k * BT_THREADS, r1 - 1);
        zv[k] = z[i];
        pv[k] = p[i];
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int i = i0 + k * BT_THREADS;
        if (i < r1) p[i] = zv[k] + beta * pv[k];
      }
    }
  }
  __syncthreads();
  return it;
}

// 1 when every row of [r0, r1) has row pointers inside [0, nnz] and every
// column inside [r0, r1) (block-diagonal pattern), else 0; workgroup-uniform.
__device__ int bt_pattern_ok(int r0, int r1, int N, int nnz, const int *__restrict__ rp,
                             const int *__restrict__ ci, double (*red)[2]) {
  double bad = 0.0, dummy = 0.0;
  if (r0 < 0 || r1 > N || r1 <= r0) bad = 1.0;
  else
    for (int i = r0 + threadIdx.x; i < r1; i += BT_THREADS) {
      const int e0 = rp[i], e1 = rp[i + 1];
      if (e0 < 0 || e1 < e0 || e1 > nnz) { bad = 1.0; continue; }
      for (int e = e0; e < e1; ++e)
        if (ci[e] < r0 || ci[e] >= r1) bad = 1.0;
    }
  bt_sum2(bad, dummy, red);
  return bad == 0.0;
}

// One system on rows [r0, r1): FP64 CG from x0 (PREC 0), or FP32 CG with
// FP64 reliable updates (PREC 1).
#define BT_W 6   // row window of the SpMV (FP64: the register budget of 1024-thread groups)
template <int PREC>
__device__ void bt_system(int r0, int r1, int N, const int *__restrict__ rp,
                          const int *__restrict__ ci, const double *__restrict__ v,
                          const float *__restrict__ v32, const double *__restrict__ dinv,
                          const float *__restrict__ dinv32, const double *__restrict__ b,
                          double *__restrict__ x, double *__restrict__ w64,
                          float *__restrict__ w32, double tol, int maxiter, double inner_tol,
                          double (*red)[2], int &it, double &rel) {
  constexpr int W = BT_W;
  it = 0;
  rel = 0.0;
  if (PREC == 0) {
    // FP64 CG from x0 (the workspace holds r, z, p, q, rhs, d)
    double *r = w64, *z = w64 + N, *p = z + N, *q = p + N;
    double *rhs = q + N, *d = rhs + N;
    double bb = 0.0, dummy = 0.0;   // ||b - A x0||^2: the base of tol and relres
    bt_rows<W, double, double>(r0, r1, rp, ci, v, x, [&](int i, double ax) {
      const double t = b[i] - ax;
      rhs[i] = t;
      bb += t * t;
    });
    bt_sum2(bb, dummy, red);
    it = bt_cg<W, double, double>(r0, r1, rp, ci, v, dinv, rhs, d, r, z, p, q, tol, maxiter, red);
    double rn = 0.0, dm = 0.0;
    for (int i = r0 + threadIdx.x; i < r1; i += BT_THREADS) x[i] += d[i];
    __syncthreads();
    bt_rows<W, double, double>(r0, r1, rp, ci, v, x, [&](int i, double ax) {
      const double t = b[i] - ax;
      rn += t * t;
    });
    bt_sum2(rn, dm, red);
    rel = bb > 0.0 ? sqrt(rn / bb) : 0.0;
  } else {
    // One FP32 CG recurrence (FP32 values, vectors and SpMV) with reliable
    // updates: whenever its residual has dropped by inner_tol since the last
    // replacement, or claims convergence, the FP32 correction d is added to
    // the FP64 x, the FP64 residual b - A x replaces the FP32 one and decides
    // convergence; the search direction is kept (no restart, so the
    // iteration count stays near the FP64 CG's).
    constexpr int U = 4;
    float *r = w32, *z = w32 + N, *p = z + N, *q = p + N, *d = q + N;
    double rn64 = 0.0, dm = 0.0;
    bt_rows<W, double, double>(r0, r1, rp, ci, v, x, [&](int i, double ax) {
      const double t = b[i] - ax;
      r[i] = (float)t;
      rn64 += t * t;
    });
    bt_sum2(rn64, dm, red);
    const double bb = rn64;   // ||b - A x0||^2: the base of tol and relres
    if (!(bb > 0.0)) return;  // x0 solves the system
    double rref = sqrt(bb), tau = 0.0, t2 = 0.0;
    for (int i = r0 + threadIdx.x; i < r1; i += BT_THREADS) {
      const float ri = r[i], zi = dinv32[i] * ri;
      z[i] = zi;
      p[i] = zi;
      d[i] = 0.0f;
      tau += (double)zi * (double)ri;
    }
    bt_sum2(tau, t2, red);
    rel = 1.0;
    bool fresh = true;   // x and rel hold the last FP64 residual
    while (it < maxiter) {
      __syncthreads();
      double sigma = bt_spmv_dot<W, float, float>(r0, r1, rp, ci, v32, p, q), dummy = 0.0;
      bt_sum2(sigma, dummy, red);
      const float alpha = (float)(tau / sigma);
      double tn = 0.0, rn = 0.0;
      for (int i0 = r0 + threadIdx.x; i0 < r1; i0 += U * BT_THREADS) {
        float pv[U], qv[U], rv[U], dv[U], di[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
          const int i = min(i0 + k * BT_THREADS, r1 - 1);
          pv[k] = p[i];
          qv[k] = q[i];
          rv[k] = r[i];
          dv[k] = d[i];
          di[k] = dinv32[i];
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
          const int i = i0 + k * BT_THREADS;
          if (i < r1) {
            d[i] = dv[k] + alpha * pv[k];
            const float ri = rv[k] - alpha * qv[k];
            r[i] = ri;
            const float zi = di[k] * ri;
            z[i] = zi;
            tn += (double)zi * (double)ri;
            rn += (double)ri * (double)ri;
          }
        }
      }
      bt_sum2(tn, rn, red);
      ++it;
      fresh = false;
      const double rnorm = sqrt(rn);
      if (!(rnorm > inner_tol * rref) || !(rnorm > tol * sqrt(bb))) {
        // reliable update: x += d, r = b - A x in FP64
        __syncthreads();
        for (int i = r0 + threadIdx.x; i < r1; i += BT_THREADS) {
          x[i] += (double)d[i];
          d[i] = 0.0f;
        }
        __syncthreads();
        rn64 = 0.0;
        bt_rows<W, double, double>(r0, r1, rp, ci, v, x, [&](int i, double ax) {
          const double t = b[i] - ax;
          r[i] = (float)t;
          rn64 += t * t;
        });
        bt_sum2(rn64, dm, red);
        fresh = true;
        rel = sqrt(rn64 / bb);
        if (!(rel >= tol)) break;   // converged (or NaN)
        rref = sqrt(rn64);
        tn = 0.0;
        for (int i = r0 + threadIdx.x; i < r1; i += BT_THREADS) {
          const float ri = r[i], zi = dinv32[i] * ri;
          z[i] = zi;
          tn += (double)zi * (double)ri;
        }
        bt_sum2(tn, dm, red);
      }
      const float beta = (float)(tn / tau);
      tau = tn;
      for (int i0 = r0 + threadIdx.x; i0 < r1; i0 += U * BT_THREADS) {
        float zv[U], pv[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
          const int i = min(i0 + k * BT_THREADS, r1 - 1);
          zv[k] = z[i];
          pv[k] = p[i];
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
          const int i = i0 + k * BT_THREADS;
          if (i < r1) p[i] = zv[k] + beta * pv[k];
        }
      }
    }
    if (!fresh) {   // out of iterations: x and relres with the last correction
      __syncthreads();
      for (int i = r0 + threadIdx.x; i < r1; i += BT_THREADS) x[i] += (double)d[i];
      __syncthreads();
      rn64 = 0.0;
      bt_rows<W, double, double>(r0, r1, rp, ci, v, x, [&](int i, double ax) {
        const double t = b[i] - ax;
        rn64 += t * t;
      });
      bt_sum2(rn64, dm, red);
      rel = sqrt(rn64 / bb);
    }
  }
}

// N: rows of the whole batch (the workspace stride); nnz: entries of the
// whole batch.  A system whose rows or pattern fall outside those bounds is
// not solved: iters = -1, relres = NaN.
template <int PREC>
__global__ __launch_bounds__(BT_THREADS) void k_batch_cg(
    int nsys, int N, int nnz, const int *__restrict__ order, const int *__restrict__ sys_ptr,
    const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v,
    const float *__restrict__ v32, const double *__restrict__ dinv,
    const float *__restrict__ dinv32, const double *__restrict__ b, double *__restrict__ x,
    double *__restrict__ w64, float *__restrict__ w32, double tol, int maxiter,
    double inner_tol, unsigned *queue, int *iters, double *relres) {
  __shared__ double red[BT_WAVES][2];
  __shared__ int s_sys;
  for (;;) {
    if (threadIdx.x == 0) {
      const unsigned t = atomicAdd(queue, 1u);
      s_sys = t < (unsigned)nsys ? (order ? order[t] : (int)t) : -1;
    }
    __syncthreads();
    const int sys = s_sys;
    __syncthreads();
    if (sys < 0) break;
    const int r0 = sys_ptr[sys], r1 = sys_ptr[sys + 1];
    int it = 0;
    double rel = 0.0;
    if (!bt_pattern_ok(r0, r1, N, nnz, rp, ci, red)) {
      it = -1;
      rel = __builtin_nan("");
    } else {
      bt_system<PREC>(r0, r1, N, rp, ci, v, v32, dinv, dinv32, b, x, w64, w32, tol, maxiter,
                      inner_tol, red, it, rel);
    }
    if (threadIdx.x == 0) {
      iters[sys] = it;
      relres[sys] = rel;
    }
    __syncthreads();
  }
}

__global__ void k_batch_prep(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                             const double *__restrict__ v, double *__restrict__ dinv,
                             float *__restrict__ dinv32, float *__restrict__ v32, int *bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int e = rp[i]; e < rp[i + 1]; ++e) {
    if (ci[e] == i) d = v[e];
    v32[e] = (float)v[e];
  }
  if (d == 0.0) atomicMin(bad, i + 1);
  dinv[i] = 1.0 / d;
  dinv32[i] = (float)(1.0 / d);
}

// FP64 Jacobi set-up only (1/a_ii; a zero or missing diagonal gives inf,
// which the CG turns into a non-finite relres for that system), and the
// work-queue counter cleared by the same launch (thread 0), so a caller
// under graph capture records kernels only.
__global__ void k_batch_prep64(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                               const double *__restrict__ v, double *__restrict__ dinv,
                               unsigned *queue) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *queue = 0u;
  if (i >= n) return;
  double d = 0.0;
  for (int e = rp[i]; e < rp[i + 1]; ++e)
    if (ci[e] == i) d = v[e];
  dinv[i] = 1.0 / d;
}

void xs_batch_prep64(hipStream_t s, int n, const int *rp, const int *ci, const double *v,
                     double *dinv, unsigned *queue) {
  hipLaunchKernelGGL(k_batch_prep64, dim3((n + 255) / 256), dim3(256), 0, s, n, rp, ci, v, dinv,
                     queue);
}

void xs_batch_prep(hipStream_t s, int n, const int *rp, const int *ci, const double *v,
                   double *dinv, float *dinv32, float *v32, int *bad) {
  hipLaunchKernelGGL(k_batch_prep, dim3((n + 255) / 256), dim3(256), 0, s, n, rp, ci, v, dinv,
                     dinv32, v32, bad);
}

void xs_batch_cg(hipStream_t s, int grid, int nsys, int N, int nnz, const int *order,
                 const int *sys_ptr, const int *rp, const int *ci, const double *v,
                 const float *v32, const double *dinv, const float *dinv32, const double *b,
                 double *x, double *w64, float *w32, double tol, int maxiter, int precision,
                 double inner_tol, unsigned *queue, int *iters, double *relres) {
  if (precision == 0)
    hipLaunchKernelGGL(k_batch_cg<0>, dim3(grid), dim3(BT_THREADS), 0, s, nsys, N, nnz, order,
                       sys_ptr, rp, ci, v, v32, dinv, dinv32, b, x, w64, w32, tol, maxiter,
                       inner_tol, queue, iters, relres);
  else
    hipLaunchKernelGGL(k_batch_cg<1>, dim3(grid), dim3(BT_THREADS), 0, s, nsys, N, nnz, order,
                       sys_ptr, rp, ci, v, v32, dinv, dinv32, b, x, w64, w32, tol, maxiter,
                       inner_tol, queue, iters, relres);
}
