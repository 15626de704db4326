// xs_device.h -- device-side reduction helpers shared by the kernel files:
// wave / workgroup sums and the "last workgroup" grid reduction that runs the
// Krylov scalar updates (do_finish) on the device.
#pragma once
#include "xs_internal.h"

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, XS_WAVE);
  return v;
}

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double red[XS_BLOCK / XS_WAVE][NV];
  const int lane = threadIdx.x & (XS_WAVE - 1), w = threadIdx.x / XS_WAVE;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[w][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < XS_BLOCK / XS_WAVE; ++q) s += red[q][k];
      v[k] = s;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ bool conv_check(xs_state *st, double resl) {
  double *sc = st->sc;
  sc[S_RESL] = resl;
  if (!isfinite(resl)) { st->done = 2; return true; }
  if (resl / sc[S_RES0] < sc[S_TOL] || resl <= sc[S_ABSTOL]) { st->done = 1; return true; }
  return false;
}

// tot[] holds NV totals followed by zeros (4 entries).
__device__ inline void do_finish(const xs_reduce &R, const double *tot, int NV) {
  xs_state *st = R.st;
  double *sc = st->sc;
  switch (R.op) {
    case F_STORE:
      for (int k = 0; k < NV; ++k) sc[R.slot + k] = tot[k];
      break;
    case F_RES0: {
      double r0 = sqrt(tot[0]);
      sc[S_RES0] = r0; sc[S_RESL] = r0;
      if (R.resvec) R.resvec[0] = r0;
      if (!isfinite(r0)) st->done = 2;
      else if (r0 == 0.0 || r0 <= sc[S_ABSTOL]) st->done = 1;
      // BiCGSTAB start (hip/Solver/xsolver_bicgstab.cpp:36-55): rhat = r0,
      // beta = rhat.r, omega = beta*gama/(alpha*beto + 1e-40), beto = beta.
      sc[S_ALPHA] = 1.0; sc[S_GAMA] = 1.0;
      sc[S_BETA] = tot[0];
      sc[S_OMEGA] = (tot[0] * 1.0) / (1.0 * 1.0 + 1e-40);
      sc[S_BETO] = tot[0];
      break;
    }
    case F_CG_INIT: {   // hip/Solver/xsolver_cg.cpp:28-39
      sc[S_TAU] = tot[0];
      double r0 = sqrt(tot[1]);
      sc[S_RES0] = r0; sc[S_RESL] = r0;
      if (R.resvec) R.resvec[0] = r0;
      if (!isfinite(r0)) st->done = 2;
      else if (r0 == 0.0 || r0 <= sc[S_ABSTOL]) st->done = 1;
      break;
    }
    case F_CG_ALPHA:    // alpha = tau / (Ap.p)   xsolver_cg.cpp:47-48
      sc[S_CGALPHA] = sc[S_TAU] / tot[0];
      break;
    case F_CG_RES: {    // tau = z.r, beta = tau/taoo, resl = ||r||  :52-58
      double taoo = sc[S_TAU];
      sc[S_TAU] = tot[0];
      sc[S_CGBETA] = tot[0] / taoo;
      double rr = (NV > 1) ? tot[1] : sc[S_SCRATCH0];
      double resl = sqrt(rr);
      int it = st->iter + 1;
      st->iter = it;
      if (R.resvec) R.resvec[it] = resl;
      conv_check(st, resl);
      break;
    }
    case F_BICG_GAMA:   // gama = beta / (rhat.u)  xsolver_bicgstab.cpp:80
      sc[S_GAMA] = sc[S_BETA] / tot[0];
      break;
    case F_BICG_ALPHA:  // alpha = (s.v)/(v.v)      :95
      sc[S_ALPHA] = tot[0] / tot[1];
      break;
    case F_BICG_RES: {  // resl = ||r|| (:101-103); next beta/omega (:50-53)
      double resl = sqrt(tot[0]);
      int it = st->iter + 1;
      st->iter = it;
      if (R.resvec) R.resvec[it] = resl;
      conv_check(st, resl);
      double beta = tot[1];
      sc[S_OMEGA] = (beta * sc[S_GAMA]) / (sc[S_ALPHA] * sc[S_BETO] + 1e-40);
      sc[S_BETO] = beta;
      sc[S_BETA] = beta;
      break;
    }
    default:
      break;
  }
}

// Block-reduce v, publish the block partial, and let the last workgroup of the
// grid combine all partials (in block order) and run the finishing op.
template <int NV>
__device__ __forceinline__ void reduce_finish(double (&v)[NV], const xs_reduce &R) {
  __shared__ int is_last;
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) R.partial[(size_t)blockIdx.x * NV + k] = v[k];
    __threadfence();
    unsigned t = atomicAdd(R.counter, 1u);
    is_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  double tot[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = 0.0;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += XS_BLOCK) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      tot[k] += __longlong_as_double((long long)__hip_atomic_load(
          reinterpret_cast<unsigned long long *>(&R.partial[(size_t)b * NV + k]), __ATOMIC_RELAXED,
          __HIP_MEMORY_SCOPE_AGENT));
  }
  block_sum<NV>(tot);
  if (threadIdx.x == 0) {
    *R.counter = 0u;
    double t4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NV; ++k) t4[k] = tot[k];
    do_finish(R, t4, NV);
  }
}

// ---------------------------------------------------------------------------
// Ready-flag loads/stores of the sync-free triangular solves: a value slot
// holds XS_SENTINEL until its row is published with an agent-scope store.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long ld_relaxed(const double *p) {
  return __hip_atomic_load(reinterpret_cast<unsigned long long *>(const_cast<double *>(p)),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Spin until the row value is published.  The wait is bounded (about a
// second of polling, far beyond any legitimate wait): on expiry the flag
// *err is raised and a quiet NaN is used, so every wave always exits and the
// solve reports the failure instead of hanging.
#define XS_SPIN_LIMIT (1u << 22)
__device__ __forceinline__ double ld_ready(const double *p, unsigned *err) {
  unsigned long long u = ld_relaxed(p);
  unsigned polls = 0;
  while (u == XS_SENTINEL) {
    if (++polls > XS_SPIN_LIMIT) {
      atomicOr(err, 1u);
      u = 0x7FF8000000000000ull;
      break;
    }
    __builtin_amdgcn_s_sleep(1);
    u = ld_relaxed(p);
  }
  return __longlong_as_double((long long)u);
}
__device__ __forceinline__ void st_ready(double *p, double v) {
  unsigned long long u = (unsigned long long)__double_as_longlong(v);
  if (u == XS_SENTINEL) u = 0x7FF8000000000000ull;
  __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), u, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
