// xs_device.h -- device-side reduction helpers shared by the kernel files:
// wave / workgroup sums and the "last workgroup" grid reduction that runs the
// Krylov scalar updates (do_finish) on the device.
#pragma once
#include "xs_internal.h"

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
// lane t's value of v in every lane (t uniform): two v_readlane, no LDS trip
__device__ __forceinline__ double wave_readlane_f64(double v, int t) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, t);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), t);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ void st_ready(double *p, double v) {
  unsigned long long u = (unsigned long long)__double_as_longlong(v);
  if (u == XS_SENTINEL) u = 0x7FF8000000000000ull;
  __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), u, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Inclusive affine scan across the 64 lanes of a wave with DPP moves:
// lane l ends with (c, m) = (c_l, m_l) o (c_{l-1}, m_{l-1}) o ... o (c_0, m_0)
// where (c, m) o (c', m') = (c + m c', m m').  Row shifts by 1, 2, 4, 8 inside
// each 16-lane row, then row_bcast:15 and row_bcast:31 across rows; lanes
// without a source take the identity (0, 1).  Six VALU steps instead of six
// ds_bpermute round trips (about 0.3 us less per dependency hop of the
// sync-free triangular solves on gfx950).
// ---------------------------------------------------------------------------
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d(double v, double ident) {
  const long long b = __double_as_longlong(v), ib = __double_as_longlong(ident);
  const int lo = __builtin_amdgcn_update_dpp((int)(ib & 0xffffffff), (int)(b & 0xffffffff), CTRL,
                                             ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(ib >> 32), (int)(b >> 32), CTRL, ROWMASK, 0xF,
                                             false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned long long)(unsigned)lo));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_f(float v, float ident) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(ident), __float_as_int(v), CTRL,
                                                    ROWMASK, 0xF, false));
}
template <int CTRL, int ROWMASK, int KC>
__device__ __forceinline__ void affine_step(double (&c)[KC], double &m) {
  const double mj = dpp_d<CTRL, ROWMASK>(m, 1.0);
#pragma unroll
  for (int q = 0; q < KC; ++q) c[q] = c[q] + m * dpp_d<CTRL, ROWMASK>(c[q], 0.0);
  m = m * mj;
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void affine_step_f(float &c, float &m) {
  const float cj = dpp_f<CTRL, ROWMASK>(c, 0.f), mj = dpp_f<CTRL, ROWMASK>(m, 1.f);
  c = c + m * cj;
  m = m * mj;
}
// KC independent right-hand sides sharing the multipliers m
template <int KC>
__device__ __forceinline__ void wave_affine_scan(double (&c)[KC], double &m) {
  affine_step<0x111, 0xF, KC>(c, m);  // row_shr:1
  affine_step<0x112, 0xF, KC>(c, m);  // row_shr:2
  affine_step<0x114, 0xF, KC>(c, m);  // row_shr:4
  affine_step<0x118, 0xF, KC>(c, m);  // row_shr:8
  affine_step<0x142, 0xA, KC>(c, m);  // row_bcast:15 into rows 1 and 3
  affine_step<0x143, 0xC, KC>(c, m);  // row_bcast:31 into rows 2 and 3
}
__device__ __forceinline__ void wave_affine_scan_f(float &c, float &m) {
  affine_step_f<0x111, 0xF>(c, m);
  affine_step_f<0x112, 0xF>(c, m);
  affine_step_f<0x114, 0xF>(c, m);
  affine_step_f<0x118, 0xF>(c, m);
  affine_step_f<0x142, 0xA>(c, m);
  affine_step_f<0x143, 0xC>(c, m);
}

// Chunk of work for workgroup `b` of `nb`.  Workgroups are dealt to the 8
// XCDs round-robin (b % 8); this gives each XCD one contiguous run of chunks,
// so rows that read neighbouring parts of a vector share one XCD's L2.  It is
// a bijection on [0, nb) whatever the dispatch does; only locality depends on
// the round-robin order.
#define XS_NXCD 8
__device__ __forceinline__ int xcd_chunk(int b, int nb) {
  const int xcd = b % XS_NXCD, q = nb / XS_NXCD, rem = nb % XS_NXCD;
  return xcd * q + (xcd < rem ? xcd : rem) + b / XS_NXCD;
}

// Work order of the sync-free triangular solves.  The grid is persistent and
// small (a few hundred workgroups): each workgroup repeatedly takes one ticket
// (one same-address atomic per ticket) and its XS_BLOCK / XS_WAVE waves solve
// the consecutive 64-row groups of that ticket, until the tickets run out.
// A wave only ever waits for rows of groups with earlier tickets, held by
// running workgroups that finish them before taking another ticket, or for
// rows in its own group: every wait ends.  Few resident waves also means
// little polling traffic: only the waves near the dependency frontier poll.
// Returns this wave's ticket slot; `base` is the first slot of the ticket.
__device__ __forceinline__ unsigned trsv_next(unsigned *tickets, unsigned &base) {
  __shared__ unsigned s_ticket;
  __syncthreads();   // every wave has read the previous ticket
  if (threadIdx.x == 0) s_ticket = atomicAdd(&tickets[0], 1u);
  __syncthreads();
  base = s_ticket * (XS_BLOCK / XS_WAVE);
  return base + threadIdx.x / XS_WAVE;
}
// Groups of a triangular solve and the rows [r0, r1) of group w.
__device__ __forceinline__ unsigned trsv_ngroups(int n, const xs_tsched &S) {
  return S.gb ? (unsigned)S.ngroups : (unsigned)((n + XS_WAVE - 1) / XS_WAVE);
}
template <bool UPPER>
__device__ __forceinline__ void trsv_rows(int n, const xs_tsched &S, unsigned w, int &r0, int &r1) {
  if (S.gb) {
    if (!UPPER) { r0 = S.gb[w]; r1 = S.gb[w + 1]; }
    else { r1 = S.gb[w]; r0 = S.gb[w + 1]; }
  } else if (!UPPER) {
    r0 = (int)w * XS_WAVE; r1 = min(n, r0 + XS_WAVE);
  } else {
    r1 = n - (int)w * XS_WAVE; r0 = max(0, r1 - XS_WAVE);
  }
}
// The workgroup that retires last resets both counters for the next launch.
__device__ __forceinline__ void trsv_retire(unsigned *tickets) {
  if (threadIdx.x == 0) {
    const unsigned f = atomicAdd(&tickets[1], 1u);
    if (f == gridDim.x - 1) {
      __hip_atomic_store(&tickets[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&tickets[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

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
    case F_BICG_ALPHA:  // alpha = (s.v)/(v.v)      :95; v = 0 (s solved exactly): 0, not 0/0
      sc[S_ALPHA] = tot[1] != 0.0 ? tot[0] / tot[1] : 0.0;
      break;
    case F_BICG_RES: {  // resl = ||r|| (:101-103); next beta/omega (:50-53)
      double resl = sqrt(tot[0]);
      int it = st->iter + 1;
      st->iter = it;
      if (R.resvec) R.resvec[it] = resl;
      conv_check(st, resl);
      double beta = tot[1];
      sc[S_OMEGA] = (beta * sc[S_GAMA]) / (sc[S_ALPHA] * sc[S_BETO] + 1e-40);
      sc[S_RHOO] = sc[S_BETO];
      sc[S_BETO] = beta;
      sc[S_BETA] = beta;
      break;
    }
    case F_BICG_REPL: {   // reliable update: the recursive residual r replaced by
                          // the true one; beta' and the next omega from its rhat.r
      const double beta = tot[1], resl = sqrt(tot[0]);
      sc[S_RESL] = resl;
      sc[S_OMEGA] = (beta * sc[S_GAMA]) / (sc[S_ALPHA] * sc[S_RHOO] + 1e-40);
      sc[S_BETO] = beta;
      sc[S_BETA] = beta;
      if (!isfinite(resl) || !isfinite(beta)) st->done = 2;
      else if (st->done == 1) st->done = 0;   // the host decides on the true residual
      break;
    }
    default:
      break;
  }
}

// Block-reduce v and combine the blocks' partials in two levels: blocks are
// grouped by XS_RED_GROUP consecutive indices; the group's highest-index block
// (dispatched last in its group) waits for the group's other partials, adds
// its own and publishes the group partial; the grid's highest-index block,
// the leader of the last group, then combines the group partials and runs the
// finishing op.  Fixed combination order (deterministic).  Partial slots hold
// XS_SENTINEL between reductions: a value is published with an agent-scope
// store (it is its own ready flag) and the reader writes the sentinel back.
// No fence and no same-address atomic per workgroup: on gfx950 a
// device-scope fence writes back the XCD's L2 and same-address atomics
// serialise (~10 ns each), which made a counter-based last-block reduction
// over 8k workgroups cost ~150 us.  Each combining step is one round trip of
// at most one slot per thread for up to 64 x 64 blocks.  Layout: block
// partials [gridDim.x][NV], then group partials [ngroups][NV].
#define XS_RED_GROUP 64
// in-wave entries a triangular-solve lane keeps in registers, and the
// off-group entries whose values it loads together before waiting
#ifndef XS_INW
#define XS_INW 8
#endif
#define XS_GATHER 4
template <int NV>
__device__ __forceinline__ void group_gather(const double *slots, unsigned m, const xs_reduce &R,
                                             double (&tot)[NV]) {
  // slots[0..m) in NV-interleaved order; thread t sums slots t, t + XS_BLOCK, ...
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = 0.0;
  for (unsigned s = threadIdx.x; s < m; s += XS_BLOCK) {
    double *p = const_cast<double *>(slots) + s;
    const unsigned long long u = ld_relaxed(p);
    const double val = u == XS_SENTINEL ? ld_ready(p, R.counter) : __longlong_as_double((long long)u);
    tot[s % NV] += val;
    __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), XS_SENTINEL, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum<NV>(tot);
}

template <int NV>
__device__ __forceinline__ void reduce_finish(double (&v)[NV], const xs_reduce &R) {
  block_sum<NV>(v);
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned g = b / XS_RED_GROUP, ng = (nb + XS_RED_GROUP - 1) / XS_RED_GROUP;
  const unsigned g0 = g * XS_RED_GROUP, g1 = min(nb, g0 + XS_RED_GROUP);
  double *gpart = R.partial + (size_t)nb * NV;
  if (b != g1 - 1) {   // not the group's leader: publish and leave
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < NV; ++k) st_ready(&R.partial[(size_t)b * NV + k], v[k]);
    }
    return;
  }
  double tot[NV];
  group_gather<NV>(R.partial + (size_t)g0 * NV, (g1 - 1 - g0) * NV, R, tot);
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] += v[k];   // valid in thread 0
  if (g != ng - 1) {
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < NV; ++k) st_ready(&gpart[(size_t)g * NV + k], tot[k]);
    }
    return;
  }
  double all[NV];
  group_gather<NV>(gpart, (ng - 1) * NV, R, all);
  if (threadIdx.x == 0) {
    double t4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NV; ++k) t4[k] = all[k] + tot[k];
    do_finish(R, t4, NV);
  }
}
