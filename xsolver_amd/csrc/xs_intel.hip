// xs_intel.hip -- kernels of the intel-family Krylov loops
// (intel/mpisolver-C-v1.cpp: cpu_cg_solve, cpu_icg_solve, cpu_icgm_solve,
// cpu_ibicgstab_solve).  The vector work is a few generic kernels: one pass
// computes up to six inner products (what the reference groups into one
// communicator_sum), a one-workgroup kernel reduces them in a fixed order and
// runs that loop's scalar step on the device, and a linear combination of up
// to four vectors applies the updates with the scalars read from the device
// state.  So every iteration is a fixed sequence of launches (capturable in
// a graph) with no host synchronisation.  small = 1e-20 in every denominator,
// as in the reference.
#include "xs_device.h"
#include "xs_intel.h"

#define IS_SMALL 1e-20
#define MD_BLOCK 256
#define MD_ROWS 4

int xs_mdot_blocks(int n) {
  int g = (n + MD_BLOCK * MD_ROWS - 1) / (MD_BLOCK * MD_ROWS);
  return g < 1 ? 1 : (g > 2048 ? 2048 : g);
}

struct md_ptrs {
  const double *x[XS_MDOT_MAX];
  const double *y[XS_MDOT_MAX];
};

template <int NP>
__global__ __launch_bounds__(MD_BLOCK) void k_mdot(int n, md_ptrs P, double *partial,
                                                   const xs_state *st) {
  if (st->done) return;
  double acc[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) acc[p] = 0.0;
  for (int i = blockIdx.x * MD_BLOCK + threadIdx.x; i < n; i += gridDim.x * MD_BLOCK) {
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p] += P.x[p][i] * P.y[p][i];
  }
  __shared__ double red[MD_BLOCK / XS_WAVE][NP];
  const int lane = threadIdx.x & (XS_WAVE - 1), w = threadIdx.x / XS_WAVE;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    double v = acc[p];
#pragma unroll
    for (int off = XS_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, XS_WAVE);
    if (lane == 0) red[w][p] = v;
  }
  __syncthreads();
  if (threadIdx.x < NP) {
    double v = 0.0;
    for (int q = 0; q < MD_BLOCK / XS_WAVE; ++q) v += red[q][threadIdx.x];
    partial[(size_t)blockIdx.x * XS_MDOT_STRIDE + threadIdx.x] = v;
  }
}

// ||r|| judged against ||r0|| + small, the iteration counted, resvec written
__device__ void is_commit(xs_state *st, double *resvec, double resl) {
  double *sc = st->sc;
  const int it = st->iter + 1;
  st->iter = it;
  if (resvec) resvec[it] = resl;
  sc[S_RESL] = resl;
  if (!isfinite(resl)) st->done = 2;
  else if (resl / (sc[S_RES0] + IS_SMALL) < sc[S_TOL] || resl <= sc[S_ABSTOL]) st->done = 1;
}

__device__ void is_res0(xs_state *st, double *resvec, double res0) {
  double *sc = st->sc;
  sc[S_RES0] = res0;
  sc[S_RESL] = res0;
  sc[IS_ONE] = 1.0;
  sc[IS_MONE] = -1.0;
  if (resvec) resvec[0] = res0;
  if (!isfinite(res0)) st->done = 2;
  else if (res0 == 0.0 || res0 <= sc[S_ABSTOL]) st->done = 1;
}

__device__ void is_step(int step, const double *t, xs_state *st, double *resvec) {
  double *sc = st->sc;
  switch (step) {
    case IST_CGL_INIT:       // cpu_cg_solve :47-51
      sc[IS_TAU] = t[0];
      is_res0(st, resvec, sqrt(t[0]));
      break;
    case IST_CGL_ALPHA: {    // :61-63
      const double a = sc[IS_TAU] / (t[0] + IS_SMALL);
      sc[IS_ALPHA] = a;
      sc[IS_NALPHA] = -a;
      break;
    }
    case IST_CGL_RES: {      // :68-88
      const double tao = t[0];
      is_commit(st, resvec, sqrt(tao));
      sc[IS_BETA] = tao / (sc[IS_TAU] + IS_SMALL);
      sc[IS_TAU] = tao;
      break;
    }
    case IST_ICG_ALPHA: {    // cpu_icg_solve :136-145
      const double a = sc[IS_TAU] / (t[0] + IS_SMALL);
      sc[IS_ALPHA] = a;
      sc[IS_NALPHA] = -a;
      sc[IS_RHO] = t[1];
      sc[IS_THETA] = t[2];
      break;
    }
    case IST_ICG_RES: {      // :148-176
      is_commit(st, resvec, sqrt(t[0]));
      const double a = sc[IS_ALPHA], taoo = sc[IS_TAU];
      const double tao = taoo - 2.0 * a * sc[IS_RHO] + a * a * sc[IS_THETA];
      sc[IS_BETA] = tao / (taoo + IS_SMALL);
      sc[IS_TAU] = tao;
      break;
    }
    case IST_ICGM_INIT:      // cpu_icgm_solve :298-306
      sc[IS_TAU] = t[1];
      is_res0(st, resvec, sqrt(t[0]));
      break;
    case IST_ICGM_MAIN: {    // :313-334 (the commit follows the updates)
      const double a = sc[IS_TAU] / (t[0] + IS_SMALL);
      sc[IS_ALPHA] = a;
      sc[IS_NALPHA] = -a;
      const double tauo = sc[IS_TAU];
      const double tau = tauo - a * t[2] - a * t[1] + a * a * t[3];
      sc[IS_TAU] = tau;
      sc[IS_BETA] = tau / (tauo + IS_SMALL);
      sc[IS_RESL] = sqrt(t[4]);
      break;
    }
    case IST_COMMIT:
      is_commit(st, resvec, sc[IS_RESL]);
      break;
    case IST_IBI_INIT:       // cpu_ibicgstab_solve :514-562
      sc[IS_ALPHA] = 1.0;
      sc[IS_ALPHAO] = 1.0;
      sc[IS_ROU] = 1.0;
      sc[IS_OMEGA] = 1.0;
      sc[IS_PI] = 0.0;
      sc[IS_TAU] = 0.0;
      sc[IS_MU] = 0.0;
      sc[IS_FAI] = t[0];
      sc[IS_SIGMA] = t[1];
      is_res0(st, resvec, sqrt(t[0]));
      break;
    case IST_IBI_S1: {       // :568-574, with the update weights of :576-609
      const double rouo = sc[IS_ROU], omega = sc[IS_OMEGA], alpha = sc[IS_ALPHA];
      const double rou = sc[IS_FAI] - omega * sc[IS_MU];
      const double delta = (rou * alpha) / (rouo + IS_SMALL);
      const double beta = delta / (omega + IS_SMALL);
      const double tau = sc[IS_SIGMA] + beta * sc[IS_TAU] - delta * sc[IS_PI];
      const double an = rou / (tau + IS_SMALL);
      sc[IS_ROU] = rou;
      sc[IS_DELTA] = delta;
      sc[IS_NDELTA] = -delta;
      sc[IS_BETA] = beta;
      sc[IS_TAU] = tau;
      sc[IS_ALPHAO] = alpha;
      sc[IS_ALPHA] = an;
      sc[IS_NALPHA] = -an;
      sc[IS_NOMEGA] = -omega;
      sc[IS_T1] = (beta * an) / (alpha + IS_SMALL);
      sc[IS_NT2] = -(an * delta);
      break;
    }
    case IST_IBI_S2: {       // :611-627
      sc[IS_FAI] = t[0];
      sc[IS_PI] = t[1];
      sc[IS_MU] = t[4];
      const double omega = t[2] / (t[3] + IS_SMALL);
      sc[IS_OMEGA] = omega;
      sc[IS_NOMEGA] = -omega;
      sc[IS_SIGMA] = t[4] - omega * t[5];
      break;
    }
    case IST_IBI_RES:        // :636-641
      is_commit(st, resvec, sqrt(t[0]));
      break;
  }
}

// Reduce the per-workgroup partials in a fixed order (one workgroup), then
// run the scalar step on thread 0.
__global__ __launch_bounds__(256) void k_istep(int np, int nb, const double *partial, int step,
                                               xs_state *st, double *resvec) {
  if (st->done) return;
  __shared__ double tot[XS_MDOT_MAX];
  __shared__ double red[256 / XS_WAVE][XS_MDOT_MAX];
  const int lane = threadIdx.x & (XS_WAVE - 1), w = threadIdx.x / XS_WAVE;
  for (int p = 0; p < np; ++p) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) v += partial[(size_t)b * XS_MDOT_STRIDE + p];
#pragma unroll
    for (int off = XS_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, XS_WAVE);
    if (lane == 0) red[w][p] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int p = 0; p < np; ++p) {
      double v = 0.0;
      for (int q = 0; q < 256 / XS_WAVE; ++q) v += red[q][p];
      tot[p] = v;
    }
    is_step(step, tot, st, resvec);
  }
}

void xs_mdot_step(hipStream_t s, int n, int np, const double *const *x, const double *const *y,
                  double *partial, int step, xs_state *st, double *resvec) {
  XS_GUARD(np >= 1 && np <= XS_MDOT_MAX && partial && st);
  md_ptrs P = {};
  for (int p = 0; p < np; ++p) {
    XS_GUARD(x[p] && y[p]);
    P.x[p] = x[p];
    P.y[p] = y[p];
  }
  const int g = xs_mdot_blocks(n);
  switch (np) {
    case 1: hipLaunchKernelGGL(k_mdot<1>, dim3(g), dim3(MD_BLOCK), 0, s, n, P, partial, st); break;
    case 2: hipLaunchKernelGGL(k_mdot<2>, dim3(g), dim3(MD_BLOCK), 0, s, n, P, partial, st); break;
    case 3: hipLaunchKernelGGL(k_mdot<3>, dim3(g), dim3(MD_BLOCK), 0, s, n, P, partial, st); break;
    case 4: hipLaunchKernelGGL(k_mdot<4>, dim3(g), dim3(MD_BLOCK), 0, s, n, P, partial, st); break;
    case 5: hipLaunchKernelGGL(k_mdot<5>, dim3(g), dim3(MD_BLOCK), 0, s, n, P, partial, st); break;
    default: hipLaunchKernelGGL(k_mdot<6>, dim3(g), dim3(MD_BLOCK), 0, s, n, P, partial, st); break;
  }
  hipLaunchKernelGGL(k_istep, dim3(1), dim3(256), 0, s, np, g, partial, step, st, resvec);
}

void xs_istep(hipStream_t s, int step, xs_state *st, double *resvec) {
  XS_GUARD(st);
  hipLaunchKernelGGL(k_istep, dim3(1), dim3(256), 0, s, 0, 0, (const double *)nullptr, step, st,
                     resvec);
}

struct lc_args {
  const double *x[4];
  int slot[4];
};

template <int NT>
__global__ __launch_bounds__(XS_BLOCK) void k_lincomb(int n, lc_args A, double *y,
                                                      const xs_state *st) {
  if (st->done) return;
  double c[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) c[t] = st->sc[A.slot[t]];
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    double v = c[0] * A.x[0][i];
#pragma unroll
    for (int t = 1; t < NT; ++t) v += c[t] * A.x[t][i];
    y[i] = v;
  }
}

void xs_lincomb(hipStream_t s, int n, int nt, const double *const *x, const int *slot, double *y,
                const xs_state *st) {
  XS_GUARD(nt >= 1 && nt <= 4 && y && st);
  lc_args A = {};
  for (int t = 0; t < nt; ++t) {
    XS_GUARD(x[t] && slot[t] >= 0 && slot[t] < XS_NSCAL);
    A.x[t] = x[t];
    A.slot[t] = slot[t];
  }
  int g = (n + XS_BLOCK * 4 - 1) / (XS_BLOCK * 4);
  g = g < 1 ? 1 : (g > 4096 ? 4096 : g);
  switch (nt) {
    case 1: hipLaunchKernelGGL(k_lincomb<1>, dim3(g), dim3(XS_BLOCK), 0, s, n, A, y, st); break;
    case 2: hipLaunchKernelGGL(k_lincomb<2>, dim3(g), dim3(XS_BLOCK), 0, s, n, A, y, st); break;
    case 3: hipLaunchKernelGGL(k_lincomb<3>, dim3(g), dim3(XS_BLOCK), 0, s, n, A, y, st); break;
    default: hipLaunchKernelGGL(k_lincomb<4>, dim3(g), dim3(XS_BLOCK), 0, s, n, A, y, st); break;
  }
}

// y += x, unconditionally (after the loop)
__global__ void k_add(int n, const double *__restrict__ x, double *__restrict__ y) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    y[i] += x[i];
}

void xs_add(hipStream_t s, int n, const double *x, double *y) {
  XS_GUARD(x && y);
  int g = (n + 255) / 256;
  g = g < 1 ? 1 : (g > 4096 ? 4096 : g);
  hipLaunchKernelGGL(k_add, dim3(g), dim3(256), 0, s, n, x, y);
}
