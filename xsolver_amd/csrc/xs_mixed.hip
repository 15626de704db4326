// xs_mixed.hip -- FP32 inner-iteration kernels for the mixed-precision solve:
// FP32 CSR / ELL SpMV, sync-free SpTRSV over factors stored as FP32, BF16 or
// FP16 (accumulated in FP32), FP32 BiCGSTAB vector updates, and the
// FP64 <-> FP32 conversions of the iterative-refinement outer loop.  Inner
// products are accumulated and reduced in FP64; every Krylov scalar stays
// FP64 on the device.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <climits>

#include "xs_device.h"
#include "xs_mixed.h"

__device__ __forceinline__ float fl(const float *p, size_t i) { return p[i]; }
__device__ __forceinline__ float fl(const __hip_bfloat16 *p, size_t i) { return __bfloat162float(p[i]); }
__device__ __forceinline__ float fl(const __half *p, size_t i) { return __half2float(p[i]); }

template <int EPI, int NV>
__device__ __forceinline__ void epi32(int r, float s, float *__restrict__ y,
                                      const float *__restrict__ b, const float *__restrict__ w,
                                      double (&acc)[NV]) {
  if (EPI == EPI_BMAX || EPI == EPI_BMAXN) {
    const float t = b[r] - s;
    y[r] = t;
    if (EPI == EPI_BMAXN) acc[0] += (double)t * (double)t;
  } else {
    y[r] = s;
    if (EPI == EPI_DOT1 || EPI == EPI_DOT2) acc[0] += (double)w[r] * (double)s;
    if (EPI == EPI_DOT2) acc[NV - 1] += (double)s * (double)s;
  }
}

// The same with the epilogue's operand (b[r] or w[r]) already loaded.
template <int EPI, int NV>
__device__ __forceinline__ void epi32_v(int r, float s, float bw, float *__restrict__ y,
                                        double (&acc)[NV]) {
  if (EPI == EPI_BMAX || EPI == EPI_BMAXN) {
    const float t = bw - s;
    y[r] = t;
    if (EPI == EPI_BMAXN) acc[0] += (double)t * (double)t;
  } else {
    y[r] = s;
    if (EPI == EPI_DOT1 || EPI == EPI_DOT2) acc[0] += (double)bw * (double)s;
    if (EPI == EPI_DOT2) acc[NV - 1] += (double)s * (double)s;
  }
}

// CSR SpMV in FP32 over the same row-block plan as the FP64 kernel.
template <int EPI>
__global__ __launch_bounds__(XS_BLOCK) void k_spmv_csr32(
    const int *__restrict__ rp, const int *__restrict__ ci, const float *__restrict__ v,
    const float *__restrict__ x, float *__restrict__ y, const int *__restrict__ blk_row,
    const float *__restrict__ b, const float *__restrict__ w, xs_reduce R) {
  __shared__ float prod[XS_SPMV_LDS];
  constexpr int NV = (EPI == EPI_DOT2) ? 2 : 1;
  constexpr bool RED = (EPI == EPI_DOT1 || EPI == EPI_DOT2 || EPI == EPI_BMAXN);
  if (R.st != nullptr && R.st->done) return;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  const int r0 = blk_row[blockIdx.x], r1 = blk_row[blockIdx.x + 1];
  if (r1 > r0) {
    const int e0 = rp[r0], e1 = rp[r1];
    const int ne = e1 - e0;
    if (ne <= XS_SPMV_LDS) {
      for (int k = threadIdx.x; k < ne; k += XS_BLOCK) prod[k] = v[e0 + k] * x[ci[e0 + k]];
      __syncthreads();
      for (int r = r0 + threadIdx.x; r < r1; r += XS_BLOCK) {
        const int a = rp[r] - e0, bnd = rp[r + 1] - e0;
        float s = 0.f;
        for (int k = a; k < bnd; ++k) s += prod[k];
        epi32<EPI, NV>(r, s, y, b, w, acc);
      }
    } else {
      const int lane = threadIdx.x & (XS_WAVE - 1);
      for (int r = r0 + (int)(threadIdx.x / XS_WAVE); r < r1; r += XS_BLOCK / XS_WAVE) {
        const int a = rp[r], bnd = rp[r + 1];
        double s = 0.0;
        for (int k = a + lane; k < bnd; k += XS_WAVE) s += (double)(v[k] * x[ci[k]]);
        s = wave_sum(s);
        if (lane == 0) epi32<EPI, NV>(r, (float)s, y, b, w, acc);
      }
    }
  }
  if (RED) reduce_finish<NV>(acc, R);
}

// Row-binned vector CSR in FP32 (the scheme of k_spmv_vec in xs_kernels.hip).
template <int L, int EPI>
__global__ __launch_bounds__(XS_BLOCK) void k_spmv_vec32(
    int n, const int *__restrict__ rp, const int *__restrict__ ci, const float *__restrict__ v,
    const float *__restrict__ x, float *__restrict__ y, const int *__restrict__ long_rows,
    int nlong, int long_thr, const float *__restrict__ b, const float *__restrict__ w,
    xs_reduce R) {
  constexpr int NV = (EPI == EPI_DOT2) ? 2 : 1;
  constexpr bool RED = (EPI == EPI_DOT1 || EPI == EPI_DOT2 || EPI == EPI_BMAXN);
  if (R.st != nullptr && R.st->done) return;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  const int nlb = (nlong + XS_BLOCK / XS_WAVE - 1) / (XS_BLOCK / XS_WAVE);
  if ((int)blockIdx.x < nlb) {
    const int lane = threadIdx.x & (XS_WAVE - 1);
    const int idx = blockIdx.x * (XS_BLOCK / XS_WAVE) + threadIdx.x / XS_WAVE;
    if (idx < nlong) {
      const int r = long_rows[idx];
      double s = 0.0;
      for (int e = rp[r] + lane; e < rp[r + 1]; e += XS_WAVE) s += (double)(v[e] * x[ci[e]]);
      s = wave_sum(s);
      if (lane == 0) epi32<EPI, NV>(r, (float)s, y, b, w, acc);
    }
  } else {
    // same chunking and load batching as the FP64 vector-CSR kernel
    constexpr int GPB = XS_BLOCK / L;
    constexpr int UR = 4;
    const int sub = threadIdx.x % L;
    const int G = (int)gridDim.x - nlb;
    const int rpc = ((n + G - 1) / G + GPB - 1) / GPB * GPB;
    const int c0 = xcd_chunk((int)blockIdx.x - nlb, G) * rpc;
    const int c1 = min(n, c0 + rpc);
    for (int row0 = c0 + (int)threadIdx.x / L; row0 < c1; row0 += GPB * UR) {
      int a[UR], e1[UR];
      bool on[UR];
#pragma unroll
      for (int u = 0; u < UR; ++u) {
        const int row = row0 + u * GPB;
        a[u] = row < c1 ? rp[row] : 0;
        e1[u] = row < c1 ? rp[row + 1] : 0;
        on[u] = row < c1 && e1[u] - a[u] <= long_thr;
      }
      int cc[UR];
      float vv[UR], s[UR];
#pragma unroll
      for (int u = 0; u < UR; ++u) {
        const int e = a[u] + sub;
        const bool in = on[u] && e < e1[u];
        cc[u] = in ? ci[e] : -1;
        vv[u] = in ? v[e] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < UR; ++u) s[u] = cc[u] >= 0 ? vv[u] * x[cc[u]] : 0.f;
#pragma unroll
      for (int u = 0; u < UR; ++u)
        if (on[u])
          for (int e = a[u] + sub + L; e < e1[u]; e += L) s[u] += v[e] * x[ci[e]];
#pragma unroll
      for (int u = 0; u < UR; ++u) {
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s[u] += __shfl_xor(s[u], off, L);
        if (sub == 0 && on[u]) epi32<EPI, NV>(row0 + u * GPB, s[u], y, b, w, acc);
      }
    }
  }
  if (RED) reduce_finish<NV>(acc, R);
}

template <int L>
static void spmv_vec32_launch(hipStream_t s, const xs_csr &A, const float *v32,
                              const xs_spmv_plan &P, const float *x, float *y, int epi,
                              const float *b, const float *w, const xs_reduce &R) {
  const int nlb = (P.nlong + XS_BLOCK / XS_WAVE - 1) / (XS_BLOCK / XS_WAVE);
  long g = ((long)A.n + XS_BLOCK / L - 1) / (XS_BLOCK / L);
  const long cap = XS_MAX_REDUCE_BLOCKS - nlb > 1 ? XS_MAX_REDUCE_BLOCKS - nlb : 1;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  const dim3 grid((unsigned)(nlb + g)), blk(XS_BLOCK);
#define XS_VEC32(E) hipLaunchKernelGGL((k_spmv_vec32<L, E>), grid, blk, 0, s, A.n, A.rp, A.ci, v32, x, y, P.long_rows, P.nlong, P.long_thr, b, w, R)
  switch (epi) {
    case EPI_Y: XS_VEC32(EPI_Y); break;
    case EPI_BMAX: XS_VEC32(EPI_BMAX); break;
    case EPI_DOT1: XS_VEC32(EPI_DOT1); break;
    case EPI_DOT2: XS_VEC32(EPI_DOT2); break;
    case EPI_BMAXN: XS_VEC32(EPI_BMAXN); break;
  }
#undef XS_VEC32
}

// ELL SpMV in FP32, two rows per thread (8-byte value and index loads).
template <int EPI, bool O16>
__global__ __launch_bounds__(XS_BLOCK) void k_spmv_ell32(
    int n, int ld, int W, const int *__restrict__ ci, const short *__restrict__ co,
    const float *__restrict__ v, const float *__restrict__ x, float *__restrict__ y,
    const float *__restrict__ b, const float *__restrict__ w, xs_reduce R) {
  constexpr int NV = (EPI == EPI_DOT2) ? 2 : 1;
  constexpr bool RED = (EPI == EPI_DOT1 || EPI == EPI_DOT2 || EPI == EPI_BMAXN);
  if (R.st != nullptr && R.st->done) return;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  const int r = 2 * (xcd_chunk((int)blockIdx.x, (int)gridDim.x) * XS_BLOCK + (int)threadIdx.x);
  if (r < n) {
    // the epilogue's operand is loaded before the slabs, off the critical path
    constexpr bool OPB = (EPI == EPI_BMAX || EPI == EPI_BMAXN);
    constexpr bool OPW = (EPI == EPI_DOT1 || EPI == EPI_DOT2);
    float e0 = 0.f, e1 = 0.f;
    if (OPB || OPW) {
      const float *op = OPB ? b : w;
      e0 = op[r];
      if (r + 1 < n) e1 = op[r + 1];
    }
    float s0 = 0.f, s1 = 0.f;
    constexpr int UK = 4;   // slabs per batch: all loads of a batch, then all gathers
    for (int k0 = 0; k0 < W; k0 += UK) {
      // branch-free loads (a slot past W re-reads slab W - 1, masked from the
      // sums) so that the batch's loads are all issued before any wait
      int2 c[UK];
      float2 a[UK];
#pragma unroll
      for (int q = 0; q < UK; ++q) {
        const size_t p = (size_t)min(k0 + q, W - 1) * ld + r;
        if (O16) {
          const short2 o = *reinterpret_cast<const short2 *>(co + p);
          c[q] = make_int2(r + o.x, r + 1 + o.y);
        } else {
          c[q] = *reinterpret_cast<const int2 *>(ci + p);
        }
        a[q] = *reinterpret_cast<const float2 *>(v + p);
      }
      float x0[UK], x1[UK];
#pragma unroll
      for (int q = 0; q < UK; ++q) { x0[q] = x[c[q].x]; x1[q] = x[c[q].y]; }
#pragma unroll
      for (int q = 0; q < UK; ++q) {
        if (k0 + q < W) { s0 += a[q].x * x0[q]; s1 += a[q].y * x1[q]; }
      }
    }
    epi32_v<EPI, NV>(r, s0, e0, y, acc);
    if (r + 1 < n) epi32_v<EPI, NV>(r + 1, s1, e1, y, acc);
  }
  if (RED) reduce_finish<NV>(acc, R);
}

void xs_launch_spmv_csr32(hipStream_t s, const xs_csr &A, const float *v32,
                          const xs_spmv_plan &P, const float *x, float *y, int epi,
                          const float *b, const float *w, const xs_reduce *red) {
  xs_reduce R{};
  if (red) R = *red;
  if (!(A.rp && A.ci && v32 && P.blk_row && x && y)) { xs_raise_launch_guard(); return; }
  if (P.vec) {
    if (P.nlong > 0 && !P.long_rows) { xs_raise_launch_guard(); return; }
    switch (P.vec_L) {
      case 2: spmv_vec32_launch<2>(s, A, v32, P, x, y, epi, b, w, R); break;
      case 4: spmv_vec32_launch<4>(s, A, v32, P, x, y, epi, b, w, R); break;
      case 8: spmv_vec32_launch<8>(s, A, v32, P, x, y, epi, b, w, R); break;
      case 16: spmv_vec32_launch<16>(s, A, v32, P, x, y, epi, b, w, R); break;
      case 32: spmv_vec32_launch<32>(s, A, v32, P, x, y, epi, b, w, R); break;
      default: spmv_vec32_launch<64>(s, A, v32, P, x, y, epi, b, w, R); break;
    }
    return;
  }
  dim3 grid(P.nblk), blk(XS_BLOCK);
  switch (epi) {
    case EPI_Y: hipLaunchKernelGGL(k_spmv_csr32<EPI_Y>, grid, blk, 0, s, A.rp, A.ci, v32, x, y, P.blk_row, b, w, R); break;
    case EPI_BMAX: hipLaunchKernelGGL(k_spmv_csr32<EPI_BMAX>, grid, blk, 0, s, A.rp, A.ci, v32, x, y, P.blk_row, b, w, R); break;
    case EPI_DOT1: hipLaunchKernelGGL(k_spmv_csr32<EPI_DOT1>, grid, blk, 0, s, A.rp, A.ci, v32, x, y, P.blk_row, b, w, R); break;
    case EPI_DOT2: hipLaunchKernelGGL(k_spmv_csr32<EPI_DOT2>, grid, blk, 0, s, A.rp, A.ci, v32, x, y, P.blk_row, b, w, R); break;
    case EPI_BMAXN: hipLaunchKernelGGL(k_spmv_csr32<EPI_BMAXN>, grid, blk, 0, s, A.rp, A.ci, v32, x, y, P.blk_row, b, w, R); break;
  }
}

// k_spmv_ell32 over the FP32 copy of A's diagonal-slot values (no index
// loads; the zero slots of rows that lack a diagonal read a clamped x)
template <int EPI>
__global__ __launch_bounds__(XS_BLOCK) void k_spmv_dia32(
    int n, int ld, int W, const xs_dia_offs doff, const float *__restrict__ v,
    const float *__restrict__ x, float *__restrict__ y, const float *__restrict__ b,
    const float *__restrict__ w, xs_reduce R) {
  constexpr int NV = (EPI == EPI_DOT2) ? 2 : 1;
  constexpr bool RED = (EPI == EPI_DOT1 || EPI == EPI_DOT2 || EPI == EPI_BMAXN);
  if (R.st != nullptr && R.st->done) return;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  const int r = 2 * (xcd_chunk((int)blockIdx.x, (int)gridDim.x) * XS_BLOCK + (int)threadIdx.x);
  if (r < n) {
    constexpr bool OPB = (EPI == EPI_BMAX || EPI == EPI_BMAXN);
    constexpr bool OPW = (EPI == EPI_DOT1 || EPI == EPI_DOT2);
    float e0 = 0.f, e1 = 0.f;
    if (OPB || OPW) {
      const float *op = OPB ? b : w;
      e0 = op[r];
      if (r + 1 < n) e1 = op[r + 1];
    }
    float s0 = 0.f, s1 = 0.f;
    constexpr int UK = 4;
#pragma unroll
    for (int k0 = 0; k0 < XS_DIA_MAX; k0 += UK) {
      if (k0 >= W) break;
      float2 a[UK];
      int c0[UK], c1[UK];
#pragma unroll
      for (int q = 0; q < UK; ++q) {
        const int k = min(k0 + q, W - 1);
        a[q] = *reinterpret_cast<const float2 *>(v + (size_t)k * ld + r);
        c0[q] = min(max(r + doff.o[k], 0), n - 1);
        c1[q] = min(max(r + 1 + doff.o[k], 0), n - 1);
      }
      float x0[UK], x1[UK];
#pragma unroll
      for (int q = 0; q < UK; ++q) { x0[q] = x[c0[q]]; x1[q] = x[c1[q]]; }
#pragma unroll
      for (int q = 0; q < UK; ++q) {
        if (k0 + q < W) { s0 += a[q].x * x0[q]; s1 += a[q].y * x1[q]; }
      }
    }
    epi32_v<EPI, NV>(r, s0, e0, y, acc);
    if (r + 1 < n) epi32_v<EPI, NV>(r + 1, s1, e1, y, acc);
  }
  if (RED) reduce_finish<NV>(acc, R);
}

void xs_launch_spmv_dia32(hipStream_t s, const xs_ell &E, const float *d32, const float *x,
                          float *y, int epi, const float *b, const float *w,
                          const xs_reduce *red) {
  xs_reduce R{};
  if (red) R = *red;
  if (!(d32 && E.nd > 0 && E.nd <= XS_DIA_MAX && E.ld % 2 == 0 && x && y)) {
    xs_raise_launch_guard();
    return;
  }
  const int threads = (E.n + 1) / 2;
  dim3 grid((threads + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
#define XS_DIAK32(E_) hipLaunchKernelGGL((k_spmv_dia32<E_>), grid, blk, 0, s, E.n, E.ld, E.nd, E.doff, d32, x, y, b, w, R)
  switch (epi) {
    case EPI_Y: XS_DIAK32(EPI_Y); break;
    case EPI_BMAX: XS_DIAK32(EPI_BMAX); break;
    case EPI_DOT1: XS_DIAK32(EPI_DOT1); break;
    case EPI_DOT2: XS_DIAK32(EPI_DOT2); break;
    case EPI_BMAXN: XS_DIAK32(EPI_BMAXN); break;
  }
#undef XS_DIAK32
}

void xs_launch_spmv_ell32(hipStream_t s, const xs_ell &E, const float *v32, const float *x,
                          float *y, int epi, const float *b, const float *w,
                          const xs_reduce *red) {
  xs_reduce R{};
  if (red) R = *red;
  if (!(E.ci && v32 && x && y)) { xs_raise_launch_guard(); return; }
  const int threads = (E.n + 1) / 2;
  dim3 grid((threads + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
#define XS_ELLK32(E_, O) hipLaunchKernelGGL((k_spmv_ell32<E_, O>), grid, blk, 0, s, E.n, E.ld, E.W, E.ci, E.co, v32, x, y, b, w, R)
#define XS_ELLE32(E_) do { if (E.co) XS_ELLK32(E_, true); else XS_ELLK32(E_, false); } while (0)
  switch (epi) {
    case EPI_Y: XS_ELLE32(EPI_Y); break;
    case EPI_BMAX: XS_ELLE32(EPI_BMAX); break;
    case EPI_DOT1: XS_ELLE32(EPI_DOT1); break;
    case EPI_DOT2: XS_ELLE32(EPI_DOT2); break;
    case EPI_BMAXN: XS_ELLE32(EPI_BMAXN); break;
  }
#undef XS_ELLE32
#undef XS_ELLK32
}

// ---------------------------------------------------------------------------
// Sync-free SpTRSV in FP32 (same algorithm as k_trsv in xs_kernels.hip):
// a 32-bit signalling-NaN sentinel marks rows not yet solved.
// ---------------------------------------------------------------------------
#define XS_SENTINEL32 0xFFA0DEADu
#define XS_SPIN_LIMIT32 (1u << 22)

__device__ __forceinline__ void st_ready32(float *p, float v) {
  unsigned u = __float_as_uint(v);
  if (u == XS_SENTINEL32) u = 0x7FC00000u;
  __hip_atomic_store(reinterpret_cast<unsigned *>(p), u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned ld_relaxed32(const float *p) {
  return __hip_atomic_load(reinterpret_cast<unsigned *>(const_cast<float *>(p)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Up to G entries e, e+dir, ... (stopping at `stop`) of a row: columns c (-1
// past the stop), values a and the current x of those columns u (possibly
// the sentinel); branch-free loads, all in flight at once (as trsv_issue in
// xs_kernels.hip).
template <int G, class TF>
__device__ __forceinline__ void trsv_issue32(int e, int stop, int dir, const int *__restrict__ ci,
                                             const TF *__restrict__ lu, const float *x, int *c,
                                             float *a, unsigned *u) {
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int eq = e + q * dir;
    const bool in = dir > 0 ? eq < stop : eq > stop;
    const int cv = ci[in ? eq : e];
    const float av = fl(lu, in ? eq : e);
    c[q] = in ? cv : -1;
    a[q] = in ? av : 0.f;
  }
#pragma unroll
  for (int q = 0; q < G; ++q) u[q] = ld_relaxed32(x + (c[q] >= 0 ? c[q] : c[0]));
}

template <int G>
__device__ __forceinline__ bool trsv_pending32(const int *c, const unsigned *u) {
  bool pend = false;
#pragma unroll
  for (int q = 0; q < G; ++q) pend |= (c[q] >= 0 && u[q] == XS_SENTINEL32);
  return pend;
}

// fixed pairwise tree over G values (G a power of two)
template <int G>
__device__ __forceinline__ float trsv_tree32(const float *p) {
  float t[G];
#pragma unroll
  for (int q = 0; q < G; ++q) t[q] = p[q];
#pragma unroll
  for (int w = G / 2; w >= 1; w /= 2) {
#pragma unroll
    for (int q = 0; q < w; ++q) t[q] = t[q] + t[q + w];
  }
  return t[0];
}

// acc minus one batch of products (0 where there is no entry), summed by the
// fixed tree, each product rounded on its own: the same roundings whatever
// the timing (as trsv_sub in xs_kernels.hip)
template <int G>
__device__ __forceinline__ float trsv_sub32(float acc, const int *c, const float *a,
                                            const unsigned *u) {
#pragma clang fp contract(off)
  float p[G];
#pragma unroll
  for (int q = 0; q < G; ++q) p[q] = c[q] >= 0 ? a[q] * __uint_as_float(u[q]) : 0.f;
  return acc - trsv_tree32<G>(p);
}

// acc - one batch of products, waiting (bounded) for the unpublished values
template <int G, class TF>
__device__ __forceinline__ float trsv_gather32(float acc, int e, int stop, int dir,
                                               const int *__restrict__ ci,
                                               const TF *__restrict__ lu, const float *x,
                                               unsigned *err) {
  unsigned u[G];
  int c[G];
  float a[G];
  trsv_issue32<G>(e, stop, dir, ci, lu, x, c, a, u);
  unsigned polls = 0;
  while (trsv_pending32<G>(c, u)) {
    if (++polls > XS_SPIN_LIMIT32) {  // bounded wait: flag it, continue with NaN
      atomicOr(err, 1u);
#pragma unroll
      for (int q = 0; q < G; ++q)
        if (u[q] == XS_SENTINEL32) u[q] = 0x7FC00000u;
      break;
    }
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int q = 0; q < G; ++q)
      if (c[q] >= 0 && u[q] == XS_SENTINEL32) u[q] = ld_relaxed32(x + c[q]);
  }
  return trsv_sub32<G>(acc, c, a, u);
}

// FP32 sync-free solve: the algorithm of k_trsv (xs_kernels.hip) with FP32
// vectors and FP32/BF16/FP16 factor values -- row groups handed out by
// tickets, chain groups by an affine scan after all external values are in,
// the other groups lane by lane with every row published as soon as it is
// final (two batches of external entries held, a bit mask of the entries
// still unpublished, every later lane re-polled per round trip).
template <bool UPPER, class TF, int G>
__global__ __launch_bounds__(XS_BLOCK) void k_trsv32(int n, const int *__restrict__ rp,
                                                     const int *__restrict__ ci,
                                                     const TF *__restrict__ lu,
                                                     const int *__restrict__ dpos,
                                                     const float *__restrict__ dinv,
                                                     const float *__restrict__ b, float *x,
                                                     unsigned *tickets, const xs_tsched S,
                                                     const xs_state *st) {
#pragma clang fp contract(off)
  if (st != nullptr && st->done) return;
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
    const int r = UPPER ? r1 - 1 - lane : r0 + lane;
    float acc = 0.f, di = 1.f;
    int eb = 0, ee = 0;   // in-wave entries [eb, ee)
    int xb = 0, xe = 0;   // entries outside the group
    int iw_c[XS_INW];     // first in-wave entries in solve order, loaded before any wait
    float iw_v[XS_INW];
#pragma unroll
    for (int q = 0; q < XS_INW; ++q) { iw_c[q] = -1; iw_v[q] = 0.f; }
    if (act) {
      acc = b[r];
      if (!UPPER) {
        xb = rp[r];
        const int end = dpos[r];
        xe = xb;
        while (xe < end && ci[xe] < r0) ++xe;
        eb = xe; ee = end;
      } else {
        di = dinv[r];
        const int beg = dpos[r] + 1;
        xe = rp[r + 1];
        xb = xe;
        while (xb > beg && ci[xb - 1] >= r1) --xb;
        eb = beg; ee = xb;
      }
#pragma unroll
      for (int q = 0; q < XS_INW; ++q) {
        const int e = UPPER ? ee - 1 - q : eb + q;
        if (q < ee - eb) { iw_c[q] = ci[e]; iw_v[q] = fl(lu, e); }
      }
    }
    // external entries in solve order: L from xb up, U from xe - 1 down
    const int dir = UPPER ? -1 : 1;
    const int estop = UPPER ? xb - 1 : xe;
    int ep = UPPER ? xe - 1 : xb;
    const int nint = ee - eb;
    const int nb = UPPER ? r + 1 : r - 1;
    const bool chain = !act || nint == 0 || (nint == 1 && iw_c[0] == nb);
    if (!S.prog && __all(chain)) {
      if (act)
        for (; UPPER ? ep > estop : ep < estop; ep += G * dir)
          acc = trsv_gather32<G>(acc, ep, estop, dir, ci, lu, x, err);
      float c = act ? di * acc : 0.f;
      float m = (act && nint == 1) ? -di * iw_v[0] : 0.f;
      wave_affine_scan_f(c, m);
      if (act) st_ready32(x + r, c);
    } else {
      // held: the first batch with an unpublished value and the one after
      // it; hc the columns (-1: none), hv the value until x of the column
      // arrives, then the product (0 for no entry), hm the waiting entries
      int hc[2 * G];
      float hv[2 * G];
      unsigned hm = 0;
      bool stalled = false, more = false;   // more: the second batch is held
      {
        int pc[G];
        float pa[G];
        unsigned pu[G];
#pragma unroll
        for (int q = 0; q < G; ++q) { pc[q] = -1; pa[q] = 0.f; pu[q] = 0u; }
        if (act) {
          for (; UPPER ? ep > estop : ep < estop; ep += G * dir) {
            trsv_issue32<G>(ep, estop, dir, ci, lu, x, pc, pa, pu);
            if (trsv_pending32<G>(pc, pu)) { stalled = true; break; }
            acc = trsv_sub32<G>(acc, pc, pa, pu);
          }
        }
#pragma unroll
        for (int q = 0; q < G; ++q) {
          const bool wait = pc[q] >= 0 && pu[q] == XS_SENTINEL32;
          hc[q] = stalled ? pc[q] : -1;
          hv[q] = (!stalled || pc[q] < 0) ? 0.f : wait ? pa[q] : pa[q] * __uint_as_float(pu[q]);
          if (stalled && wait) hm |= 1u << q;
        }
        const int e2 = ep + G * dir;
        more = stalled && (UPPER ? e2 > estop : e2 < estop);
        if (more) trsv_issue32<G>(e2, estop, dir, ci, lu, x, pc, pa, pu);
#pragma unroll
        for (int q = 0; q < G; ++q) {
          const bool wait = pc[q] >= 0 && pu[q] == XS_SENTINEL32;
          hc[G + q] = more ? pc[q] : -1;
          hv[G + q] = (!more || pc[q] < 0) ? 0.f : wait ? pa[q] : pa[q] * __uint_as_float(pu[q]);
          if (more && wait) hm |= 1u << (G + q);
        }
      }
      float ain = 0.f;   // the in-wave products, subtracted in solve order
      int left = act ? nint - XS_INW : 0;
      int enext = UPPER ? ee - 1 - XS_INW : eb + XS_INW;
      for (int k = 0; k < nrows; ++k) {
        // the turn check reads only hm (no memory result: a wait for a load
        // would also wait for the previous turn's publishing store)
        unsigned polls = 0;
        bool expired = false;
        while (__builtin_amdgcn_readlane((int)hm, k)) {
          if (++polls > XS_SPIN_LIMIT32) {   // bounded wait: flag it, continue with NaN
            if (lane == k) atomicOr(err, 1u);
            expired = true;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
          if (hm && lane >= k) {
            unsigned pu[2 * G];
#pragma unroll
            for (int q = 0; q < 2 * G; ++q)
              pu[q] = ((hm >> q) & 1u) ? ld_relaxed32(x + hc[q]) : XS_SENTINEL32;
#pragma unroll
            for (int q = 0; q < 2 * G; ++q)
              if (pu[q] != XS_SENTINEL32) {
                hv[q] = hv[q] * __uint_as_float(pu[q]);
                hm &= ~(1u << q);
              }
          }
        }
        float vk = 0.f;
        if (lane == k) {
          if (stalled) {
            acc -= trsv_tree32<G>(hv);
            if (more) acc -= trsv_tree32<G>(hv + G);   // else all zeros: acc is unchanged
            for (ep += 2 * G * dir; UPPER ? ep > estop : ep < estop; ep += G * dir)
              acc = trsv_gather32<G>(acc, ep, estop, dir, ci, lu, x, err);
          }
          vk = expired ? __uint_as_float(0x7FC00000u) : (acc + ain) * di;
          st_ready32(x + r, vk);
        }
        const float xk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vk), k));
        const int rk = UPPER ? r1 - 1 - k : r0 + k;
        if (iw_c[0] == rk) {
          ain -= iw_v[0] * xk;
#pragma unroll
          for (int q = 0; q + 1 < XS_INW; ++q) { iw_c[q] = iw_c[q + 1]; iw_v[q] = iw_v[q + 1]; }
          if (left > 0) {
            iw_c[XS_INW - 1] = ci[enext];
            iw_v[XS_INW - 1] = fl(lu, enext);
            enext += UPPER ? -1 : 1;
            --left;
          } else {
            iw_c[XS_INW - 1] = -1;
          }
        }
      }
    }
  }
  }
  trsv_retire(tickets);
}

// One Jacobi sweep of an FP32/BF16/FP16 factor with FP32 vectors (the
// reference's iterative GPU ILU apply, see k_tri_sweep in xs_kernels.hip).
template <bool UPPER, class TF>
__global__ __launch_bounds__(XS_BLOCK) void k_tri_sweep32(int n, const int *__restrict__ rp,
                                                          const int *__restrict__ ci,
                                                          const TF *__restrict__ lu,
                                                          const int *__restrict__ dpos,
                                                          const float *__restrict__ dinv,
                                                          const float *__restrict__ b,
                                                          const float *__restrict__ yo,
                                                          float *__restrict__ yn,
                                                          const xs_state *st) {
  if (st != nullptr && st->done) return;
  const int i = blockIdx.x * XS_BLOCK + threadIdx.x;
  if (i >= n) return;
  float r = 0.f;
  if (yo) {
    const int e0 = UPPER ? dpos[i] + 1 : rp[i], e1 = UPPER ? rp[i + 1] : dpos[i];
    for (int e = e0; e < e1; ++e) r += fl(lu, e) * yo[ci[e]];
  }
  yn[i] = UPPER ? dinv[i] * (b[i] - r) : b[i] - r;
}

void xs_tri_sweep32(hipStream_t s, bool upper, int n, const int *rp, const int *ci, const void *lu,
                    int ftype, const int *dpos, const float *dinv, const float *b, const float *yo,
                    float *yn, const xs_state *st) {
  if (!(rp && ci && lu && dpos && b && yn && (!upper || dinv))) { xs_raise_launch_guard(); return; }
  const dim3 grid((n + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
#define XS_SW(TF)                                                                                 \
  do {                                                                                            \
    const TF *f = static_cast<const TF *>(lu);                                                    \
    if (upper) hipLaunchKernelGGL((k_tri_sweep32<true, TF>), grid, blk, 0, s, n, rp, ci, f, dpos, dinv, b, yo, yn, st); \
    else hipLaunchKernelGGL((k_tri_sweep32<false, TF>), grid, blk, 0, s, n, rp, ci, f, dpos, dinv, b, yo, yn, st); \
  } while (0)
  if (ftype == XS_BF16) XS_SW(__hip_bfloat16);
  else if (ftype == XS_F16) XS_SW(__half);
  else XS_SW(float);
#undef XS_SW
}

// FP32 Jacobi sweep over the ELL copy of a factor part (FP32 values); the
// arithmetic of k_tri_sweep32 entry by entry, FUSED as in the FP64 kernel.
// pairs of factor values at i, i + 1 (i even: 8- or 4-byte aligned), converted
// exactly as fl() converts one
__device__ __forceinline__ float2 fl2(const float *p, size_t i) {
  return *reinterpret_cast<const float2 *>(p + i);
}
__device__ __forceinline__ float2 fl2(const __hip_bfloat16 *p, size_t i) {
  const unsigned u = *reinterpret_cast<const unsigned *>(p + i);   // [i] low, [i + 1] high
  return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xFFFF0000u));
}
__device__ __forceinline__ float2 fl2(const __half *p, size_t i) {
  return __half22float2(*reinterpret_cast<const __half2 *>(p + i));
}

// Two rows per thread (r, r + 1; ld even): one 4-byte offset load and one
// paired value load per slab cover both rows, half the load instructions of
// a row per thread (which matters most for 2-byte BF16 / FP16 values).  Per
// row the arithmetic of the one-row form, slab by slab.  SX (fused L pass
// only): the right-hand side is BiCGSTAB's s = r - gama u, formed from b = r
// and ru = u at every row read and written to sv (k_bicg_s32's expression).
template <bool UPPER, bool FUSED, bool O16, typename TF, bool SX = false>
__global__ __launch_bounds__(XS_BLOCK) void k_tri_sweep_ell32(int n, int ld, int W,
                                                              const int *__restrict__ eci,
                                                              const short *__restrict__ eco,
                                                              const TF *__restrict__ ev,
                                                              const int *__restrict__ ww,
                                                              const float *__restrict__ dinv,
                                                              const float *__restrict__ b,
                                                              const float *__restrict__ yo,
                                                              float *__restrict__ yn,
                                                              const xs_state *st,
                                                              const float *__restrict__ ru = nullptr,
                                                              float *__restrict__ sv = nullptr) {
  static_assert(!SX || (FUSED && !UPPER), "s is formed in the fused L pass only");
  if (st != nullptr && st->done) return;
  const int r = 2 * (xcd_chunk((int)blockIdx.x, (int)gridDim.x) * XS_BLOCK + (int)threadIdx.x);
  if (r >= n) return;
  const bool two = r + 1 < n;
  const float gama = SX ? (float)st->sc[S_GAMA] : 0.f;
  // the rows' own operands, loaded before the slabs (off the critical path)
  const float b0 = SX ? fmaf(-gama, ru[r], b[r]) : b[r];
  const float b1 = two ? (SX ? fmaf(-gama, ru[r + 1], b[r + 1]) : b[r + 1]) : 0.f;
  if (SX) {
    sv[r] = b0;
    if (two) sv[r + 1] = b1;
  }
  const float d0 = UPPER ? dinv[r] : 1.f, d1 = (UPPER && two) ? dinv[r + 1] : 1.f;
  W = ww ? ww[r / XS_WAVE] : W;   // slabs past the group's longest row are padding
  float s0 = 0.f, s1 = 0.f;
  for (int k0 = 0; k0 < W; k0 += 4) {
    int c0[4], c1[4];
    float a0[4], a1[4], y0[4], y1[4];
    // branch-free loads: a slot past W re-reads slab W - 1 and a padding
    // entry gathers the row's own values, all masked from the sums, so that
    // the batch's loads are issued before any wait
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool in = k0 + q < W;
      const size_t p = (size_t)min(k0 + q, W - 1) * ld + r;
      if (O16) {
        const short2 o = *reinterpret_cast<const short2 *>(eco + p);
        c0[q] = (!in || o.x == XS_OFF16_PAD) ? -1 : r + o.x;
        c1[q] = (!in || o.y == XS_OFF16_PAD || !two) ? -1 : r + 1 + o.y;
      } else {
        const int2 cc = *reinterpret_cast<const int2 *>(eci + p);
        c0[q] = in ? cc.x : -1;
        c1[q] = (in && two) ? cc.y : -1;
      }
      const float2 av = fl2(ev, p);
      a0[q] = av.x;
      a1[q] = av.y;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int g0 = c0[q] >= 0 ? c0[q] : r, g1 = c1[q] >= 0 ? c1[q] : r;
      if (SX) {
        y0[q] = fmaf(-gama, ru[g0], b[g0]) - 0.f;
        y1[q] = fmaf(-gama, ru[g1], b[g1]) - 0.f;
      } else if (FUSED) {
        y0[q] = UPPER ? dinv[g0] * (b[g0] - 0.f) : b[g0] - 0.f;
        y1[q] = UPPER ? dinv[g1] * (b[g1] - 0.f) : b[g1] - 0.f;
      } else {
        y0[q] = yo[g0];
        y1[q] = yo[g1];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (c0[q] >= 0) s0 += a0[q] * y0[q];
      if (c1[q] >= 0) s1 += a1[q] * y1[q];
    }
  }
  yn[r] = UPPER ? d0 * (b0 - s0) : b0 - s0;
  if (two) yn[r + 1] = UPPER ? d1 * (b1 - s1) : b1 - s1;
}

// BiCGSTAB's p = r + beta (p - omega u) in FP32 (beta in sc[S_OMEGA], omega
// in sc[S_ALPHA]): one expression shared by the update kernel and the L pass
// that forms p, so both round alike
__device__ __forceinline__ float bicg_p32_val(float r, float p, float u, float beta, float omega) {
  return r + beta * (p - omega * u);
}

// k_tri_sweep_ell32 over the diagonal-slot copy of a factor part (E.dvlo in
// the factor storage type, E.nd slots): two rows per thread, one paired
// value load per slot and no index loads; the zero slots of rows that lack a
// diagonal read a clamped column and add zero products.
// SX: s = b - gama ru formed here (b = r, ru = u), written to sv; PX: p =
// b + beta (pin - omega ru) formed here (b = r, ru = u), written to sv
template <bool UPPER, bool FUSED, typename TF, bool SX = false, bool PX = false>
__global__ __launch_bounds__(XS_BLOCK) void k_tri_sweep_dia32(int n, int ld, int nd,
                                                              const xs_dia_offs doff,
                                                              const TF *__restrict__ ev,
                                                              const float *__restrict__ dinv,
                                                              const float *__restrict__ b,
                                                              const float *__restrict__ yo,
                                                              float *__restrict__ yn,
                                                              const xs_state *st,
                                                              const float *__restrict__ ru = nullptr,
                                                              float *__restrict__ sv = nullptr,
                                                              const float *__restrict__ pin = nullptr) {
  static_assert(!SX || (FUSED && !UPPER), "s is formed in the fused L pass only");
  static_assert(!PX || (FUSED && !UPPER && !SX), "p is formed in the fused L pass only");
  if (st != nullptr && st->done) return;
  const int r = 2 * (xcd_chunk((int)blockIdx.x, (int)gridDim.x) * XS_BLOCK + (int)threadIdx.x);
  if (r >= n) return;
  const bool two = r + 1 < n;
  const float gama = SX ? (float)st->sc[S_GAMA] : 0.f;
  const float beta = PX ? (float)st->sc[S_OMEGA] : 0.f, omega = PX ? (float)st->sc[S_ALPHA] : 0.f;
  const float b0 = SX ? fmaf(-gama, ru[r], b[r]) : PX ? bicg_p32_val(b[r], pin[r], ru[r], beta, omega) : b[r];
  const float b1 = two ? (SX ? fmaf(-gama, ru[r + 1], b[r + 1])
                         : PX ? bicg_p32_val(b[r + 1], pin[r + 1], ru[r + 1], beta, omega) : b[r + 1])
                       : 0.f;
  if (SX || PX) {
    sv[r] = b0;
    if (two) sv[r + 1] = b1;
  }
  const float d0 = UPPER ? dinv[r] : 1.f, d1 = (UPPER && two) ? dinv[r + 1] : 1.f;
  float s0 = 0.f, s1 = 0.f;
#pragma unroll
  for (int k0 = 0; k0 < XS_DIA_MAX; k0 += 4) {
    if (k0 >= nd) break;
    int c0[4], c1[4];
    float a0[4], a1[4], y0[4], y1[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = min(k0 + q, nd - 1);
      const float2 av = fl2(ev, (size_t)k * ld + r);
      a0[q] = av.x;
      a1[q] = av.y;
      c0[q] = min(max(r + doff.o[k], 0), n - 1);
      c1[q] = min(max(r + 1 + doff.o[k], 0), n - 1);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int g0 = c0[q], g1 = c1[q];
      if (SX) {
        y0[q] = fmaf(-gama, ru[g0], b[g0]) - 0.f;
        y1[q] = fmaf(-gama, ru[g1], b[g1]) - 0.f;
      } else if (PX) {
        y0[q] = bicg_p32_val(b[g0], pin[g0], ru[g0], beta, omega) - 0.f;
        y1[q] = bicg_p32_val(b[g1], pin[g1], ru[g1], beta, omega) - 0.f;
      } else if (FUSED) {
        y0[q] = UPPER ? dinv[g0] * (b[g0] - 0.f) : b[g0] - 0.f;
        y1[q] = UPPER ? dinv[g1] * (b[g1] - 0.f) : b[g1] - 0.f;
      } else {
        y0[q] = yo[g0];
        y1[q] = yo[g1];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (k0 + q < nd) {
        s0 += a0[q] * y0[q];
        s1 += a1[q] * y1[q];
      }
    }
  }
  yn[r] = UPPER ? d0 * (b0 - s0) : b0 - s0;
  if (two) yn[r + 1] = UPPER ? d1 * (b1 - s1) : b1 - s1;
}

static bool tell_dia32(const xs_tell &E) {
  return E.nd > 0 && E.nd <= XS_DIA_MAX && E.dvlo != nullptr;
}

void xs_tri_sweep_ell32(hipStream_t s, bool upper, bool fused, int n, const xs_tell &E, int ftype,
                        const float *dinv, const float *b, const float *yo, float *yn,
                        const xs_state *st) {
  if (tell_dia32(E)) {
    if (!(E.ld >= n && E.ld % 2 == 0 && b && yn && (!upper || dinv) && (fused || yo))) {
      xs_raise_launch_guard();
      return;
    }
    const int pairs = (n + 1) / 2;
    const dim3 grid((pairs + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
#define XS_SWD32(U, FU, T) hipLaunchKernelGGL((k_tri_sweep_dia32<U, FU, T>), grid, blk, 0, s, n, E.ld, E.nd, E.doff, static_cast<const T *>(E.dvlo), dinv, b, yo, yn, st)
#define XS_SWT(U, FU) do { if (ftype == XS_BF16) XS_SWD32(U, FU, __hip_bfloat16); else if (ftype == XS_F16) XS_SWD32(U, FU, __half); else XS_SWD32(U, FU, float); } while (0)
    if (upper) { if (fused) XS_SWT(true, true); else XS_SWT(true, false); }
    else { if (fused) XS_SWT(false, true); else XS_SWT(false, false); }
#undef XS_SWT
#undef XS_SWD32
    return;
  }
  if (!((E.W == 0 || (E.ci && E.vlo)) && E.ld >= n && E.ld % 2 == 0 && b && yn &&
        (!upper || dinv) && (fused || yo))) {
    xs_raise_launch_guard();
    return;
  }
  const int pairs = (n + 1) / 2;
  const dim3 grid((pairs + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
#define XS_SWK32(U, FU, O, T) hipLaunchKernelGGL((k_tri_sweep_ell32<U, FU, O, T>), grid, blk, 0, s, n, E.ld, E.W, E.ci, E.co, static_cast<const T *>(E.vlo), E.ww, dinv, b, yo, yn, st)
#define XS_SWT32(U, FU, O) do { if (ftype == XS_BF16) XS_SWK32(U, FU, O, __hip_bfloat16); else if (ftype == XS_F16) XS_SWK32(U, FU, O, __half); else XS_SWK32(U, FU, O, float); } while (0)
#define XS_SWE32(U, FU) do { if (E.co) XS_SWT32(U, FU, true); else XS_SWT32(U, FU, false); } while (0)
  if (upper) { if (fused) XS_SWE32(true, true); else XS_SWE32(true, false); }
  else { if (fused) XS_SWE32(false, true); else XS_SWE32(false, false); }
#undef XS_SWE32
#undef XS_SWT32
#undef XS_SWK32
}

void xs_tri_sweep_ell32_s(hipStream_t s, int n, const xs_tell &E, int ftype, const float *r,
                          const float *u, float *sv, float *yn, const xs_state *st) {
  if (tell_dia32(E)) {
    if (!(E.ld >= n && E.ld % 2 == 0 && r && u && sv && yn && st)) {
      xs_raise_launch_guard();
      return;
    }
    const int pairs = (n + 1) / 2;
    const dim3 grid((pairs + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
#define XS_SWS(T) hipLaunchKernelGGL((k_tri_sweep_dia32<false, true, T, true>), grid, blk, 0, s, n, E.ld, E.nd, E.doff, static_cast<const T *>(E.dvlo), nullptr, r, nullptr, yn, st, u, sv)
    if (ftype == XS_BF16) XS_SWS(__hip_bfloat16);
    else if (ftype == XS_F16) XS_SWS(__half);
    else XS_SWS(float);
#undef XS_SWS
    return;
  }
  if (!((E.W == 0 || (E.ci && E.vlo)) && E.ld >= n && E.ld % 2 == 0 && r && u && sv && yn && st)) {
    xs_raise_launch_guard();
    return;
  }
  const int pairs = (n + 1) / 2;
  const dim3 grid((pairs + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
#define XS_SWS32(O, T) hipLaunchKernelGGL((k_tri_sweep_ell32<false, true, O, T, true>), grid, blk, 0, s, n, E.ld, E.W, E.ci, E.co, static_cast<const T *>(E.vlo), E.ww, nullptr, r, nullptr, yn, st, u, sv)
#define XS_SWT32S(O) do { if (ftype == XS_BF16) XS_SWS32(O, __hip_bfloat16); else if (ftype == XS_F16) XS_SWS32(O, __half); else XS_SWS32(O, float); } while (0)
  if (E.co) XS_SWT32S(true); else XS_SWT32S(false);
#undef XS_SWT32S
#undef XS_SWS32
}

bool xs_tell32_dia(const xs_tell &E) { return tell_dia32(E); }

void xs_tri_sweep_ell32_p(hipStream_t s, int n, const xs_tell &E, int ftype, const float *r,
                          const float *u, const float *pin, float *pout, float *yn,
                          const xs_state *st) {
  if (!(tell_dia32(E) && E.ld >= n && E.ld % 2 == 0 && r && u && pin && pout && pin != pout && yn &&
        st)) {
    xs_raise_launch_guard();
    return;
  }
  const int pairs = (n + 1) / 2;
  const dim3 grid((pairs + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
#define XS_SWP(T) hipLaunchKernelGGL((k_tri_sweep_dia32<false, true, T, false, true>), grid, blk, 0, s, n, E.ld, E.nd, E.doff, static_cast<const T *>(E.dvlo), nullptr, r, nullptr, yn, st, u, pout, pin)
  if (ftype == XS_BF16) XS_SWP(__hip_bfloat16);
  else if (ftype == XS_F16) XS_SWP(__half);
  else XS_SWP(float);
#undef XS_SWP
}

__global__ void k_fill_u32(unsigned *p, int n, unsigned val) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = val;
}

static inline int vgrid(int n) {
  int g = (n + XS_BLOCK * 4 - 1) / (XS_BLOCK * 4);
  if (g > XS_MAX_REDUCE_BLOCKS) g = XS_MAX_REDUCE_BLOCKS;
  return g < 1 ? 1 : g;
}

void xs_trsv32(hipStream_t s, bool upper, int n, const int *rp, const int *ci, const void *lu,
               int ftype, const int *dpos, const float *dinv, const float *b, float *x,
               unsigned *tickets, const xs_tsched &S, const xs_state *st) {
  if (!(rp && ci && lu && dpos && b && x && tickets) || (upper && !dinv) || S.max_wg <= 0 ||
      (S.gb && S.ngroups <= 0)) {
    xs_raise_launch_guard();
    return;
  }
  hipLaunchKernelGGL(k_fill_u32, dim3(vgrid(n)), dim3(XS_BLOCK), 0, s,
                     reinterpret_cast<unsigned *>(x), n, XS_SENTINEL32);
  const int nw = S.gb ? S.ngroups : (n + XS_WAVE - 1) / XS_WAVE;
  const int tk = (nw + 3) / 4;
  dim3 grid(tk < S.max_wg ? tk : S.max_wg), blk(XS_BLOCK);
#define TRSV32G(U, T, G) hipLaunchKernelGGL((k_trsv32<U, T, G>), grid, blk, 0, s, n, rp, ci, (const T *)lu, dpos, dinv, b, x, tickets, S, st)
#define TRSV32(U, T) do { if (S.gather == 16) TRSV32G(U, T, 16); else TRSV32G(U, T, XS_GATHER); } while (0)
  if (ftype == XS_F32) { if (upper) TRSV32(true, float); else TRSV32(false, float); }
  else if (ftype == XS_BF16) { if (upper) TRSV32(true, __hip_bfloat16); else TRSV32(false, __hip_bfloat16); }
  else { if (upper) TRSV32(true, __half); else TRSV32(false, __half); }
#undef TRSV32
#undef TRSV32G
}

// ---------------------------------------------------------------------------
// conversions and FP32 BiCGSTAB updates
// ---------------------------------------------------------------------------
__global__ void k_cvt_factor(int nnz, const double *__restrict__ src, void *dst, int ftype) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const double v = src[e];
  if (ftype == XS_F32) reinterpret_cast<float *>(dst)[e] = (float)v;
  else if (ftype == XS_BF16) reinterpret_cast<__hip_bfloat16 *>(dst)[e] = __float2bfloat16((float)v);
  else reinterpret_cast<__half *>(dst)[e] = __float2half((float)v);
}
void xs_cvt_factor(hipStream_t s, int nnz, const double *src, void *dst, int ftype) {
  hipLaunchKernelGGL(k_cvt_factor, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, src, dst, ftype);
}

__global__ void k_d2f(int n, const double *__restrict__ a, float *__restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (float)a[i];
}
void xs_d2f(hipStream_t s, int n, const double *a, float *b) {
  hipLaunchKernelGGL(k_d2f, dim3(vgrid(n)), dim3(XS_BLOCK), 0, s, n, a, b);
}

// r32 = r64 / ||r64|| (norm from sc[slot])
__global__ void k_d2f_scaled(int n, const double *__restrict__ a, float *__restrict__ b,
                             const xs_state *sref, int slot) {
  const double inv = 1.0 / sqrt(sref->sc[slot]);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    b[i] = (float)(a[i] * inv);
}
void xs_d2f_scaled(hipStream_t s, int n, const double *a, float *b, const xs_state *sref, int slot) {
  hipLaunchKernelGGL(k_d2f_scaled, dim3(vgrid(n)), dim3(XS_BLOCK), 0, s, n, a, b, sref, slot);
}

// x64 += ||r64|| * d32
__global__ void k_add_f2d(int n, const float *__restrict__ d, double *__restrict__ x,
                          const xs_state *sref, int slot) {
  const double sc = sqrt(sref->sc[slot]);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    x[i] += sc * (double)d[i];
}
void xs_add_f2d(hipStream_t s, int n, const float *d, double *x, const xs_state *sref, int slot) {
  hipLaunchKernelGGL(k_add_f2d, dim3(vgrid(n)), dim3(XS_BLOCK), 0, s, n, d, x, sref, slot);
}

// x64 += sigma d32, d32 = 0 (sigma = sqrt(sc[slot]))
__global__ void k_flush_f2d(int n, float *__restrict__ d, double *__restrict__ x,
                            const xs_state *sref, int slot) {
  const double sc = sqrt(sref->sc[slot]);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += sc * (double)d[i];
    d[i] = 0.f;
  }
}
void xs_flush_f2d(hipStream_t s, int n, float *d, double *x, const xs_state *sref, int slot) {
  hipLaunchKernelGGL(k_flush_f2d, dim3(vgrid(n)), dim3(XS_BLOCK), 0, s, n, d, x, sref, slot);
}

// r32 = r64 / sigma (sigma = sqrt(sref->sc[slot])) with r32.r32 and rhat.r32
// reduced into R (the replacement of BiCGSTAB's recursive residual)
__global__ __launch_bounds__(XS_BLOCK) void k_replace32(int n, const double *__restrict__ r64,
                                                        float *__restrict__ r,
                                                        const float *__restrict__ rh,
                                                        const xs_state *sref, int slot,
                                                        xs_reduce R) {
  const double inv = 1.0 / sqrt(sref->sc[slot]);
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    const float ri = (float)(r64[i] * inv);
    r[i] = ri;
    acc[0] += (double)ri * (double)ri;
    acc[1] += (double)rh[i] * (double)ri;
  }
  reduce_finish<2>(acc, R);
}
void xs_replace32(hipStream_t s, int n, const double *r64, float *r, const float *rh,
                  const xs_state *sref, int slot, const xs_reduce &red) {
  hipLaunchKernelGGL(k_replace32, dim3(vgrid(n)), dim3(XS_BLOCK), 0, s, n, r64, r, rh, sref, slot,
                     red);
}

__global__ void k_zero32(int n, float *__restrict__ a) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0.f;
}
void xs_zero32(hipStream_t s, int n, float *a) {
  hipLaunchKernelGGL(k_zero32, dim3(vgrid(n)), dim3(XS_BLOCK), 0, s, n, a);
}
__global__ void k_copy32(int n, const float *__restrict__ a, float *__restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = a[i];
}
void xs_copy32(hipStream_t s, int n, const float *a, float *b) {
  hipLaunchKernelGGL(k_copy32, dim3(vgrid(n)), dim3(XS_BLOCK), 0, s, n, a, b);
}
__global__ void k_scale32(int n, const float *__restrict__ d, const float *__restrict__ a,
                          float *__restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = d[i] * a[i];
}
void xs_scale32(hipStream_t s, int n, const float *d, const float *a, float *b) {
  hipLaunchKernelGGL(k_scale32, dim3(vgrid(n)), dim3(XS_BLOCK), 0, s, n, d, a, b);
}

__global__ __launch_bounds__(XS_BLOCK) void k_bicg_p32(int n, const float *__restrict__ r,
                                                       const float *__restrict__ u,
                                                       float *__restrict__ p,
                                                       const float *__restrict__ dinv,
                                                       float *__restrict__ z, xs_state *st) {
  if (st->done) return;
  const float omega = (float)st->sc[S_OMEGA], alpha = (float)st->sc[S_ALPHA];
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    const float pi = bicg_p32_val(r[i], p[i], u[i], omega, alpha);
    p[i] = pi;
    if (dinv) z[i] = dinv[i] * pi;
  }
}
void xs_bicg_p32(hipStream_t s, int n, const float *r, const float *u, float *p,
                 const float *dinv, float *z, xs_state *st) {
  hipLaunchKernelGGL(k_bicg_p32, dim3(vgrid(n)), dim3(XS_BLOCK), 0, s, n, r, u, p, dinv, z, st);
}

__global__ __launch_bounds__(XS_BLOCK) void k_bicg_s32(int n, const float *__restrict__ r,
                                                       const float *__restrict__ u,
                                                       float *__restrict__ sv,
                                                       const float *__restrict__ dinv,
                                                       float *__restrict__ z1, xs_state *st) {
  if (st->done) return;
  const float gama = (float)st->sc[S_GAMA];
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    const float si = fmaf(-gama, u[i], r[i]);
    sv[i] = si;
    if (dinv) z1[i] = dinv[i] * si;
  }
}
void xs_bicg_s32(hipStream_t s, int n, const float *r, const float *u, float *sv,
                 const float *dinv, float *z1, xs_state *st) {
  hipLaunchKernelGGL(k_bicg_s32, dim3(vgrid(n)), dim3(XS_BLOCK), 0, s, n, r, u, sv, dinv, z1, st);
}

__global__ __launch_bounds__(XS_BLOCK) void k_bicg_xr32(int n, float *__restrict__ x,
                                                        float *__restrict__ r,
                                                        const float *__restrict__ z,
                                                        const float *__restrict__ z1,
                                                        const float *__restrict__ sv,
                                                        const float *__restrict__ v,
                                                        const float *__restrict__ rh, xs_reduce R) {
  xs_state *st = R.st;
  if (st->done) return;
  const float gama = (float)st->sc[S_GAMA], alpha = (float)st->sc[S_ALPHA];
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    x[i] = x[i] + gama * z[i] + alpha * z1[i];
    const float ri = sv[i] - alpha * v[i];
    r[i] = ri;
    acc[0] += (double)ri * (double)ri;
    acc[1] += (double)rh[i] * (double)ri;
  }
  reduce_finish<2>(acc, R);
}
void xs_bicg_xr32(hipStream_t s, int n, float *x, float *r, const float *z, const float *z1,
                  const float *sv, const float *v, const float *rh, const xs_reduce &red) {
  hipLaunchKernelGGL(k_bicg_xr32, dim3(vgrid(n)), dim3(XS_BLOCK), 0, s, n, x, r, z, z1, sv, v, rh, red);
}

__global__ void k_diag_f32(int n, const double *__restrict__ d, float *__restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = (float)d[i];
}
void xs_vec_d2f(hipStream_t s, int n, const double *d, float *f) {
  hipLaunchKernelGGL(k_diag_f32, dim3((n + 255) / 256), dim3(256), 0, s, n, d, f);
}
