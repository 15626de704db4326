// xs_kernels.hip -- gfx950 kernels of the preconditioned Krylov path.
//
// All kernels are memory-bound.  Design notes:
//  * 64-lane wavefronts, 256-thread workgroups; reductions use cross-lane
//    shuffles inside a wave, LDS across the 4 waves of a workgroup, and a
//    self-resetting "last workgroup" ticket across the grid, so every scalar of
//    the Krylov recurrences (alpha, beta, omega, residual norm, convergence)
//    is computed on the device and never crosses to the host per iteration.
//  * Partial sums are combined in block-index order: results are
//    deterministic for a fixed launch shape.
//  * Every kernel that belongs to a Krylov iteration checks st->done first, so
//    iterations launched ahead of the host's convergence check are no-ops.
#include "xs_internal.h"

#include <algorithm>
#include <atomic>
#include <climits>

// Launch guard: a launcher that finds a required pointer null (a set-up
// ordering bug) does not launch; it raises this flag, which the host driver
// turns into an error instead of letting a kernel touch address 0.
static std::atomic<int> g_launch_guard{0};
int xs_take_launch_guard() { return g_launch_guard.exchange(0); }
void xs_raise_launch_guard() { g_launch_guard.store(1); }

#include "xs_device.h"

// Per-row SpMV epilogue (see xs_spmv_epi).
template <int EPI, int NV>
__device__ __forceinline__ void epilogue(int r, double s, double *__restrict__ y,
                                         const double *__restrict__ b,
                                         const double *__restrict__ w, double (&acc)[NV]) {
  if (EPI == EPI_BMAX || EPI == EPI_BMAXN) {
    const double t = b[r] - s;
    y[r] = t;
    if (EPI == EPI_BMAXN) acc[0] += t * t;
  } else {
    y[r] = s;
    if (EPI == EPI_DOT1 || EPI == EPI_DOT2) acc[0] += w[r] * s;
    if (EPI == EPI_DOT2) acc[NV - 1] += s * s;
  }
}

// The same with the epilogue's operand (b[r] or w[r]) already loaded.
template <int EPI, int NV>
__device__ __forceinline__ void epilogue_v(int r, double s, double bw, double *__restrict__ y,
                                           double (&acc)[NV]) {
  if (EPI == EPI_BMAX || EPI == EPI_BMAXN) {
    const double t = bw - s;
    y[r] = t;
    if (EPI == EPI_BMAXN) acc[0] += t * t;
  } else {
    y[r] = s;
    if (EPI == EPI_DOT1 || EPI == EPI_DOT2) acc[0] += bw * s;
    if (EPI == EPI_DOT2) acc[NV - 1] += s * s;
  }
}

// ---------------------------------------------------------------------------
// CSR SpMV: LDS-staged row blocks ("stream" rows of one nnz chunk through LDS
// with fully coalesced loads of val/colidx, then reduce each row from LDS);
// rows too long for the LDS tile fall back to one wavefront per row.
// Reference kernels replaced: matMultCSR / bmAxCSR
// (hip/HostMatrix/DeviceMatrix/DeviceMatrixCSR.cpp:59-84, one thread per row).
// ---------------------------------------------------------------------------
template <int EPI>
__global__ __launch_bounds__(XS_BLOCK) void k_spmv_csr(
    const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v,
    const double *__restrict__ x, double *__restrict__ y, const int *__restrict__ blk_row,
    const double *__restrict__ b, const double *__restrict__ w, xs_reduce R) {
  __shared__ double prod[XS_SPMV_LDS];
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
        double s = 0.0;
        for (int k = a; k < bnd; ++k) s += prod[k];
        epilogue<EPI, NV>(r, s, y, b, w, acc);
      }
    } else {
      const int lane = threadIdx.x & (XS_WAVE - 1);
      for (int r = r0 + (int)(threadIdx.x / XS_WAVE); r < r1; r += XS_BLOCK / XS_WAVE) {
        const int a = rp[r], bnd = rp[r + 1];
        double s = 0.0;
        for (int k = a + lane; k < bnd; k += XS_WAVE) s += v[k] * x[ci[k]];
        s = wave_sum(s);
        if (lane == 0) epilogue<EPI, NV>(r, s, y, b, w, acc);
      }
    }
  }
  if (RED) reduce_finish<NV>(acc, R);
}

// Row-binned vector CSR: rows of length <= long_thr are reduced by groups of
// L lanes (L = the average row length rounded up to a power of two), each
// group reading its row's entries with consecutive lanes (coalesced across
// the wave) and combining them with xor shuffles inside the group; the
// longer rows (listed at set-up) go one wavefront each, in the first
// workgroups.  At most XS_MAX_REDUCE_BLOCKS workgroups (few partials for the
// fused dot epilogues), each owning one contiguous chunk of rows placed by
// xcd_chunk; a group takes XS_VEC_UR rows per step and issues the row
// pointers, then the first entry of every row, then the x gathers together,
// so the three dependent loads of a row overlap those of the other rows.
#define XS_VEC_UR 4
template <int L, int EPI>
__global__ __launch_bounds__(XS_BLOCK) void k_spmv_vec(
    int n, const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v,
    const double *__restrict__ x, double *__restrict__ y, const int *__restrict__ long_rows,
    int nlong, int long_thr, const double *__restrict__ b, const double *__restrict__ w,
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
      for (int e = rp[r] + lane; e < rp[r + 1]; e += XS_WAVE) s += v[e] * x[ci[e]];
      s = wave_sum(s);
      if (lane == 0) epilogue<EPI, NV>(r, s, y, b, w, acc);
    }
  } else {
    constexpr int GPB = XS_BLOCK / L;  // row groups per workgroup
    constexpr int UR = XS_VEC_UR;
    const int sub = threadIdx.x % L;
    const int G = (int)gridDim.x - nlb;
    const int rpc = ((n + G - 1) / G + GPB - 1) / GPB * GPB;  // rows per chunk
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
      double vv[UR], s[UR];
#pragma unroll
      for (int u = 0; u < UR; ++u) {
        const int e = a[u] + sub;
        const bool in = on[u] && e < e1[u];
        cc[u] = in ? ci[e] : -1;
        vv[u] = in ? v[e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < UR; ++u) s[u] = cc[u] >= 0 ? vv[u] * x[cc[u]] : 0.0;
#pragma unroll
      for (int u = 0; u < UR; ++u)
        if (on[u])
          for (int e = a[u] + sub + L; e < e1[u]; e += L) s[u] += v[e] * x[ci[e]];
#pragma unroll
      for (int u = 0; u < UR; ++u) {
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s[u] += __shfl_xor(s[u], off, L);
        if (sub == 0 && on[u]) epilogue<EPI, NV>(row0 + u * GPB, s[u], y, b, w, acc);
      }
    }
  }
  if (RED) reduce_finish<NV>(acc, R);
}

template <int L>
static void spmv_vec_launch(hipStream_t s, const xs_csr &A, const xs_spmv_plan &P,
                            const double *x, double *y, int epi, const double *b,
                            const double *w, const xs_reduce &R) {
  const int nlb = (P.nlong + XS_BLOCK / XS_WAVE - 1) / (XS_BLOCK / XS_WAVE);
  long g = ((long)A.n + XS_BLOCK / L - 1) / (XS_BLOCK / L);
  const long cap = std::max(1, XS_MAX_REDUCE_BLOCKS - nlb);
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  const dim3 grid((unsigned)(nlb + g)), blk(XS_BLOCK);
#define XS_VEC(E) hipLaunchKernelGGL((k_spmv_vec<L, E>), grid, blk, 0, s, A.n, A.rp, A.ci, A.v, x, y, P.long_rows, P.nlong, P.long_thr, b, w, R)
  switch (epi) {
    case EPI_Y: XS_VEC(EPI_Y); break;
    case EPI_BMAX: XS_VEC(EPI_BMAX); break;
    case EPI_DOT1: XS_VEC(EPI_DOT1); break;
    case EPI_DOT2: XS_VEC(EPI_DOT2); break;
    case EPI_BMAXN: XS_VEC(EPI_BMAXN); break;
  }
#undef XS_VEC
}

// ELL SpMV over column-major slabs v[k*ld + row]; two rows per thread with
// 16-byte loads, the slabs taken four at a time (all index and value loads
// of a batch issued, then all x gathers) and the 512-row blocks placed by
// xcd_chunk.  Padding uses the row's own index and a zero value.
// Reference: matMultELL / bmAxELL (DeviceMatrixELL.cpp:3-34), width fixed at
// 16 there; here the width is the true maximum row length.
#define XS_ELL_UK 4
template <int EPI, bool O16>
__global__ __launch_bounds__(XS_BLOCK) void k_spmv_ell(
    int n, int ld, int W, const int *__restrict__ ci, const short *__restrict__ co,
    const double *__restrict__ v, const double *__restrict__ x, double *__restrict__ y,
    const double *__restrict__ b, const double *__restrict__ w, xs_reduce R) {
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
    double e[2] = {0.0, 0.0};
    if (OPB || OPW) {
      const double *op = OPB ? b : w;
      e[0] = op[r];
      if (r + 1 < n) e[1] = op[r + 1];
    }
    double s0 = 0.0, s1 = 0.0;
    for (int k0 = 0; k0 < W; k0 += XS_ELL_UK) {
      // branch-free loads (a slot past W re-reads slab W - 1, masked from the
      // sums) so that the batch's loads are all issued before any wait
      int2 c[XS_ELL_UK];
      double2 a[XS_ELL_UK];
#pragma unroll
      for (int q = 0; q < XS_ELL_UK; ++q) {
        const size_t p = (size_t)min(k0 + q, W - 1) * ld + r;
        if (O16) {   // 16-bit offsets: 4 bytes of indices for the two rows
          const short2 o = *reinterpret_cast<const short2 *>(co + p);
          c[q] = make_int2(r + o.x, r + 1 + o.y);
        } else {
          c[q] = *reinterpret_cast<const int2 *>(ci + p);
        }
        a[q] = *reinterpret_cast<const double2 *>(v + p);
      }
      double x0[XS_ELL_UK], x1[XS_ELL_UK];
#pragma unroll
      for (int q = 0; q < XS_ELL_UK; ++q) { x0[q] = x[c[q].x]; x1[q] = x[c[q].y]; }
#pragma unroll
      for (int q = 0; q < XS_ELL_UK; ++q) {
        if (k0 + q < W) { s0 += a[q].x * x0[q]; s1 += a[q].y * x1[q]; }
      }
    }
    double s[2] = {s0, s1};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int rr = r + q;
      if (rr < n) epilogue_v<EPI, NV>(rr, s[q], e[q], y, acc);
    }
  }
  if (RED) reduce_finish<NV>(acc, R);
}

// SpMV over the diagonal-slot copy: k_spmv_ell without index loads.  Slot k
// of row r reads x at r + doff.o[k] (clamped into [0, n) for the zero-valued
// slots of rows that lack that diagonal), so the sums run over the same
// entries in the same column order as the ELL rows, with zero products in
// place of absent entries (as the ELL padding adds them at the row's end).
template <int EPI>
__global__ __launch_bounds__(XS_BLOCK) void k_spmv_dia(
    int n, int ld, int W, const xs_dia_offs doff, const double *__restrict__ v,
    const double *__restrict__ x, double *__restrict__ y, const double *__restrict__ b,
    const double *__restrict__ w, xs_reduce R) {
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
    double e[2] = {0.0, 0.0};
    if (OPB || OPW) {
      const double *op = OPB ? b : w;
      e[0] = op[r];
      if (r + 1 < n) e[1] = op[r + 1];
    }
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int k0 = 0; k0 < XS_DIA_MAX; k0 += XS_ELL_UK) {
      if (k0 >= W) break;
      double2 a[XS_ELL_UK];
      int c0[XS_ELL_UK], c1[XS_ELL_UK];
#pragma unroll
      for (int q = 0; q < XS_ELL_UK; ++q) {
        const int k = min(k0 + q, W - 1);
        a[q] = *reinterpret_cast<const double2 *>(v + (size_t)k * ld + r);
        c0[q] = min(max(r + doff.o[k], 0), n - 1);
        c1[q] = min(max(r + 1 + doff.o[k], 0), n - 1);
      }
      double x0[XS_ELL_UK], x1[XS_ELL_UK];
#pragma unroll
      for (int q = 0; q < XS_ELL_UK; ++q) { x0[q] = x[c0[q]]; x1[q] = x[c1[q]]; }
#pragma unroll
      for (int q = 0; q < XS_ELL_UK; ++q) {
        if (k0 + q < W) { s0 += a[q].x * x0[q]; s1 += a[q].y * x1[q]; }
      }
    }
    double s[2] = {s0, s1};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int rr = r + q;
      if (rr < n) epilogue_v<EPI, NV>(rr, s[q], e[q], y, acc);
    }
  }
  if (RED) reduce_finish<NV>(acc, R);
}

static xs_reduce no_reduce() {
  xs_reduce r{};
  return r;
}

void xs_launch_spmv_csr(hipStream_t s, const xs_csr &A, const xs_spmv_plan &P,
                        const double *x, double *y, int epi, const double *b,
                        const double *w, const xs_reduce *red) {
  XS_GUARD(A.rp && A.ci && A.v && P.blk_row && P.nblk > 0 && x && y);
  XS_GUARD(!(epi == EPI_BMAX || epi == EPI_BMAXN) || b);
  XS_GUARD(!(epi == EPI_DOT1 || epi == EPI_DOT2) || w);
  XS_GUARD(!(epi == EPI_DOT1 || epi == EPI_DOT2 || epi == EPI_BMAXN) || (red && red->partial && red->counter && red->st));
  xs_reduce R = red ? *red : no_reduce();
  if (P.vec) {
    XS_GUARD(P.nlong == 0 || P.long_rows);
    switch (P.vec_L) {
      case 2: spmv_vec_launch<2>(s, A, P, x, y, epi, b, w, R); break;
      case 4: spmv_vec_launch<4>(s, A, P, x, y, epi, b, w, R); break;
      case 8: spmv_vec_launch<8>(s, A, P, x, y, epi, b, w, R); break;
      case 16: spmv_vec_launch<16>(s, A, P, x, y, epi, b, w, R); break;
      case 32: spmv_vec_launch<32>(s, A, P, x, y, epi, b, w, R); break;
      default: spmv_vec_launch<64>(s, A, P, x, y, epi, b, w, R); break;
    }
    return;
  }
  dim3 grid(P.nblk), blk(XS_BLOCK);
  switch (epi) {
    case EPI_Y: hipLaunchKernelGGL(k_spmv_csr<EPI_Y>, grid, blk, 0, s, A.rp, A.ci, A.v, x, y, P.blk_row, b, w, R); break;
    case EPI_BMAX: hipLaunchKernelGGL(k_spmv_csr<EPI_BMAX>, grid, blk, 0, s, A.rp, A.ci, A.v, x, y, P.blk_row, b, w, R); break;
    case EPI_DOT1: hipLaunchKernelGGL(k_spmv_csr<EPI_DOT1>, grid, blk, 0, s, A.rp, A.ci, A.v, x, y, P.blk_row, b, w, R); break;
    case EPI_DOT2: hipLaunchKernelGGL(k_spmv_csr<EPI_DOT2>, grid, blk, 0, s, A.rp, A.ci, A.v, x, y, P.blk_row, b, w, R); break;
    case EPI_BMAXN: hipLaunchKernelGGL(k_spmv_csr<EPI_BMAXN>, grid, blk, 0, s, A.rp, A.ci, A.v, x, y, P.blk_row, b, w, R); break;
  }
}

void xs_launch_spmv_ell(hipStream_t s, const xs_ell &E, const double *x, double *y,
                        int epi, const double *b, const double *w, const xs_reduce *red) {
  XS_GUARD(E.ci && E.v && E.ld % 2 == 0 && E.ld >= E.n && x && y);
  XS_GUARD(!(epi == EPI_BMAX || epi == EPI_BMAXN) || b);
  XS_GUARD(!(epi == EPI_DOT1 || epi == EPI_DOT2) || w);
  XS_GUARD(!(epi == EPI_DOT1 || epi == EPI_DOT2 || epi == EPI_BMAXN) || (red && red->partial && red->counter && red->st));
  xs_reduce R = red ? *red : no_reduce();
  int threads = (E.n + 1) / 2;
  dim3 grid((threads + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
  if (E.dv && E.nd > 0 && E.nd <= XS_DIA_MAX) {
#define XS_DIAK(E_) hipLaunchKernelGGL((k_spmv_dia<E_>), grid, blk, 0, s, E.n, E.ld, E.nd, E.doff, E.dv, x, y, b, w, R)
    switch (epi) {
      case EPI_Y: XS_DIAK(EPI_Y); break;
      case EPI_BMAX: XS_DIAK(EPI_BMAX); break;
      case EPI_DOT1: XS_DIAK(EPI_DOT1); break;
      case EPI_DOT2: XS_DIAK(EPI_DOT2); break;
      case EPI_BMAXN: XS_DIAK(EPI_BMAXN); break;
    }
#undef XS_DIAK
    return;
  }
#define XS_ELLK(E_, O) hipLaunchKernelGGL((k_spmv_ell<E_, O>), grid, blk, 0, s, E.n, E.ld, E.W, E.ci, E.co, E.v, x, y, b, w, R)
#define XS_ELLE(E_) do { if (E.co) XS_ELLK(E_, true); else XS_ELLK(E_, false); } while (0)
  switch (epi) {
    case EPI_Y: XS_ELLE(EPI_Y); break;
    case EPI_BMAX: XS_ELLE(EPI_BMAX); break;
    case EPI_DOT1: XS_ELLE(EPI_DOT1); break;
    case EPI_DOT2: XS_ELLE(EPI_DOT2); break;
    case EPI_BMAXN: XS_ELLE(EPI_BMAXN); break;
  }
#undef XS_ELLE
#undef XS_ELLK
}

// --- analysis kernels -------------------------------------------------------
// blk_row[k] = first row whose first nonzero lies at or after k*chunk.
__global__ void k_build_plan(int n, const int *__restrict__ rp, int chunk, int nblk,
                             int *__restrict__ blk_row) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nblk) return;
  if (k == nblk) { blk_row[k] = n; return; }
  const long long target = (long long)k * chunk;
  int lo = 0, hi = n;  // lower_bound over rp[0..n)
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (rp[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  blk_row[k] = lo;
}

// Default nnz budget per block: about 256 rows at the average row length, a
// multiple of 256, between 512 and 2048 so a straddling row still fits.
int xs_default_chunk(const xs_csr &A) {
  double avg = A.n ? (double)A.nnz / A.n : 1.0;
  int chunk = (int)(avg * XS_BLOCK);
  chunk = ((chunk + XS_BLOCK - 1) / XS_BLOCK) * XS_BLOCK;
  if (chunk < 512) chunk = 512;
  if (chunk > 2048) chunk = 2048;
  return chunk;
}

// blk must hold (nnz + chunk - 1) / chunk + 1 entries.
void xs_build_plan(hipStream_t s, const xs_csr &A, xs_spmv_plan &P, int *blk, int chunk) {
  XS_GUARD(A.rp && blk && chunk > 0);
  P.chunk = chunk;
  P.nblk = (A.nnz + chunk - 1) / chunk;
  if (P.nblk < 1) P.nblk = 1;
  P.blk_row = blk;
  int threads = P.nblk + 1;
  hipLaunchKernelGGL(k_build_plan, dim3((threads + 255) / 256), dim3(256), 0, s, A.n, A.rp,
                     chunk, P.nblk, P.blk_row);
}

// Grid-stride maxima with one atomic per workgroup (same-address atomics
// serialise, so a per-wave atomic over millions of rows costs ~0.5 ms).
__device__ __forceinline__ int block_max(int m) {
  __shared__ int wm[XS_BLOCK / XS_WAVE];
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, XS_WAVE));
  __syncthreads();
  if ((threadIdx.x & (XS_WAVE - 1)) == 0) wm[threadIdx.x / XS_WAVE] = m;
  __syncthreads();
  m = wm[0];
  for (int k = 1; k < XS_BLOCK / XS_WAVE; ++k) m = max(m, wm[k]);
  return m;
}
#define XS_MAX_GRID 512

__global__ __launch_bounds__(XS_BLOCK) void k_max_row_len(int n, const int *__restrict__ rp, int *out) {
  int m = 0;
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK)
    m = max(m, rp[i + 1] - rp[i]);
  m = block_max(m);
  if (threadIdx.x == 0) atomicMax(out, m);
}

int xs_max_row_len(hipStream_t s, const xs_csr &A, int *d_tmp) {
  if (!A.rp || !d_tmp) { xs_raise_launch_guard(); return 0; }
  int h = 0;
  (void)hipMemsetAsync(d_tmp, 0, sizeof(int), s);
  hipLaunchKernelGGL(k_max_row_len, dim3(std::max(1, std::min((A.n + XS_BLOCK - 1) / XS_BLOCK, XS_MAX_GRID))),
                     dim3(XS_BLOCK), 0, s, A.n, A.rp, d_tmp);
  (void)hipMemcpyAsync(&h, d_tmp, sizeof(int), hipMemcpyDeviceToHost, s);
  (void)hipStreamSynchronize(s);
  return h;
}

__global__ void k_build_ell(int n, int ld, int W, const int *__restrict__ rp,
                            const int *__restrict__ ci, const double *__restrict__ v,
                            int *__restrict__ eci, double *__restrict__ ev) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ld) return;
  int a = 0, len = 0;
  if (r < n) { a = rp[r]; len = rp[r + 1] - a; }
  const int own = r < n ? r : 0;
  for (int k = 0; k < W; ++k) {
    const size_t p = (size_t)k * ld + r;
    if (k < len) { eci[p] = ci[a + k]; ev[p] = v[a + k]; }
    else { eci[p] = own; ev[p] = 0.0; }
  }
}

__global__ __launch_bounds__(XS_BLOCK) void k_ell_max_offset(int n, int ld, int W,
                                                             const int *__restrict__ eci,
                                                             int *out) {
  int m = 0;
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK)
    for (int k = 0; k < W; ++k) {
      const int c = eci[(size_t)k * ld + i];
      if (c >= 0) m = max(m, abs(c - i));
    }
  m = block_max(m);
  if (threadIdx.x == 0) atomicMax(out, m);
}

int xs_ell_max_offset(hipStream_t s, int n, int ld, int W, const int *eci, int *d_tmp) {
  if (!eci || !d_tmp) { xs_raise_launch_guard(); return -1; }
  int h = 0;
  if (hipMemsetAsync(d_tmp, 0, sizeof(int), s) != hipSuccess) return -1;
  hipLaunchKernelGGL(k_ell_max_offset, dim3(std::max(1, std::min((n + XS_BLOCK - 1) / XS_BLOCK, XS_MAX_GRID))),
                     dim3(XS_BLOCK), 0, s, n, ld, W, eci, d_tmp);
  if (hipMemcpyAsync(&h, d_tmp, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -1;
  return h;
}

__global__ void k_ell_build_offsets(int n, int ld, int W, const int *__restrict__ eci,
                                    short *__restrict__ eco) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ld) return;
  for (int k = 0; k < W; ++k) {
    const size_t p = (size_t)k * ld + i;
    const int c = eci[p];
    eco[p] = (short)(i >= n ? (n - 1) - i : c < 0 ? XS_OFF16_PAD : c - i);
  }
}

void xs_ell_build_offsets(hipStream_t s, int n, int ld, int W, const int *eci, short *eco) {
  XS_GUARD(eci && eco && n > 0 && ld - n <= XS_OFF16_MAX);
  hipLaunchKernelGGL(k_ell_build_offsets, dim3((ld + 255) / 256), dim3(256), 0, s, n, ld, W, eci, eco);
}

void xs_build_ell(hipStream_t s, const xs_csr &A, int W, xs_ell &E) {
  XS_GUARD(A.rp && A.ci && A.v && E.ci && E.v);
  hipLaunchKernelGGL(k_build_ell, dim3((E.ld + 255) / 256), dim3(256), 0, s, A.n, E.ld, W, A.rp,
                     A.ci, A.v, E.ci, E.v);
  if (E.dv) xs_build_dia(s, A, E);
}

// Distinct offsets column - row, gathered into t[0 .. XS_DIA_MAX) (INT_MIN:
// free slot); t[XS_DIA_MAX] = 1 when there are more, or a row repeats a
// column.  A slot only ever changes from INT_MIN to its offset, so a plain
// read that finds the offset is final and a CAS claims a free slot.  Each
// workgroup collects into an LDS table first and merges its (at most
// XS_DIA_MAX) offsets into t at the end: the global table sees a handful of
// atomics per workgroup instead of a read per entry.
__device__ __forceinline__ bool dia_insert(int *tab, int o) {
  for (int k = 0; k < XS_DIA_MAX; ++k) {
    int cur = tab[k];
    if (cur == INT_MIN) cur = atomicCAS(tab + k, INT_MIN, o);
    if (cur == INT_MIN || cur == o) return true;
  }
  return false;
}

// part: 0 all entries, 1 strictly lower, 2 strictly upper (dpos: pivots)
__global__ __launch_bounds__(256) void k_dia_offsets(int n, const int *__restrict__ rp,
                                                     const int *__restrict__ ci,
                                                     const int *__restrict__ dpos, int part,
                                                     int *t) {
  __shared__ int lt[XS_DIA_MAX + 1];
  __shared__ int done;
  if (threadIdx.x <= XS_DIA_MAX) lt[threadIdx.x] = threadIdx.x < XS_DIA_MAX ? INT_MIN : 0;
  // a workgroup that starts after another has found more offsets than slots
  // has nothing to add (the flag only goes from 0 to 1; a stale 0 costs time)
  if (threadIdx.x == 0) done = __hip_atomic_load(t + XS_DIA_MAX, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (done) return;   // the same value for the whole workgroup
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    if (lt[XS_DIA_MAX]) break;   // this workgroup has its answer: no diagonal slots
    const int e0 = part == 2 ? dpos[r] + 1 : rp[r], e1 = part == 1 ? dpos[r] : rp[r + 1];
    for (int e = e0; e < e1; ++e) {
      const int o = ci[e] - r;
      if ((e > e0 && ci[e] == ci[e - 1]) || !dia_insert(lt, o)) {
        lt[XS_DIA_MAX] = 1;
        break;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < XS_DIA_MAX && lt[threadIdx.x] != INT_MIN) {
    if (!dia_insert(t, lt[threadIdx.x])) atomicOr(t + XS_DIA_MAX, 1);
  } else if (threadIdx.x == XS_DIA_MAX && lt[XS_DIA_MAX]) {
    atomicOr(t + XS_DIA_MAX, 1);
  }
}

static int dia_offsets(hipStream_t s, int n, const int *rp, const int *ci, const int *dpos,
                       int part, int *d_tmp, xs_dia_offs &out) {
  if (!(rp && ci && d_tmp && (part == 0 || dpos)) || n <= 0) return -1;
  int h[XS_DIA_MAX + 1];
  for (int k = 0; k < XS_DIA_MAX; ++k) h[k] = INT_MIN;
  h[XS_DIA_MAX] = 0;
  if (hipMemcpyAsync(d_tmp, h, sizeof(h), hipMemcpyHostToDevice, s) != hipSuccess) return -1;
  // the first rows first: a wide stencil (27 points) already shows more
  // offsets than slots there, and the pass over every row is skipped
  const int ns = std::min(n, 4096);
  hipLaunchKernelGGL(k_dia_offsets, dim3((ns + 255) / 256), dim3(256), 0, s, ns, rp, ci, dpos,
                     part, d_tmp);
  if (hipMemcpyAsync(h, d_tmp, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -1;
  if (h[XS_DIA_MAX]) return 0;
  if (ns < n) {
    hipLaunchKernelGGL(k_dia_offsets, dim3(std::min(4096, (n + 255) / 256)), dim3(256), 0, s, n,
                       rp, ci, dpos, part, d_tmp);
    if (hipMemcpyAsync(h, d_tmp, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return -1;
    if (h[XS_DIA_MAX]) return 0;
  }
  int nd = 0;
  for (int k = 0; k < XS_DIA_MAX; ++k)
    if (h[k] != INT_MIN) out.o[nd++] = h[k];
  std::sort(out.o, out.o + nd);
  for (int k = nd; k < XS_DIA_MAX; ++k) out.o[k] = 0;
  return nd;
}

int xs_dia_offsets(hipStream_t s, const xs_csr &A, int *d_tmp, xs_dia_offs &out) {
  return dia_offsets(s, A.n, A.rp, A.ci, nullptr, 0, d_tmp, out);
}

int xs_tri_dia_offsets(hipStream_t s, const xs_ilu &F, bool upper, int *d_tmp, xs_dia_offs &out) {
  return dia_offsets(s, F.n, F.rp, F.ci, F.dpos, upper ? 2 : 1, d_tmp, out);
}

// Diagonal slots of a factor part from its ELL copy (slab-major, read
// coalesced): slab j of row r holds column eci[j ld + r] (-1: padding) with
// value ev (and, for U, u_ij / u_jj in evs, the values of the fused first
// pass); each goes to the slot of its offset, the others stay 0.
template <bool UPPER>
__global__ void k_tell_to_dia(int n, int ld, int W, int nd, const xs_dia_offs doff,
                              const int *__restrict__ eci, const double *__restrict__ ev,
                              const double *__restrict__ evs, double *__restrict__ dv,
                              double *__restrict__ dvs) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ld) return;
  double slot[XS_DIA_MAX], sc[XS_DIA_MAX];
#pragma unroll
  for (int k = 0; k < XS_DIA_MAX; ++k) { slot[k] = 0.0; sc[k] = 0.0; }
  if (r < n)
    for (int j = 0; j < W; ++j) {
      const size_t p = (size_t)j * ld + r;
      const int c = eci[p];
      if (c < 0) continue;
      const int o = c - r;
      const double v = ev[p];
      const double vs = UPPER ? evs[p] : 0.0;
#pragma unroll
      for (int k = 0; k < XS_DIA_MAX; ++k)
        if (k < nd && doff.o[k] == o) {
          slot[k] = v;
          if (UPPER) sc[k] = vs;
        }
    }
#pragma unroll
  for (int k = 0; k < XS_DIA_MAX; ++k)
    if (k < nd) {
      dv[(size_t)k * ld + r] = slot[k];
      if (UPPER) dvs[(size_t)k * ld + r] = sc[k];
    }
}

// from the ELL copy, which must be current (xs_build_tell, and for U
// xs_tell_scale_cols)
void xs_build_tri_dia(hipStream_t s, const xs_ilu &F, bool upper, xs_tell &E) {
  XS_GUARD(E.ci && E.v && E.dv && E.nd > 0 && E.nd <= XS_DIA_MAX && E.ld >= F.n &&
           (!upper || (E.dvs && E.vs)));
  const dim3 grid((E.ld + 255) / 256), blk(256);
  if (upper)
    hipLaunchKernelGGL(k_tell_to_dia<true>, grid, blk, 0, s, F.n, E.ld, E.W, E.nd, E.doff, E.ci,
                       E.v, E.vs, E.dv, E.dvs);
  else
    hipLaunchKernelGGL(k_tell_to_dia<false>, grid, blk, 0, s, F.n, E.ld, E.W, E.nd, E.doff, E.ci,
                       E.v, nullptr, E.dv, nullptr);
}

__global__ void k_build_dia(int n, int ld, int nd, const xs_dia_offs doff,
                            const int *__restrict__ rp, const int *__restrict__ ci,
                            const double *__restrict__ v, double *__restrict__ dv) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ld) return;
  double slot[XS_DIA_MAX];
#pragma unroll
  for (int k = 0; k < XS_DIA_MAX; ++k) slot[k] = 0.0;
  if (r < n)
    for (int e = rp[r]; e < rp[r + 1]; ++e) {
      const int o = ci[e] - r;
#pragma unroll
      for (int k = 0; k < XS_DIA_MAX; ++k)
        if (k < nd && doff.o[k] == o) slot[k] = v[e];
    }
#pragma unroll
  for (int k = 0; k < XS_DIA_MAX; ++k)
    if (k < nd) dv[(size_t)k * ld + r] = slot[k];
}

void xs_build_dia(hipStream_t s, const xs_csr &A, xs_ell &E) {
  XS_GUARD(A.rp && A.ci && A.v && E.dv && E.nd > 0 && E.nd <= XS_DIA_MAX && E.ld >= A.n);
  hipLaunchKernelGGL(k_build_dia, dim3((E.ld + 255) / 256), dim3(256), 0, s, A.n, E.ld, E.nd,
                     E.doff, A.rp, A.ci, A.v, E.dv);
}

// Transposed SpMV y = A^T x by a gather over an explicitly transposed CSR is
// done by the host driver (it builds A^T once); this scatter form is kept for
// small setup-time uses only.
__global__ void k_spmv_t_scatter(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                 const double *__restrict__ v, const double *__restrict__ x,
                                 double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = x[i];
  for (int e = rp[i]; e < rp[i + 1]; ++e) atomicAdd(&y[ci[e]], v[e] * xi);
}

void xs_launch_spmv_t(hipStream_t s, const xs_csr &A, const double *x, double *y) {
  XS_GUARD(A.rp && A.ci && A.v && x && y);
  (void)hipMemsetAsync(y, 0, sizeof(double) * A.n, s);
  hipLaunchKernelGGL(k_spmv_t_scatter, dim3((A.n + 255) / 256), dim3(256), 0, s, A.n, A.rp, A.ci,
                     A.v, x, y);
}

// ---------------------------------------------------------------------------
// Vector kernels
// ---------------------------------------------------------------------------
static inline int vec_grid(int n) {
  int g = (n + XS_BLOCK * 4 - 1) / (XS_BLOCK * 4);
  if (g > XS_MAX_REDUCE_BLOCKS) g = XS_MAX_REDUCE_BLOCKS;
  if (g < 1) g = 1;
  return g;
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long val) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = val;
}

void xs_fill_sentinel(hipStream_t s, double *x, int n) {
  XS_GUARD(x);
  hipLaunchKernelGGL(k_fill_u64, dim3(vec_grid(n)), dim3(XS_BLOCK), 0, s,
                     reinterpret_cast<unsigned long long *>(x), n, XS_SENTINEL);
}

__global__ void k_init_state(xs_state *st, double tol, double abstol, int maxiter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int k = 0; k < XS_NSCAL; ++k) st->sc[k] = 0.0;
    st->sc[S_TOL] = tol;
    st->sc[S_ABSTOL] = abstol;
    st->done = 0;
    st->iter = 0;
    st->maxiter = maxiter;
  }
}

void xs_init_state(hipStream_t s, xs_state *st, double tol, double abstol, int maxiter) {
  XS_GUARD(st);
  hipLaunchKernelGGL(k_init_state, dim3(1), dim3(64), 0, s, st, tol, abstol, maxiter);
}

// Generic dot kernel: tot0 = a.b, tot1 = c.d (c may be null)
template <int NV>
__global__ __launch_bounds__(XS_BLOCK) void k_dot(int n, const double *__restrict__ a,
                                                  const double *__restrict__ b,
                                                  const double *__restrict__ c,
                                                  const double *__restrict__ d, xs_reduce R) {
  if (R.st != nullptr && R.st->done) return;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    acc[0] += a[i] * b[i];
    if (NV > 1) acc[NV - 1] += c[i] * d[i];
  }
  reduce_finish<NV>(acc, R);
}

void xs_launch_dot(hipStream_t s, int n, const double *a, const double *b, const double *c,
                   const double *d, const xs_reduce &red) {
  XS_GUARD(a && b && red.partial && red.counter && red.st && (!c || d));
  if (c) hipLaunchKernelGGL(k_dot<2>, dim3(vec_grid(n)), dim3(XS_BLOCK), 0, s, n, a, b, c, d, red);
  else hipLaunchKernelGGL(k_dot<1>, dim3(vec_grid(n)), dim3(XS_BLOCK), 0, s, n, a, b, c, d, red);
}

// BiCGSTAB p = r + omega (p - alpha u) [+ z = dinv o p]
// (DeviceVector.cpp:55-60 kernel1HIP; jacobiHIP :17-23)
// BiCGSTAB's direction update for one entry, p = r + omega (p - alpha u)
// (the reference's names for beta and omega), written with explicit fused
// multiply-adds so that every kernel forming it rounds the same way.
__device__ __forceinline__ double bicg_p_val(double r, double p, double u, double omega,
                                             double alpha) {
  return fma(omega, fma(-alpha, u, p), r);
}

__global__ __launch_bounds__(XS_BLOCK) void k_bicg_p(int n, const double *__restrict__ r,
                                                     const double *__restrict__ u,
                                                     double *__restrict__ p,
                                                     const double *__restrict__ dinv,
                                                     double *__restrict__ z, xs_state *st) {
  if (st->done) return;
  const double omega = st->sc[S_OMEGA], alpha = st->sc[S_ALPHA];
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    const double pi = bicg_p_val(r[i], p[i], u[i], omega, alpha);
    p[i] = pi;
    if (dinv) z[i] = dinv[i] * pi;
  }
}

void xs_bicg_p(hipStream_t s, int n, const double *r, const double *u, double *p,
               const double *dinv, double *z, xs_state *st) {
  XS_GUARD(r && u && p && st && (!dinv || z));
  hipLaunchKernelGGL(k_bicg_p, dim3(vec_grid(n)), dim3(XS_BLOCK), 0, s, n, r, u, p, dinv, z, st);
}

// s = r - gama u [+ z1 = dinv o s]   (kernel2HIP, DeviceVector.cpp:61-66)
__global__ __launch_bounds__(XS_BLOCK) void k_bicg_s(int n, const double *__restrict__ r,
                                                     const double *__restrict__ u,
                                                     double *__restrict__ sv,
                                                     const double *__restrict__ dinv,
                                                     double *__restrict__ z1, xs_state *st) {
  if (st->done) return;
  const double gama = st->sc[S_GAMA];
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    const double si = fma(-gama, u[i], r[i]);
    sv[i] = si;
    if (dinv) z1[i] = dinv[i] * si;
  }
}

void xs_bicg_s(hipStream_t s, int n, const double *r, const double *u, double *sv,
               const double *dinv, double *z1, xs_state *st) {
  XS_GUARD(r && u && sv && st && (!dinv || z1));
  hipLaunchKernelGGL(k_bicg_s, dim3(vec_grid(n)), dim3(XS_BLOCK), 0, s, n, r, u, sv, dinv, z1, st);
}

// x += gama z + alpha z1; r = s - alpha v; dots r.r and rhat.r
// (kernel3HIP DeviceVector.cpp:67-72, xsolver_bicgstab.cpp:96-103 and :50)
__global__ __launch_bounds__(XS_BLOCK) void k_bicg_xr(int n, double *__restrict__ x,
                                                      double *__restrict__ r,
                                                      const double *__restrict__ z,
                                                      const double *__restrict__ z1,
                                                      const double *__restrict__ sv,
                                                      const double *__restrict__ v,
                                                      const double *__restrict__ rh, xs_reduce R) {
  xs_state *st = R.st;
  if (st->done) return;
  const double gama = st->sc[S_GAMA], alpha = st->sc[S_ALPHA];
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    x[i] = x[i] + gama * z[i] + alpha * z1[i];
    const double ri = sv[i] - alpha * v[i];
    r[i] = ri;
    acc[0] += ri * ri;
    acc[1] += rh[i] * ri;
  }
  reduce_finish<2>(acc, R);
}

void xs_bicg_xr(hipStream_t s, int n, double *x, double *r, const double *z, const double *z1,
                const double *sv, const double *v, const double *rh, const xs_reduce &red) {
  XS_GUARD(x && r && z && z1 && sv && v && rh && red.partial && red.counter && red.st);
  hipLaunchKernelGGL(k_bicg_xr, dim3(vec_grid(n)), dim3(XS_BLOCK), 0, s, n, x, r, z, z1, sv, v, rh, red);
}

// CG: x += alpha p; r -= alpha Ap; [z = dinv o r; dots z.r, r.r] or [dot r.r]
// (hip/Solver/xsolver_cg.cpp:49-58)
template <bool JAC>
__global__ __launch_bounds__(XS_BLOCK) void k_cg_xr(int n, double *__restrict__ x,
                                                    double *__restrict__ r,
                                                    const double *__restrict__ p,
                                                    const double *__restrict__ Ap,
                                                    const double *__restrict__ dinv,
                                                    double *__restrict__ z, xs_reduce R) {
  xs_state *st = R.st;
  if (st->done) return;
  const double alpha = st->sc[S_CGALPHA];
  constexpr int NV = JAC ? 2 : 1;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    if (JAC) {
      const double zi = dinv[i] * ri;
      z[i] = zi;
      acc[0] += zi * ri;
    }
    acc[NV - 1] += ri * ri;
  }
  reduce_finish<NV>(acc, R);
}

void xs_cg_xr(hipStream_t s, int n, double *x, double *r, const double *p, const double *Ap,
              const double *dinv, double *z, const xs_reduce &red) {
  XS_GUARD(x && r && p && Ap && (!dinv || z) && red.partial && red.counter && red.st);
  if (dinv) hipLaunchKernelGGL(k_cg_xr<true>, dim3(vec_grid(n)), dim3(XS_BLOCK), 0, s, n, x, r, p, Ap, dinv, z, red);
  else hipLaunchKernelGGL(k_cg_xr<false>, dim3(vec_grid(n)), dim3(XS_BLOCK), 0, s, n, x, r, p, Ap, dinv, z, red);
}

void xs_cg_zr_dot(hipStream_t s, int n, const double *z, const double *r, const xs_reduce &red) {
  XS_GUARD(z && r && red.partial && red.counter && red.st);
  hipLaunchKernelGGL(k_dot<1>, dim3(vec_grid(n)), dim3(XS_BLOCK), 0, s, n, z, r,
                     (const double *)nullptr, (const double *)nullptr, red);
}

// p = z + beta p  (xsolver_cg.cpp:54-55)
__global__ __launch_bounds__(XS_BLOCK) void k_cg_p(int n, double *__restrict__ p,
                                                   const double *__restrict__ z, xs_state *st) {
  if (st->done) return;
  const double beta = st->sc[S_CGBETA];
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK)
    p[i] = beta * p[i] + z[i];
}

void xs_cg_p(hipStream_t s, int n, double *p, const double *z, xs_state *st) {
  XS_GUARD(p && z && st);
  hipLaunchKernelGGL(k_cg_p, dim3(vec_grid(n)), dim3(XS_BLOCK), 0, s, n, p, z, st);
}

__global__ void k_scale(int n, const double *__restrict__ d, const double *__restrict__ x,
                        double *__restrict__ y) {
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) y[i] = d[i] * x[i];
}
void xs_scale(hipStream_t s, int n, const double *d, const double *x, double *y) {
  XS_GUARD(d && x && y);
  hipLaunchKernelGGL(k_scale, dim3(vec_grid(n)), dim3(XS_BLOCK), 0, s, n, d, x, y);
}
__global__ void k_copy(int n, const double *__restrict__ x, double *__restrict__ y) {
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) y[i] = x[i];
}
void xs_copy(hipStream_t s, int n, const double *x, double *y) {
  XS_GUARD(x && y);
  hipLaunchKernelGGL(k_copy, dim3(vec_grid(n)), dim3(XS_BLOCK), 0, s, n, x, y);
}

// ---------------------------------------------------------------------------
// Jacobi: dinv_i = 1/a_ii (small = 0, hip/Precond/precon_jacobi.h:18 and
// DeviceVector.cpp:2-15).  A missing/zero pivot is reported through *bad
// (smallest offending row + 1), where the reference would exit
// (hip/HostVector/HostVector.cpp:88-92).
// ---------------------------------------------------------------------------
__global__ void k_jacobi_dinv(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                              const double *__restrict__ v, double *__restrict__ dinv, int *bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int e = rp[i]; e < rp[i + 1]; ++e)
    if (ci[e] == i) d = v[e];
  if (d == 0.0) atomicMin(bad, i + 1);
  dinv[i] = 1.0 / d;
}

void xs_jacobi_dinv(hipStream_t s, const xs_csr &A, double *dinv, int *d_bad) {
  XS_GUARD(A.rp && A.ci && A.v && dinv && d_bad);
  hipLaunchKernelGGL(k_jacobi_dinv, dim3((A.n + 255) / 256), dim3(256), 0, s, A.n, A.rp, A.ci, A.v,
                     dinv, d_bad);
}

// ---------------------------------------------------------------------------
// ParILU(0): Chow--Patel fixed-point sweeps over pattern(A) with the factors
// stored in place of A's values (lower part = L without its unit diagonal,
// diagonal and upper part = U).  One thread per nonzero, as the reference's
// ilu0_ij_aw (hip/Precond/gpuparilu.cpp:46-102):
//   s = a_ij - sum_{k < min(i,j)} l_ik u_kj   (merge of L row i and U column j)
//   i > j : l_ij = s / u_jj  (|u_jj| < 1e-10 -> +-1e-10)
//   i <= j: u_ij = s         (|u_ii| < 1e-10 -> +-1e-10)
//   non-finite results keep the previous value.
// async = true updates in place (the reference's behaviour); false reads
// lu_in and writes lu_out (deterministic Jacobi sweep).  With sc, the sweep
// factors the symmetrically scaled matrix S A S (s = |a_ii|^{-1/2}).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(XS_BLOCK) void k_parilu_sweep(
    int nnz, const int *__restrict__ rowidx, const int *__restrict__ rp, const int *__restrict__ ci,
    const double *__restrict__ a, const int *__restrict__ dpos, const int *__restrict__ ucp,
    const int *__restrict__ uri, const int *__restrict__ upos, const double *lu_in,
    double *lu_out, const double *__restrict__ sc) {
  const int k = blockIdx.x * XS_BLOCK + threadIdx.x;
  if (k >= nnz) return;
  const int i = rowidx[k], j = ci[k];
  const int m = i < j ? i : j;
  double s = sc ? a[k] * sc[i] * sc[j] : a[k];
  int il = rp[i];
  const int le = dpos[i];
  int iu = ucp[j];
  const int ue = ucp[j + 1];
  while (il < le && iu < ue) {
    const int kl = ci[il], ku = uri[iu];
    if (kl >= m || ku >= m) break;
    if (kl == ku) {
      s -= lu_in[il] * lu_in[upos[iu]];
      ++il; ++iu;
    } else if (kl < ku) {
      ++il;
    } else {
      ++iu;
    }
  }
  const double old = lu_in[k];
  double nv;
  if (i > j) {
    double d = lu_in[dpos[j]];
    if (fabs(d) < 1e-10) d = d >= 0.0 ? 1e-10 : -1e-10;
    nv = s / d;
  } else {
    if (i == j && fabs(s) < 1e-10) s = s >= 0.0 ? 1e-10 : -1e-10;
    nv = s;
  }
  lu_out[k] = isfinite(nv) ? nv : old;
}

void xs_parilu_sweep(hipStream_t s, const xs_csr &A, xs_ilu &F, double *lu_out, bool async,
                     const double *sc) {
  XS_GUARD(A.rp && A.ci && A.v && F.rowidx && F.dpos && F.ucp && F.uri && F.upos && F.lu);
  XS_GUARD(async || lu_out);
  double *out = async ? F.lu : lu_out;
  hipLaunchKernelGGL(k_parilu_sweep, dim3((A.nnz + XS_BLOCK - 1) / XS_BLOCK), dim3(XS_BLOCK), 0, s,
                     A.nnz, F.rowidx, A.rp, A.ci, A.v, F.dpos, F.ucp, F.uri, F.upos, F.lu, out, sc);
}

// Update lists of the ParILU(0) sweep.  The products entry (i, j) subtracts
// depend only on the pattern, so the merge of L's row i with U's column j is
// done once per pattern: one pass counts the matches (and records where u_jj
// sits), a scan turns the counts into offsets, a second pass writes the
// matched factor positions in the merge order.
template <bool FILL>
__global__ __launch_bounds__(XS_BLOCK) void k_parilu_pairs(
    int nnz, const int *__restrict__ rowidx, const int *__restrict__ rp,
    const int *__restrict__ ci, const int *__restrict__ dpos, const int *__restrict__ ucp,
    const int *__restrict__ uri, const int *__restrict__ upos, int *__restrict__ cnt,
    const int *__restrict__ pp, int *__restrict__ pl, int *__restrict__ pu, int *__restrict__ pd) {
  const int k = blockIdx.x * XS_BLOCK + threadIdx.x;
  if (k >= nnz) return;
  const int i = rowidx[k], j = ci[k];
  const int m = i < j ? i : j;
  int il = rp[i], iu = ucp[j], c = 0;
  const int le = dpos[i], ue = ucp[j + 1];
  const int base = FILL ? pp[k] : 0;
  while (il < le && iu < ue) {
    const int kl = ci[il], ku = uri[iu];
    if (kl >= m || ku >= m) break;
    if (kl == ku) {
      if (FILL) { pl[base + c] = il; pu[base + c] = upos[iu]; }
      ++c; ++il; ++iu;
    } else if (kl < ku) {
      ++il;
    } else {
      ++iu;
    }
  }
  if (!FILL) {
    cnt[k] = c;
    pd[k] = i > j ? dpos[j] : (i == j ? -2 : -1);
  }
}

void xs_parilu_pairs(hipStream_t s, const xs_csr &A, const xs_ilu &F, int *cnt, const int *pp,
                     int *pl, int *pu, int *pd) {
  XS_GUARD(A.rp && A.ci && F.rowidx && F.dpos && F.ucp && F.uri && F.upos);
  const dim3 grid((A.nnz + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
  if (!pl) {
    XS_GUARD(cnt && pd);
    hipLaunchKernelGGL(k_parilu_pairs<false>, grid, blk, 0, s, A.nnz, F.rowidx, A.rp, A.ci,
                       F.dpos, F.ucp, F.uri, F.upos, cnt, nullptr, nullptr, nullptr, pd);
  } else {
    XS_GUARD(pp && pu);
    hipLaunchKernelGGL(k_parilu_pairs<true>, grid, blk, 0, s, A.nnz, F.rowidx, A.rp, A.ci,
                       F.dpos, F.ucp, F.uri, F.upos, nullptr, pp, pl, pu, nullptr);
  }
}

// start values in sweep order: a_k s_i s_j (a_k without sc) at cmap[k] of
// as, lc and lc2
__global__ void k_parilu_start(int nnz, const int *__restrict__ rowidx,
                               const int *__restrict__ ci, const double *__restrict__ a,
                               const double *__restrict__ sc, const int *__restrict__ cmap,
                               double *__restrict__ as, double *__restrict__ lc,
                               double *__restrict__ lc2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const double v = sc ? a[k] * sc[rowidx[k]] * sc[ci[k]] : a[k];
  const int p = cmap[k];
  as[p] = v;
  lc[p] = v;
  if (lc2) lc2[p] = v;
}

void xs_parilu_start(hipStream_t s, const xs_csr &A, const xs_ilu &F, const double *sc,
                     double *lc2) {
  XS_GUARD(A.ci && A.v && F.rowidx && F.cmap && F.as && F.lc);
  hipLaunchKernelGGL(k_parilu_start, dim3((A.nnz + 255) / 256), dim3(256), 0, s, A.nnz,
                     F.rowidx, A.ci, A.v, sc, F.cmap, F.as, F.lc, lc2);
}

// factor values back in CSR order, scaled back as k_pt_rescale does
__global__ void k_parilu_finish(int nnz, const int *__restrict__ rowidx,
                                const int *__restrict__ ci, const double *__restrict__ sc,
                                const int *__restrict__ cmap, const double *__restrict__ lc,
                                double *__restrict__ lu) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  double v = lc[cmap[k]];
  if (sc) {
    const int i = rowidx[k], j = ci[k];
    if (j < i) v *= sc[j] / sc[i];
    else v /= sc[i] * sc[j];
  }
  lu[k] = v;
}

void xs_parilu_finish(hipStream_t s, const xs_csr &A, const xs_ilu &F, const double *sc,
                      const double *lc) {
  XS_GUARD(A.ci && F.rowidx && F.cmap && F.lu && lc);
  hipLaunchKernelGGL(k_parilu_finish, dim3((A.nnz + 255) / 256), dim3(256), 0, s, A.nnz,
                     F.rowidx, A.ci, sc, F.cmap, lc, F.lu);
}

// --- ParILU(0) over diagonal slots -----------------------------------------
__global__ void k_parilu_dia_map(int n, int ld, int nd, const xs_dia_offs doff,
                                 const int *__restrict__ rp, const int *__restrict__ ci,
                                 unsigned char *__restrict__ mask, int *__restrict__ dmap) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  unsigned m = 0;
  for (int e = rp[r]; e < rp[r + 1]; ++e) {
    const int o = ci[e] - r;
    int slot = -1;
#pragma unroll
    for (int k = 0; k < XS_DIA_MAX; ++k)
      if (k < nd && doff.o[k] == o) slot = k;
    if (slot < 0) continue;   // not reached: the offsets were detected on this pattern
    m |= 1u << slot;
    dmap[e] = slot * ld + r;
  }
  mask[r] = (unsigned char)m;
}

void xs_parilu_dia_map(hipStream_t s, const xs_csr &A, xs_ilu &F) {
  XS_GUARD(A.rp && A.ci && F.pmask && F.dmap && F.pnd > 0 && F.pnd <= XS_DIA_MAX && F.pld >= A.n);
  hipLaunchKernelGGL(k_parilu_dia_map, dim3((A.n + 255) / 256), dim3(256), 0, s, A.n, F.pld, F.pnd,
                     F.poff, A.rp, A.ci, F.pmask, F.dmap);
}

void xs_parilu_dia_start(hipStream_t s, const xs_csr &A, const xs_ilu &F, const double *sc,
                         double *lc2) {
  XS_GUARD(A.ci && A.v && F.rowidx && F.dmap && F.das && F.dlc);
  hipLaunchKernelGGL(k_parilu_start, dim3((A.nnz + 255) / 256), dim3(256), 0, s, A.nnz,
                     F.rowidx, A.ci, A.v, sc, F.dmap, F.das, F.dlc, lc2);
}

void xs_parilu_dia_finish(hipStream_t s, const xs_csr &A, const xs_ilu &F, const double *sc,
                          const double *lc) {
  XS_GUARD(A.ci && F.rowidx && F.dmap && F.lu && lc);
  hipLaunchKernelGGL(k_parilu_finish, dim3((A.nnz + 255) / 256), dim3(256), 0, s, A.nnz,
                     F.rowidx, A.ci, sc, F.dmap, lc, F.lu);
}

// One ParILU(0) sweep over the diagonal slots, a row per thread: for each of
// the row's entries the arithmetic of k_parilu_sweep_pairs -- the start
// value minus the products l(i, k) u(k, j) over the pattern in ascending k,
// then the pivot guard and division (below the diagonal) or the diagonal's
// guard -- reading the slots of the row and of the rows k and j instead of
// index lists.  A pair counts only where both of its entries are in the
// pattern (the row masks), as the list version's pattern merge.
__global__ __launch_bounds__(XS_BLOCK) void k_parilu_dia_sweep(
    int n, int ld, int nd, const xs_dia_offs doff, const xs_dia_pairs P,
    const unsigned char *__restrict__ mask, const double *__restrict__ as, const double *lc_in,
    double *lc_out) {
  const int i = xcd_chunk((int)blockIdx.x, (int)gridDim.x) * XS_BLOCK + (int)threadIdx.x;
  if (i >= n) return;
  const unsigned m = mask[i];
#pragma unroll
  for (int sl = 0; sl < XS_DIA_MAX; ++sl) {
    if (sl >= nd || !((m >> sl) & 1u) || P.kind[sl] == 3) continue;
    const size_t p = (size_t)sl * ld + i;
    double v = as[p];
    for (int q = 0; q < P.np[sl]; ++q) {
      const int a = P.pl[sl][q], b = P.pu[sl][q];
      if (!((m >> a) & 1u)) continue;
      const int k = i + doff.o[a];
      if (!((mask[k] >> b) & 1u)) continue;
      v -= lc_in[(size_t)a * ld + i] * lc_in[(size_t)b * ld + k];
    }
    double nv;
    if (P.kind[sl] == 0) {
      double dd = lc_in[(size_t)P.diag * ld + i + doff.o[sl]];
      if (fabs(dd) < 1e-10) dd = dd >= 0.0 ? 1e-10 : -1e-10;
      nv = v / dd;
    } else {
      if (P.kind[sl] == 1 && fabs(v) < 1e-10) v = v >= 0.0 ? 1e-10 : -1e-10;
      nv = v;
    }
    lc_out[p] = isfinite(nv) ? nv : lc_in[p];
  }
}

void xs_parilu_dia_sweep(hipStream_t s, const xs_ilu &F, const double *lc_in, double *lc_out) {
  XS_GUARD(F.pmask && F.das && lc_in && lc_out && F.pnd > 0 && F.pnd <= XS_DIA_MAX);
  hipLaunchKernelGGL(k_parilu_dia_sweep, dim3((F.n + XS_BLOCK - 1) / XS_BLOCK), dim3(XS_BLOCK), 0,
                     s, F.n, F.pld, F.pnd, F.poff, F.ppairs, F.pmask, F.das, lc_in, lc_out);
}

// One sweep over the update lists: the arithmetic of k_parilu_sweep (same
// products in the same order, same pivot guard), without the merge, over the
// active entries only (an entry above the diagonal without products is its
// start value at every sweep).  Values are in sweep order, so entry t's
// start value, list data and result are all read / written coalesced.  The
// previous value is read only when the new one is not finite.
__global__ __launch_bounds__(XS_BLOCK) void k_parilu_sweep_pairs(
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
    double dd = lc_in[d];
    if (fabs(dd) < 1e-10) dd = dd >= 0.0 ? 1e-10 : -1e-10;
    nv = s / dd;
  } else {
    if (d == -2 && fabs(s) < 1e-10) s = s >= 0.0 ? 1e-10 : -1e-10;
    nv = s;
  }
  lc_out[t] = isfinite(nv) ? nv : lc_in[t];
}

void xs_parilu_sweep_pairs(hipStream_t s, const xs_ilu &F, const double *lc_in, double *lc_out) {
  XS_GUARD(F.app && F.apd && F.pl && F.pu && F.as && lc_in && lc_out);
  if (F.nact == 0) return;
  hipLaunchKernelGGL(k_parilu_sweep_pairs, dim3((F.nact + XS_BLOCK - 1) / XS_BLOCK), dim3(XS_BLOCK),
                     0, s, F.nact, F.as, F.app, F.pl, F.pu, F.apd, lc_in, lc_out);
}

__global__ void k_parilu_active_lists(int nact, int nnz, const int *__restrict__ act,
                                      const int *__restrict__ pp, const int *__restrict__ pd,
                                      const int *__restrict__ cmap, int *__restrict__ app,
                                      int *__restrict__ apd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nact) {
    const int k = act[t], d = pd[k];
    app[t] = pp[k];
    apd[t] = d >= 0 ? cmap[d] : d;
  }
  if (t == 0) app[nact] = pp[nnz];
}
void xs_parilu_active_lists(hipStream_t s, int nact, int nnz, const int *act, const int *pp,
                            const int *pd, const int *cmap, int *app, int *apd) {
  XS_GUARD(pp && pd && cmap && app && apd && (act || nact == 0));
  hipLaunchKernelGGL(k_parilu_active_lists, dim3((std::max(nact, 1) + 255) / 256), dim3(256), 0, s,
                     nact, nnz, act, pp, pd, cmap, app, apd);
}

__global__ void k_parilu_cmap(int nnz, int nact, const int *__restrict__ flag,
                              const int *__restrict__ pos, int *__restrict__ cmap) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) cmap[k] = flag[k] ? pos[k] : nact + k - pos[k];
}
void xs_parilu_cmap(hipStream_t s, int nnz, int nact, const int *flag, const int *pos, int *cmap) {
  XS_GUARD(flag && pos && cmap);
  hipLaunchKernelGGL(k_parilu_cmap, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, nact, flag, pos,
                     cmap);
}

__global__ void k_remap_idx(int n, const int *__restrict__ cmap, int *__restrict__ idx) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) idx[e] = cmap[idx[e]];
}
void xs_remap_idx(hipStream_t s, int n, const int *cmap, int *idx) {
  XS_GUARD(cmap && (idx || n == 0));
  if (n <= 0) return;
  hipLaunchKernelGGL(k_remap_idx, dim3((n + 255) / 256), dim3(256), 0, s, n, cmap, idx);
}

__global__ void k_parilu_active_flags(int nnz, const int *__restrict__ cnt,
                                      const int *__restrict__ pd, int *__restrict__ flag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) flag[k] = (cnt[k] > 0 || pd[k] != -1) ? 1 : 0;
}
void xs_parilu_active_flags(hipStream_t s, int nnz, const int *cnt, const int *pd, int *flag) {
  XS_GUARD(cnt && pd && flag);
  hipLaunchKernelGGL(k_parilu_active_flags, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, cnt, pd, flag);
}
__global__ void k_compact_flagged(int nnz, const int *__restrict__ flag, const int *__restrict__ pos,
                                  int *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz && flag[k]) out[pos[k]] = k;
}
void xs_compact_flagged(hipStream_t s, int nnz, const int *flag, const int *pos, int *out) {
  XS_GUARD(flag && pos && out);
  hipLaunchKernelGGL(k_compact_flagged, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, flag, pos, out);
}

void xs_ilu_copy_values(hipStream_t s, const xs_csr &A, xs_ilu &F) {
  XS_GUARD(A.v && F.lu);
  (void)hipMemcpyAsync(F.lu, A.v, sizeof(double) * A.nnz, hipMemcpyDeviceToDevice, s);
}

__global__ void k_ilu_dinv(int n, const double *__restrict__ lu, const int *__restrict__ dpos,
                           double *__restrict__ dinv, int *bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = lu[dpos[i]];
  if (d == 0.0 || !isfinite(d)) atomicMin(bad, i + 1);
  dinv[i] = 1.0 / d;
}

void xs_ilu_dinv(hipStream_t s, xs_ilu &F, int *d_bad) {
  XS_GUARD(F.lu && F.dpos && F.dinv && d_bad);
  hipLaunchKernelGGL(k_ilu_dinv, dim3((F.n + 255) / 256), dim3(256), 0, s, F.n, F.lu, F.dpos,
                     F.dinv, d_bad);
}

// ---------------------------------------------------------------------------
// Sync-free triangular solves (no level sets).
//
// Work is handed out in 64-row groups, one row per lane: each workgroup takes
// a ticket from a global counter and its waves solve the groups of that
// ticket (tickets map to groups in dependency-level order when a schedule is
// given, in row order otherwise), so a wave only ever waits on rows of waves
// that started before it or sit in its workgroup: every wait ends.
// The solution vector itself carries the per-row ready flag: it is
// pre-filled with a signalling-NaN sentinel, each row's value is published
// with an agent-scope atomic store, and a dependent row polls (s_sleep
// between polls) with agent-scope atomic loads until the sentinel is gone;
// all pending values of a row are re-polled together, one round trip per
// poll.
// The diagonal is unit when dinv is null (the L factor, and L^T in the
// transposed solve) and 1/u_ii otherwise (U, and U^T).
// Lanes run along the solve order (upper solves: lane l holds row r1-1-l), so
// dependencies inside a wave always point to lower lanes.  They are resolved
// in registers: a DPP affine scan when each row depends at most on the
// previous one (x_i = c_i + m_i x_{i-1}), a 64-step broadcast loop otherwise.
// The last workgroup to retire resets the tickets for the next call.
// Reference kernels replaced: Lcsr_trsv_sync_free / Ucsc_trsv_sync_free
// (hip/Precond/gpuparilu.cpp:257-391): there one wave per row spins on a
// separate volatile int flag; the backward solve scatters with double atomics.
// ---------------------------------------------------------------------------
// Up to G entries e, e+dir, ... (stopping at `stop`) of a row: their columns
// c (-1 past the stop), values a and the current x of those columns u (which
// may still hold the sentinel).  Branch-free loads (entry e is in range: a
// slot past the stop re-reads it and is masked by c = -1), so that they are
// all in flight at once.
template <int G>
__device__ __forceinline__ void trsv_issue(int e, int stop, int dir, const int *__restrict__ ci,
                                           const double *__restrict__ lu, const double *x, int *c,
                                           double *a, unsigned long long *u) {
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int eq = e + q * dir;
    const bool in = dir > 0 ? eq < stop : eq > stop;
    const int cv = ci[in ? eq : e];
    const double av = lu[in ? eq : e];
    c[q] = in ? cv : -1;
    a[q] = in ? av : 0.0;
  }
#pragma unroll
  for (int q = 0; q < G; ++q) u[q] = ld_relaxed(x + (c[q] >= 0 ? c[q] : c[0]));
}

template <int G>
__device__ __forceinline__ bool trsv_pending(const int *c, const unsigned long long *u) {
  bool pend = false;
#pragma unroll
  for (int q = 0; q < G; ++q) pend |= (c[q] >= 0 && u[q] == XS_SENTINEL);
  return pend;
}

// re-poll the values still holding the sentinel, all together, until none
// is left (bounded: past the spin limit the flag is raised and NaN used)
template <int G>
__device__ __forceinline__ void trsv_poll(const int *c, unsigned long long *u, const double *x,
                                          unsigned *err) {
  unsigned polls = 0;
  while (trsv_pending<G>(c, u)) {
    if (++polls > XS_SPIN_LIMIT) {
      atomicOr(err, 1u);
#pragma unroll
      for (int q = 0; q < G; ++q)
        if (u[q] == XS_SENTINEL) u[q] = 0x7FF8000000000000ull;
      break;
    }
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int q = 0; q < G; ++q)
      if (c[q] >= 0 && u[q] == XS_SENTINEL) u[q] = ld_relaxed(x + c[q]);
  }
}

// Sum of G values (G a power of two) by a fixed pairwise tree: log2(G)
// dependent additions instead of G.
template <int G>
__device__ __forceinline__ double trsv_tree(const double *p) {
  double t[G];
#pragma unroll
  for (int q = 0; q < G; ++q) t[q] = p[q];
#pragma unroll
  for (int w = G / 2; w >= 1; w /= 2) {
#pragma unroll
    for (int q = 0; q < w; ++q) t[q] = t[q] + t[q + w];
  }
  return t[0];
}

// acc minus one batch of products (0 where there is no entry), the batch
// summed by the fixed tree; batches always start at the same entries (the
// row's first external entry in solve order, then every G), so every row's
// value is the same sequence of roundings whatever the timing.  Products are
// rounded on their own (no fused multiply-add): the lane-by-lane path forms
// some of them ahead of the subtraction and must round them the same way.
template <int G>
__device__ __forceinline__ double trsv_sub(double acc, const int *c, const double *a,
                                           const unsigned long long *u) {
#pragma clang fp contract(off)
  double p[G];
#pragma unroll
  for (int q = 0; q < G; ++q) p[q] = c[q] >= 0 ? a[q] * __longlong_as_double((long long)u[q]) : 0.0;
  return acc - trsv_tree<G>(p);
}

// acc - sum of lu[e] * x[ci[e]] over up to G entries from e, waiting for the
// values that are not yet published
template <int G>
__device__ __forceinline__ double trsv_gather(double acc, int e, int stop, int dir,
                                               const int *__restrict__ ci,
                                               const double *__restrict__ lu, const double *x,
                                               unsigned *err) {
  unsigned long long u[G];
  int c[G];
  double a[G];
  trsv_issue<G>(e, stop, dir, ci, lu, x, c, a, u);
  trsv_poll<G>(c, u, x, err);
  return trsv_sub<G>(acc, c, a, u);
}

template <bool UPPER, int G>
__global__ __launch_bounds__(XS_BLOCK) void k_trsv(int n, const int *__restrict__ rp,
                                                   const int *__restrict__ ci,
                                                   const double *__restrict__ lu,
                                                   const int *__restrict__ dpos,
                                                   const double *__restrict__ dinv,
                                                   const double *__restrict__ b, double *x,
                                                   unsigned *tickets, const xs_tsched S,
                                                   const xs_state *st) {
#pragma clang fp contract(off)
  if (st != nullptr && st->done) return;
  unsigned *err = tickets + 4;  // wait-expired flag, 4 words past the tickets
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
    double acc = 0.0, di = 1.0;
    int eb = 0, ee = 0;  // in-wave entries [eb, ee)
    int xb = 0, xe = 0;  // entries outside the group
    // the first XS_INW in-wave entries in solve order, loaded before any wait
    // so that no load sits on the in-wave chain
    int iw_c[XS_INW];
    double iw_v[XS_INW];
#pragma unroll
    for (int q = 0; q < XS_INW; ++q) { iw_c[q] = -1; iw_v[q] = 0.0; }
    if (act) {
      acc = b[r];
      if (dinv) di = dinv[r];
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
    }
    // external entries in solve order: L from xb up, U from xe - 1 down
    const int dir = UPPER ? -1 : 1;
    const int estop = UPPER ? xb - 1 : xe;
    int ep = UPPER ? xe - 1 : xb;
    const int nint = ee - eb;
    const int nb = UPPER ? r + 1 : r - 1;
    const bool chain = !act || nint == 0 || (nint == 1 && iw_c[0] == nb);
    if (!S.prog && __all(chain)) {
      // every row depends at most on the one before it: wait for all
      // external values, then x_r = c + m * x_nb composed along the lanes
      if (act)
        for (; UPPER ? ep > estop : ep < estop; ep += G * dir)
          acc = trsv_gather<G>(acc, ep, estop, dir, ci, lu, x, err);
      double c[1] = {act ? di * acc : 0.0};
      double m = (act && nint == 1) ? -di * iw_v[0] : 0.0;
      wave_affine_scan<1>(c, m);
      if (act) st_ready(x + r, c[0]);
    } else {
      // Rows depend on several earlier rows of the group: they are finished
      // one lane at a time, and each is published as soon as it is final,
      // so the groups that wait on it go on without waiting for the rest of
      // this group.  Up front every lane subtracts its external products in
      // order up to the first batch holding a value not yet published (that
      // batch stays in registers); at its turn the lane waits for the rest.
      // held: the first batch with an unpublished value and the batch after
      // it (2G entries), so that a lane's turn rarely issues new loads.  A
      // held entry keeps its column (hc, -1: none) and in hv its value until
      // x of that column arrives, then the product (0 for no entry); hm has
      // a bit per entry still waiting.  Products are rounded on their own
      // (no fused multiply-add, as in trsv_sub), so an entry subtracts the
      // same bits whether its product was formed up front or held.
      int hc[2 * G];
      double hv[2 * G];
      unsigned hm = 0;
      bool stalled = false, more = false;   // more: the second batch is held
      {
        int pc[G];
        double pa[G];
        unsigned long long pu[G];
#pragma unroll
        for (int q = 0; q < G; ++q) { pc[q] = -1; pa[q] = 0.0; pu[q] = 0; }
        if (act) {
          for (; UPPER ? ep > estop : ep < estop; ep += G * dir) {
            trsv_issue<G>(ep, estop, dir, ci, lu, x, pc, pa, pu);
            if (trsv_pending<G>(pc, pu)) { stalled = true; break; }
            acc = trsv_sub<G>(acc, pc, pa, pu);
          }
        }
#pragma unroll
        for (int q = 0; q < G; ++q) {
          const bool wait = pc[q] >= 0 && pu[q] == XS_SENTINEL;
          hc[q] = stalled ? pc[q] : -1;
          hv[q] = (!stalled || pc[q] < 0) ? 0.0
                  : wait ? pa[q] : pa[q] * __longlong_as_double((long long)pu[q]);
          if (stalled && wait) hm |= 1u << q;
        }
        const int e2 = ep + G * dir;
        more = stalled && (UPPER ? e2 > estop : e2 < estop);
        if (more) trsv_issue<G>(e2, estop, dir, ci, lu, x, pc, pa, pu);
#pragma unroll
        for (int q = 0; q < G; ++q) {
          const bool wait = pc[q] >= 0 && pu[q] == XS_SENTINEL;
          hc[G + q] = more ? pc[q] : -1;
          hv[G + q] = (!more || pc[q] < 0) ? 0.0
                      : wait ? pa[q] : pa[q] * __longlong_as_double((long long)pu[q]);
          if (more && wait) hm |= 1u << (G + q);
        }
      }
      double ain = 0.0;   // the in-wave products, subtracted in solve order
      int left = act ? nint - XS_INW : 0;
      int enext = UPPER ? ee - 1 - XS_INW : eb + XS_INW;
      // The turn check reads only hm (a register refreshed where polls
      // return), so a turn whose values are known to be there waits for no
      // memory result: on this target a wait for any load also waits for
      // the previous turn's publishing store.
      for (int k = 0; k < nrows; ++k) {
        // while lane k's held batches still have an unpublished value, every
        // lane from k on re-polls its own in the same round trip, so that
        // they rarely wait at their own turn: a round trip (~2.5 us on the
        // 100^3 ParILUT factors) is worth more than the polls it carries --
        // re-polling only the next 4 lanes was 8% slower there, and the same
        // on the 7-point factors
        unsigned polls = 0;
        bool expired = false;
        while (__builtin_amdgcn_readlane((int)hm, k)) {
          if (++polls > XS_SPIN_LIMIT) {   // bounded wait: flag it, continue with NaN
            if (lane == k) atomicOr(err, 1u);
            expired = true;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
          if (hm && lane >= k) {
            unsigned long long pu[2 * G];
#pragma unroll
            for (int q = 0; q < 2 * G; ++q)
              pu[q] = ((hm >> q) & 1u) ? ld_relaxed(x + hc[q]) : XS_SENTINEL;
#pragma unroll
            for (int q = 0; q < 2 * G; ++q)
              if (pu[q] != XS_SENTINEL) {
                hv[q] = hv[q] * __longlong_as_double((long long)pu[q]);
                hm &= ~(1u << q);
              }
          }
        }
        double vk = 0.0;
        if (lane == k) {
          if (stalled) {
            acc -= trsv_tree<G>(hv);
            if (more) acc -= trsv_tree<G>(hv + G);   // else all zeros: acc is unchanged
            for (ep += 2 * G * dir; UPPER ? ep > estop : ep < estop; ep += G * dir)
              acc = trsv_gather<G>(acc, ep, estop, dir, ci, lu, x, err);
          }
          vk = expired ? __longlong_as_double(0x7FF8000000000000ll) : (acc + ain) * di;
          st_ready(x + r, vk);
        }
        const double xk = wave_readlane_f64(vk, k);
        const int rk = UPPER ? r1 - 1 - k : r0 + k;
        if (iw_c[0] == rk) {
          ain -= iw_v[0] * xk;
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
    }
  }
  }
  trsv_retire(tickets);
}

// x must hold the sentinel in every entry when the kernel starts
// (xs_fill_sentinel); the kernel leaves no sentinel behind.
static void trsv_launch(hipStream_t s, bool upper, int n, const int *rp, const int *ci,
                        const double *lu, const int *dpos, const double *dinv, const double *b,
                        double *x, unsigned *tickets, const xs_tsched &S, const xs_state *st) {
  XS_GUARD(rp && ci && lu && dpos && b && x && tickets && S.max_wg > 0 && (!S.gb || S.ngroups > 0));
  const int nw = S.gb ? S.ngroups : (n + XS_WAVE - 1) / XS_WAVE;
  const int tk = (nw + (XS_BLOCK / XS_WAVE) - 1) / (XS_BLOCK / XS_WAVE);
  dim3 grid(tk < S.max_wg ? tk : S.max_wg), blk(XS_BLOCK);
#define XS_TRK(U, G) hipLaunchKernelGGL((k_trsv<U, G>), grid, blk, 0, s, n, rp, ci, lu, dpos, dinv, b, x, tickets, S, st)
  if (S.gather == 16) { if (upper) XS_TRK(true, 16); else XS_TRK(false, 16); }
  else { if (upper) XS_TRK(true, XS_GATHER); else XS_TRK(false, XS_GATHER); }
#undef XS_TRK
}

void xs_trsv_lower(hipStream_t s, const xs_ilu &F, const double *b, double *x, unsigned *tickets,
                   const xs_state *st) {
  trsv_launch(s, false, F.n, F.rp, F.ci, F.lu, F.dpos, nullptr, b, x, tickets, xs_ilu_tsched(F, false), st);
}
void xs_trsv_upper(hipStream_t s, const xs_ilu &F, const double *b, double *x, unsigned *tickets,
                   const xs_state *st) {
  trsv_launch(s, true, F.n, F.rp, F.ci, F.lu, F.dpos, F.dinv, b, x, tickets, xs_ilu_tsched(F, true), st);
}
// Transposed factors (F^T = U^T | L^T on the transposed pattern): the lower
// solve divides by u_ii, the upper one has a unit diagonal; consecutive
// 64-row groups in solve order.
void xs_trsv_lower_t(hipStream_t s, const xs_ilu &F, const double *b, double *x,
                     unsigned *tickets, const xs_state *st) {
  xs_tsched S;
  S.gb = F.tlgb;
  S.ngroups = F.tlgroups;
  S.order = F.tlsched;
  S.max_wg = F.trsv_wg;
  S.gather = xs_ilu_gather(F);
  S.prog = F.trsv_prog >= 0 ? F.trsv_prog : (S.gather == 16 ? 1 : 0);
  trsv_launch(s, false, F.n, F.trp, F.tci, F.tlu, F.tdpos, F.dinv, b, x, tickets, S, st);
}
void xs_trsv_upper_t(hipStream_t s, const xs_ilu &F, const double *b, double *x,
                     unsigned *tickets, const xs_state *st) {
  xs_tsched S;
  S.gb = F.tugb;
  S.ngroups = F.tugroups;
  S.order = F.tusched;
  S.max_wg = F.trsv_wg;
  S.gather = xs_ilu_gather(F);
  S.prog = F.trsv_prog >= 0 ? F.trsv_prog : (S.gather == 16 ? 1 : 0);
  trsv_launch(s, true, F.n, F.trp, F.tci, F.tlu, F.tdpos, nullptr, b, x, tickets, S, st);
}

// ---------------------------------------------------------------------------
// Jacobi sweeps over ELL copies of the factor parts (the arithmetic of
// k_tri_sweep, entry by entry in column order, so the results are identical
// except in the fused U pass, which reads u_ij / u_jj (to rounding); coalesced
// slab loads instead of a thread walking its CSR row).
// ---------------------------------------------------------------------------
// strictly-lower / strictly-upper entry counts: maxima into w[0], w[1]
__global__ __launch_bounds__(XS_BLOCK) void k_tri_widths(int n, const int *__restrict__ rp,
                                                         const int *__restrict__ dpos, int *w) {
  int lo = 0, up = 0;
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    lo = max(lo, dpos[i] - rp[i]);
    up = max(up, rp[i + 1] - dpos[i] - 1);
  }
  lo = block_max(lo);
  up = block_max(up);
  if (threadIdx.x == 0) { atomicMax(&w[0], lo); atomicMax(&w[1], up); }
}

int xs_tri_widths(hipStream_t s, const xs_ilu &F, int *d_tmp, int *wl, int *wu) {
  if (!F.rp || !F.dpos || !d_tmp) { xs_raise_launch_guard(); return 1; }
  int h[2] = {0, 0};
  if (hipMemsetAsync(d_tmp, 0, 2 * sizeof(int), s) != hipSuccess) return 1;
  hipLaunchKernelGGL(k_tri_widths, dim3(std::max(1, std::min((F.n + XS_BLOCK - 1) / XS_BLOCK, XS_MAX_GRID))),
                     dim3(XS_BLOCK), 0, s, F.n, F.rp, F.dpos, d_tmp);
  if (hipMemcpyAsync(h, d_tmp, 2 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return 1;
  *wl = h[0];
  *wu = h[1];
  return 0;
}

template <bool UPPER>
__global__ void k_build_tell(int n, int ld, int W, const int *__restrict__ rp,
                             const int *__restrict__ ci, const double *__restrict__ lu,
                             const int *__restrict__ dpos, int *__restrict__ eci,
                             double *__restrict__ ev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ld) return;
  int a = 0, len = 0;
  if (i < n) {
    a = UPPER ? dpos[i] + 1 : rp[i];
    len = UPPER ? rp[i + 1] - dpos[i] - 1 : dpos[i] - rp[i];
  }
  for (int k = 0; k < W; ++k) {
    const size_t p = (size_t)k * ld + i;
    if (k < len) { eci[p] = ci[a + k]; ev[p] = lu[a + k]; }
    else { eci[p] = -1; ev[p] = 0.0; }
  }
}

void xs_build_tell(hipStream_t s, const xs_ilu &F, bool upper, xs_tell &E) {
  XS_GUARD(F.rp && F.ci && F.lu && F.dpos && (E.W == 0 || (E.ci && E.v)) && E.ld >= F.n);
  if (E.W == 0) return;
  const dim3 grid((E.ld + 255) / 256), blk(256);
  if (upper)
    hipLaunchKernelGGL(k_build_tell<true>, grid, blk, 0, s, F.n, E.ld, E.W, F.rp, F.ci, F.lu, F.dpos, E.ci, E.v);
  else
    hipLaunchKernelGGL(k_build_tell<false>, grid, blk, 0, s, F.n, E.ld, E.W, F.rp, F.ci, F.lu, F.dpos, E.ci, E.v);
}

template <bool UPPER>
__global__ __launch_bounds__(XS_BLOCK) void k_tri_wave_widths(int n, const int *__restrict__ rp,
                                                              const int *__restrict__ dpos,
                                                              int *__restrict__ ww) {
  const int i = blockIdx.x * XS_BLOCK + threadIdx.x;
  int len = 0;
  if (i < n) len = UPPER ? rp[i + 1] - dpos[i] - 1 : dpos[i] - rp[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) len = max(len, __shfl_xor(len, off, XS_WAVE));
  if ((threadIdx.x & (XS_WAVE - 1)) == 0 && i < n) ww[i / XS_WAVE] = len;
}

void xs_tri_wave_widths(hipStream_t s, const xs_ilu &F, bool upper, int *ww) {
  XS_GUARD(F.rp && F.dpos && ww);
  const dim3 grid((F.n + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
  if (upper) hipLaunchKernelGGL(k_tri_wave_widths<true>, grid, blk, 0, s, F.n, F.rp, F.dpos, ww);
  else hipLaunchKernelGGL(k_tri_wave_widths<false>, grid, blk, 0, s, F.n, F.rp, F.dpos, ww);
}

__global__ void k_tell_scale_cols(int W, int ld, const int *__restrict__ eci,
                                  const double *__restrict__ ev, const double *__restrict__ dinv,
                                  double *__restrict__ vs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ld) return;
  for (int k = 0; k < W; ++k) {
    const size_t p = (size_t)k * ld + i;
    const int c = eci[p];
    vs[p] = c >= 0 ? ev[p] * dinv[c] : 0.0;
  }
}
void xs_tell_scale_cols(hipStream_t s, const xs_tell &E, const double *dinv) {
  XS_GUARD((E.W == 0 || (E.ci && E.v && E.vs)) && dinv);
  if (E.W == 0) return;
  hipLaunchKernelGGL(k_tell_scale_cols, dim3((E.ld + 255) / 256), dim3(256), 0, s, E.W, E.ld, E.ci,
                     E.v, dinv, E.vs);
}

// yn = D (b - T yo), T given by its ELL slabs; FUSED: yo = b (L) or D b (U),
// i.e. sweeps 1 and 2 from y = 0 in one pass.  SX (fused L pass only): the
// right-hand side is BiCGSTAB's s = r - gama u, formed here from b = r and
// ru = u at every row it is read and written to sv for the later kernels
// (the s update of k_bicg_s, same expression).  Slabs four at a time: all
// index/value loads of a batch, then all gathers.  Rows in 256-row blocks
// placed by xcd_chunk.
template <bool UPPER, bool FUSED, bool O16, bool SX = false>
__global__ __launch_bounds__(XS_BLOCK) void k_tri_sweep_ell(int n, int ld, int W,
                                                            const int *__restrict__ eci,
                                                            const short *__restrict__ eco,
                                                            const double *__restrict__ ev,
                                                            const int *__restrict__ ww,
                                                            const double *__restrict__ dinv,
                                                            const double *__restrict__ b,
                                                            const double *__restrict__ yo,
                                                            double *__restrict__ yn,
                                                            const xs_state *st,
                                                            const double *__restrict__ ru = nullptr,
                                                            double *__restrict__ sv = nullptr) {
  static_assert(!SX || (FUSED && !UPPER), "s is formed in the fused L pass only");
  if (st != nullptr && st->done) return;
  const int i = xcd_chunk((int)blockIdx.x, (int)gridDim.x) * XS_BLOCK + (int)threadIdx.x;
  if (i >= n) return;
  const double gama = SX ? st->sc[S_GAMA] : 0.0;
  // the row's own operands, loaded before the slabs (off the critical path);
  // the slabs past the longest row of the row's 64-row group are all padding
  const double bi = SX ? fma(-gama, ru[i], b[i]) : b[i];
  if (SX) sv[i] = bi;
  const double di = UPPER ? dinv[i] : 1.0;
  W = ww ? ww[i / XS_WAVE] : W;
  double r = 0.0;
  for (int k0 = 0; k0 < W; k0 += 4) {
    // branch-free loads: a slot past W re-reads slab W - 1 and a padding
    // entry gathers the row's own value, both masked from the sum, so that
    // the batch's loads are all issued before any wait
    int c[4];
    double a[4], y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const size_t p = (size_t)min(k0 + q, W - 1) * ld + i;
      if (O16) {
        const int o = (int)eco[p];
        c[q] = (k0 + q >= W || o == XS_OFF16_PAD) ? -1 : i + o;
      } else {
        c[q] = k0 + q < W ? eci[p] : -1;
      }
      a[q] = ev[p];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // FUSED: the previous iterate is D b (U) or b (L); for U the slab
      // values are u_ij / u_jj (E.vs), so b is gathered alone
      const int cc = c[q] >= 0 ? c[q] : i;
      y[q] = SX ? fma(-gama, ru[cc], b[cc]) - 0.0 : FUSED ? b[cc] - 0.0 : yo[cc];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (c[q] >= 0) r += a[q] * y[q];
  }
  yn[i] = UPPER ? di * (bi - r) : bi - r;
}

// k_tri_sweep_ell over the diagonal-slot copy (E.nd slots, offsets by
// value): no index loads; slot k of row i reads the iterate at i + o_k
// (clamped for the zero slots of rows that lack the diagonal).  Same
// entries in the same column order as the ELL rows, the zero slots adding
// zero products.
template <bool UPPER, bool FUSED, bool SX = false>
__global__ __launch_bounds__(XS_BLOCK) void k_tri_sweep_dia(int n, int ld, int nd,
                                                            const xs_dia_offs doff,
                                                            const double *__restrict__ ev,
                                                            const double *__restrict__ dinv,
                                                            const double *__restrict__ b,
                                                            const double *__restrict__ yo,
                                                            double *__restrict__ yn,
                                                            const xs_state *st,
                                                            const double *__restrict__ ru = nullptr,
                                                            double *__restrict__ sv = nullptr) {
  static_assert(!SX || (FUSED && !UPPER), "s is formed in the fused L pass only");
  if (st != nullptr && st->done) return;
  const int i = xcd_chunk((int)blockIdx.x, (int)gridDim.x) * XS_BLOCK + (int)threadIdx.x;
  if (i >= n) return;
  const double gama = SX ? st->sc[S_GAMA] : 0.0;
  const double bi = SX ? fma(-gama, ru[i], b[i]) : b[i];
  if (SX) sv[i] = bi;
  const double di = UPPER ? dinv[i] : 1.0;
  double r = 0.0;
#pragma unroll
  for (int k0 = 0; k0 < XS_DIA_MAX; k0 += 4) {
    if (k0 >= nd) break;
    double a[4], y[4];
    int cc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = min(k0 + q, nd - 1);
      a[q] = ev[(size_t)k * ld + i];
      cc[q] = min(max(i + doff.o[k], 0), n - 1);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      y[q] = SX ? fma(-gama, ru[cc[q]], b[cc[q]]) - 0.0 : FUSED ? b[cc[q]] - 0.0 : yo[cc[q]];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (k0 + q < nd) r += a[q] * y[q];
  }
  yn[i] = UPPER ? di * (bi - r) : bi - r;
}

// k_tri_sweep_ell_px over the diagonal-slot copy of L
__global__ __launch_bounds__(XS_BLOCK) void k_tri_sweep_dia_px(int n, int ld, int nd,
                                                               const xs_dia_offs doff,
                                                               const double *__restrict__ ev,
                                                               const double *__restrict__ r,
                                                               const double *__restrict__ u,
                                                               const double *__restrict__ pin,
                                                               double *__restrict__ pout,
                                                               double *__restrict__ yn,
                                                               const xs_state *st) {
  if (st->done) return;
  const int i = xcd_chunk((int)blockIdx.x, (int)gridDim.x) * XS_BLOCK + (int)threadIdx.x;
  if (i >= n) return;
  const double omega = st->sc[S_OMEGA], alpha = st->sc[S_ALPHA];
  const double bi = bicg_p_val(r[i], pin[i], u[i], omega, alpha);
  pout[i] = bi;
  double acc = 0.0;
#pragma unroll
  for (int k0 = 0; k0 < XS_DIA_MAX; k0 += 4) {
    if (k0 >= nd) break;
    double a[4], y[4];
    int cc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = min(k0 + q, nd - 1);
      a[q] = ev[(size_t)k * ld + i];
      cc[q] = min(max(i + doff.o[k], 0), n - 1);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) y[q] = bicg_p_val(r[cc[q]], pin[cc[q]], u[cc[q]], omega, alpha) - 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (k0 + q < nd) acc += a[q] * y[q];
  }
  yn[i] = bi - acc;
}

static bool tell_dia(const xs_tell &E) { return E.nd > 0 && E.nd <= XS_DIA_MAX && E.dv; }

void xs_tri_sweep_ell(hipStream_t s, const xs_ilu &F, bool upper, bool fused, const double *b,
                      const double *yo, double *yn, const xs_state *st) {
  const xs_tell &E = upper ? F.tu : F.tl;
  if (tell_dia(E)) {
    XS_GUARD(E.ld >= F.n && b && yn && (!upper || (F.dinv && E.dvs)) && (fused || yo));
    const dim3 grid((F.n + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
    const double *ev = (upper && fused) ? E.dvs : E.dv;
#define XS_SWD(U, FU) hipLaunchKernelGGL((k_tri_sweep_dia<U, FU>), grid, blk, 0, s, F.n, E.ld, E.nd, E.doff, ev, F.dinv, b, yo, yn, st)
    if (upper) { if (fused) XS_SWD(true, true); else XS_SWD(true, false); }
    else { if (fused) XS_SWD(false, true); else XS_SWD(false, false); }
#undef XS_SWD
    return;
  }
  XS_GUARD((E.W == 0 || (E.ci && E.v)) && E.ld >= F.n && b && yn && (!upper || F.dinv) &&
           (fused || yo));
  const dim3 grid((F.n + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
  XS_GUARD(!(upper && fused) || E.W == 0 || E.vs);
  const double *ev = (upper && fused) ? E.vs : E.v;
#define XS_SWK(U, FU, O) hipLaunchKernelGGL((k_tri_sweep_ell<U, FU, O>), grid, blk, 0, s, F.n, E.ld, E.W, E.ci, E.co, ev, E.ww, F.dinv, b, yo, yn, st)
#define XS_SWE(U, FU) do { if (E.co) XS_SWK(U, FU, true); else XS_SWK(U, FU, false); } while (0)
  if (upper) { if (fused) XS_SWE(true, true); else XS_SWE(true, false); }
  else { if (fused) XS_SWE(false, true); else XS_SWE(false, false); }
#undef XS_SWE
#undef XS_SWK
}

// The fused first L pass with BiCGSTAB's direction update folded in: the
// right-hand side p_j = r_j + omega (pin_j - alpha u_j) (bicg_p_val, as in
// k_bicg_p) is formed at every row read, written to pout for the row itself;
// pin and pout are different vectors (rows read their neighbours' pin).
// Otherwise the arithmetic of k_tri_sweep_ell<false, true, O16>.
template <bool O16>
__global__ __launch_bounds__(XS_BLOCK) void k_tri_sweep_ell_px(int n, int ld, int W,
                                                               const int *__restrict__ eci,
                                                               const short *__restrict__ eco,
                                                               const double *__restrict__ ev,
                                                               const int *__restrict__ ww,
                                                               const double *__restrict__ r,
                                                               const double *__restrict__ u,
                                                               const double *__restrict__ pin,
                                                               double *__restrict__ pout,
                                                               double *__restrict__ yn,
                                                               const xs_state *st) {
  if (st->done) return;
  const int i = xcd_chunk((int)blockIdx.x, (int)gridDim.x) * XS_BLOCK + (int)threadIdx.x;
  if (i >= n) return;
  const double omega = st->sc[S_OMEGA], alpha = st->sc[S_ALPHA];
  const double bi = bicg_p_val(r[i], pin[i], u[i], omega, alpha);
  pout[i] = bi;
  W = ww ? ww[i / XS_WAVE] : W;
  double acc = 0.0;
  for (int k0 = 0; k0 < W; k0 += 4) {
    int c[4];
    double a[4], y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const size_t p = (size_t)min(k0 + q, W - 1) * ld + i;
      if (O16) {
        const int o = (int)eco[p];
        c[q] = (k0 + q >= W || o == XS_OFF16_PAD) ? -1 : i + o;
      } else {
        c[q] = k0 + q < W ? eci[p] : -1;
      }
      a[q] = ev[p];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int cc = c[q] >= 0 ? c[q] : i;
      y[q] = bicg_p_val(r[cc], pin[cc], u[cc], omega, alpha) - 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (c[q] >= 0) acc += a[q] * y[q];
  }
  yn[i] = bi - acc;
}

void xs_tri_sweep_ell_p(hipStream_t s, const xs_ilu &F, const double *r, const double *u,
                        const double *pin, double *pout, double *yn, const xs_state *st) {
  const xs_tell &E = F.tl;
  if (tell_dia(E)) {
    XS_GUARD(E.ld >= F.n && r && u && pin && pout && pin != pout && yn && st);
    const dim3 grid((F.n + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
    hipLaunchKernelGGL(k_tri_sweep_dia_px, grid, blk, 0, s, F.n, E.ld, E.nd, E.doff, E.dv, r, u,
                       pin, pout, yn, st);
    return;
  }
  XS_GUARD((E.W == 0 || (E.ci && E.v)) && E.ld >= F.n && r && u && pin && pout && pin != pout &&
           yn && st);
  const dim3 grid((F.n + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
  if (E.co)
    hipLaunchKernelGGL((k_tri_sweep_ell_px<true>), grid, blk, 0, s, F.n, E.ld, E.W, E.ci, E.co,
                       E.v, E.ww, r, u, pin, pout, yn, st);
  else
    hipLaunchKernelGGL((k_tri_sweep_ell_px<false>), grid, blk, 0, s, F.n, E.ld, E.W, E.ci, E.co,
                       E.v, E.ww, r, u, pin, pout, yn, st);
}

void xs_tri_sweep_ell_s(hipStream_t s, const xs_ilu &F, const double *r, const double *u,
                        double *sv, double *yn, const xs_state *st) {
  const xs_tell &E = F.tl;
  if (tell_dia(E)) {
    XS_GUARD(E.ld >= F.n && r && u && sv && yn && st);
    const dim3 grid((F.n + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
    hipLaunchKernelGGL((k_tri_sweep_dia<false, true, true>), grid, blk, 0, s, F.n, E.ld, E.nd,
                       E.doff, E.dv, F.dinv, r, nullptr, yn, st, u, sv);
    return;
  }
  XS_GUARD((E.W == 0 || (E.ci && E.v)) && E.ld >= F.n && r && u && sv && yn && st);
  const dim3 grid((F.n + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
  if (E.co)
    hipLaunchKernelGGL((k_tri_sweep_ell<false, true, true, true>), grid, blk, 0, s, F.n, E.ld, E.W,
                       E.ci, E.co, E.v, E.ww, F.dinv, r, nullptr, yn, st, u, sv);
  else
    hipLaunchKernelGGL((k_tri_sweep_ell<false, true, false, true>), grid, blk, 0, s, F.n, E.ld, E.W,
                       E.ci, E.co, E.v, E.ww, F.dinv, r, nullptr, yn, st, u, sv);
}

// ---------------------------------------------------------------------------
// Iterative application of one ILU factor: the reference's GPU ILU apply
// (DeviceMatrixCSR::Lsolve_iter / Usolve_iter,
// hip/HostMatrix/DeviceMatrix/DeviceMatrixCSR.cpp:281-304, kernel pairs
// LMultCSR + lsolveiterCSR / UMultCSR + usolveiterCSR :14-58), one fused
// kernel per sweep:  y_new = D (b - T_off y_old), D = I for L, 1/u_ii for U;
// y_old == null is the first sweep from y = 0 (y_new = D b).  Off-diagonal
// sums in ascending column order as the reference.  Thread per row.
// ---------------------------------------------------------------------------
template <bool UPPER>
__global__ __launch_bounds__(XS_BLOCK) void k_tri_sweep(int n, const int *__restrict__ rp,
                                                        const int *__restrict__ ci,
                                                        const double *__restrict__ lu,
                                                        const int *__restrict__ dpos,
                                                        const double *__restrict__ dinv,
                                                        const double *__restrict__ b,
                                                        const double *__restrict__ yo,
                                                        double *__restrict__ yn,
                                                        const xs_state *st) {
  if (st != nullptr && st->done) return;
  const int i = blockIdx.x * XS_BLOCK + threadIdx.x;
  if (i >= n) return;
  double r = 0.0;
  if (yo) {
    const int e0 = UPPER ? dpos[i] + 1 : rp[i], e1 = UPPER ? rp[i + 1] : dpos[i];
    for (int e = e0; e < e1; ++e) r += lu[e] * yo[ci[e]];
  }
  yn[i] = UPPER ? dinv[i] * (b[i] - r) : b[i] - r;
}

void xs_tri_sweep(hipStream_t s, const xs_ilu &F, bool upper, const double *b, const double *yo,
                  double *yn, const xs_state *st) {
  XS_GUARD(F.rp && F.ci && F.lu && F.dpos && b && yn && (!upper || F.dinv));
  const dim3 grid((F.n + XS_BLOCK - 1) / XS_BLOCK), blk(XS_BLOCK);
  if (upper)
    hipLaunchKernelGGL(k_tri_sweep<true>, grid, blk, 0, s, F.n, F.rp, F.ci, F.lu, F.dpos, F.dinv, b, yo, yn, st);
  else
    hipLaunchKernelGGL(k_tri_sweep<false>, grid, blk, 0, s, F.n, F.rp, F.ci, F.lu, F.dpos, F.dinv, b, yo, yn, st);
}

// values of a transposed pattern: dst[perm[k]] = src[k]
__global__ void k_permute_vals(int nnz, const int *__restrict__ perm, const double *__restrict__ src,
                               double *__restrict__ dst) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) dst[perm[k]] = src[k];
}
void xs_permute_vals(hipStream_t s, int nnz, const int *perm, const double *src, double *dst) {
  XS_GUARD(perm && src && dst);
  hipLaunchKernelGGL(k_permute_vals, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, perm, src, dst);
}

// dL/dA_ij = -lambda_i x_j on the pattern of A (thread per row)
__global__ void k_grad_vals(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                            const double *__restrict__ lam, const double *__restrict__ x,
                            double *__restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double li = lam[i];
  for (int e = rp[i]; e < rp[i + 1]; ++e) g[e] = -li * x[ci[e]];
}
void xs_grad_vals(hipStream_t s, const xs_csr &A, const double *lam, const double *x, double *g) {
  XS_GUARD(A.rp && A.ci && lam && x && g);
  hipLaunchKernelGGL(k_grad_vals, dim3((A.n + 255) / 256), dim3(256), 0, s, A.n, A.rp, A.ci, lam, x, g);
}
