// xs_parilut.hip -- ParILUT on the device (threshold-based incomplete LU by
// fixed-point sweeps over a changing pattern).
//
// Structure of the reference's parilut_clean (hip/Precond/parilut.cpp:621-1206),
// redesigned for gfx950: all factors live in one combined CSR F (row i: strict
// lower part = L without its unit diagonal, the pivot, the strict upper part
// = U), so no L/U splits, no CSC transposes and no SpGEMM library calls.
// Per outer step:
//   1. candidate pattern C_i = A_i u F_i u (union over k in L_i of U_k incl.
//      the pivot): one wavefront per row inserts the columns into an LDS hash
//      set, counts (pass 1), then writes them sorted (pass 2);
//   2. embed: existing entries keep their values, new entries take
//      (a_ij - S_ij)/u_jj below the diagonal, a_ij - S_ij on/above it, where
//      S_ij = sum_{k in L_i, k < min(i,j)} l_ik u_kj (embedA_intoB_L2/U2);
//   3. one Jacobi sweep over C (LU_sweep);
//   4. per factor, drop entries below the value of rank (count - keep) in
//      ascending |v| (8-bit radix select on the IEEE bit patterns, all on the
//      device) and below droptol * ||a_i||_2; the pivot is always kept
//      (find_kth_element + thresold_remove2);
//   5. one Jacobi sweep over the kept pattern.
// The sweeps read the symmetrically scaled values S A S, S = |diag A|^{-1/2};
// the caller scales A's values in and the factors back out.
#include <climits>

#include "xs_internal.h"
#include "xs_parilut.h"

#define PT_HASH 2048          // LDS hash slots per wavefront
#define PT_WAVES 4            // wavefronts per workgroup
#define PT_EMPTY (-1)

__device__ __forceinline__ int pt_wave_sum_i(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, XS_WAVE);
  return v;
}

__device__ __forceinline__ void pt_insert(int *tab, int c, int *ovf) {
  unsigned h = ((unsigned)c * 2654435761u) & (PT_HASH - 1);
  for (int probe = 0; probe < PT_HASH; ++probe) {
    const int old = atomicCAS(&tab[h], PT_EMPTY, c);
    if (old == PT_EMPTY || old == c) return;
    h = (h + 1) & (PT_HASH - 1);
  }
  atomicOr(ovf, 1);
}

// Pass 1 (fill == false): cnt[i] = |C_i|.  Pass 2 (fill == true): write C_i
// sorted into cci[crp[i]..] and the row index into crow.
template <bool FILL>
__global__ __launch_bounds__(XS_WAVE *PT_WAVES) void k_pt_symbolic(
    int n, const int *__restrict__ arp, const int *__restrict__ aci, const int *__restrict__ frp,
    const int *__restrict__ fci, const int *__restrict__ fdpos, int *__restrict__ cnt,
    const int *__restrict__ crp, int *__restrict__ cci, int *__restrict__ crow, int *ovf) {
  __shared__ int tab_all[PT_WAVES][PT_HASH];
  const int lane = threadIdx.x & (XS_WAVE - 1), wv = threadIdx.x / XS_WAVE;
  const int i = blockIdx.x * PT_WAVES + wv;
  int *tab = tab_all[wv];
  int *keys = tab;   // compacted in place (see below)
  for (int q = lane; q < PT_HASH; q += XS_WAVE) tab[q] = PT_EMPTY;
  __syncthreads();
  if (i < n) {
    for (int e = arp[i] + lane; e < arp[i + 1]; e += XS_WAVE) pt_insert(tab, aci[e], ovf);
    for (int e = frp[i] + lane; e < frp[i + 1]; e += XS_WAVE) pt_insert(tab, fci[e], ovf);
    // union of the upper parts (incl. pivot) of the rows k in L_i (a set:
    // order is free).  Up to 64 L entries at once: lane t holds the bounds of
    // row k_t; then the first 64 entries of four such rows are loaded
    // together, the next four while the current ones are inserted.
    constexpr int D = 4;
    const int lb = frp[i], le = fdpos[i];
    for (int c0 = lb; c0 < le; c0 += XS_WAVE) {
      const int nl = min(XS_WAVE, le - c0);
      int ub = 0, ue = 0;
      if (lane < nl) {
        const int k = fci[c0 + lane];
        ub = fdpos[k];
        ue = frp[k + 1];
      }
      int qb[D], qe[D], jj[D], nqb[D], nqe[D], njj[D];
      auto load_group = [&](int g0, int (&gb)[D], int (&ge)[D], int (&gj)[D]) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const bool ok = g0 + d < nl;
          gb[d] = ok ? __shfl(ub, g0 + d, XS_WAVE) : 0;
          ge[d] = ok ? __shfl(ue, g0 + d, XS_WAVE) : 0;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const int q = gb[d] + lane;
          gj[d] = q < ge[d] ? fci[q] : -1;
        }
      };
      load_group(0, qb, qe, jj);
      for (int g0 = 0; g0 < nl; g0 += D) {
        if (g0 + D < nl) load_group(g0 + D, nqb, nqe, njj);
#pragma unroll
        for (int d = 0; d < D; ++d) {
          if (jj[d] >= 0) pt_insert(tab, jj[d], ovf);
          for (int q = qb[d] + XS_WAVE + lane; q < qe[d]; q += XS_WAVE) pt_insert(tab, fci[q], ovf);
        }
#pragma unroll
        for (int d = 0; d < D; ++d) { qb[d] = nqb[d]; qe[d] = nqe[d]; jj[d] = njj[d]; }
      }
    }
  }
  __syncthreads();
  if (i < n) {
    // compact the keys (slot order) to the front of the same table, then
    // rank-sort them.  In place is safe: a chunk's 64 slots are all read
    // (one wave-wide load) before its keys are stored, and they go to
    // positions base + before <= q0 + lane, all read already.
    int base = 0;
    for (int q0 = 0; q0 < PT_HASH; q0 += XS_WAVE) {
      const int c = tab[q0 + lane];
      const bool has = c != PT_EMPTY;
      const unsigned long long m = __ballot(has);
      const int before = __popcll(m & ((1ull << lane) - 1ull));
      if (has) keys[base + before] = c;
      base += __popcll(m);
    }
    const int u = base;
    if (!FILL) {
      if (lane == 0) cnt[i] = u;
    } else {
      const int off = crp[i];
      for (int a = lane; a < u; a += XS_WAVE) {
        const int key = keys[a];
        int rank = 0;
        for (int b = 0; b < u; ++b) rank += keys[b] < key;
        cci[off + rank] = key;
        crow[off + rank] = i;
      }
    }
  }
}

// value of F at (row, col) or 0 (binary search in [a, b))
__device__ __forceinline__ int pt_find(const int *__restrict__ ci, int a, int b, int col) {
  while (a < b) {
    const int m = (a + b) >> 1;
    if (ci[m] < col) a = m + 1;
    else b = m;
  }
  return a;
}

__device__ __forceinline__ double pt_a(const int *__restrict__ arp, const int *__restrict__ aci,
                                       const double *__restrict__ av,
                                       const double *__restrict__ sc, int i, int j) {
  const int a = arp[i], b = arp[i + 1];
  const int p = pt_find(aci, a, b, j);
  return (p < b && aci[p] == j) ? av[p] * sc[i] * sc[j] : 0.0;
}

// S_ij = sum_{k in L_i, k < min(i,j)} l_ik u_kj over factor (rp, ci, v, dpos)
__device__ __forceinline__ double pt_partial(const int *__restrict__ rp, const int *__restrict__ ci,
                                             const double *__restrict__ v,
                                             const int *__restrict__ dpos, int i, int j) {
  const int m = i < j ? i : j;
  double s = 0.0;
  for (int e = rp[i]; e < dpos[i]; ++e) {
    const int k = ci[e];
    if (k >= m) break;
    const int a = dpos[k], b = rp[k + 1];
    const int p = pt_find(ci, a, b, j);
    if (p < b && ci[p] == j) s += v[e] * v[p];
  }
  return s;
}

__device__ __forceinline__ double pt_clamp(double d) {
  if (fabs(d) < 1e-10) d = d >= 0.0 ? 1e-10 : -1e-10;
  return d;
}

// pivot positions of a sorted CSR pattern
__global__ void k_pt_dpos(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                          int *__restrict__ dpos, int *bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = rp[i], b = rp[i + 1];
  const int p = pt_find(ci, a, b, i);
  if (p < b && ci[p] == i) dpos[i] = p;
  else { dpos[i] = a; atomicOr(bad, 1); }
}

// ---------------------------------------------------------------------------
// Row-wise forms of the embed and the sweep: one wave per pattern row, the
// row's columns in LDS, and the partial sums S_ij = sum_{k in L_i, k <
// min(i,j)} l_ik u_kj accumulated in LDS by walking, in ascending k, the
// strictly upper part of F's row k for every k of row i's L part -- each U
// row read once (coalesced) per L entry instead of a binary search per
// product.  Contributions reach S_ij in ascending k, the order of
// pt_partial, so the results are identical.  Rows longer than PT_MAXR fall
// back to the per-entry form.
// ---------------------------------------------------------------------------
#define PT_ROWW 4
#define PT_MAXR 256
#define PT_RTAB 512           // slots of a row's column table (load <= 1/2)

// Column table of a pattern row held in LDS: slot t + 1 of cols[t] at the
// multiplicative hash of cols[t], linear probing, 0 = empty.  A lookup takes
// one or two probes where a binary search over the row takes log2(len) + 1
// dependent LDS reads.
__device__ __forceinline__ unsigned pt_rhash(int c) {
  return ((unsigned)c * 2654435761u) >> (32 - 9);
}
static_assert(PT_RTAB == 512 && 2 * PT_MAXR <= PT_RTAB, "pt_rhash yields 9 bits");

__device__ __forceinline__ void pt_row_table(const int *cols, int len, int *tab, int lane) {
  for (int q = lane; q < PT_RTAB; q += XS_WAVE) tab[q] = 0;
  for (int t = lane; t < len; t += XS_WAVE) {
    unsigned h = pt_rhash(cols[t]);
    // the table has room for every column (len <= PT_MAXR = PT_RTAB / 2)
    while (atomicCAS(&tab[h], 0, t + 1) != 0) h = (h + 1) & (PT_RTAB - 1);
  }
}

// slot of column j in the row, or -1
__device__ __forceinline__ int pt_row_slot(const int *cols, const int *tab, int j) {
  unsigned h = pt_rhash(j);
  for (;;) {
    const int t = tab[h];
    if (t == 0) return -1;
    if (cols[t - 1] == j) return t - 1;
    h = (h + 1) & (PT_RTAB - 1);
  }
}

__device__ __forceinline__ void pt_row_partial(int i, const int *__restrict__ frp,
                                               const int *__restrict__ fci,
                                               const double *__restrict__ fv,
                                               const int *__restrict__ fdpos,
                                               const int *cols, int *tab, double *S, int len,
                                               int lane) {
  constexpr int D = 4;   // U rows whose first 64 entries are loaded together
  for (int t = lane; t < len; t += XS_WAVE) S[t] = 0.0;
  pt_row_table(cols, len, tab, lane);
  const int lb = frp[i], le = fdpos[i];
  for (int c0 = lb; c0 < le; c0 += XS_WAVE) {
    // up to 64 L entries of the row at once: lane t holds entry c0 + t, its
    // value and the bounds of the strict upper part of row k = its column
    const int cnt = min(XS_WAVE, le - c0);
    int ub = 0, ue = 0;
    double lv = 0.0;
    if (lane < cnt) {
      const int k = fci[c0 + lane];
      lv = fv[c0 + lane];
      ub = fdpos[k] + 1;
      ue = frp[k + 1];
    }
    // groups of D rows; the next group's first 64 entries are loaded while
    // the current group's products are added
    int qb[D], qe[D], jj[D], nqb[D], nqe[D], njj[D];
    double lk[D], uv[D], nlk[D], nuv[D];
    auto load_group = [&](int g0, int (&gb)[D], int (&ge)[D], double (&gl)[D], int (&gj)[D],
                          double (&gu)[D]) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const bool ok = g0 + d < cnt;
        gb[d] = ok ? __shfl(ub, g0 + d, XS_WAVE) : 0;
        ge[d] = ok ? __shfl(ue, g0 + d, XS_WAVE) : 0;
        gl[d] = ok ? __shfl(lv, g0 + d, XS_WAVE) : 0.0;
      }
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const int q = gb[d] + lane;
        gj[d] = q < ge[d] ? fci[q] : -1;
        gu[d] = q < ge[d] ? fv[q] : 0.0;
      }
    };
    load_group(0, qb, qe, lk, jj, uv);
    for (int g0 = 0; g0 < cnt; g0 += D) {
      if (g0 + D < cnt) load_group(g0 + D, nqb, nqe, nlk, njj, nuv);
      // updates strictly in ascending k: all of row d (first chunk from
      // registers, the rest from memory) before row d + 1
#pragma unroll
      for (int d = 0; d < D; ++d) {
        for (int q0 = qb[d]; q0 < qe[d]; q0 += XS_WAVE) {
          int j;
          double u;
          if (q0 == qb[d]) { j = jj[d]; u = uv[d]; }
          else {
            const int q = q0 + lane;
            j = q < qe[d] ? fci[q] : -1;
            u = q < qe[d] ? fv[q] : 0.0;
          }
          if (j < 0) continue;
          const int t = pt_row_slot(cols, tab, j);
          if (t >= 0) S[t] += lk[d] * u;
        }
      }
#pragma unroll
      for (int d = 0; d < D; ++d) {
        qb[d] = nqb[d]; qe[d] = nqe[d]; jj[d] = njj[d]; lk[d] = nlk[d]; uv[d] = nuv[d];
      }
    }
  }
}

// a_ij of the scaled matrix on the row's slots (0 where A has no entry):
// the value pt_a returns, without a search per entry
__device__ __forceinline__ void pt_row_scatter_a(int i, const int *__restrict__ arp,
                                                 const int *__restrict__ aci,
                                                 const double *__restrict__ av,
                                                 const double *__restrict__ sc, const int *cols,
                                                 const int *tab, double *Av, int len, int lane) {
  for (int t = lane; t < len; t += XS_WAVE) Av[t] = 0.0;
  for (int e = arp[i] + lane; e < arp[i + 1]; e += XS_WAVE) {
    const int j = aci[e];
    const int t = pt_row_slot(cols, tab, j);
    if (t >= 0) Av[t] = av[e] * sc[i] * sc[j];
  }
}

__global__ __launch_bounds__(XS_WAVE * PT_ROWW) void k_pt_sweep_rows(
    int n, const int *__restrict__ rp, const int *__restrict__ ci, const int *__restrict__ dpos,
    const double *__restrict__ vin, double *__restrict__ vout, const int *__restrict__ arp,
    const int *__restrict__ aci, const double *__restrict__ av, const double *__restrict__ sc) {
  __shared__ int cols_s[PT_ROWW][PT_MAXR];
  __shared__ int tab_s[PT_ROWW][PT_RTAB];
  __shared__ double S_s[PT_ROWW][PT_MAXR];
  __shared__ double A_s[PT_ROWW][PT_MAXR];
  const int w = threadIdx.x / XS_WAVE, lane = threadIdx.x & (XS_WAVE - 1);
  const int i = blockIdx.x * PT_ROWW + w;
  if (i >= n) return;   // wave-uniform, no workgroup barrier below
  const int a = rp[i], len = rp[i + 1] - a;
  int *cols = cols_s[w];
  double *S = S_s[w], *Av = A_s[w];
  const bool rows = len <= PT_MAXR;
  if (rows) {
    for (int t = lane; t < len; t += XS_WAVE) cols[t] = ci[a + t];
    pt_row_partial(i, rp, ci, vin, dpos, cols, tab_s[w], S, len, lane);
    pt_row_scatter_a(i, arp, aci, av, sc, cols, tab_s[w], Av, len, lane);
  }
  for (int t = lane; t < len; t += XS_WAVE) {
    const int e = a + t, j = ci[e];
    const double part = rows ? S[t] : pt_partial(rp, ci, vin, dpos, i, j);
    const double s = (rows ? Av[t] : pt_a(arp, aci, av, sc, i, j)) - part;
    double r;
    if (i > j) r = s / pt_clamp(vin[dpos[j]]);
    else r = (i == j) ? pt_clamp(s) : s;
    vout[e] = isfinite(r) ? r : vin[e];
  }
}

__global__ __launch_bounds__(XS_WAVE * PT_ROWW) void k_pt_embed_rows(
    int n, const int *__restrict__ crp, const int *__restrict__ cci, const int *__restrict__ arp,
    const int *__restrict__ aci, const double *__restrict__ av, const double *__restrict__ sc,
    const int *__restrict__ frp, const int *__restrict__ fci, const double *__restrict__ fv,
    const int *__restrict__ fdpos, double *__restrict__ cv) {
  __shared__ int cols_s[PT_ROWW][PT_MAXR];
  __shared__ int tab_s[PT_ROWW][PT_RTAB];
  __shared__ double S_s[PT_ROWW][PT_MAXR];
  __shared__ double A_s[PT_ROWW][PT_MAXR];
  __shared__ int F_s[PT_ROWW][PT_MAXR];
  const int w = threadIdx.x / XS_WAVE, lane = threadIdx.x & (XS_WAVE - 1);
  const int i = blockIdx.x * PT_ROWW + w;
  if (i >= n) return;
  const int a = crp[i], len = crp[i + 1] - a;
  int *cols = cols_s[w], *Fp = F_s[w];
  double *S = S_s[w], *Av = A_s[w];
  const bool rows = len <= PT_MAXR;
  const int fa = frp[i], fb = frp[i + 1];
  if (rows) {
    for (int t = lane; t < len; t += XS_WAVE) cols[t] = cci[a + t];
    pt_row_partial(i, frp, fci, fv, fdpos, cols, tab_s[w], S, len, lane);
    pt_row_scatter_a(i, arp, aci, av, sc, cols, tab_s[w], Av, len, lane);
    // positions of the old factor's entries of row i on the row's slots
    for (int t = lane; t < len; t += XS_WAVE) Fp[t] = -1;
    for (int e = fa + lane; e < fb; e += XS_WAVE) {
      const int t = pt_row_slot(cols, tab_s[w], fci[e]);
      if (t >= 0) Fp[t] = e;
    }
  }
  for (int t = lane; t < len; t += XS_WAVE) {
    const int j = cci[a + t];
    int p;
    if (rows) {
      p = Fp[t];
    } else {
      p = pt_find(fci, fa, fb, j);
      if (!(p < fb && fci[p] == j)) p = -1;
    }
    double val;
    if (p >= 0) {
      val = fv[p];
    } else {
      const double part = rows ? S[t] : pt_partial(frp, fci, fv, fdpos, i, j);
      const double r = (rows ? Av[t] : pt_a(arp, aci, av, sc, i, j)) - part;
      val = (i > j) ? r / pt_clamp(fv[fdpos[j]]) : r;
    }
    cv[a + t] = isfinite(val) ? val : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Radix select of both thresholds at once (find_kth_element,
// hip/Precond/par_ilut_hip/select_kth.h:501-634, applied per factor as in
// hip/Precond/parilut.cpp:1008-1019): the value of rank k (0-based) in
// ascending magnitude over L (slot 0: its strict entries and its n unit
// diagonals, value 1) and over U (slot 1: its entries with the pivots).  The
// magnitudes are those of the unscaled factors (the sweeps run on S A S):
// |l_ij| = |l~_ij (s_j / s_i)|, |u_ij| = |u~_ij / (s_i s_j)|, the expressions
// the compaction compares against the thresholds.  |v| bit patterns are
// ordered like the values; passes of 8-bit digits from the most significant.
// Once the bin of rank k holds at most PT_SEL_CAP entries, the next pass
// collects them and the rank is found among them in one block; later passes
// return at once.  State per slot p: sel[4p + 0] prefix bits, + 1 remaining
// rank, + 2 mask of fixed bits, + 3 mode (0 histogram, 1 collect, 2 done).
// ---------------------------------------------------------------------------
#define PT_SEL_CAP 1024
static_assert(XS_BLOCK == 256, "k_pt_hist clears its 512 bins with two stores per thread");

// magnitude of entry (i, j) of the unscaled factor from its scaled value v
__device__ __forceinline__ double pt_mag(int i, int j, double v, const double *__restrict__ sc) {
  return j < i ? fabs(v * (sc[j] / sc[i])) : fabs(v / (sc[i] * sc[j]));
}

__global__ void k_pt_sel_init(const unsigned long long *__restrict__ cnt2, long long keepL,
                              long long keepU, unsigned long long *sel, unsigned *hist,
                              unsigned *lcnt, double *thr) {
  for (int q = threadIdx.x; q < 512; q += blockDim.x) hist[q] = 0;
  if (threadIdx.x < 2) {
    const int p = threadIdx.x;
    const long long tot = (long long)cnt2[p], keep = p ? keepU : keepL;
    sel[4 * p + 0] = 0;
    sel[4 * p + 1] = tot > keep ? (unsigned long long)(tot - keep) : 0ull;
    sel[4 * p + 2] = 0;
    sel[4 * p + 3] = tot > keep ? 0ull : 2ull;
    lcnt[p] = 0;
    thr[p] = 0.0;   // nothing dropped unless the select finishes
  }
}

__global__ __launch_bounds__(XS_BLOCK) void k_pt_hist(int nnz, const int *__restrict__ row,
                                                      const int *__restrict__ ci,
                                                      const double *__restrict__ v,
                                                      const double *__restrict__ sc, int shift,
                                                      const unsigned long long *sel,
                                                      unsigned *__restrict__ hist,
                                                      unsigned *lcnt,
                                                      unsigned long long *__restrict__ list) {
  const unsigned long long m0 = sel[3], m1 = sel[7];
  if (m0 == 2 && m1 == 2) return;
  __shared__ unsigned h[512];
  h[threadIdx.x] = 0;
  h[threadIdx.x + XS_BLOCK] = 0;
  __syncthreads();
  const unsigned long long pre[2] = {sel[0], sel[4]}, msk[2] = {sel[2], sel[6]};
  const unsigned long long mode[2] = {m0, m1};
  auto consider = [&](int p, double mag) {
    if (mode[p] == 2) return;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(mag);
    if ((bits & msk[p]) != pre[p]) return;
    if (mode[p] == 1) {
      const unsigned q = atomicAdd(&lcnt[p], 1u);
      if (q < PT_SEL_CAP) list[p * PT_SEL_CAP + q] = bits;
    } else {
      atomicAdd(&h[p * 256 + ((bits >> shift) & 0xFF)], 1u);
    }
  };
  for (int e = blockIdx.x * XS_BLOCK + threadIdx.x; e < nnz; e += gridDim.x * XS_BLOCK) {
    const int i = row[e], j = ci[e];
    const double mag = pt_mag(i, j, v[e], sc);
    if (j == i) consider(0, 1.0);    // L's unit diagonal
    consider(j < i ? 0 : 1, mag);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 512; q += XS_BLOCK)
    if (h[q]) atomicAdd(&hist[q], h[q]);
}

// one block of 256 threads per slot: pick the digit of rank k (or the rank
// among the collected entries) and advance the state
__global__ __launch_bounds__(256) void k_pt_pick(int shift, unsigned long long *sel,
                                                 unsigned *hist, unsigned *lcnt,
                                                 const unsigned long long *__restrict__ list,
                                                 double *thr) {
  const int p = blockIdx.x, t = threadIdx.x;
  unsigned long long *st = sel + 4 * p;
  const unsigned long long mode = st[3];
  if (mode == 2) return;
  __shared__ unsigned long long vals[PT_SEL_CAP];
  __shared__ unsigned cum[256];
  const unsigned long long k = st[1];
  __syncthreads();
  if (mode == 1) {
    const int c = (int)min(lcnt[p], (unsigned)PT_SEL_CAP);
    for (int q = t; q < c; q += 256) vals[q] = list[p * PT_SEL_CAP + q];
    __syncthreads();
    for (int q = t; q < c; q += 256) {
      const unsigned long long x = vals[q];
      unsigned long long less = 0, eq = 0;
      for (int r = 0; r < c; ++r) {
        less += vals[r] < x;
        eq += vals[r] == x;
      }
      // every thread that holds the value of rank k writes that same value
      if (less <= k && k < less + eq) st[0] = x;
    }
    __syncthreads();
    if (t == 0) {
      st[2] = ~0ull;
      st[3] = 2;
      thr[p] = __longlong_as_double((long long)st[0]);
    }
    return;
  }
  // inclusive scan of the 256 bin counts
  const unsigned hc = hist[p * 256 + t];
  cum[t] = hc;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const unsigned add = t >= off ? cum[t - off] : 0u;
    __syncthreads();
    cum[t] += add;
    __syncthreads();
  }
  const unsigned long long below = cum[t] - hc;
  if (hc && below <= k && k < below + hc) {
    st[0] |= (unsigned long long)t << shift;
    st[1] = k - below;
    st[2] |= 0xFFull << shift;
    if (shift == 0) {
      st[3] = 2;
      thr[p] = __longlong_as_double((long long)st[0]);
    } else if (hc <= PT_SEL_CAP) {
      st[3] = 1;
      lcnt[p] = 0;
    }
  }
  hist[p * 256 + t] = 0;
}

// entries of L with its unit diagonal (col <= row) -> cnt2[0] and of U with
// the pivots (col >= row) -> cnt2[1]: the counts the selection ranks over
// (hip/Precond/parilut.cpp:35-63 counts each factor's diagonal)
__global__ __launch_bounds__(XS_BLOCK) void k_pt_count_parts(int nnz, const int *__restrict__ row,
                                                             const int *__restrict__ ci,
                                                             unsigned long long *cnt2) {
  unsigned long long a = 0, b = 0;
  for (int e = blockIdx.x * XS_BLOCK + threadIdx.x; e < nnz; e += gridDim.x * XS_BLOCK) {
    const int i = row[e], j = ci[e];
    a += j <= i;
    b += j >= i;
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, XS_WAVE);
    b += __shfl_xor(b, off, XS_WAVE);
  }
  if ((threadIdx.x & (XS_WAVE - 1)) == 0) {
    atomicAdd(&cnt2[0], a);
    atomicAdd(&cnt2[1], b);
  }
}

// keep predicate for the compaction
// (thresold_remove_nnz2, hip/Precond/parilut.cpp:303-323: |v| >= threshold
// survives, the diagonal always does; plus the drop tolerance on the same
// unscaled magnitude)
__device__ __forceinline__ bool pt_keep(int i, int j, double v, double tl, double tu,
                                        double dt, const double *__restrict__ sc) {
  if (i == j) return true;
  const double a = pt_mag(i, j, v, sc);
  if (a < dt) return false;
  return j < i ? a >= tl : a >= tu;
}

// per-row kept counts and the compaction: a wave per row, kept entries
// placed in column order by ballot
__global__ __launch_bounds__(XS_BLOCK) void k_pt_keep_count(
    int n, const int *__restrict__ crp, const int *__restrict__ cci,
    const double *__restrict__ cv, const double *thr, const double *__restrict__ dtol,
    const double *__restrict__ sc, int *__restrict__ cnt) {
  const int lane = threadIdx.x & (XS_WAVE - 1);
  const int i = blockIdx.x * (XS_BLOCK / XS_WAVE) + threadIdx.x / XS_WAVE;
  if (i >= n) return;
  const double tl = thr[0], tu = thr[1], dt = dtol[i];
  int c = 0;
  for (int e = crp[i] + lane; e < crp[i + 1]; e += XS_WAVE) c += pt_keep(i, cci[e], cv[e], tl, tu, dt, sc);
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, XS_WAVE);
  if (lane == 0) cnt[i] = c;
}

__global__ __launch_bounds__(XS_BLOCK) void k_pt_keep_write(
    int n, const int *__restrict__ crp, const int *__restrict__ cci,
    const double *__restrict__ cv, const double *thr, const double *__restrict__ dtol,
    const double *__restrict__ sc, const int *__restrict__ frp, int *__restrict__ fci,
    double *__restrict__ fv, int *__restrict__ frow) {
  const int lane = threadIdx.x & (XS_WAVE - 1);
  const int i = blockIdx.x * (XS_BLOCK / XS_WAVE) + threadIdx.x / XS_WAVE;
  if (i >= n) return;
  const double tl = thr[0], tu = thr[1], dt = dtol[i];
  const unsigned long long below = (1ull << lane) - 1ull;
  int w = frp[i];
  const int re = crp[i + 1];
  for (int e0 = crp[i]; e0 < re; e0 += XS_WAVE) {
    const int e = e0 + lane;
    int j = 0;
    double x = 0.0;
    bool keep = false;
    if (e < re) {
      j = cci[e];
      x = cv[e];
      keep = pt_keep(i, j, x, tl, tu, dt, sc);
    }
    const unsigned long long m = __ballot(keep);
    if (keep) {
      const int q = w + __popcll(m & below);
      fci[q] = j;
      fv[q] = x;
      frow[q] = i;
    }
    w += __popcll(m);
  }
}

// threshold value from the select state (0 when nothing is dropped)
__global__ void k_pt_thr(const unsigned long long *sel, int use, double *thr_slot) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *thr_slot = use ? __longlong_as_double((long long)sel[0]) : 0.0;
}

// droptol * ||a_i||_2 of the unscaled matrix (ilutp_final's tnorm,
// hip/Precond/ilut_diag_in_csr.cpp:67-68)
__global__ void k_pt_dtol(int n, const int *__restrict__ arp, const double *__restrict__ av,
                          double droptol, double *__restrict__ dtol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int e = arp[i]; e < arp[i + 1]; ++e) s += av[e] * av[e];
  dtol[i] = droptol * sqrt(s);
}

// ---------------------------------------------------------------------------
// exclusive scan of n ints into out[0..n] (out[n] = total); 3 phases
// ---------------------------------------------------------------------------
#define SCAN_B 1024
__global__ void k_scan_blocks(int n, const int *__restrict__ in, int *__restrict__ out,
                              int *__restrict__ bsum) {
  __shared__ int s[SCAN_B];
  const int g = blockIdx.x * SCAN_B + threadIdx.x;
  s[threadIdx.x] = g < n ? in[g] : 0;
  __syncthreads();
  for (int off = 1; off < SCAN_B; off <<= 1) {
    const int t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (g < n) out[g + 1] = s[threadIdx.x];
  if (threadIdx.x == SCAN_B - 1) bsum[blockIdx.x] = s[threadIdx.x];
}
// exclusive scan of the nb block sums in place, one workgroup of SCAN_B
// threads: each thread owns a contiguous chunk, the chunk totals are scanned
// in LDS, then every chunk is rewritten with its offset
__global__ __launch_bounds__(SCAN_B) void k_scan_sums(int nb, int *__restrict__ bsum) {
  __shared__ int s[SCAN_B];
  const int per = (nb + SCAN_B - 1) / SCAN_B;
  const int b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
  int tot = 0;
  for (int b = b0; b < b1; ++b) tot += bsum[b];
  s[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < SCAN_B; off <<= 1) {
    const int t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int acc = s[threadIdx.x] - tot;   // exclusive offset of this chunk
  for (int b = b0; b < b1; ++b) {
    const int t = bsum[b];
    bsum[b] = acc;
    acc += t;
  }
}
__global__ void k_scan_add(int n, int *__restrict__ out, const int *__restrict__ bsum) {
  const int g = blockIdx.x * SCAN_B + threadIdx.x;
  if (g < n) out[g + 1] += bsum[blockIdx.x];
  if (g == 0) out[0] = 0;
}

// ---------------------------------------------------------------------------
// scaling helpers
// ---------------------------------------------------------------------------
__global__ void k_pt_scale_vec(int n, const int *__restrict__ arp, const int *__restrict__ aci,
                               const double *__restrict__ av, double *__restrict__ sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int e = arp[i]; e < arp[i + 1]; ++e)
    if (aci[e] == i) d = av[e];
  sc[i] = d != 0.0 ? 1.0 / sqrt(fabs(d)) : 1.0;
}

// combined factor (values on a CSR pattern with row index array) scaled by
// l_ij *= s_j / s_i below the diagonal, u_ij /= s_i s_j on and above it
// (unscale = true), or the matrix values multiplied by s_i s_j (unscale = false).
__global__ void k_pt_rescale(int nnz, const int *__restrict__ row, const int *__restrict__ ci,
                             double *__restrict__ v, const double *__restrict__ sc, int unscale) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const int i = row[e], j = ci[e];
  if (!unscale) v[e] *= sc[i] * sc[j];
  else if (j < i) v[e] *= sc[j] / sc[i];
  else v[e] /= sc[i] * sc[j];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void xs_scan_excl(hipStream_t s, int n, const int *in, int *out, int *bsum) {
  const int nb = (n + SCAN_B - 1) / SCAN_B;
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(SCAN_B), 0, s, n, in, out, bsum);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_B), 0, s, nb, bsum);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(SCAN_B), 0, s, n, out, bsum);
}
int xs_scan_bsum_len(int n) { return (n + SCAN_B - 1) / SCAN_B + 1; }

void xs_pt_symbolic(hipStream_t s, bool fill, int n, const int *arp, const int *aci,
                    const int *frp, const int *fci, const int *fdpos, int *cnt, const int *crp,
                    int *cci, int *crow, int *ovf) {
  dim3 grid((n + PT_WAVES - 1) / PT_WAVES), blk(XS_WAVE * PT_WAVES);
  if (fill)
    hipLaunchKernelGGL(k_pt_symbolic<true>, grid, blk, 0, s, n, arp, aci, frp, fci, fdpos, cnt, crp, cci, crow, ovf);
  else
    hipLaunchKernelGGL(k_pt_symbolic<false>, grid, blk, 0, s, n, arp, aci, frp, fci, fdpos, cnt, crp, cci, crow, ovf);
}

void xs_pt_dpos(hipStream_t s, int n, const int *rp, const int *ci, int *dpos, int *bad) {
  hipLaunchKernelGGL(k_pt_dpos, dim3((n + 255) / 256), dim3(256), 0, s, n, rp, ci, dpos, bad);
}

void xs_pt_embed(hipStream_t s, int n, const int *crp, const int *cci, const int *arp,
                 const int *aci, const double *av, const double *sc, const int *frp,
                 const int *fci, const double *fv, const int *fdpos, double *cv) {
  hipLaunchKernelGGL(k_pt_embed_rows, dim3((n + PT_ROWW - 1) / PT_ROWW), dim3(XS_WAVE * PT_ROWW), 0,
                     s, n, crp, cci, arp, aci, av, sc, frp, fci, fv, fdpos, cv);
}

void xs_pt_sweep(hipStream_t s, int n, const int *rp, const int *ci, const int *dpos,
                 const double *vin, double *vout, const int *arp, const int *aci,
                 const double *av, const double *sc) {
  hipLaunchKernelGGL(k_pt_sweep_rows, dim3((n + PT_ROWW - 1) / PT_ROWW), dim3(XS_WAVE * PT_ROWW), 0,
                     s, n, rp, ci, dpos, vin, vout, arp, aci, av, sc);
}

void xs_pt_select(hipStream_t s, int nnz, const int *row, const int *ci, const double *v,
                  const double *sc, const unsigned long long *cnt2, long long keepL, long long keepU,
                  unsigned long long *sel, unsigned *hist, unsigned *lcnt,
                  unsigned long long *list, double *thr) {
  hipLaunchKernelGGL(k_pt_sel_init, dim3(1), dim3(256), 0, s, cnt2, keepL, keepU, sel, hist, lcnt,
                     thr);
  int g = (nnz + XS_BLOCK * 16 - 1) / (XS_BLOCK * 16);
  if (g > 2048) g = 2048;
  if (g < 1) g = 1;
  for (int shift = 56; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(k_pt_hist, dim3(g), dim3(XS_BLOCK), 0, s, nnz, row, ci, v, sc, shift, sel,
                       hist, lcnt, list);
    hipLaunchKernelGGL(k_pt_pick, dim3(2), dim3(256), 0, s, shift, sel, hist, lcnt, list, thr);
  }
}

int xs_pt_select_words() { return 8; }
int xs_pt_list_len() { return 2 * PT_SEL_CAP; }

void xs_pt_count_parts(hipStream_t s, int nnz, const int *row, const int *ci,
                       unsigned long long *cnt2) {
  int g = (nnz + XS_BLOCK * 8 - 1) / (XS_BLOCK * 8);
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  (void)hipMemsetAsync(cnt2, 0, 2 * sizeof(unsigned long long), s);
  hipLaunchKernelGGL(k_pt_count_parts, dim3(g), dim3(XS_BLOCK), 0, s, nnz, row, ci, cnt2);
}

void xs_pt_dtol(hipStream_t s, int n, const int *arp, const double *av, double droptol,
                double *dtol) {
  hipLaunchKernelGGL(k_pt_dtol, dim3((n + 255) / 256), dim3(256), 0, s, n, arp, av, droptol, dtol);
}

void xs_pt_keep(hipStream_t s, bool write, int n, const int *crp, const int *cci,
                const double *cv, const double *thr, const double *dtol, const double *sc,
                int *cnt, const int *frp, int *fci, double *fv, int *frow) {
  const int rows = XS_BLOCK / XS_WAVE, g = (n + rows - 1) / rows;
  if (!write)
    hipLaunchKernelGGL(k_pt_keep_count, dim3(g), dim3(XS_BLOCK), 0, s, n, crp, cci, cv, thr, dtol,
                       sc, cnt);
  else
    hipLaunchKernelGGL(k_pt_keep_write, dim3(g), dim3(XS_BLOCK), 0, s, n, crp, cci, cv, thr, dtol,
                       sc, frp, fci, fv, frow);
}

void xs_pt_scale_vec(hipStream_t s, int n, const int *arp, const int *aci, const double *av,
                     double *sc) {
  hipLaunchKernelGGL(k_pt_scale_vec, dim3((n + 255) / 256), dim3(256), 0, s, n, arp, aci, av, sc);
}

void xs_pt_rescale(hipStream_t s, int nnz, const int *row, const int *ci, double *v,
                   const double *sc, int unscale) {
  hipLaunchKernelGGL(k_pt_rescale, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, row, ci, v, sc,
                     unscale);
}

__global__ void k_pt_rowidx(int n, const int *__restrict__ rp, int *__restrict__ row) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int e = rp[i]; e < rp[i + 1]; ++e) row[e] = i;
}
void xs_pt_rowidx(hipStream_t s, int n, const int *rp, int *row) {
  hipLaunchKernelGGL(k_pt_rowidx, dim3((n + 255) / 256), dim3(256), 0, s, n, rp, row);
}
