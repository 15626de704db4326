// xs_split.hip -- pattern analysis on the device: the ILU split's upper
// triangle by columns (CSC) and CSR transposes, from one stable sort of
// (column, position) pairs by column.
//
// The reference splits A into L (CSR) and U (CSC) on the GPU with a count
// pass and hipsparseDcsr2csc (hip/Precond/par_ilu_hip/Asplit_gpu.cpp:4-49,
// 62-161).  Here the combined factor keeps A's CSR pattern, and only the
// upper triangle's column view (ucp / uri / upos: column j lists the rows i
// <= j with an entry (i, j), ascending, and the entry's CSR position) is
// built, by a least-significant-digit radix sort written for this library:
// 4-bit digits, each thread ranks a chunk of consecutive pairs, so the sort
// is stable and every kernel is a plain data-parallel pass (no look-back,
// no waits between workgroups).  Pairs enter in row-major order, so the
// stable sort by column leaves each column's rows ascending -- the order the
// host loop of CSR -> CSC gives.
#include "xs_device.h"
#include "xs_internal.h"
#include "xs_parilut.h"

#define RS_CHUNK 32    // pairs ranked by one thread
#define RS_BLOCK 256
#define RS_DIGITS 16   // 4-bit digits

static size_t rs_threads(int m) { return ((size_t)m + RS_CHUNK - 1) / RS_CHUNK; }

// digit counts of each thread's chunk: hist[d * T + t]
__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(int m, int T, int shift,
                                                      const unsigned *__restrict__ key,
                                                      int *__restrict__ hist) {
  __shared__ int h[RS_DIGITS][RS_BLOCK];
  const int t = blockIdx.x * RS_BLOCK + threadIdx.x;
  if (t >= T) return;   // no barrier below: each thread owns column threadIdx.x of h
#pragma unroll
  for (int d = 0; d < RS_DIGITS; ++d) h[d][threadIdx.x] = 0;
  const size_t a = (size_t)t * RS_CHUNK;
  const size_t b = a + RS_CHUNK < (size_t)m ? a + RS_CHUNK : (size_t)m;
  for (size_t q = a; q < b; ++q) h[(key[q] >> shift) & (RS_DIGITS - 1)][threadIdx.x]++;
#pragma unroll
  for (int d = 0; d < RS_DIGITS; ++d) hist[(size_t)d * T + t] = h[d][threadIdx.x];
}

// stable scatter: pair q of thread t goes to off[d * T + t] + (pairs of digit
// d before q in t's chunk)
__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(int m, int T, int shift,
                                                         const unsigned *__restrict__ key,
                                                         const int *__restrict__ val,
                                                         const int *__restrict__ off,
                                                         unsigned *__restrict__ key2,
                                                         int *__restrict__ val2) {
  __shared__ int h[RS_DIGITS][RS_BLOCK];
  const int t = blockIdx.x * RS_BLOCK + threadIdx.x;
  if (t >= T) return;
#pragma unroll
  for (int d = 0; d < RS_DIGITS; ++d) h[d][threadIdx.x] = off[(size_t)d * T + t];
  const size_t a = (size_t)t * RS_CHUNK;
  const size_t b = a + RS_CHUNK < (size_t)m ? a + RS_CHUNK : (size_t)m;
  for (size_t q = a; q < b; ++q) {
    const unsigned k = key[q];
    const int p = h[(k >> shift) & (RS_DIGITS - 1)][threadIdx.x]++;
    key2[p] = k;
    val2[p] = val[q];
  }
}

size_t xs_colsort_hist_len(int m) { return (size_t)RS_DIGITS * rs_threads(m) + 1; }

int xs_colsort(hipStream_t s, int m, int ncol, unsigned *key, int *val, unsigned *key2, int *val2,
               int *hist, int *off, int *bsum) {
  if (m <= 0) return 0;
  const int T = (int)rs_threads(m);
  const int H = RS_DIGITS * T;
  int bits = 1;
  while (bits < 32 && (1u << bits) < (unsigned)ncol) ++bits;
  const int grid = (T + RS_BLOCK - 1) / RS_BLOCK;
  int pass = 0;
  for (int shift = 0; shift < bits; shift += 4, ++pass) {
    unsigned *ki = pass & 1 ? key2 : key, *ko = pass & 1 ? key : key2;
    int *vi = pass & 1 ? val2 : val, *vo = pass & 1 ? val : val2;
    hipLaunchKernelGGL(k_rs_hist, dim3(grid), dim3(RS_BLOCK), 0, s, m, T, shift, ki, hist);
    xs_scan_excl(s, H, hist, off, bsum);
    hipLaunchKernelGGL(k_rs_scatter, dim3(grid), dim3(RS_BLOCK), 0, s, m, T, shift, ki, vi, off,
                       ko, vo);
  }
  return pass & 1;
}

// cp[j] = first position of a key >= j in the sorted keys, j = 0..ncol
// (every cp entry is written by exactly one thread: q's thread covers the
// columns after key[q - 1] up to key[q], thread m those after the last key)
__global__ void k_key_ptr(int m, int ncol, const unsigned *__restrict__ key, int *__restrict__ cp) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > m) return;
  const int lo = q == 0 ? 0 : (int)key[q - 1] + 1;
  const int hi = q == m ? ncol : min((int)key[q], ncol);
  for (int j = lo; j <= hi; ++j) cp[j] = q;
}
void xs_key_ptr(hipStream_t s, int m, int ncol, const unsigned *skey, int *cp) {
  hipLaunchKernelGGL(k_key_ptr, dim3(m / 256 + 1), dim3(256), 0, s, m, ncol, skey, cp);
}

// entries on and right of the diagonal: count per row (pivot positions dpos)
__global__ void k_upper_count(int n, const int *__restrict__ rp, const int *__restrict__ dpos,
                              int *__restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cnt[i] = rp[i + 1] - dpos[i];
}
void xs_upper_count(hipStream_t s, int n, const int *rp, const int *dpos, int *cnt) {
  hipLaunchKernelGGL(k_upper_count, dim3((n + 255) / 256), dim3(256), 0, s, n, rp, dpos, cnt);
}

// (column, CSR position) pairs of the upper entries in row-major order
__global__ void k_upper_pairs(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                              const int *__restrict__ dpos, const int *__restrict__ uoff,
                              unsigned *__restrict__ key, int *__restrict__ val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int q = uoff[i];
  for (int e = dpos[i]; e < rp[i + 1]; ++e, ++q) {
    key[q] = (unsigned)ci[e];
    val[q] = e;
  }
}
void xs_upper_pairs(hipStream_t s, int n, const int *rp, const int *ci, const int *dpos,
                    const int *uoff, unsigned *key, int *val) {
  hipLaunchKernelGGL(k_upper_pairs, dim3((n + 255) / 256), dim3(256), 0, s, n, rp, ci, dpos, uoff,
                     key, val);
}

// (column, position) pairs of every entry: the position is the entry itself
__global__ void k_all_pairs(int nnz, const int *__restrict__ ci, unsigned *__restrict__ key,
                            int *__restrict__ val) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nnz) {
    key[e] = (unsigned)ci[e];
    val[e] = e;
  }
}
void xs_all_pairs(hipStream_t s, int nnz, const int *ci, unsigned *key, int *val) {
  hipLaunchKernelGGL(k_all_pairs, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, ci, key, val);
}

// out[q] = row of the entry at sorted position q; perm[pos[q]] = q when perm
__global__ void k_sorted_rows(int m, const int *__restrict__ rowidx, const int *__restrict__ pos,
                              int *__restrict__ out, int *__restrict__ perm) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const int e = pos[q];
  out[q] = rowidx[e];
  if (perm) perm[e] = q;
}
void xs_sorted_rows(hipStream_t s, int m, const int *rowidx, const int *pos, int *out, int *perm) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_sorted_rows, dim3((m + 255) / 256), dim3(256), 0, s, m, rowidx, pos, out,
                     perm);
}

// pattern checks of a device CSR (row pointers already checked monotone on
// the host, so every row's range lies in [0, nnz)): *flag |= 1 for a column
// out of range or not strictly ascending within its row, |= 2 for a row
// without its diagonal
__global__ void k_check_pattern(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                int *flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int f = 0, prev = -1;
  bool d = false;
  for (int e = rp[i]; e < rp[i + 1]; ++e) {
    const int j = ci[e];
    if (j < 0 || j >= n || j <= prev) f |= 1;
    d |= j == i;
    prev = j;
  }
  if (!d) f |= 2;
  if (f) atomicOr(flag, f);
}
void xs_check_pattern(hipStream_t s, int n, const int *rp, const int *ci, int *flag) {
  hipLaunchKernelGGL(k_check_pattern, dim3((n + 255) / 256), dim3(256), 0, s, n, rp, ci, flag);
}
