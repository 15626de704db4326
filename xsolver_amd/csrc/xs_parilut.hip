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

#include "xs_device.h"
#include "xs_parilut.h"

#define PT_HASH 2048          // LDS hash slots per wavefront
#define PT_WAVES 4            // wavefronts per workgroup
#define PT_EMPTY (-1)

__device__ __forceinline__ int pt_wave_sum_i(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, XS_WAVE);
  return v;
}

// insert column c (when v) into a row's 2048-slot set; *ovf |= 1 when the
// set is full.  At most one CAS per lane; the probe loop runs only when some
// lane's first slot holds another column.
__device__ int pt_insert_from(int *tab, int c, unsigned h, int *ovf) {
  for (int probe = 1; probe < PT_HASH; ++probe) {
    const int old = atomicCAS(&tab[h], PT_EMPTY, c);
    if (old == PT_EMPTY || old == c) return 0;
    h = (h + 1) & (PT_HASH - 1);
  }
  atomicOr(ovf, 1);
  return 1;
}

__device__ __forceinline__ void pt_insert(int *tab, int c, bool v, int *ovf) {
  const unsigned h = ((unsigned)c * 2654435761u) & (PT_HASH - 1);
  // most inserts repeat a column already in its first slot: a plain read
  // settles those, the CAS is left to the others
  int old = tab[h];
  if (v && old != c) old = atomicCAS(&tab[h], PT_EMPTY, c);
  const bool more = v && old != PT_EMPTY && old != c;
  if (__any(more))
    if (more) pt_insert_from(tab, c, (h + 1) & (PT_HASH - 1), ovf);
}

// The products l_ik u_kj of row i's L entries [lb, le) of F (columns k,
// values l_ik) with the upper parts [fdpos[k] + off, frp[k + 1]) of the rows
// k, as one load-balanced stream: the parts of up to 64 L entries laid end to
// end (lane t holds entry t's part and the inclusive scan inc of the part
// lengths), 128 products per step whatever the part lengths, each lane
// taking positions p0 + 64 c + lane (c < PT_SPOS), two steps in flight.  The
// parts a step spans are found by ballot (the owner of position p is the
// number of lanes with inc <= p) and visited in a wave-uniform loop that
// hands each lane its part's bounds and l_ik by readlane -- no LDS traffic.
// fn(j, l_ik u_kj, first, valid) per lane and position, for every lane
// (valid false: no product there, j a column of the matrix all the same);
// value 0 without VALS; first marks a part's entry at fdpos[k] + off.
// Every lane runs every step.
__device__ __forceinline__ double pt_readlane(double v, int t) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, t);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), t);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

#define PT_SPOS 2   // stream positions per lane per step

// One block of the stream (up to 64 L entries) in two phases: begin() takes
// the block's per-lane inputs (or load_begin() loads them), scans the part
// lengths and issues the first step's loads; finish() runs the steps.  A
// kernel that already holds the first block's inputs begins it before its
// own LDS set-up, so the first loads are in flight meanwhile.
template <bool VALS>
struct pt_stream_t {
  static constexpr int P = PT_SPOS, STEP = P * XS_WAVE;
  struct step {
    int j[P];
    double l[P], u[P];
    bool f[P], v[P];     // first entry of its part; a product at this position
  };
  const int *__restrict__ fci;
  const double *__restrict__ fv;
  int lane, c0, inc, tot, st, base;
  double lv;
  step A, B;

  // lane t < cnt: L entry c0 + t, its value lv_ (l_ik) and its part [ub,
  // ub + len) of (fci, fv); lanes t >= cnt hold no part
  __device__ __forceinline__ void begin(int c0_, int cnt, int ub, int len, double lv_) {
    c0 = c0_;
    lv = lv_;
    if (lane >= cnt) len = 0;
    inc = len;
#pragma unroll
    for (int o = 1; o < XS_WAVE; o <<= 1) {
      const int y = __shfl_up(inc, o, XS_WAVE);
      if (lane >= o) inc += y;
    }
    tot = __builtin_amdgcn_readlane(inc, XS_WAVE - 1);
    st = inc - len;            // first stream position of this lane's part
    base = ub - st;            // q = base + p for the positions of this part
    issue(0, A);
  }

  // block c0_ of the L entries [c0_, le): loads clamped to the block's last
  // entry, so they issue unconditionally
  __device__ __forceinline__ void load_begin(int c0_, int le, int off, const int *__restrict__ frp,
                                             const int *__restrict__ fdpos) {
    const int cnt = min(XS_WAVE, le - c0_);
    const int ce = c0_ + min(lane, cnt - 1);
    const int k = fci[ce];
    const double l = VALS ? fv[ce] : 0.0;
    const int ub = fdpos[k] + off;
    begin(c0_, cnt, ub, frp[k + 1] - ub, l);
  }

  // locate the step's positions and issue their loads, for every lane and
  // whether or not the step holds products (a position past the end loads
  // entry c0, so no load sits behind a branch)
  __device__ __forceinline__ void issue(int p0, step &z) {
    const int pe = min(p0 + STEP, tot) - 1;
    // owners of the first and the last position of the step (uniform; no
    // iteration once p0 is past the end)
    const int o0 = __popcll(__ballot(inc <= p0));
    const int o1 = __popcll(__ballot(inc <= pe));
    int q[P], sp[P];
#pragma unroll
    for (int c = 0; c < P; ++c) {
      q[c] = c0;
      sp[c] = -1;
      z.l[c] = 0.0;
    }
    for (int t = o0; t <= o1; ++t) {
      // parts in ascending order: each lane keeps the last part starting at
      // or before its position (an empty part shares its start with the
      // next one, which overrides it)
      const int ts = __builtin_amdgcn_readlane(st, t);
      const int tb = __builtin_amdgcn_readlane(base, t);
      const double tl = VALS ? pt_readlane(lv, t) : 0.0;
#pragma unroll
      for (int c = 0; c < P; ++c) {
        const int p = p0 + c * XS_WAVE + lane;
        if (p >= ts) { q[c] = tb + p; z.l[c] = tl; sp[c] = ts; }
      }
    }
#pragma unroll
    for (int c = 0; c < P; ++c) {
      const int p = p0 + c * XS_WAVE + lane;
      z.v[c] = p < tot;
      const int qq = z.v[c] ? q[c] : c0;
      z.j[c] = fci[qq];
      z.u[c] = VALS ? fv[qq] : 0.0;
      // first: the part's first entry (the pivot u_kk when the parts start
      // at fdpos[k])
      z.f[c] = p == sp[c];
    }
  }

  template <typename Fn>
  __device__ __forceinline__ void run(const step &z, Fn &fn) {
#pragma unroll
    for (int c = 0; c < P; ++c) fn(z.j[c], z.l[c] * z.u[c], z.f[c], z.v[c]);
  }

  // two steps in flight, in two register sets used in turn (no copies of
  // registers whose loads are pending): the next step's loads are issued
  // before the current step's products are added
  template <typename Fn>
  __device__ __forceinline__ void finish(Fn &fn) {
    for (int p0 = 0; p0 < tot; p0 += 2 * STEP) {
      issue(p0 + STEP, B);
      run(A, fn);
      issue(p0 + 2 * STEP, A);
      run(B, fn);
    }
  }
};

template <bool VALS, typename Fn>
__device__ __forceinline__ void pt_stream(int lb, int le, int off, const int *__restrict__ frp,
                                          const int *__restrict__ fci,
                                          const double *__restrict__ fv,
                                          const int *__restrict__ fdpos, int lane, Fn &&fn) {
  pt_stream_t<VALS> str{fci, fv, lane};
  for (int c0 = lb; c0 < le; c0 += XS_WAVE) {
    str.load_begin(c0, le, off, frp, fdpos);
    str.finish(fn);
  }
}

// Pass 1 (fill == false): cnt[i] = |C_i| for every row.  Pass 2 (fill ==
// true), for the rows with more than minlen candidates only: C_i sorted into
// cci[crp[i]..], the row index into crow and the pivot position into cdpos
// (*ovf |= 2 for a row without its diagonal).  The 2048-slot hash takes any
// row up to 1024 candidates (*ovf |= 1 beyond).
template <bool FILL>
__global__ __launch_bounds__(XS_WAVE *PT_WAVES) void k_pt_symbolic(
    int n, const int *__restrict__ arp, const int *__restrict__ aci, const int *__restrict__ frp,
    const int *__restrict__ fci, const int *__restrict__ fdpos, int *__restrict__ cnt,
    const int *__restrict__ crp, int *__restrict__ cci, int *__restrict__ crow,
    int *__restrict__ cdpos, int minlen, int *ovf, const int *__restrict__ rows = nullptr,
    const int *__restrict__ nrows = nullptr) {
  __shared__ int tab_all[PT_WAVES][PT_HASH];
  const int lane = threadIdx.x & (XS_WAVE - 1), wv = threadIdx.x / XS_WAVE;
  const int w = xcd_chunk(blockIdx.x, gridDim.x) * PT_WAVES + wv;
  // with a row list (counts only): the w-th listed row, for w < *nrows
  const int i = rows ? (w < *nrows ? rows[w] : n) : w;
  const bool act = i < n && (!FILL || crp[i + 1] - crp[i] > minlen);   // wave-uniform
  int *tab = tab_all[wv];
  int *keys = tab;   // compacted in place (see below)
  if (act) {
    // the union of the upper parts (incl. pivot) of the rows k in L_i (a set:
    // order is free): its first block begun from the row's first 64 entries
    // before the table is set up
    const int fa = frp[i], fb = frp[i + 1], fd = fdpos[i], nlf = fd - fa;
    const int fj = fci[fb > fa ? fa + min(lane, fb - fa - 1) : 0];
    const int kf = lane < nlf ? fj : i;
    const int dkf = fdpos[kf], rkf = frp[kf + 1];
    pt_stream_t<false> str{fci, nullptr, lane};
    if (nlf > 0) str.begin(fa, min(XS_WAVE, nlf), dkf, rkf - dkf, 0.0);
    for (int q = lane; q < PT_HASH; q += XS_WAVE) tab[q] = PT_EMPTY;
    for (int e = arp[i] + lane; e < arp[i + 1]; e += XS_WAVE) pt_insert(tab, aci[e], true, ovf);
    pt_insert(tab, fj, lane < fb - fa, ovf);
    for (int e = fa + XS_WAVE + lane; e < fb; e += XS_WAVE) pt_insert(tab, fci[e], true, ovf);
    auto add = [&](int j, double, bool, bool v) { pt_insert(tab, j, v, ovf); };
    if (nlf > 0) {
      str.finish(add);
      for (int c0 = fa + XS_WAVE; c0 < fd; c0 += XS_WAVE) {
        str.load_begin(c0, fd, 0, frp, fdpos);
        str.finish(add);
      }
    }
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
      bool piv = false;
      for (int a = lane; a < u; a += XS_WAVE) {
        const int key = keys[a];
        int rank = 0;
        for (int b = 0; b < u; ++b) rank += keys[b] < key;
        cci[off + rank] = key;
        crow[off + rank] = i;
        if (key == i) {
          cdpos[i] = off + rank;
          piv = true;
        }
      }
      if (!__any(piv) && lane == 0) {
        cdpos[i] = off;
        atomicOr(ovf, 2);
      }
    }
  }
}

// Candidate counts in a TS-slot set per wave (TS < PT_HASH: a quarter of
// the LDS, so more waves per CU, and a quarter of the clearing), counted as
// the inserts that claim an empty slot (no compaction pass).  A row whose
// set fills up is listed in longrows and counted again by k_pt_symbolic's
// 2048-slot set (k_pt_count_long).
template <int TS>
__device__ __forceinline__ void pt_insert_cnt(int *tab, int c, bool v, int &nnew, bool &full) {
  unsigned h = ((unsigned)c * 2654435761u) & (TS - 1);
  int old = tab[h];
  if (v && old != c) {
    old = atomicCAS(&tab[h], PT_EMPTY, c);
    if (old == PT_EMPTY) ++nnew;
  }
  const bool more = v && old != PT_EMPTY && old != c;
  if (__any(more)) {
    if (more) {
      int probe = 1;
      for (; probe < TS; ++probe) {
        h = (h + 1) & (TS - 1);
        const int o = atomicCAS(&tab[h], PT_EMPTY, c);
        if (o == PT_EMPTY) { ++nnew; break; }
        if (o == c) break;
      }
      if (probe == TS) full = true;
    }
  }
}

template <int TS>
__global__ __launch_bounds__(XS_WAVE *PT_WAVES) void k_pt_count(
    int n, const int *__restrict__ arp, const int *__restrict__ aci, const int *__restrict__ frp,
    const int *__restrict__ fci, const int *__restrict__ fdpos, int *__restrict__ cnt,
    int *__restrict__ longrows, int *nlong) {
  __shared__ int tab_all[PT_WAVES][TS];
  const int lane = threadIdx.x & (XS_WAVE - 1), wv = threadIdx.x / XS_WAVE;
  const int i = xcd_chunk(blockIdx.x, gridDim.x) * PT_WAVES + wv;
  if (i >= n) return;   // wave-uniform; the waves share no barrier
  int *tab = tab_all[wv];
  int nnew = 0;
  bool full = false;
  const int fa = frp[i], fb = frp[i + 1], fd = fdpos[i], nlf = fd - fa;
  const int fj = fci[fb > fa ? fa + min(lane, fb - fa - 1) : 0];
  const int kf = lane < nlf ? fj : i;
  const int dkf = fdpos[kf], rkf = frp[kf + 1];
  pt_stream_t<false> str{fci, nullptr, lane};
  if (nlf > 0) str.begin(fa, min(XS_WAVE, nlf), dkf, rkf - dkf, 0.0);
  for (int q = lane; q < TS; q += XS_WAVE) tab[q] = PT_EMPTY;
  for (int e = arp[i] + lane; e < arp[i + 1]; e += XS_WAVE) pt_insert_cnt<TS>(tab, aci[e], true, nnew, full);
  pt_insert_cnt<TS>(tab, fj, lane < fb - fa, nnew, full);
  for (int e = fa + XS_WAVE + lane; e < fb; e += XS_WAVE) pt_insert_cnt<TS>(tab, fci[e], true, nnew, full);
  auto add = [&](int j, double, bool, bool v) { pt_insert_cnt<TS>(tab, j, v, nnew, full); };
  if (nlf > 0) {
    str.finish(add);
    for (int c0 = fa + XS_WAVE; c0 < fd; c0 += XS_WAVE) {
      str.load_begin(c0, fd, 0, frp, fdpos);
      str.finish(add);
    }
  }
  const int u = pt_wave_sum_i(nnew);
  const bool any_full = __any(full);
  if (lane == 0) {
    if (any_full) {
      cnt[i] = 0;
      longrows[atomicAdd(nlong, 1)] = i;
    } else {
      cnt[i] = u;
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

// pivot positions of a sorted CSR pattern (*bad |= 2 for a row without)
__global__ void k_pt_dpos(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                          int *__restrict__ dpos, int *bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = rp[i], b = rp[i + 1];
  const int p = pt_find(ci, a, b, i);
  if (p < b && ci[p] == i) dpos[i] = p;
  else { dpos[i] = a; atomicOr(bad, 2); }
}

// ---------------------------------------------------------------------------
// Row-wise forms of the embed and the sweep: one wave per pattern row.  The
// row's columns go into an LDS hash table (multiplicative hash, linear
// probing, load <= 1/2) whose slots also hold the row's accumulators, so a
// product reaches its accumulator in one or two probes.  The partial sums
// S_ij = sum_{k in L_i, k < min(i,j)} l_ik u_kj come from pt_stream over the
// strictly upper parts of the rows k of L_i: products outside the row's
// pattern are dropped, the others added to their slots with LDS atomics (two
// products of one step may share a slot).  Rows longer than MAXR fall back to
// the per-entry form.  Rows are spread over the XCDs in contiguous chunks, so
// the rows k a chunk reads stay in one XCD's L2.
// ---------------------------------------------------------------------------
constexpr int pt_log2(int v) { return v <= 1 ? 0 : 1 + pt_log2(v >> 1); }

template <int RTAB>
__device__ __forceinline__ unsigned pt_rhash(int c) {
  return ((unsigned)c * 2654435761u) >> (32 - pt_log2(RTAB));
}

// Slot of column j, inserted if absent (when v; -1 otherwise).  The table
// has room (len <= RTAB / 2).  At most one CAS per lane; the probe loop runs
// only when some lane's first slot holds another column.
template <int RTAB>
__device__ int pt_slot_insert_from(int *key, int j, unsigned h) {
  for (;;) {
    const int old = atomicCAS(&key[h], PT_EMPTY, j);
    if (old == PT_EMPTY || old == j) return (int)h;
    h = (h + 1) & (RTAB - 1);
  }
}

template <int RTAB>
__device__ __forceinline__ int pt_slot_insert_if(int *key, int j, bool v) {
  const unsigned h = pt_rhash<RTAB>(j);
  int old = key[h];   // a column already in its first slot needs no CAS
  if (v && old != j) old = atomicCAS(&key[h], PT_EMPTY, j);
  const bool more = v && old != PT_EMPTY && old != j;
  int r = v ? (int)h : -1;
  if (__any(more))
    if (more) r = pt_slot_insert_from<RTAB>(key, j, (h + 1) & (RTAB - 1));
  return r;
}

template <int RTAB>
__device__ __forceinline__ int pt_slot_insert(int *key, int j) {
  return pt_slot_insert_if<RTAB>(key, j, true);
}

// Slot of column j, or -1.  The first two probes are read together; the
// probe loop runs only when some lane found two other columns there.
template <int RTAB>
__device__ int pt_slot_find_from(const int *key, int j, unsigned h) {
  for (;;) {
    const int k = key[h];
    if (k == j) return (int)h;
    if (k == PT_EMPTY) return -1;
    h = (h + 1) & (RTAB - 1);
  }
}

template <int RTAB>
__device__ __forceinline__ int pt_slot_find(const int *key, int j) {
  const unsigned h = pt_rhash<RTAB>(j), h1 = (h + 1) & (RTAB - 1);
  const int k0 = key[h], k1 = key[h1];
  // an empty first slot ends the search (no deletions: j was never placed
  // beyond it)
  int r = k0 == j ? (int)h : (k0 != PT_EMPTY && k1 == j) ? (int)h1 : -1;
  const bool more = k0 != j && k0 != PT_EMPTY && k1 != j && k1 != PT_EMPTY;
  if (__any(more))
    if (more) r = pt_slot_find_from<RTAB>(key, j, (h + 2) & (RTAB - 1));
  return r;
}

// Selection key of entry (i, j) from its scaled value v (s_i = si, s_j =
// sj): the bit pattern of the unscaled factor's magnitude, |l_ij| =
// |l~_ij (s_j / s_i)| or |u_ij| = |u~_ij / (s_i s_j)| (non-negative, so the
// patterns order like the values), with the factor in the top bit (0 for L,
// j < i; 1 for U, j >= i).
#define PT_KEY_U (1ull << 63)
__device__ __forceinline__ unsigned long long pt_key(int i, int j, double v, double si,
                                                     double sj) {
  const double m = j < i ? fabs(v * (sj / si)) : fabs(v / (si * sj));
  return (unsigned long long)__double_as_longlong(m) | (j >= i ? PT_KEY_U : 0ull);
}

// rows of at most MAXR entries per wave and waves per workgroup (12 bytes of
// LDS per slot, two slots per entry)
template <int MAXR> struct pt_rows { static constexpr int W = MAXR <= 256 ? 4 : 2; };

// One sweep over a pattern (with keys: also each entry's selection key,
// pt_key).  a_ij enters after the products: each slot holds
// sum l_ik u_kj, the lanes of A's entries subtract a_ij from their slots and
// every value reads the negation -- a_ij - S_ij and 0 - S_ij exactly (IEEE
// rounding is symmetric), without an LDS array for A.  The loads that do not
// depend on LDS (the row's columns, the pivots u_jj of its L part, A's first
// 64 entries) are issued together at the start.
template <int MAXR>
__global__ __launch_bounds__(XS_WAVE * pt_rows<MAXR>::W) void k_pt_sweep_rows(
    int n, const int *__restrict__ rp, const int *__restrict__ ci, const int *__restrict__ dpos,
    const double *__restrict__ vin, double *__restrict__ vout, const int *__restrict__ arp,
    const int *__restrict__ aci, const double *__restrict__ av, const double *__restrict__ sc,
    unsigned long long *__restrict__ keys) {
  constexpr int W = pt_rows<MAXR>::W, RTAB = 2 * MAXR, NM = MAXR / XS_WAVE;
  __shared__ int key_s[W][RTAB];
  __shared__ double S_s[W][RTAB];
  const int w = threadIdx.x / XS_WAVE, lane = threadIdx.x & (XS_WAVE - 1);
  const int i = xcd_chunk(blockIdx.x, gridDim.x) * W + w;
  if (i >= n) return;   // wave-uniform, no workgroup barrier below
  const int a = rp[i], len = rp[i + 1] - a;
  const double si = sc[i];
  if (len > MAXR) {
    for (int t = lane; t < len; t += XS_WAVE) {
      const int e = a + t, j = ci[e];
      const double s = pt_a(arp, aci, av, sc, i, j) - pt_partial(rp, ci, vin, dpos, i, j);
      double r;
      if (i > j) r = s / pt_clamp(vin[dpos[j]]);
      else r = (i == j) ? pt_clamp(s) : s;
      const double o = isfinite(r) ? r : vin[e];
      vout[e] = o;
      if (keys) keys[e] = pt_key(i, j, o, si, sc[j]);
    }
    return;
  }
  const int nl = dpos[i] - a, ab = arp[i], ae = arp[i + 1];
  // the row's columns; for its L entries j = k: the pivot positions dpos[k]
  // (u_kk divides the entry; the product parts start after it), the row
  // ends rp[k + 1] and the values l_ik -- loads with clamped indices, so
  // they issue together
  int cj[NM], dk[NM];
  double piv[NM], scj[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const int t = lane + m * XS_WAVE;
    cj[m] = t < len ? ci[a + t] : -1;
  }
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const int t = lane + m * XS_WAVE;
    dk[m] = dpos[t < nl ? cj[m] : i];
  }
  const int rk0 = rp[(lane < nl ? cj[0] : i) + 1];
  const double lv0 = vin[len > 0 ? a + min(lane, len - 1) : 0];
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const int t = lane + m * XS_WAVE;
    piv[m] = t < nl ? vin[dk[m]] : 1.0;
    scj[m] = (keys && cj[m] >= 0) ? sc[cj[m]] : 1.0;
  }
  const int aj = ab + lane < ae ? aci[ab + lane] : -1;
  const double aval = aj >= 0 ? av[ab + lane] * si * sc[aj] : 0.0;
  // the product stream's first block, begun before the table is built
  pt_stream_t<true> str{ci, vin, lane};
  if (nl > 0) str.begin(a, min(XS_WAVE, nl), dk[0] + 1, rk0 - (dk[0] + 1), lv0);
  int *key = key_s[w];
  double *S = S_s[w];
  for (int q = lane; q < RTAB; q += XS_WAVE) key[q] = PT_EMPTY;
  int hs[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    hs[m] = -1;
    if (cj[m] >= 0) {
      const int h = pt_slot_insert<RTAB>(key, cj[m]);
      S[h] = 0.0;
      hs[m] = h;
    }
  }
  const int ha = aj >= 0 ? pt_slot_find<RTAB>(key, aj) : -1;
  auto add = [&](int j, double x, bool, bool v) {
    const int h = pt_slot_find<RTAB>(key, j);
    if (v && h >= 0) atomicAdd(&S[h], x);
  };
  if (nl > 0) {
    str.finish(add);
    for (int c0 = a + XS_WAVE; c0 < a + nl; c0 += XS_WAVE) {
      str.load_begin(c0, a + nl, 1, rp, dpos);
      str.finish(add);
    }
  }
  if (ha >= 0) S[ha] = S[ha] - aval;
  for (int e = ab + XS_WAVE + lane; e < ae; e += XS_WAVE) {   // rows of A beyond 64 entries
    const int j = aci[e];
    const int h = pt_slot_find<RTAB>(key, j);
    if (h >= 0) S[h] = S[h] - av[e] * si * sc[j];
  }
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    if (cj[m] >= 0) {
      const int e = a + lane + m * XS_WAVE, j = cj[m];
      const double s = -S[hs[m]];
      double r;
      if (i > j) r = s / pt_clamp(piv[m]);
      else r = (i == j) ? pt_clamp(s) : s;
      const double o = isfinite(r) ? r : vin[e];
      vout[e] = o;
      if (keys) keys[e] = pt_key(i, j, o, si, scj[m]);
    }
  }
}

// The embed for the rows of more than minlen candidates (pt_fill_embed takes
// the others): per entry, the old factor's value where it has the entry,
// else the residual form with binary searches (embedA_intoB_L2 / _U2,
// hip/Precond/parilut.cpp).  A wave per row.
__global__ __launch_bounds__(XS_BLOCK) void k_pt_embed_long(
    int n, int minlen, const int *__restrict__ crp, const int *__restrict__ cci,
    const int *__restrict__ arp, const int *__restrict__ aci, const double *__restrict__ av,
    const double *__restrict__ sc, const int *__restrict__ frp, const int *__restrict__ fci,
    const double *__restrict__ fv, const int *__restrict__ fdpos, double *__restrict__ cv) {
  const int lane = threadIdx.x & (XS_WAVE - 1);
  const int i = blockIdx.x * (XS_BLOCK / XS_WAVE) + threadIdx.x / XS_WAVE;
  if (i >= n) return;
  const int a = crp[i], len = crp[i + 1] - a;
  if (len <= minlen) return;
  const int fa = frp[i], fb = frp[i + 1];
  for (int t = lane; t < len; t += XS_WAVE) {
    const int j = cci[a + t];
    const int p = pt_find(fci, fa, fb, j);
    double val;
    if (p < fb && fci[p] == j) {
      val = fv[p];
    } else {
      const double r = pt_a(arp, aci, av, sc, i, j) - pt_partial(frp, fci, fv, fdpos, i, j);
      val = (i > j) ? r / pt_clamp(fv[fdpos[j]]) : r;
    }
    cv[a + t] = isfinite(val) ? val : 0.0;
  }
}

// Candidate pattern and embedded values in one pass, for rows of at most
// MAXR candidates (their count from pass 1 of k_pt_symbolic): the row's
// columns -- A's, the old factor's and the union of the upper parts with
// pivots of the rows k in L_i -- go into an LDS hash table whose slots
// accumulate S_ij = sum_{k in L_i, k < min(i,j)} l_ik u_kj from the same
// product stream (pivots excluded).  A's lanes subtract a_ij, every slot
// takes r = a_ij - S_ij as the negation (exact), then (r / u_jj) below the
// diagonal and r on and above it, and the old factor's lanes put back their
// values (embedA_intoB_L2 / _U2, hip/Precond/parilut.cpp).  The keys are
// compacted with their values and rank-sorted into cci / cv, with crow and
// the pivot position (*ovf |= 2 for a row without its diagonal).
template <int MAXR>
__global__ __launch_bounds__(XS_WAVE * pt_rows<MAXR>::W) void k_pt_fill_embed(
    int n, const int *__restrict__ arp, const int *__restrict__ aci, const double *__restrict__ av,
    const double *__restrict__ sc, const int *__restrict__ frp, const int *__restrict__ fci,
    const double *__restrict__ fv, const int *__restrict__ fdpos, const int *__restrict__ crp,
    int *__restrict__ cci, int *__restrict__ crow, double *__restrict__ cv,
    int *__restrict__ cdpos, int *ovf) {
  constexpr int W = pt_rows<MAXR>::W, RTAB = 2 * MAXR;
  __shared__ int key_s[W][RTAB];
  __shared__ double S_s[W][RTAB];
  const int w = threadIdx.x / XS_WAVE, lane = threadIdx.x & (XS_WAVE - 1);
  const int i = xcd_chunk(blockIdx.x, gridDim.x) * W + w;
  if (i >= n) return;   // wave-uniform, no workgroup barrier below
  const int off = crp[i], u = crp[i + 1] - off;
  if (u > MAXR) return;  // k_pt_symbolic<true> + k_pt_embed_long
  const int ab = arp[i], ae = arp[i + 1], fa = frp[i], fb = frp[i + 1], fd = fdpos[i];
  const double si = sc[i];
  // A's and the old factor's first 64 entries of the row, loaded together,
  // with what the product stream's first block needs of the old factor's L
  // entries k (pivot positions, row ends)
  const int nlf = fd - fa;
  const int aj = ab + lane < ae ? aci[ab + lane] : -1;
  const double aval = aj >= 0 ? av[ab + lane] * si * sc[aj] : 0.0;
  const int fj = fa + lane < fb ? fci[fa + lane] : -1;
  const double fval = fj >= 0 ? fv[fa + lane] : 0.0;
  const int kf = lane < nlf ? fj : i;
  const int dkf = fdpos[kf], rkf = frp[kf + 1];
  pt_stream_t<true> str{fci, fv, lane};
  if (nlf > 0) str.begin(fa, min(XS_WAVE, nlf), dkf, rkf - dkf, fval);
  int *key = key_s[w];
  double *S = S_s[w];
  for (int q = lane; q < RTAB; q += XS_WAVE) {
    key[q] = PT_EMPTY;
    S[q] = 0.0;
  }
  const int ha = aj >= 0 ? pt_slot_insert<RTAB>(key, aj) : -1;
  for (int e = ab + XS_WAVE + lane; e < ae; e += XS_WAVE) pt_slot_insert<RTAB>(key, aci[e]);
  const int hf = fj >= 0 ? pt_slot_insert<RTAB>(key, fj) : -1;
  for (int e = fa + XS_WAVE + lane; e < fb; e += XS_WAVE) pt_slot_insert<RTAB>(key, fci[e]);
  auto add = [&](int j, double x, bool first, bool v) {
    const int h = pt_slot_insert_if<RTAB>(key, j, v);
    if (v && !first) atomicAdd(&S[h], x);
  };
  if (nlf > 0) {
    str.finish(add);
    for (int c0 = fa + XS_WAVE; c0 < fd; c0 += XS_WAVE) {
      str.load_begin(c0, fd, 0, frp, fdpos);
      str.finish(add);
    }
  }
  // S_ij - a_ij on A's slots; every value then reads the negation
  if (ha >= 0) S[ha] = S[ha] - aval;
  for (int e = ab + XS_WAVE + lane; e < ae; e += XS_WAVE) {
    const int j = aci[e];
    const int h = pt_slot_find<RTAB>(key, j);
    S[h] = S[h] - av[e] * si * sc[j];
  }
  for (int q = lane; q < RTAB; q += XS_WAVE) {
    const int j = key[q];
    if (j == PT_EMPTY) continue;
    const double r = -S[q];
    const double val = (i > j) ? r / pt_clamp(fv[fdpos[j]]) : r;
    S[q] = isfinite(val) ? val : 0.0;
  }
  // the old factor's entries keep their values
  if (hf >= 0) S[hf] = isfinite(fval) ? fval : 0.0;
  for (int e = fa + XS_WAVE + lane; e < fb; e += XS_WAVE) {
    const int h = pt_slot_find<RTAB>(key, fci[e]);
    const double x = fv[e];
    S[h] = isfinite(x) ? x : 0.0;
  }
  // compact keys and values to the front of their tables (in place: a
  // chunk's slots are all read before any of its entries is stored, at
  // positions <= q0 + lane), then rank-sort
  int base = 0;
  for (int q0 = 0; q0 < RTAB; q0 += XS_WAVE) {
    const int c = key[q0 + lane];
    const double v = S[q0 + lane];
    const bool has = c != PT_EMPTY;
    const unsigned long long m = __ballot(has);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (has) {
      key[base + before] = c;
      S[base + before] = v;
    }
    base += __popcll(m);
  }
  if (base != u) {   // the count pass saw another set: leave the row alone
    if (lane == 0) atomicOr(ovf, 1);
    return;
  }
  bool piv = false;
  for (int q = lane; q < base; q += XS_WAVE) {
    const int kk = key[q];
    int rank = 0;
    for (int b = 0; b < base; ++b) rank += key[b] < kk;
    cci[off + rank] = kk;
    crow[off + rank] = i;
    cv[off + rank] = S[q];
    if (kk == i) {
      cdpos[i] = off + rank;
      piv = true;
    }
  }
  if (!__any(piv) && lane == 0) {
    cdpos[i] = off;
    atomicOr(ovf, 2);
  }
}

// row-length statistics of a pattern from its per-row counts: stats[0] the
// longest row, stats[1 + b] the rows longer than 128 << b (b < 3)
__global__ __launch_bounds__(XS_BLOCK) void k_pt_len_stats(int n, const int *__restrict__ cnt,
                                                           unsigned *stats) {
  unsigned mx = 0, o[3] = {0, 0, 0};
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    const unsigned c = (unsigned)cnt[i];
    mx = max(mx, c);
#pragma unroll
    for (int b = 0; b < 3; ++b) o[b] += c > (128u << b);
  }
  for (int off = 32; off > 0; off >>= 1) {
    mx = max(mx, (unsigned)__shfl_xor((int)mx, off, XS_WAVE));
#pragma unroll
    for (int b = 0; b < 3; ++b) o[b] += __shfl_xor(o[b], off, XS_WAVE);
  }
  if ((threadIdx.x & (XS_WAVE - 1)) == 0) {
    atomicMax(&stats[0], mx);
#pragma unroll
    for (int b = 0; b < 3; ++b)
      if (o[b]) atomicAdd(&stats[1 + b], o[b]);
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
// the compaction compares against the thresholds, read from the entries'
// selection keys (pt_key, written by the sweep over the candidates) plus the
// n unit diagonals of L, counted at once.  |v| bit patterns are ordered like
// the values; passes of 8-bit digits from the most significant.
// Once the bin of rank k holds at most PT_SEL_CAP entries, the next pass
// collects them and the rank is found among them in one block; later passes
// return at once.  State per slot p: sel[4p + 0] prefix bits, + 1 remaining
// rank, + 2 mask of fixed bits, + 3 mode (0 histogram, 1 collect, 2 done).
// ---------------------------------------------------------------------------
#define PT_SEL_CAP 1024
static_assert(XS_BLOCK == 256, "k_pt_hist clears its 512 bins with two stores per thread");


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

__global__ __launch_bounds__(XS_BLOCK) void k_pt_hist(int nnz,
                                                      const unsigned long long *__restrict__ key,
                                                      int nunit, int shift,
                                                      const unsigned long long *sel,
                                                      unsigned *__restrict__ hist,
                                                      unsigned *lcnt,
                                                      unsigned long long *__restrict__ list) {
  const unsigned long long m0 = sel[3], m1 = sel[7];
  if (m0 == 2 && m1 == 2) return;
  __shared__ unsigned h[512];
  __shared__ unsigned ubase;
  h[threadIdx.x] = 0;
  h[threadIdx.x + XS_BLOCK] = 0;
  __syncthreads();
  const unsigned long long pre[2] = {sel[0], sel[4]}, msk[2] = {sel[2], sel[6]};
  const unsigned long long mode[2] = {m0, m1};
  for (int e = blockIdx.x * XS_BLOCK + threadIdx.x; e < nnz; e += gridDim.x * XS_BLOCK) {
    const unsigned long long k = key[e];
    const int p = (int)(k >> 63);
    const unsigned long long bits = k & ~PT_KEY_U;
    if (mode[p] == 2 || (bits & msk[p]) != pre[p]) continue;
    if (mode[p] == 1) {
      const unsigned q = atomicAdd(&lcnt[p], 1u);
      if (q < PT_SEL_CAP) list[p * PT_SEL_CAP + q] = bits;
    } else {
      atomicAdd(&h[p * 256 + ((bits >> shift) & 0xFF)], 1u);
    }
  }
  // L's nunit unit diagonals (value 1), counted by the first workgroup
  const unsigned long long one = (unsigned long long)__double_as_longlong(1.0);
  if (blockIdx.x == 0 && m0 != 2 && (one & msk[0]) == pre[0]) {
    if (m0 == 1) {
      if (threadIdx.x == 0) ubase = atomicAdd(&lcnt[0], (unsigned)nunit);
      __syncthreads();
      for (int q = threadIdx.x; q < nunit; q += XS_BLOCK)
        if (ubase + q < PT_SEL_CAP) list[ubase + q] = one;
    } else if (threadIdx.x == 0) {
      atomicAdd(&h[(one >> shift) & 0xFF], (unsigned)nunit);
    }
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
// (hip/Precond/parilut.cpp:35-63 counts each factor's diagonal).  From the
// pivot positions of a sorted pattern whose rows all hold their diagonal:
// row i has dpos_i - rp_i + 1 entries with col <= i and rp_{i+1} - dpos_i
// with col >= i.  One atomic pair per workgroup.
__global__ __launch_bounds__(XS_BLOCK) void k_pt_count_parts(int n, const int *__restrict__ rp,
                                                             const int *__restrict__ dpos,
                                                             unsigned long long *cnt2) {
  __shared__ unsigned long long red[2][XS_BLOCK / XS_WAVE];
  unsigned long long a = 0, b = 0;
  for (int i = blockIdx.x * XS_BLOCK + threadIdx.x; i < n; i += gridDim.x * XS_BLOCK) {
    const int d = dpos[i];
    a += (unsigned long long)(d - rp[i] + 1);
    b += (unsigned long long)(rp[i + 1] - d);
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, XS_WAVE);
    b += __shfl_xor(b, off, XS_WAVE);
  }
  const int w = threadIdx.x / XS_WAVE;
  if ((threadIdx.x & (XS_WAVE - 1)) == 0) {
    red[0][w] = a;
    red[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < XS_BLOCK / XS_WAVE; ++q) {
      a += red[0][q];
      b += red[1][q];
    }
    atomicAdd(&cnt2[0], a);
    atomicAdd(&cnt2[1], b);
  }
}

// keep predicate for the compaction
// (thresold_remove_nnz2, hip/Precond/parilut.cpp:303-323: |v| >= threshold
// survives, the diagonal always does; plus the drop tolerance on the same
// unscaled magnitude), on the selection key of the entry
__device__ __forceinline__ bool pt_keep(bool diag, unsigned long long k, double tl, double tu,
                                        double dt) {
  if (diag) return true;
  const double a = __longlong_as_double((long long)(k & ~PT_KEY_U));
  if (a < dt) return false;
  return (k & PT_KEY_U) ? a >= tu : a >= tl;
}

// per-row kept counts and the compaction: a wave per row, kept entries
// placed in column order by ballot
__global__ __launch_bounds__(XS_BLOCK) void k_pt_keep_count(
    int n, const int *__restrict__ crp, const int *__restrict__ cdpos,
    const unsigned long long *__restrict__ key, const double *thr,
    const double *__restrict__ dtol, int *__restrict__ cnt) {
  const int lane = threadIdx.x & (XS_WAVE - 1);
  const int i = blockIdx.x * (XS_BLOCK / XS_WAVE) + threadIdx.x / XS_WAVE;
  if (i >= n) return;
  const double tl = thr[0], tu = thr[1], dt = dtol[i];
  const int dp = cdpos[i];
  int c = 0;
  for (int e = crp[i] + lane; e < crp[i + 1]; e += XS_WAVE) c += pt_keep(e == dp, key[e], tl, tu, dt);
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, XS_WAVE);
  if (lane == 0) cnt[i] = c;
}

__global__ __launch_bounds__(XS_BLOCK) void k_pt_keep_write(
    int n, const int *__restrict__ crp, const int *__restrict__ cci, const int *__restrict__ cdpos,
    const double *__restrict__ cv, const unsigned long long *__restrict__ key, const double *thr,
    const double *__restrict__ dtol, const int *__restrict__ frp, int *__restrict__ fci,
    double *__restrict__ fv, int *__restrict__ frow, int *__restrict__ fdpos, int *bad) {
  const int lane = threadIdx.x & (XS_WAVE - 1);
  const int i = blockIdx.x * (XS_BLOCK / XS_WAVE) + threadIdx.x / XS_WAVE;
  if (i >= n) return;
  const double tl = thr[0], tu = thr[1], dt = dtol[i];
  const int cd = cdpos[i];
  const unsigned long long below = (1ull << lane) - 1ull;
  int w = frp[i], dp = -1;
  const int re = crp[i + 1];
  for (int e0 = crp[i]; e0 < re; e0 += XS_WAVE) {
    const int e = e0 + lane;
    int j = 0;
    double x = 0.0;
    bool keep = false;
    if (e < re) {
      j = cci[e];
      x = cv[e];
      keep = pt_keep(e == cd, key[e], tl, tu, dt);
    }
    const unsigned long long m = __ballot(keep);
    const int q = w + __popcll(m & below);
    if (keep) {
      fci[q] = j;
      fv[q] = x;
      frow[q] = i;
    }
    // the pivot (always kept): its position, wave-uniform
    const unsigned long long md = __ballot(keep && j == i);
    if (md) dp = w + __popcll(m & ((1ull << (__ffsll((long long)md) - 1)) - 1ull));
    w += __popcll(m);
  }
  if (lane == 0) {
    fdpos[i] = dp >= 0 ? dp : frp[i];
    if (dp < 0) atomicOr(bad, 2);
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

__global__ void k_pt_fill_i32(int *p, int n, int v) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) p[k] = v;
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

void xs_pt_symbolic(hipStream_t s, int n, const int *arp, const int *aci, const int *frp,
                    const int *fci, const int *fdpos, int *cnt, int *ovf, int *longrows,
                    int *nlong) {
  dim3 grid((n + PT_WAVES - 1) / PT_WAVES), blk(XS_WAVE * PT_WAVES);
  if (!longrows || !nlong) {   // every row in the 2048-slot set
    hipLaunchKernelGGL(k_pt_symbolic<false>, grid, blk, 0, s, n, arp, aci, frp, fci, fdpos, cnt,
                       nullptr, nullptr, nullptr, nullptr, 0, ovf, nullptr, nullptr);
    return;
  }
  // 512-slot sets first (a full set lists its row), then the listed rows in
  // the 2048-slot set; the second grid covers every row, the waves past the
  // list's length exit at once
  hipLaunchKernelGGL(k_pt_fill_i32, dim3(1), dim3(64), 0, s, nlong, 1, 0);
  hipLaunchKernelGGL(k_pt_count<512>, grid, blk, 0, s, n, arp, aci, frp, fci, fdpos, cnt, longrows,
                     nlong);
  hipLaunchKernelGGL(k_pt_symbolic<false>, grid, blk, 0, s, n, arp, aci, frp, fci, fdpos, cnt,
                     nullptr, nullptr, nullptr, nullptr, 0, ovf, longrows, nlong);
}

void xs_pt_dpos(hipStream_t s, int n, const int *rp, const int *ci, int *dpos, int *bad) {
  hipLaunchKernelGGL(k_pt_dpos, dim3((n + 255) / 256), dim3(256), 0, s, n, rp, ci, dpos, bad);
}

template <int MAXR>
static dim3 pt_row_grid(int n) {
  return dim3((n + pt_rows<MAXR>::W - 1) / pt_rows<MAXR>::W);
}

void xs_pt_fill_embed(hipStream_t s, int n, const int *arp, const int *aci, const double *av,
                      const double *sc, const int *frp, const int *fci, const double *fv,
                      const int *fdpos, const int *crp, int *cci, int *crow, double *cv,
                      int *cdpos, int *ovf, int maxr, bool has_long) {
#define PT_FILL(M)                                                                              \
  hipLaunchKernelGGL(k_pt_fill_embed<M>, pt_row_grid<M>(n), dim3(XS_WAVE * pt_rows<M>::W), 0, s, \
                     n, arp, aci, av, sc, frp, fci, fv, fdpos, crp, cci, crow, cv, cdpos, ovf)
  if (maxr <= 128) PT_FILL(128);
  else if (maxr <= 256) PT_FILL(256);
  else PT_FILL(512);
#undef PT_FILL
  if (!has_long) return;
  const int m = maxr <= 128 ? 128 : maxr <= 256 ? 256 : 512;
  hipLaunchKernelGGL(k_pt_symbolic<true>, dim3((n + PT_WAVES - 1) / PT_WAVES),
                     dim3(XS_WAVE * PT_WAVES), 0, s, n, arp, aci, frp, fci, fdpos, nullptr, crp,
                     cci, crow, cdpos, m, ovf);
  const int rows = XS_BLOCK / XS_WAVE;
  hipLaunchKernelGGL(k_pt_embed_long, dim3((n + rows - 1) / rows), dim3(XS_BLOCK), 0, s, n, m, crp,
                     cci, arp, aci, av, sc, frp, fci, fv, fdpos, cv);
}

void xs_pt_sweep(hipStream_t s, int n, const int *rp, const int *ci, const int *dpos,
                 const double *vin, double *vout, const int *arp, const int *aci,
                 const double *av, const double *sc, int maxr, unsigned long long *keys) {
#define PT_SWEEP(M)                                                                              \
  hipLaunchKernelGGL(k_pt_sweep_rows<M>, pt_row_grid<M>(n), dim3(XS_WAVE * pt_rows<M>::W), 0, s, \
                     n, rp, ci, dpos, vin, vout, arp, aci, av, sc, keys)
  if (maxr <= 128) PT_SWEEP(128);
  else if (maxr <= 256) PT_SWEEP(256);
  else PT_SWEEP(512);
#undef PT_SWEEP
}

void xs_pt_len_stats(hipStream_t s, int n, const int *cnt, unsigned *stats) {
  int g = (n + XS_BLOCK * 8 - 1) / (XS_BLOCK * 8);
  if (g > 1024) g = 1024;
  if (g < 1) g = 1;
  (void)hipMemsetAsync(stats, 0, 4 * sizeof(unsigned), s);
  hipLaunchKernelGGL(k_pt_len_stats, dim3(g), dim3(XS_BLOCK), 0, s, n, cnt, stats);
}

bool xs_pt_rows_over(int maxr, const unsigned *stats) {
  return stats[maxr <= 128 ? 1 : maxr <= 256 ? 2 : 3] != 0;
}

int xs_pt_pick_rows(int n, const unsigned *stats) {
  // the smallest table that leaves at most n / 1024 rows to the per-entry form
  const unsigned few = (unsigned)(n / 1024);
  if (stats[1] <= few) return 128;
  if (stats[2] <= few) return 256;
  return 512;
}

void xs_pt_select(hipStream_t s, int nnz, const unsigned long long *keys, int nunit,
                  const unsigned long long *cnt2, long long keepL, long long keepU,
                  unsigned long long *sel, unsigned *hist, unsigned *lcnt,
                  unsigned long long *list, double *thr) {
  hipLaunchKernelGGL(k_pt_sel_init, dim3(1), dim3(256), 0, s, cnt2, keepL, keepU, sel, hist, lcnt,
                     thr);
  int g = (nnz + XS_BLOCK * 16 - 1) / (XS_BLOCK * 16);
  if (g > 2048) g = 2048;
  if (g < 1) g = 1;
  for (int shift = 56; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(k_pt_hist, dim3(g), dim3(XS_BLOCK), 0, s, nnz, keys, nunit, shift, sel,
                       hist, lcnt, list);
    hipLaunchKernelGGL(k_pt_pick, dim3(2), dim3(256), 0, s, shift, sel, hist, lcnt, list, thr);
  }
}

int xs_pt_select_words() { return 8; }
int xs_pt_list_len() { return 2 * PT_SEL_CAP; }

void xs_pt_count_parts(hipStream_t s, int n, const int *rp, const int *dpos,
                       unsigned long long *cnt2) {
  int g = (n + XS_BLOCK * 8 - 1) / (XS_BLOCK * 8);
  if (g > 512) g = 512;
  if (g < 1) g = 1;
  (void)hipMemsetAsync(cnt2, 0, 2 * sizeof(unsigned long long), s);
  hipLaunchKernelGGL(k_pt_count_parts, dim3(g), dim3(XS_BLOCK), 0, s, n, rp, dpos, cnt2);
}

void xs_pt_dtol(hipStream_t s, int n, const int *arp, const double *av, double droptol,
                double *dtol) {
  hipLaunchKernelGGL(k_pt_dtol, dim3((n + 255) / 256), dim3(256), 0, s, n, arp, av, droptol, dtol);
}

void xs_pt_keep(hipStream_t s, bool write, int n, const int *crp, const int *cci,
                const int *cdpos, const double *cv, const unsigned long long *keys,
                const double *thr, const double *dtol, int *cnt, const int *frp, int *fci,
                double *fv, int *frow, int *fdpos, int *bad) {
  const int rows = XS_BLOCK / XS_WAVE, g = (n + rows - 1) / rows;
  if (!write)
    hipLaunchKernelGGL(k_pt_keep_count, dim3(g), dim3(XS_BLOCK), 0, s, n, crp, cdpos, keys, thr,
                       dtol, cnt);
  else
    hipLaunchKernelGGL(k_pt_keep_write, dim3(g), dim3(XS_BLOCK), 0, s, n, crp, cci, cdpos, cv, keys,
                       thr, dtol, frp, fci, fv, frow, fdpos, bad);
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
