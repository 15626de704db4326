// xs_host.cpp -- host driver of the gfx950 Krylov library: contexts, matrix
// upload and analysis, preconditioner set-up, device-resident Krylov loops and
// the extern "C" entry points declared in include/xsolver_c.h.
//
// Reference call sequence mirrored: hip/xsolver_c.cpp:41-333 (mat_setup ->
// ksp_settype -> ksp_setoption -> pc_settype -> solve -> free), solver loops
// hip/Solver/xsolver_{cg,bicgstab,igcr}.cpp, GMRES from
// intel/mpisolver-C-v1.cpp:933-1247.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xsolver_c.h"
#include "xs_internal.h"
#include "xs_multi.h"
#include "xs_parilut.h"
#include "xs_ic.h"
#include "xs_solvers.h"

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
#define XS_HIP(call)                                                    \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess) {                                             \
      snprintf(xs_errbuf, sizeof(xs_errbuf), "%s: %s", #call,           \
               hipGetErrorString(e_));                                  \
      return XS_ERR_DEVICE;                                             \
    }                                                                   \
  } while (0)

static thread_local char xs_errbuf[512];

template <class T>
static int dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  XS_HIP(hipMalloc((void **)p, sizeof(T) * count));
  return XS_OK;
}
template <class T>
static void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

static bool streq_any(const char *s, std::initializer_list<const char *> names) {
  for (const char *n : names)
    if (strcmp(s, n) == 0) return true;
  return false;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
xs_ctx::xs_ctx() {}

static void free_precond(xs_ctx *c) {
  dfree(c->jac_dinv);
  xs_ilu &F = c->ilu;
  dfree(F.lu); dfree(F.dpos); dfree(F.dinv); dfree(F.ucp); dfree(F.uri); dfree(F.upos);
  dfree(F.rowidx); dfree(F.rp_own); dfree(F.ci_own); dfree(F.trp); dfree(F.tci);
  dfree(F.tdpos); dfree(F.tlu); dfree(F.scale); dfree(F.tperm);
  dfree(F.tlgb); dfree(F.tugb); dfree(F.tlsched); dfree(F.tusched);
  dfree(F.lsched); dfree(F.usched); dfree(F.lgb); dfree(F.ugb);
  dfree(F.pp); dfree(F.pl); dfree(F.pu); dfree(F.pd); dfree(F.as); dfree(F.act);
  dfree(F.app); dfree(F.apd); dfree(F.cmap); dfree(F.lc); dfree(F.mpos);
  dfree(F.pmask); dfree(F.dmap); dfree(F.das); dfree(F.dlc); dfree(F.dlc2);
  dfree(F.tl.ci); dfree(F.tl.v); dfree(F.tl.vlo); dfree(F.tl.co); dfree(F.tl.ww);
  dfree(F.tu.ci); dfree(F.tu.v); dfree(F.tu.vlo); dfree(F.tu.co); dfree(F.tu.ww); dfree(F.tu.vs);
  dfree(F.tl.dv); dfree(F.tu.dv); dfree(F.tu.dvs); dfree(F.tl.dvlo); dfree(F.tu.dvlo);
  c->tell_ready = c->tell32_ready = false;
  F = xs_ilu();
  F.trsv_wg = c->trsv_wg;
  F.trsv_prog = c->trsv_prog;
  F.trsv_gather = c->trsv_gather;
  dfree(c->lu_tmp);
  c->pc_ready = false;
}

static void free_matrix(xs_ctx *c) {
  if (c->own_matrix) { dfree(c->A.rp); dfree(c->A.ci); dfree(c->A.v); }
  c->A = xs_csr();
  dfree(c->plan.blk_row);
  dfree(c->plan.long_rows);
  c->plan = xs_spmv_plan();
  dfree(c->ell.ci); dfree(c->ell.v); dfree(c->ell.co); dfree(c->ell.dv);
  c->ell = xs_ell();
  dfree(c->At.rp); dfree(c->At.ci); dfree(c->At.v);
  c->At = xs_csr();
  dfree(c->planT.blk_row);
  dfree(c->planT.long_rows);
  c->planT = xs_spmv_plan();
  dfree(c->perm_At);
  c->have_At = false;
  c->h_rp.clear(); c->h_ci.clear();
  free_precond(c);
  xs_mixed_free(c);
}

static void free_work(xs_ctx *c) {
  dfree(c->work);
  c->work_n = 0;
  c->work_vecs = 0;
}

int xs_ctx_ensure_work(xs_ctx *c, int nvec) {
  size_t need = (size_t)nvec * c->A.n;
  if (c->work && c->work_vecs >= nvec && c->work_n == (size_t)c->A.n) return XS_OK;
  free_work(c);
  int rc = dalloc(&c->work, need);
  if (rc) return rc;
  c->work_n = c->A.n;
  c->work_vecs = nvec;
  return XS_OK;
}

double *xs_ctx_vec(xs_ctx *c, int k) { return c->work + (size_t)k * c->work_n; }

// Base device buffers of a context (state, counters, flags); allocated once at
// creation so every later launch finds them.
static int alloc_base(xs_ctx *c) {
  int rc;
  if ((rc = dalloc(&c->st, 1))) return rc;
  if ((rc = dalloc(&c->counters, 64))) return rc;
  XS_HIP(hipMemset(c->counters, 0, sizeof(unsigned) * 64));
  if ((rc = dalloc(&c->d_int, 16))) return rc;
  XS_HIP(hipMemset(c->d_int, 0, sizeof(int) * 16));
  // pinned: [0] the solve's state, [1] the inner solve's (mixed precision);
  // a device-to-host copy into pageable memory goes through a staging blit
  XS_HIP(hipHostMalloc((void **)&c->h_state, 2 * sizeof(xs_state), hipHostMallocDefault));
  return XS_OK;
}

// Reduction partials sized for the largest grid that reduces (>= the SpMV
// row-block count).
static int ensure_common(xs_ctx *c) {
  if (!c->st || !c->counters || !c->d_int) return XS_ERR_STATE;
  size_t pblocks = std::max<size_t>(XS_MAX_REDUCE_BLOCKS, (size_t)c->plan.nblk);
  pblocks = std::max<size_t>(pblocks, (size_t)c->planT.nblk);
  // ELL launch: two rows per thread
  pblocks = std::max<size_t>(pblocks, ((size_t)c->A.n + 2 * XS_BLOCK - 1) / (2 * XS_BLOCK));
  // vector CSR: long-row workgroups + at most XS_MAX_REDUCE_BLOCKS
  for (const xs_spmv_plan *P : {&c->plan, &c->planT})
    pblocks = std::max<size_t>(pblocks, (size_t)(P->nlong + 3) / 4 + XS_MAX_REDUCE_BLOCKS);
  if (pblocks > c->partial_blocks) {
    dfree(c->partial);
    // up to 4 values per block, then per group of 64 blocks (two-level combine)
    const size_t slots = (pblocks + pblocks / 64 + 1) * 4;
    int rc = dalloc(&c->partial, slots);
    if (rc) return rc;
    c->partial_blocks = pblocks;
    // reduction slots hold the sentinel between reductions
    xs_fill_sentinel(c->stream, c->partial, (int)slots);
    XS_HIP(hipStreamSynchronize(c->stream));
  }
  return XS_OK;
}

xs_reduce xs_ctx_reduce(xs_ctx *c, int op, int slot) {
  xs_reduce r{};
  r.partial = c->partial;
  r.counter = c->counters;   // counters[0]: reductions; [2..3]: trsv tickets
  r.st = c->st;
  r.resvec = c->resvec_d;
  r.op = op;
  r.slot = slot;
  return r;
}

// ---------------------------------------------------------------------------
// matrix setup and analysis
// ---------------------------------------------------------------------------
static int analyse_pattern(xs_ctx *c) {
  // Host copy of the pattern (needed for the ILU column structure).  Checks
  // sorted columns within rows (the reference's GPU factorisations assume it,
  // hip/Precond/par_ilu_hip/Asplit_gpu.cpp:15-25).
  const std::vector<int> &rp = c->h_rp;
  const std::vector<int> &ci = c->h_ci;
  const int n = rp.empty() ? 0 : (int)rp.size() - 1;
  c->sorted = true;
  c->has_diag = true;
  for (int i = 0; i < n && c->sorted; i++) {
    bool d = false;
    for (int e = rp[i]; e < rp[i + 1]; e++) {
      if (e > rp[i] && ci[e] <= ci[e - 1]) { c->sorted = false; break; }
      if (ci[e] < 0 || ci[e] >= n) { c->sorted = false; break; }
      if (ci[e] == i) d = true;
    }
    if (!d) c->has_diag = false;
  }
  return XS_OK;
}

// (Re)build the CSR row-block plan of M with the given nnz budget per block.
static int make_plan(xs_ctx *c, const xs_csr &M, xs_spmv_plan &P, int chunk) {
  dfree(P.blk_row);   // the vector-CSR fields are kept
  P.nblk = 0;
  P.chunk = 0;
  const int nblk = std::max(1, (M.nnz + chunk - 1) / chunk);
  int rc = dalloc(&P.blk_row, (size_t)nblk + 1);
  if (rc) return rc;
  xs_build_plan(c->stream, M, P, P.blk_row, chunk);
  return XS_OK;
}

// Row-binned vector-CSR parameters from the host row pointer: lanes per row
// = average row length rounded up to a power of two (2..64); rows longer than
// 8 such groups are listed for one wavefront each (lanes and the multiple:
// the autotuner's, when given).
static int make_vec(xs_ctx *c, const std::vector<int> &hrp, xs_spmv_plan &P, int lanes = 0,
                    int long_mult = 0) {
  (void)c;
  const int n = (int)hrp.size() - 1;
  const double avg = n > 0 ? (double)hrp[n] / n : 1.0;
  int L = 2;
  while (L < avg && L < XS_WAVE) L <<= 1;
  if (lanes) L = lanes;               // the autotuner's choice
  P.vec_L = L;
  P.long_thr = (long_mult ? long_mult : 8) * L;
  std::vector<int> lr;
  for (int i = 0; i < n; i++)
    if (hrp[i + 1] - hrp[i] > P.long_thr) lr.push_back(i);
  dfree(P.long_rows);
  P.nlong = (int)lr.size();
  if (P.nlong) {
    int rc = dalloc(&P.long_rows, lr.size());
    if (rc) return rc;
    XS_HIP(hipMemcpy(P.long_rows, lr.data(), sizeof(int) * lr.size(), hipMemcpyHostToDevice));
  }
  return XS_OK;
}

// Rows regular enough for the vector-CSR kernel by default: no long rows and
// no row longer than two lane groups (the LDS row blocks cope better with
// irregular lengths).
static bool regular_rows(const std::vector<int> &hrp, const xs_spmv_plan &P) {
  const int n = (int)hrp.size() - 1;
  int mx = 0;
  for (int i = 0; i < n; i++) mx = std::max(mx, hrp[i + 1] - hrp[i]);
  return P.nlong == 0 && mx <= 2 * P.vec_L;
}

// ELL pays off over the CSR kernels when the padding to the longest row stays
// small: max row length * n <= 1.2 nnz.
static bool ell_fits(const std::vector<int> &hrp) {
  const int n = (int)hrp.size() - 1;
  if (n <= 0) return false;
  int mx = 0;
  for (int i = 0; i < n; i++) mx = std::max(mx, hrp[i + 1] - hrp[i]);
  return mx > 0 && (double)mx * n <= 1.2 * (double)hrp[n];
}

// 16-bit column offsets beside the int32 columns of an ELL layout when every
// entry is within XS_OFF16_MAX of its row (the kernels then read 2 instead of
// 4 index bytes per entry); *co stays null otherwise.
static int make_offsets(xs_ctx *c, int n, int ld, int W, const int *eci, short **co) {
  dfree(*co);
  if (W == 0 || n == 0) return XS_OK;
  const int m = xs_ell_max_offset(c->stream, n, ld, W, eci, c->d_int);
  if (m < 0) return XS_ERR_DEVICE;
  if (m > XS_OFF16_MAX) return XS_OK;
  int rc;
  if ((rc = dalloc(co, (size_t)W * ld))) return rc;
  xs_ell_build_offsets(c->stream, n, ld, W, eci, *co);
  return XS_OK;
}

static int make_ell(xs_ctx *c) {
  int rc;
  xs_csr &A = c->A;
  if (c->ell.v) return XS_OK;
  c->mixed_ready = c->mixed_mat_ready = false;   // no FP32 copy of the new ELL yet
  int W = xs_max_row_len(c->stream, A, c->d_int);
  c->ell.n = A.n;
  c->ell.W = W;
  c->ell.ld = (A.n + 511) / 512 * 512;
  if ((rc = dalloc(&c->ell.ci, (size_t)W * c->ell.ld))) return rc;
  if ((rc = dalloc(&c->ell.v, (size_t)W * c->ell.ld))) return rc;
  xs_build_ell(c->stream, A, W, c->ell);
  if ((rc = make_offsets(c, A.n, c->ell.ld, W, c->ell.ci, &c->ell.co))) return rc;
  // diagonal slots for the FP64 SpMV when A has few distinct offsets
  if (!c->spmv_dia) return XS_OK;
  xs_dia_offs off{};
  const int nd = xs_dia_offsets(c->stream, A, c->d_int, off);
  if (nd < 0) return XS_ERR_DEVICE;
  if (nd == 0) return XS_OK;
  if ((rc = dalloc(&c->ell.dv, (size_t)nd * c->ell.ld))) return rc;
  c->ell.nd = nd;
  c->ell.doff = off;
  xs_build_dia(c->stream, A, c->ell);
  return XS_OK;
}

static int build_spmv(xs_ctx *c) {
  int rc;
  xs_csr &A = c->A;
  if (!A.rp || !A.ci || !A.v || !c->d_int) return XS_ERR_STATE;
  if ((rc = make_plan(c, A, c->plan, xs_default_chunk(A)))) return rc;
  if ((rc = make_vec(c, c->h_rp, c->plan))) return rc;
  c->plan.vec = c->spmv_vec < 0 ? (regular_rows(c->h_rp, c->plan) ? 1 : 0) : c->spmv_vec;
  c->ell_auto = false;
  if (!c->use_ell && c->spmv_vec < 0 && ell_fits(c->h_rp)) {
    c->use_ell = 1;   // CSR input, ELL kernel: padding under 20% of nnz
    c->ell_auto = true;
  }
  if (c->use_ell && (rc = make_ell(c))) return rc;
  XS_HIP(hipStreamSynchronize(c->stream));
  if (xs_take_launch_guard()) return XS_ERR_STATE;
  return XS_OK;
}

// ---------------------------------------------------------------------------
// SpMV autotuner (no reference counterpart).  Times the LDS-staged CSR
// kernel at several nnz budgets per workgroup (the row-block size: rows that
// do not fit the LDS tile of their block go one wavefront per row) and the
// ELL kernel when its padding is modest, on the context's own matrix, and
// keeps the fastest.  Results are cached per matrix signature in the text
// file named by XS_TUNE_CACHE (one "key chunk ell us" line per matrix).
// ---------------------------------------------------------------------------
static std::string tune_key(const xs_ctx *c) {
  // n, nnz and an FNV-1a hash of about 4096 sampled row lengths
  const int n = c->A.n;
  unsigned long long h = 1469598103934665603ull;
  const int step = std::max(1, n / 4096);
  for (int i = 0; i < n; i += step) {
    h ^= (unsigned long long)(c->h_rp[i + 1] - c->h_rp[i]);
    h *= 1099511628211ull;
  }
  char buf[96];
  snprintf(buf, sizeof(buf), "spmv:%d:%d:%016llx", n, c->A.nnz, h);
  return buf;
}

// XS_TUNE_CACHE: one line per tuned key, "key v0 v1 ...", the last line of
// a key winning
static bool tune_cache_get(const std::string &key, std::vector<double> &vals, size_t count) {
  const char *path = getenv("XS_TUNE_CACHE");
  if (!path) return false;
  FILE *f = fopen(path, "r");
  if (!f) return false;
  char line[512];
  bool hit = false;
  while (fgets(line, sizeof(line), f)) {
    char k[128];
    int used = 0;
    if (sscanf(line, "%127s%n", k, &used) != 1 || key != k) continue;
    std::vector<double> v;
    const char *q = line + used;
    double x;
    int m = 0;
    while (sscanf(q, "%lf%n", &x, &m) == 1) { v.push_back(x); q += m; }
    if (v.size() == count) { vals = v; hit = true; }
  }
  fclose(f);
  return hit;
}

static void tune_cache_put(const std::string &key, const std::vector<double> &vals) {
  const char *path = getenv("XS_TUNE_CACHE");
  if (!path) return;
  FILE *f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s", key.c_str());
  for (double v : vals) fprintf(f, " %.6g", v);
  fprintf(f, "\n");
  fclose(f);
}

// mean microseconds of one y = A x with the current format settings
static int time_spmv(xs_ctx *c, const double *x, double *y, int reps, double *us) {
  for (int r = 0; r < 3; r++) xs_op_spmv(c, x, y, EPI_Y, nullptr, nullptr, nullptr);
  XS_HIP(hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < reps; r++) xs_op_spmv(c, x, y, EPI_Y, nullptr, nullptr, nullptr);
  XS_HIP(hipEventRecord(c->ev1, c->stream));
  XS_HIP(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  XS_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *us = 1000.0 * ms / reps;
  return XS_OK;
}

// mean microseconds of one k-column ELL SpMM (Y = A X), cw columns per thread
static int time_spmm(xs_ctx *c, int k, const double *X, double *Y, int cw, int reps,
                     double *us) {
  for (int r = 0; r < 2; r++)
    xs_mspmm_ell(c->stream, c->ell, k, X, Y, EPI_Y, nullptr, nullptr, nullptr, nullptr, cw);
  XS_HIP(hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < reps; r++)
    xs_mspmm_ell(c->stream, c->ell, k, X, Y, EPI_Y, nullptr, nullptr, nullptr, nullptr, cw);
  XS_HIP(hipEventRecord(c->ev1, c->stream));
  XS_HIP(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  XS_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *us = 1000.0 * ms / reps;
  return XS_OK;
}

// mean microseconds of a lower plus an upper exact triangular solve (the
// solution vector refilled with the sentinel before each, as a solve does)
static int time_trsv(xs_ctx *c, const double *b, double *x, int reps, double *us) {
  const int n = c->A.n;
  auto both = [&]() {
    xs_fill_sentinel(c->stream, x, n);
    xs_trsv_lower(c->stream, c->ilu, b, x, c->counters + 2, nullptr);
    xs_fill_sentinel(c->stream, x, n);
    xs_trsv_upper(c->stream, c->ilu, b, x, c->counters + 4, nullptr);
  };
  both();
  XS_HIP(hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < reps; r++) both();
  XS_HIP(hipEventRecord(c->ev1, c->stream));
  XS_HIP(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  XS_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *us = 1000.0 * ms / reps;
  return xs_take_launch_guard() ? XS_ERR_STATE : XS_OK;
}

// n, nnz and an FNV-1a hash of about 4096 sampled row lengths of a pattern
static std::string shape_key(const char *kind, const int *hrp, int n, int nnz) {
  unsigned long long h = 1469598103934665603ull;
  const int step = std::max(1, n / 4096);
  for (int i = 0; i < n; i += step) {
    h ^= (unsigned long long)(hrp[i + 1] - hrp[i]);
    h *= 1099511628211ull;
  }
  char buf[128];
  snprintf(buf, sizeof(buf), "%s:%d:%d:%016llx", kind, n, nnz, h);
  return buf;
}

static int ensure_trsv_sched(xs_ctx *c, bool exact = false);

// The triangular solves of a set-up factor: persistent workgroups, gather
// width and group mode, each combination timed on a lower plus an upper
// solve; the fastest is kept (and cached by the factor's shape).
static int tune_trsv(xs_ctx *c) {
  int rc;
  if ((rc = ensure_trsv_sched(c, true))) return rc;
  xs_ilu &F = c->ilu;
  std::vector<int> hrp(F.n + 1);
  XS_HIP(hipStreamSynchronize(c->stream));
  XS_HIP(hipMemcpy(hrp.data(), F.rp, sizeof(int) * (F.n + 1), hipMemcpyDeviceToHost));
  const std::string key = shape_key("trsv", hrp.data(), F.n, F.nnz);
  std::vector<double> tv;
  int bwg = F.trsv_wg, bga = xs_ilu_gather(F), bpr = F.trsv_prog;
  double bus = 0.0;
  if (tune_cache_get(key, tv, 4)) {
    bwg = (int)tv[0]; bga = (int)tv[1]; bpr = (int)tv[2]; bus = tv[3];
  } else {
    double *b = xs_ctx_vec(c, 0), *x = xs_ctx_vec(c, 1);
    XS_HIP(hipMemsetAsync(b, 0, sizeof(double) * c->A.n, c->stream));
    bus = 1e30;
    static const int wgs[] = {128, 256, 512, 1024}, gathers[] = {4, 16};
    for (int wg : wgs)
      for (int ga : gathers)
        for (int pr = 0; pr < 2; ++pr) {
          F.trsv_wg = wg;
          F.trsv_gather = ga;
          F.trsv_prog = pr;
          double us;
          if ((rc = time_trsv(c, b, x, 3, &us))) return rc;
          if (us < bus) { bus = us; bwg = wg; bga = ga; bpr = pr; }
        }
    tune_cache_put(key, {(double)bwg, (double)bga, (double)bpr, bus});
  }
  c->trsv_wg = F.trsv_wg = bwg;
  c->trsv_gather = F.trsv_gather = bga;
  c->trsv_prog = F.trsv_prog = bpr;
  c->tune_trsv_us = bus;
  return XS_OK;
}

// The autotuner: for the SpMV, the LDS row-block size, the row-binned
// vector-CSR kernel's lanes per row and long-row threshold, and the ELL copy
// (with its k-RHS SpMM's columns per thread); for a set-up factor, the
// triangular solves (tune_trsv).  Every candidate is timed on this matrix;
// choices are cached in XS_TUNE_CACHE by the pattern's shape.
extern "C" int xs_autotune(xs_ctx *c) {
  if (!c) return XS_ERR_ARG;
  if (!c->mat_ready) return XS_ERR_STATE;
  int rc;
  if ((rc = xs_ctx_ensure_work(c, XS_NWORK))) return rc;
  const std::string key = tune_key(c);
  int best_chunk = 0, best_ell = 0, best_L = 0, best_mult = 0, best_cw = 0;
  double best_us = 0.0;
  std::vector<double> tv;
  struct keep_prof {  // no kernel timing while tuning; restored on every path
    xs_ctx *c;
    int kind;
    ~keep_prof() { c->prof_kind = kind; }
  } kp{c, c->prof_kind};
  c->prof_kind = 0;
  if (tune_cache_get(key, tv, 6)) {
    best_chunk = (int)tv[0]; best_ell = (int)tv[1]; best_us = tv[2];
    best_L = (int)tv[3]; best_mult = (int)tv[4]; best_cw = (int)tv[5];
  } else {
    double *x = xs_ctx_vec(c, 0), *y = xs_ctx_vec(c, 1);
    XS_HIP(hipMemsetAsync(x, 0, sizeof(double) * c->A.n, c->stream));
    const int reps = 20;
    best_us = 1e30;
    c->use_ell = 0;
    c->plan.vec = 0;
    static const int chunks[] = {512, 1024, 1536, 2048, 3072, 4096};
    for (int ch : chunks) {
      if ((rc = make_plan(c, c->A, c->plan, ch))) return rc;
      double us;
      if ((rc = time_spmv(c, x, y, reps, &us))) return rc;
      if (us < best_us) { best_us = us; best_chunk = ch; best_ell = 0; }
    }
    {  // row-binned vector CSR (chunk reported as 0): lanes per row around
       // the average row length, long rows past 4, 8 or 16 lane groups
      if ((rc = make_vec(c, c->h_rp, c->plan))) return rc;
      const int L0 = c->plan.vec_L;
      const int lanes[] = {L0 / 2, L0, 2 * L0}, mults[] = {4, 8, 16};
      c->plan.vec = 1;
      for (int L : lanes) {
        if (L < 2 || L > XS_WAVE) continue;
        for (int m : mults) {
          if ((rc = make_vec(c, c->h_rp, c->plan, L, m))) return rc;
          double us;
          if ((rc = time_spmv(c, x, y, reps, &us))) return rc;
          if (us < best_us) { best_us = us; best_chunk = 0; best_ell = 0; best_L = L; best_mult = m; }
        }
      }
      c->plan.vec = 0;
    }
    // ELL when padding stays under 50%
    int W = c->ell.v ? c->ell.W : xs_max_row_len(c->stream, c->A, c->d_int);
    if ((double)W * c->A.n <= 1.5 * c->A.nnz) {
      if ((rc = make_ell(c))) return rc;
      c->use_ell = 1;
      double us;
      if ((rc = time_spmv(c, x, y, reps, &us))) return rc;
      if (us < best_us) { best_us = us; best_ell = 1; }
      // the k-RHS SpMM over the same copy: columns per thread at k = 16
      const int k = 16;
      double *X = nullptr, *Y = nullptr;
      const size_t nk = (size_t)c->A.n * k;
      if ((rc = dalloc(&X, nk))) return rc;
      if ((rc = dalloc(&Y, nk))) { dfree(X); return rc; }
      XS_HIP(hipMemsetAsync(X, 0, sizeof(double) * nk, c->stream));
      double best_mm = 1e30;
      static const int cws[] = {2, 4};
      for (int cw : cws) {
        double mm;
        if ((rc = time_spmm(c, k, X, Y, cw, 5, &mm))) { dfree(X); dfree(Y); return rc; }
        if (mm < best_mm) { best_mm = mm; best_cw = cw; }
      }
      XS_HIP(hipStreamSynchronize(c->stream));
      dfree(X);
      dfree(Y);
    }
    tune_cache_put(key, {(double)best_chunk, (double)best_ell, best_us, (double)best_L,
                         (double)best_mult, (double)best_cw});
  }
  if ((rc = make_plan(c, c->A, c->plan, best_chunk > 0 ? best_chunk : xs_default_chunk(c->A))))
    return rc;
  if ((rc = make_vec(c, c->h_rp, c->plan, best_L, best_mult))) return rc;
  c->plan.vec = (best_chunk == 0 && !best_ell) ? 1 : 0;
  c->spmv_vec = c->plan.vec;
  c->use_ell = best_ell;
  if (best_ell && (rc = make_ell(c))) return rc;
  c->spmm_cw = best_cw;
  c->tune_chunk = best_chunk;
  c->tune_us = best_us;
  c->tune_L = c->plan.vec_L;
  c->tune_long_thr = c->plan.long_thr;
  c->tuned = true;
  if (c->pc_ready && xs_pc_factored(c->pc) && (rc = tune_trsv(c))) return rc;
  XS_HIP(hipStreamSynchronize(c->stream));
  if (xs_take_launch_guard()) return XS_ERR_STATE;
  return ensure_common(c);
}

static int setup_common(xs_ctx *c, int n, int nnz, int use_ell) {
  c->mixed_ready = c->mixed_mat_ready = false;
  c->use_ell = use_ell ? 1 : 0;
  c->A.n = n;
  c->A.nnz = nnz;
  int rc = build_spmv(c);
  if (rc) return rc;
  c->mat_ready = true;
  c->pc_ready = false;
  return ensure_common(c);
}

extern "C" int xs_create(xs_ctx **out) {
  if (!out) return XS_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    snprintf(xs_errbuf, sizeof(xs_errbuf), "no HIP device");
    *out = nullptr;
    return XS_ERR_DEVICE;
  }
  xs_ctx *c = new xs_ctx();
  if (hipGetDevice(&c->device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    *out = nullptr;
    return XS_ERR_DEVICE;
  }
  (void)hipEventCreate(&c->ev0);
  (void)hipEventCreate(&c->ev1);
  int rc = alloc_base(c);
  if (rc) {
    xs_destroy(c);
    *out = nullptr;
    return rc;
  }
  *out = c;
  return XS_OK;
}

extern "C" int xs_destroy(xs_ctx *c) {
  if (!c) return XS_OK;
  (void)hipStreamSynchronize(c->stream);
  free_matrix(c);
  free_work(c);
  dfree(c->st); dfree(c->counters); dfree(c->d_int); dfree(c->partial); dfree(c->resvec_d);
  if (c->h_state) (void)hipHostFree(c->h_state);
  xs_solvers_free(c);
  xs_mixed_free(c);
  for (auto &e : c->prof_ev) (void)hipEventDestroy(e);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return XS_OK;
}

extern "C" const char *xs_status_string(int s) {
  switch (s) {
    case XS_OK: return "ok";
    case XS_NOT_CONVERGED: return "maximum iterations reached";
    case XS_ERR_ARG: return "invalid argument";
    case XS_ERR_STATE: return "call out of order";
    case XS_ERR_UNSUPPORTED: return "unsupported";
    case XS_ERR_DEVICE: return xs_errbuf[0] ? xs_errbuf : "device error";
    case XS_ERR_BREAKDOWN: return "breakdown (non-finite residual)";
    case XS_ERR_PIVOT: return "zero or missing pivot";
    case XS_ERR_MEMORY: return "out of memory";
    case XS_ERR_TOO_LARGE: return "input larger than the fixed workspace";
  }
  return "unknown status";
}

extern "C" int xs_mat_setup(xs_ctx *c, int n, const int *rowptr, const int *colidx,
                            const double *val, const int *advanced_pam) {
  if (!c || n <= 0 || !rowptr || !colidx || !val) return XS_ERR_ARG;
  int use_ell = 0;
  if (advanced_pam) {
    if (advanced_pam[0] == 1) return XS_ERR_UNSUPPORTED;  // MPI distribution
    if (advanced_pam[2] != 1) return XS_ERR_UNSUPPORTED;  // CPU backend not provided
    if (advanced_pam[1] == 1) use_ell = 1;
    else if (advanced_pam[1] != 0) return XS_ERR_UNSUPPORTED;
  }
  const int base = rowptr[0];
  if (base != 0 && base != 1) return XS_ERR_ARG;
  const int nnz = rowptr[n] - base;
  if (nnz <= 0) return XS_ERR_ARG;
  (void)hipStreamSynchronize(c->stream);
  free_matrix(c);
  c->h_rp.resize(n + 1);
  c->h_ci.resize(nnz);
  std::vector<double> hv(val, val + nnz);
  for (int i = 0; i <= n; i++) c->h_rp[i] = rowptr[i] - base;
  for (int e = 0; e < nnz; e++) c->h_ci[e] = colidx[e] - base;
  for (int i = 0; i < n; i++)
    if (c->h_rp[i + 1] < c->h_rp[i]) return XS_ERR_ARG;
  analyse_pattern(c);
  if (!c->sorted) {
    // sort each row by column (values follow); rejects out-of-range columns
    for (int i = 0; i < n; i++) {
      int a = c->h_rp[i], b = c->h_rp[i + 1];
      std::vector<std::pair<int, double>> row;
      for (int e = a; e < b; e++) {
        if (c->h_ci[e] < 0 || c->h_ci[e] >= n) return XS_ERR_ARG;
        row.push_back({c->h_ci[e], hv[e]});
      }
      std::sort(row.begin(), row.end(),
                [](const std::pair<int, double> &x, const std::pair<int, double> &y) { return x.first < y.first; });
      for (int e = a; e < b; e++) { c->h_ci[e] = row[e - a].first; hv[e] = row[e - a].second; }
    }
    analyse_pattern(c);
    if (!c->sorted) {
      snprintf(xs_errbuf, sizeof(xs_errbuf), "duplicate column index within a row");
      return XS_ERR_ARG;
    }
  }
  int rc;
  c->own_matrix = true;
  if ((rc = dalloc(&c->A.rp, n + 1))) return rc;
  if ((rc = dalloc(&c->A.ci, nnz))) return rc;
  if ((rc = dalloc(&c->A.v, nnz))) return rc;
  XS_HIP(hipMemcpy(c->A.rp, c->h_rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
  XS_HIP(hipMemcpy(c->A.ci, c->h_ci.data(), sizeof(int) * nnz, hipMemcpyHostToDevice));
  XS_HIP(hipMemcpy(c->A.v, hv.data(), sizeof(double) * nnz, hipMemcpyHostToDevice));
  return setup_common(c, n, nnz, use_ell);
}

extern "C" int xs_mat_setup_device(xs_ctx *c, int n, int nnz, const int *d_rowptr,
                                   const int *d_colidx, const double *d_val, int use_ell) {
  if (!c || n <= 0 || nnz <= 0 || !d_rowptr || !d_colidx || !d_val) return XS_ERR_ARG;
  (void)hipStreamSynchronize(c->stream);
  free_matrix(c);
  c->own_matrix = false;
  c->A.rp = const_cast<int *>(d_rowptr);
  c->A.ci = const_cast<int *>(d_colidx);
  c->A.v = const_cast<double *>(d_val);
  // only the row pointers come to the host (the SpMV plan's row bins); the
  // column checks run on the device
  c->h_rp.resize(n + 1);
  c->h_ci.clear();
  XS_HIP(hipMemcpy(c->h_rp.data(), d_rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  if (c->h_rp[0] != 0 || c->h_rp[n] != nnz) return XS_ERR_ARG;
  for (int i = 0; i < n; i++)
    if (c->h_rp[i + 1] < c->h_rp[i]) return XS_ERR_ARG;
  if (!c->d_int) return XS_ERR_STATE;
  int flag = 0;
  XS_HIP(hipMemsetAsync(c->d_int, 0, sizeof(int), c->stream));
  xs_check_pattern(c->stream, n, d_rowptr, d_colidx, c->d_int);
  XS_HIP(hipMemcpyAsync(&flag, c->d_int, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  XS_HIP(hipStreamSynchronize(c->stream));
  c->sorted = !(flag & 1);
  c->has_diag = !(flag & 2);
  if (!c->sorted) return XS_ERR_ARG;
  return setup_common(c, n, nnz, use_ell);
}

extern "C" int xs_mat_update_values(xs_ctx *c, const double *val, int on_device) {
  if (!c || !c->mat_ready || !val) return XS_ERR_STATE;
  if (!c->own_matrix) return XS_ERR_STATE;
  XS_HIP(hipMemcpy(c->A.v, val, sizeof(double) * c->A.nnz,
                   on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
  if (c->use_ell) xs_build_ell(c->stream, c->A, c->ell.W, c->ell);
  if (c->have_At) xs_permute_vals(c->stream, c->A.nnz, c->perm_At, c->A.v, c->At.v);
  if (!c->pc_reuse) {   // else the factors of the earlier values stay
    c->pc_ready = false;
    c->tell_ready = c->tell32_ready = false;
  }
  c->mixed_ready = c->mixed_mat_ready = false;   // the FP32 copy of A follows the new values
  XS_HIP(hipStreamSynchronize(c->stream));
  return XS_OK;
}

extern "C" int xs_ksp_settype(xs_ctx *c, const char *t) {
  if (!c || !t) return XS_ERR_ARG;
  if (streq_any(t, {"Bicgstab", "bicgstab", "BICGSTAB"})) c->ksp = XS_KSP_BICGSTAB;
  else if (streq_any(t, {"Igcr", "igcr", "IGCR", "Gcr", "gcr", "GCR"})) c->ksp = XS_KSP_GCR;
  else if (streq_any(t, {"Cg", "cg", "CG", "Cgm", "cgm", "CGM"})) c->ksp = XS_KSP_CG;
  // the rest of the intel family (intel/mpisolver-C-v1.cpp)
  else if (streq_any(t, {"cg_left", "CG_LEFT"})) c->ksp = XS_KSP_CG_LEFT;
  else if (streq_any(t, {"Icg", "icg", "ICG"})) c->ksp = XS_KSP_ICG;
  else if (streq_any(t, {"Icgm", "icgm", "ICGM"})) c->ksp = XS_KSP_ICGM;
  else if (streq_any(t, {"Ibicgstab", "ibicgstab", "IBICGSTAB"})) c->ksp = XS_KSP_IBICGSTAB;
  else if (streq_any(t, {"gcr_left", "GCR_LEFT"})) c->ksp = XS_KSP_GCR_LEFT;
  else if (streq_any(t, {"igcr_left", "IGCR_LEFT"})) c->ksp = XS_KSP_IGCR_LEFT;
  else if (streq_any(t, {"Gmres", "gmres", "GMRES"})) c->ksp = XS_KSP_GMRES;
  else if (streq_any(t, {"Fgmres", "fgmres", "FGMRES"})) c->ksp = XS_KSP_FGMRES;
  else return XS_ERR_ARG;
  return XS_OK;
}

extern "C" int xs_ksp_set_absolute_tol(xs_ctx *c, double a) {
  if (!c) return XS_ERR_ARG;
  c->abstol = a;
  return XS_OK;
}

extern "C" int xs_ksp_setoption(xs_ctx *c, double tol, int maxiter, int restart) {
  if (!c || maxiter < 0 || restart < 0) return XS_ERR_ARG;
  c->tol = tol;
  c->maxiter = maxiter;
  c->restart = restart;
  return XS_OK;
}

extern "C" int xs_pc_settype(xs_ctx *c, const char *t) {
  if (!c || !t) return XS_ERR_ARG;
  int pc;
  if (streq_any(t, {"Jacobi", "jacobi", "JACOBI"})) pc = XS_PC_JACOBI;
  else if (streq_any(t, {"Ilu0", "ilu0", "ILU0", "Ilu0_mkl", "ilu0_mkl", "ILU0_MKL"})) pc = XS_PC_ILU0;
  else if (streq_any(t, {"Ilut", "ilut", "ILUT", "Ilut_mkl", "ilut_mkl", "ILUT_MKL"})) pc = XS_PC_ILUT;
  else if (streq_any(t, {"None", "none", "NONE"})) pc = XS_PC_NONE;
  else if (streq_any(t, {"Ic0", "ic0", "IC0"})) pc = XS_PC_IC0;
  else if (streq_any(t, {"Ict", "ict", "ICT"})) return XS_ERR_UNSUPPORTED;
  else return XS_ERR_ARG;
  if (pc != c->pc) free_precond(c);
  c->pc = pc;
  c->pc_ready = false;
  return XS_OK;
}

extern "C" int xs_pc_setilut(xs_ctx *c, double tol, int max_fill) {
  if (!c || tol < 0 || max_fill <= 0) return XS_ERR_ARG;
  c->ilut_tol = tol;
  c->ilut_fill = max_fill;
  c->pc_ready = false;
  return XS_OK;
}

extern "C" int xs_set_option(xs_ctx *c, const char *name, double value) {
  if (!c || !name) return XS_ERR_ARG;
  std::string k(name);
  if (k == "parilu_sweeps") { if (value < 0) return XS_ERR_ARG; c->parilu_sweeps = (int)value; c->pc_ready = false; }
  else if (k == "parilu_async") { c->parilu_async = value != 0; c->pc_ready = false; }
  else if (k == "check_every") { if (value < 1) return XS_ERR_ARG; c->check_every = (int)value; }
  else if (k == "fuse_p") { c->fuse_p = value != 0; }
  else if (k == "use_graph") { if (value < -1 || value > 1) return XS_ERR_ARG; c->use_graph = (int)value; }
  else if (k == "spmv_kernel") {   // -1 = automatic, 0 = LDS row blocks, 1 = vector CSR
    if (value != 0 && value != 1 && value != -1) return XS_ERR_ARG;
    c->spmv_vec = (int)value;
    if (c->ell_auto && value >= 0) { c->use_ell = 0; c->ell_auto = false; }
    c->mixed_ready = c->mixed_mat_ready = false;   // the FP32 SpMV follows the format
    if (c->mat_ready) {
      c->plan.vec = c->spmv_vec < 0 ? (regular_rows(c->h_rp, c->plan) ? 1 : 0) : c->spmv_vec;
      c->planT.vec = c->spmv_vec < 0 ? c->plan.vec : c->spmv_vec;
    }
  }
  else if (k == "trsv_workgroups") {   // persistent workgroups of the SpTRSV
    if (value < 1 || value > 65536) return XS_ERR_ARG;
    c->trsv_wg = (int)value;
    c->ilu.trsv_wg = c->trsv_wg;
  }
  else if (k == "spmv_dia") {   // 1: diagonal-slot copy for the FP64 SpMV when possible
    if (value < 0 || value > 1) return XS_ERR_ARG;
    c->spmv_dia = value != 0;
    if (c->ell.dv && !c->spmv_dia) { dfree(c->ell.dv); c->ell.nd = 0; }
    c->mixed_ready = c->mixed_mat_ready = false;   // the FP32 SpMV follows the format
  }
  else if (k == "sweep_dia") {   // 1: diagonal-slot copies of L and U for the FP64 sweeps
    if (value < 0 || value > 1) return XS_ERR_ARG;
    c->sweep_dia = value != 0;
    c->ilu.tell_shape = false;   // rebuilt (with or without them) before the next sweep
    c->tell_ready = c->tell32_ready = false;
  }
  else if (k == "parilu_dia") {   // 1: ParILU(0) sweeps over diagonal slots when possible
    if (value < 0 || value > 1) return XS_ERR_ARG;
    if ((value != 0) != c->parilu_dia) { c->parilu_dia = value != 0; free_precond(c); }
  }
  else if (k == "trsv_progressive") {   // -1 auto, 0 scan chain groups, 1 lane by lane
    if (value < -1 || value > 1) return XS_ERR_ARG;
    c->trsv_prog = (int)value;
    c->ilu.trsv_prog = c->trsv_prog;
  }
  else if (k == "trsv_gather") {   // 0 by factor width, 4 or 16 off-group entries per poll
    if (value != 0 && value != 4 && value != 16) return XS_ERR_ARG;
    c->trsv_gather = (int)value;
    c->ilu.trsv_gather = c->trsv_gather;
  }
  else if (k == "spmm_cw") {   // 0 auto, 2 or 4 columns per thread of the k-RHS ELL SpMM
    if (value != 0 && value != 2 && value != 4) return XS_ERR_ARG;
    c->spmm_cw = (int)value;
  }
  else if (k == "trsv_schedule") { c->trsv_schedule = value != 0; c->pc_ready = false; free_precond(c); }
  else if (k == "lsolve_sweeps") { if (value < -1) return XS_ERR_ARG; c->lsweeps = (int)value; c->mixed_ready = false; }
  else if (k == "usolve_sweeps") { if (value < -1) return XS_ERR_ARG; c->usweeps = (int)value; c->mixed_ready = false; }
  else if (k == "refactor") { c->pc_ready = false; c->mixed_ready = false; c->tell_ready = c->tell32_ready = false; }
  else if (k == "factor_storage") { if (value < 0 || value > 2) return XS_ERR_ARG; c->factor_type = (int)value; c->mixed_ready = false; }
  else if (k == "mixed_reliable") { c->mixed_reliable = value != 0; }
  else if (k == "inner_tol") {   // 0 = automatic (sqrt of the solve's rtol, at least 1e-6)
    if (!(value >= 0 && value < 1)) return XS_ERR_ARG;
    c->inner_tol = value;
  }
  else if (k == "parilu_scale") { c->parilu_scale = value != 0; c->pc_ready = false; }
  else if (k == "pc_reuse") { c->pc_reuse = value != 0; }
  else if (k == "parilut_steps") { if (value < 1) return XS_ERR_ARG; c->parilut_steps = (int)value; c->pc_ready = false; }
  else if (k == "parilut_small_sets") { c->parilut_small_sets = value != 0; c->pc_ready = false; }
  else if (k == "parilut_rows") {
    if (value != 0 && value != 128 && value != 256 && value != 512) return XS_ERR_ARG;
    c->parilut_rows = (int)value;
    c->pc_ready = false;
  }
  else if (k == "profile") { c->prof_kind = (int)value; for (int i = 0; i < 4; i++) { c->prof_ms[i] = 0; c->prof_cnt[i] = 0; } }
  else if (k == "precision") { if (value != 0 && value != 1) return XS_ERR_ARG; c->precision = (int)value; }
  else if (k == "use_ell") {
    if (!c->mat_ready) { c->use_ell = value != 0; return XS_OK; }
    int want = value != 0;
    if (want && !c->ell.v) {
      int rc = make_ell(c);
      if (rc) return rc;
      XS_HIP(hipStreamSynchronize(c->stream));
    }
    c->use_ell = want;
    c->ell_auto = false;
    c->mixed_ready = c->mixed_mat_ready = false;
  } else return XS_ERR_ARG;
  return XS_OK;
}

// ---------------------------------------------------------------------------
// preconditioner set-up
// ---------------------------------------------------------------------------
// SpTRSV work order (no reference counterpart).  The sync-free solve hands
// out 64-row groups in ticket order and a resident wave spins until its
// dependencies are published; with groups taken in row order, the resident
// waves (at most ~8k on the chip) fill up with groups whose dependencies are
// far from ready -- on the 128^3 7-point factors only ~32 planes fit, so
// the solve takes ~2x its dependency depth in hops.  Ordering the groups by
// dependency level (longest chain of group dependencies below them, from
// the factor pattern once per pattern) keeps the resident waves on the
// dependency frontier.  Every dependency of a group has a smaller level, so
// it still holds an earlier ticket: the no-deadlock argument is unchanged.
// This is an order for the tickets only -- no level-synchronous execution.
// Row groups of a triangular solve in solve order (L: ascending rows, U:
// descending): at most XS_WAVE rows, cut at the last chain break within
// reach -- a row that does not depend on the row before it in solve order --
// so that groups follow the matrix's own chains (grid lines of 100 rows give
// groups of 64 + 36 rows instead of groups straddling lines, each waiting on
// the one before it).  gb holds the G + 1 boundaries in solve order.
static void trsv_groups(int n, const int *rp, const int *ci, const int *dpos, bool upper,
                        std::vector<int> &gb) {
  auto brk = [&](int q) {   // position q does not depend on position q - 1
    if (!upper) return !(dpos[q] > rp[q] && ci[dpos[q] - 1] == q - 1);
    const int i = n - 1 - q;
    return !(dpos[i] + 1 < rp[i + 1] && ci[dpos[i] + 1] == i + 1);
  };
  gb.assign(1, upper ? n : 0);
  int p = 0;
  while (p < n) {
    const int reach = std::min(n, p + XS_WAVE);
    int cut = reach;
    if (reach < n) {
      for (int q = reach; q > p + 1; --q)
        if (brk(q)) { cut = q; break; }
    }
    p = cut;
    gb.push_back(upper ? n - p : p);
  }
}

// Dependency level of each group (longest chain of group dependencies below
// it) and the groups sorted by level: the ticket order of the solve.
static void group_order(int n, const int *rp, const int *ci, const int *dpos, bool upper,
                        const std::vector<int> &gb, std::vector<int> &order) {
  const int G = (int)gb.size() - 1;
  std::vector<int> rg(n), lev(G, 0);
  for (int g = 0; g < G; ++g) {
    const int r0 = upper ? gb[g + 1] : gb[g], r1 = upper ? gb[g] : gb[g + 1];
    for (int r = r0; r < r1; ++r) rg[r] = g;
  }
  int maxlev = 0;
  for (int g = 0; g < G; ++g) {
    const int r0 = upper ? gb[g + 1] : gb[g], r1 = upper ? gb[g] : gb[g + 1];
    int L = 0;
    for (int r = r0; r < r1; ++r) {
      if (!upper) {
        for (int e = rp[r]; e < dpos[r]; ++e) {
          const int col = ci[e];
          if (col >= r0) break;
          L = std::max(L, lev[rg[col]] + 1);
        }
      } else {
        for (int e = rp[r + 1] - 1; e > dpos[r]; --e) {
          const int col = ci[e];
          if (col < r1) break;
          L = std::max(L, lev[rg[col]] + 1);
        }
      }
    }
    lev[g] = L;
    maxlev = std::max(maxlev, L);
  }
  std::vector<int> start(maxlev + 2, 0);
  for (int g = 0; g < G; ++g) start[lev[g] + 1]++;
  for (int l = 0; l <= maxlev; ++l) start[l + 1] += start[l];
  order.resize(G);
  for (int g = 0; g < G; ++g) order[start[lev[g]]++] = g;
}

// host copies of a factor pattern -> row groups (c->ilu.lgb / ugb) and, with
// trsv_schedule, their level order (c->ilu.lsched / usched)
static int build_trsv_schedules(xs_ctx *c, const int *rp, const int *ci, const int *dpos) {
  xs_ilu &F = c->ilu;
  dfree(F.lsched);
  dfree(F.usched);
  dfree(F.lgb);
  dfree(F.ugb);
  F.lgroups = F.ugroups = 0;
  const int n = c->A.n;
  std::vector<int> gb, ord;
  int rc;
  for (int part = 0; part < 2; ++part) {
    const bool upper = part == 1;
    trsv_groups(n, rp, ci, dpos, upper, gb);
    const int G = (int)gb.size() - 1;
    int *&dgb = upper ? F.ugb : F.lgb;
    if ((rc = dalloc(&dgb, (size_t)G + 1))) return rc;
    XS_HIP(hipMemcpy(dgb, gb.data(), sizeof(int) * (G + 1), hipMemcpyHostToDevice));
    (upper ? F.ugroups : F.lgroups) = G;
    if (!c->trsv_schedule) continue;
    group_order(n, rp, ci, dpos, upper, gb, ord);
    int *&ds = upper ? F.usched : F.lsched;
    if ((rc = dalloc(&ds, (size_t)G))) return rc;
    XS_HIP(hipMemcpy(ds, ord.data(), sizeof(int) * G, hipMemcpyHostToDevice));
  }
  return XS_OK;
}

// The row groups of a factor pattern (ILU(0) / IC(0) on A's pattern,
// ParILUT's own), on the first solve that applies the factor by the exact
// triangular solve (sweep counts < 0): a host pass over a copy of the
// pattern.
static int ensure_trsv_sched(xs_ctx *c, bool exact) {
  xs_ilu &F = c->ilu;
  if (!F.sched_pending || (!exact && c->lsweeps >= 0 && c->usweeps >= 0)) return XS_OK;
  const int n = F.n, nnz = F.nnz;
  std::vector<int> hrp(n + 1), hci(nnz), hdp(n);
  XS_HIP(hipStreamSynchronize(c->stream));
  XS_HIP(hipMemcpy(hrp.data(), F.rp, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  XS_HIP(hipMemcpy(hci.data(), F.ci, sizeof(int) * nnz, hipMemcpyDeviceToHost));
  XS_HIP(hipMemcpy(hdp.data(), F.dpos, sizeof(int) * n, hipMemcpyDeviceToHost));
  const int rc = build_trsv_schedules(c, hrp.data(), hci.data(), hdp.data());
  if (rc == XS_OK) F.sched_pending = false;
  return rc;
}

// Update lists of the ParILU(0) / ParIC(0) sweeps (once per pattern, on
// the device): count the products of every entry, scan, fill.
// ParILU(0) over diagonal slots when A has at most XS_DIA_MAX distinct
// offsets (stencils): the slot pairs of every offset's products (host, from
// the offsets alone), the row masks and the CSR -> slot map (device).  Left
// off (F.pnd = 0: the update lists are used) otherwise.
static int build_parilu_dia(xs_ctx *c) {
  xs_ilu &F = c->ilu;
  xs_dia_offs off{};
  const int nd = xs_dia_offsets(c->stream, c->A, c->d_int, off);
  if (nd < 0) return XS_ERR_DEVICE;
  if (nd == 0) return XS_OK;
  xs_dia_pairs P{};
  P.diag = -1;
  for (int s = 0; s < nd; ++s)
    if (off.o[s] == 0) P.diag = s;
  if (P.diag < 0) return XS_OK;
  for (int s = 0; s < nd; ++s) {
    const int o = off.o[s];
    P.kind[s] = o < 0 ? 0 : o == 0 ? 1 : 2;
    P.np[s] = 0;
    for (int a = 0; a < nd; ++a) {   // offsets ascending: l's column ascending
      const int o1 = off.o[a];
      if (o1 >= std::min(o, 0)) break;
      for (int b = 0; b < nd; ++b)
        if (off.o[b] == o - o1) {
          P.pl[s][P.np[s]] = (signed char)a;
          P.pu[s][P.np[s]] = (signed char)b;
          P.np[s]++;
        }
    }
    if (P.kind[s] == 2 && P.np[s] == 0) P.kind[s] = 3;
  }
  const int n = c->A.n;
  const size_t len = (size_t)nd * n;
  int rc;
  if ((rc = dalloc(&F.pmask, n)) || (rc = dalloc(&F.dmap, c->A.nnz)) || (rc = dalloc(&F.das, len)) ||
      (rc = dalloc(&F.dlc, len)))
    return rc;
  XS_HIP(hipMemsetAsync(F.das, 0, sizeof(double) * len, c->stream));
  XS_HIP(hipMemsetAsync(F.dlc, 0, sizeof(double) * len, c->stream));
  F.pnd = nd;
  F.pld = n;
  F.poff = off;
  F.ppairs = P;
  xs_parilu_dia_map(c->stream, c->A, F);
  XS_HIP(hipStreamSynchronize(c->stream));
  return xs_take_launch_guard() ? XS_ERR_STATE : XS_OK;
}

static int build_parilu_lists(xs_ctx *c, bool ic) {
  xs_ilu &F = c->ilu;
  const int nnz = c->A.nnz;
  int rc, *cnt = nullptr, *bsum = nullptr, np = 0;
  auto done = [&](int r) { dfree(cnt); dfree(bsum); return r; };
  if ((rc = dalloc(&cnt, nnz)) || (rc = dalloc(&bsum, xs_scan_bsum_len(nnz))) ||
      (rc = dalloc(&F.pd, nnz)) || (rc = dalloc(&F.pp, (size_t)nnz + 1)))
    return done(rc);
  auto pairs = ic ? xs_ic_pairs : xs_parilu_pairs;
  pairs(c->stream, c->A, F, cnt, nullptr, nullptr, nullptr, F.pd);
  xs_scan_excl(c->stream, nnz, cnt, F.pp, bsum);
  if (hipMemcpyAsync(&np, F.pp + nnz, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return done(XS_ERR_DEVICE);
  if ((rc = dalloc(&F.pl, (size_t)std::max(np, 1))) || (rc = dalloc(&F.pu, (size_t)std::max(np, 1))))
    return done(rc);
  pairs(c->stream, c->A, F, nullptr, F.pp, F.pl, F.pu, nullptr);
  // active entries (a sweep can change them): flag, scan, compact
  int *flag = nullptr, *pos = nullptr;
  auto done2 = [&](int r) { dfree(flag); dfree(pos); return done(r); };
  if ((rc = dalloc(&flag, nnz)) || (rc = dalloc(&pos, (size_t)nnz + 1))) return done2(rc);
  xs_parilu_active_flags(c->stream, nnz, cnt, F.pd, flag);
  xs_scan_excl(c->stream, nnz, flag, pos, bsum);
  int na = 0;
  if (hipMemcpyAsync(&na, pos + nnz, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return done2(XS_ERR_DEVICE);
  if ((rc = dalloc(&F.act, (size_t)std::max(na, 1)))) return done2(rc);
  xs_compact_flagged(c->stream, nnz, flag, pos, F.act);
  F.nact = na;
  if ((rc = dalloc(&F.app, (size_t)na + 1)) || (rc = dalloc(&F.apd, (size_t)std::max(na, 1))) ||
      (rc = dalloc(&F.cmap, nnz)))
    return done2(rc);
  // sweep order: the lists' factor and pivot positions point into it
  xs_parilu_cmap(c->stream, nnz, na, flag, pos, F.cmap);
  xs_parilu_active_lists(c->stream, na, nnz, F.act, F.pp, F.pd, F.cmap, F.app, F.apd);
  xs_remap_idx(c->stream, np, F.cmap, F.pl);
  xs_remap_idx(c->stream, np, F.cmap, F.pu);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return done2(XS_ERR_DEVICE);
  if (xs_take_launch_guard()) return done2(XS_ERR_STATE);
  F.lists = ic ? 2 : 1;
  if (!ic && c->parilu_dia && (rc = build_parilu_dia(c))) return done2(rc);
  return done2(XS_OK);
}

// Column view of m (column, position) pairs on the device (xs_split.hip):
// a stable sort by column, the column pointers cp (ncol + 1), rows[q] =
// rowidx[pos[q]], the sorted positions (pos, when given) and perm[pos[q]] = q
// (when given).  key / val are overwritten.
static int device_columns(xs_ctx *c, int m, int ncol, unsigned *key, int *val, const int *rowidx,
                          int *cp, int *rows, int *pos, int *perm) {
  hipStream_t s = c->stream;
  unsigned *key2 = nullptr;
  int *val2 = nullptr, *hist = nullptr, *off = nullptr, *bsum = nullptr;
  auto done = [&](int r) {
    dfree(key2); dfree(val2); dfree(hist); dfree(off); dfree(bsum);
    return r;
  };
  const size_t H = xs_colsort_hist_len(m);
  int rc;
  if ((rc = dalloc(&key2, (size_t)std::max(m, 1))) || (rc = dalloc(&val2, (size_t)std::max(m, 1))) ||
      (rc = dalloc(&hist, H)) || (rc = dalloc(&off, H)) ||
      (rc = dalloc(&bsum, (size_t)xs_scan_bsum_len((int)(H - 1)))))
    return done(rc);
  const int which = xs_colsort(s, m, ncol, key, val, key2, val2, hist, off, bsum);
  const unsigned *sk = which ? key2 : key;
  const int *sv = which ? val2 : val;
  xs_key_ptr(s, m, ncol, sk, cp);
  xs_sorted_rows(s, m, rowidx, sv, rows, perm);
  if (pos && m > 0 &&
      hipMemcpyAsync(pos, sv, sizeof(int) * (size_t)m, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return done(XS_ERR_DEVICE);
  // the scratch is freed below: the sort must be done with it
  if (hipStreamSynchronize(s) != hipSuccess) return done(XS_ERR_DEVICE);
  if (xs_take_launch_guard()) return done(XS_ERR_STATE);
  return done(XS_OK);
}

// CSR transpose on the device: trp / tci (rows of the transpose sorted) and
// perm[e] = position of entry e in the transpose; rowidx = row of each entry.
static int device_transpose(xs_ctx *c, int n, int nnz, const int *ci, const int *rowidx, int *trp,
                            int *tci, int *perm) {
  unsigned *key = nullptr;
  int *val = nullptr, rc;
  if ((rc = dalloc(&key, (size_t)std::max(nnz, 1))) || (rc = dalloc(&val, (size_t)std::max(nnz, 1)))) {
    dfree(key);
    return rc;
  }
  xs_all_pairs(c->stream, nnz, ci, key, val);
  rc = device_columns(c, nnz, n, key, val, rowidx, trp, tci, nullptr, perm);
  dfree(key);
  dfree(val);
  return rc;
}

static int setup_ilu_pattern(xs_ctx *c, bool ic) {
  // Combined-factor layout on A's pattern plus the CSC of the upper triangle
  // used by the ParILU sweep, built on the device (the reference builds L CSR
  // / U CSC with a count pass and hipsparseDcsr2csc,
  // hip/Precond/par_ilu_hip/Asplit_gpu.cpp:4-49, 62-161).  A's rows are
  // sorted by this point (xs_mat_setup sorts them, xs_mat_setup_device
  // rejects unsorted ones).
  const int n = c->A.n, nnz = c->A.nnz;
  hipStream_t s = c->stream;
  xs_ilu &F = c->ilu;
  F.n = n; F.nnz = nnz; F.rp = c->A.rp; F.ci = c->A.ci;
  int rc;
  if ((rc = dalloc(&F.lu, nnz))) return rc;
  if ((rc = dalloc(&F.dpos, n))) return rc;
  if ((rc = dalloc(&F.dinv, n))) return rc;
  if ((rc = dalloc(&F.rowidx, nnz))) return rc;
  if ((rc = dalloc(&F.ucp, n + 1))) return rc;
  XS_HIP(hipMemsetAsync(c->d_int, 0, sizeof(int), s));
  xs_pt_rowidx(s, n, c->A.rp, F.rowidx);
  xs_pt_dpos(s, n, c->A.rp, c->A.ci, F.dpos, c->d_int);
  int bad = 0;
  XS_HIP(hipMemcpyAsync(&bad, c->d_int, sizeof(int), hipMemcpyDeviceToHost, s));
  XS_HIP(hipStreamSynchronize(s));
  if (bad) return XS_ERR_PIVOT;
  {
    int *cnt = nullptr, *uoff = nullptr, *bsum = nullptr, *val = nullptr;
    unsigned *key = nullptr;
    auto done = [&](int r) {
      dfree(cnt); dfree(uoff); dfree(bsum); dfree(val); dfree(key);
      return r;
    };
    if ((rc = dalloc(&cnt, n)) || (rc = dalloc(&uoff, n + 1)) ||
        (rc = dalloc(&bsum, xs_scan_bsum_len(n))))
      return done(rc);
    xs_upper_count(s, n, c->A.rp, F.dpos, cnt);
    xs_scan_excl(s, n, cnt, uoff, bsum);
    int nu = 0;
    if (hipMemcpyAsync(&nu, uoff + n, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return done(XS_ERR_DEVICE);
    if ((rc = dalloc(&F.uri, nu)) || (rc = dalloc(&F.upos, nu)) || (rc = dalloc(&key, nu)) ||
        (rc = dalloc(&val, nu)))
      return done(rc);
    xs_upper_pairs(s, n, c->A.rp, c->A.ci, F.dpos, uoff, key, val);
    if ((rc = device_columns(c, nu, n, key, val, F.rowidx, F.ucp, F.uri, F.upos, nullptr)))
      return done(rc);
    done(XS_OK);
  }
  if (ic) {
    // every entry needs its mirror: U = D L~^T lives on A's upper pattern
    if ((rc = dalloc(&F.mpos, nnz))) return rc;
    XS_HIP(hipMemsetAsync(c->d_int, 0, sizeof(int), c->stream));
    xs_ic_mirror(c->stream, c->A, F.rowidx, F.mpos, c->d_int);
    int asym = 0;
    XS_HIP(hipMemcpyAsync(&asym, c->d_int, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    XS_HIP(hipStreamSynchronize(c->stream));
    if (asym) {
      snprintf(xs_errbuf, sizeof(xs_errbuf), "IC0: the matrix pattern is not symmetric");
      return XS_ERR_UNSUPPORTED;
    }
  }
  if ((rc = build_parilu_lists(c, ic))) return rc;
  // the SpTRSV row groups and their order: on the first solve that applies
  // the factor by exact triangular solves (the sweep applies use none)
  F.sched_pending = true;
  return XS_OK;
}

// ParILU(0) (hip/Precond/gpuparilu.cpp:150-165, 20 sweeps by default).  With
// parilu_scale (default) the sweeps run on S A S, S = |diag A|^{-1/2}, and the
// factors are scaled back; ILU(0) is unchanged by the scaling, but from the
// reference's unscaled start the sweeps converge slowly or diverge on the
// 27-point lognormal systems.
static int factor_parilu(xs_ctx *c) {
  xs_ilu &F = c->ilu;
  if (F.lu && F.lists != 1) free_precond(c);   // a pattern set up for ParIC(0)
  if (!F.lu) {
    int rc = setup_ilu_pattern(c, false);
    if (rc) return rc;
  }
  if (!c->parilu_async && !c->lu_tmp) {
    int rc = dalloc(&c->lu_tmp, c->A.nnz);
    if (rc) return rc;
  }
  const double *sc = nullptr;
  if (c->parilu_scale) {
    if (!F.scale) {
      int rc = dalloc(&F.scale, c->A.n);
      if (rc) return rc;
    }
    xs_pt_scale_vec(c->stream, c->A.n, c->A.rp, c->A.ci, c->A.v, F.scale);
    sc = F.scale;
  }
  if (F.pnd > 0) {
    // the same over diagonal slots (no index lists)
    if (!c->parilu_async && !F.dlc2) {
      int rc = dalloc(&F.dlc2, (size_t)F.pnd * F.pld);
      if (rc) return rc;
      XS_HIP(hipMemsetAsync(F.dlc2, 0, sizeof(double) * (size_t)F.pnd * F.pld, c->stream));
    }
    double *cur = F.dlc, *nxt = c->parilu_async ? F.dlc : F.dlc2;
    xs_parilu_dia_start(c->stream, c->A, F, sc, c->parilu_async ? nullptr : F.dlc2);
    for (int s = 0; s < c->parilu_sweeps; s++) {
      xs_parilu_dia_sweep(c->stream, F, cur, nxt);
      std::swap(cur, nxt);
    }
    xs_parilu_dia_finish(c->stream, c->A, F, sc, cur);
  } else if (F.pp) {
    // start values: the (scaled) matrix in sweep order, in both buffers of a
    // deterministic run; entries off the active list keep them.  The last
    // iterate goes back to CSR order, scaled back, in one pass.
    int rc;
    if ((!F.as && (rc = dalloc(&F.as, c->A.nnz))) || (!F.lc && (rc = dalloc(&F.lc, c->A.nnz))))
      return rc;
    double *cur = F.lc, *nxt = c->parilu_async ? F.lc : c->lu_tmp;
    xs_parilu_start(c->stream, c->A, F, sc, c->parilu_async ? nullptr : c->lu_tmp);
    for (int s = 0; s < c->parilu_sweeps; s++) {
      xs_parilu_sweep_pairs(c->stream, F, cur, nxt);
      std::swap(cur, nxt);
    }
    xs_parilu_finish(c->stream, c->A, F, sc, cur);
  } else {
    xs_ilu_copy_values(c->stream, c->A, F);
    if (sc) xs_pt_rescale(c->stream, c->A.nnz, F.rowidx, c->A.ci, F.lu, F.scale, 0);
    for (int s = 0; s < c->parilu_sweeps; s++) {
      if (c->parilu_async) {
        xs_parilu_sweep(c->stream, c->A, F, nullptr, true, sc);
      } else {
        xs_parilu_sweep(c->stream, c->A, F, c->lu_tmp, false, sc);
        std::swap(F.lu, c->lu_tmp);
      }
    }
    if (sc) xs_pt_rescale(c->stream, c->A.nnz, F.rowidx, c->A.ci, F.lu, F.scale, 1);
  }
  XS_HIP(hipMemsetAsync(c->d_int, 0x7f, sizeof(int), c->stream));
  xs_ilu_dinv(c->stream, F, c->d_int);
  int bad = 0;
  XS_HIP(hipMemcpyAsync(&bad, c->d_int, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  XS_HIP(hipStreamSynchronize(c->stream));
  if (bad != 0x7f7f7f7f) return XS_ERR_PIVOT;
  return XS_OK;
}

// ParIC(0) (no reference counterpart on the GPU; the fixed point is the
// reference's IC(0), hip/Precond/ic0.cpp:29-67): parilu_sweeps sweeps of the
// lower triangle (asynchronous in place, or deterministic double-buffered),
// on S A S with parilu_scale, then the combined factors L~ | D L~^T on A's
// pattern (kernels in xs_ic.hip).
static int factor_paric(xs_ctx *c) {
  xs_ilu &F = c->ilu;
  if (F.lu && F.lists != 2) free_precond(c);   // a pattern set up for ParILU(0)
  int rc;
  if (!F.lu && (rc = setup_ilu_pattern(c, true))) {
    free_precond(c);
    return rc;
  }
  if (!c->parilu_async && !c->lu_tmp && (rc = dalloc(&c->lu_tmp, c->A.nnz))) return rc;
  const double *sc = nullptr;
  if (c->parilu_scale) {
    if (!F.scale && (rc = dalloc(&F.scale, c->A.n))) return rc;
    xs_pt_scale_vec(c->stream, c->A.n, c->A.rp, c->A.ci, c->A.v, F.scale);
    sc = F.scale;
  }
  if ((!F.as && (rc = dalloc(&F.as, c->A.nnz))) || (!F.lc && (rc = dalloc(&F.lc, c->A.nnz))))
    return rc;
  double *cur = F.lc, *nxt = c->parilu_async ? F.lc : c->lu_tmp;
  xs_ic_start(c->stream, c->A, F, sc, c->parilu_async ? nullptr : c->lu_tmp);
  for (int s = 0; s < c->parilu_sweeps; s++) {
    xs_ic_sweep_pairs(c->stream, F, cur, nxt);
    std::swap(cur, nxt);
  }
  xs_ic_finish(c->stream, c->A, F, F.mpos, sc, cur);
  XS_HIP(hipMemsetAsync(c->d_int, 0x7f, sizeof(int), c->stream));
  xs_ilu_dinv(c->stream, F, c->d_int);
  int bad = 0;
  XS_HIP(hipMemcpyAsync(&bad, c->d_int, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  XS_HIP(hipStreamSynchronize(c->stream));
  if (bad != 0x7f7f7f7f) return XS_ERR_PIVOT;
  return XS_OK;
}

// ParILUT (hip/Precond/parilut.cpp:621-1206, precon_gpuilut.cpp:5: 10 outer
// steps): kernels in xs_parilut.hip.  Fill factor = ilut_fill (final keep
// rate per factor relative to A's strict parts), relative drop tolerance =
// ilut_tol.  The factor pattern is owned by the context (rp_own/ci_own).
int xs_parilut_factor(xs_ctx *c) {
  hipStream_t s = c->stream;
  const int n = c->A.n, nnz = c->A.nnz;
  xs_ilu &F = c->ilu;
  // free a previous factor
  dfree(F.lu); dfree(F.dpos); dfree(F.dinv); dfree(F.rowidx); dfree(F.rp_own); dfree(F.ci_own);
  dfree(F.ucp); dfree(F.uri); dfree(F.upos);
  dfree(F.pp); dfree(F.pl); dfree(F.pu); dfree(F.pd); dfree(F.as); dfree(F.act);   // ILU(0) lists only
  dfree(F.pmask); dfree(F.dmap); dfree(F.das); dfree(F.dlc); dfree(F.dlc2);
  F.pnd = 0;
  dfree(F.app); dfree(F.apd); dfree(F.cmap); dfree(F.lc); dfree(F.mpos);
  dfree(F.lsched); dfree(F.usched); dfree(F.lgb); dfree(F.ugb);
  F.lists = 0;
  F.lgroups = F.ugroups = 0;
  F.sched_pending = false;
  F.nact = 0;
  F.tell_shape = false;   // a new pattern: the sweep copies are re-shaped
  int rc;
  double *dtol = nullptr, *thr = nullptr;
  int *cnt = nullptr, *bsum = nullptr, *ovf = nullptr, *lrows = nullptr, *nlong = nullptr;
  unsigned long long *sel = nullptr, *cnt2 = nullptr, *list = nullptr, *ckey = nullptr;
  unsigned *hist = nullptr, *lcnt = nullptr, *lstat = nullptr;
  int *crp = nullptr, *cci = nullptr, *crow = nullptr, *cdpos = nullptr;
  double *cv = nullptr, *cv2 = nullptr, *fv2 = nullptr;
  auto cleanup = [&]() {
    dfree(dtol); dfree(thr); dfree(cnt); dfree(bsum); dfree(ovf); dfree(sel); dfree(cnt2);
    dfree(lrows); dfree(nlong);
    dfree(list); dfree(hist); dfree(lcnt); dfree(crp); dfree(cci); dfree(crow); dfree(cdpos);
    dfree(cv); dfree(cv2); dfree(fv2); dfree(lstat); dfree(ckey);
  };
#define PT_TRY(x) do { if ((rc = (x))) { cleanup(); return rc; } } while (0)
  PT_TRY(F.scale ? XS_OK : dalloc(&F.scale, n));
  PT_TRY(dalloc(&dtol, n));
  PT_TRY(dalloc(&thr, 2));
  PT_TRY(dalloc(&cnt, n));
  PT_TRY(dalloc(&bsum, xs_scan_bsum_len(n)));
  PT_TRY(dalloc(&ovf, 1));
  if (c->parilut_small_sets) {
    PT_TRY(dalloc(&lrows, n));
    PT_TRY(dalloc(&nlong, 1));
  }
  PT_TRY(dalloc(&sel, xs_pt_select_words()));
  PT_TRY(dalloc(&cnt2, 2));
  PT_TRY(dalloc(&list, xs_pt_list_len()));
  PT_TRY(dalloc(&hist, 512));
  PT_TRY(dalloc(&lcnt, 2));
  PT_TRY(dalloc(&lstat, 8));
  PT_TRY(dalloc(&crp, n + 1));
  PT_TRY(dalloc(&cdpos, n));
  // the factor ends near fill x nnz(A) entries (the keep counts): its
  // buffers are sized for that at once; the candidate buffers grow only
  // (to twice the need), so few steps pay for an allocation
  const size_t fcap = (size_t)((c->ilut_fill + 0.25) * (double)nnz) + (size_t)n;
  PT_TRY(dalloc(&F.rp_own, n + 1));
  PT_TRY(dalloc(&F.ci_own, fcap));
  PT_TRY(dalloc(&F.lu, fcap));
  PT_TRY(dalloc(&F.rowidx, fcap));
  PT_TRY(dalloc(&fv2, fcap));
  PT_TRY(dalloc(&F.dpos, n));
  PT_TRY(dalloc(&F.dinv, n));
  size_t cap_c = 0, cap_ci = fcap, cap_lu = fcap, cap_row = fcap, cap_v2 = fcap;
  auto grow = [](size_t need, size_t &cap, auto **pp, size_t factor) -> int {
    if (need <= cap) return XS_OK;
    dfree(*pp);
    const size_t want = need * factor;
    const int r = dalloc(pp, want);
    cap = r == XS_OK ? want : 0;
    return r;
  };
  XS_HIP(hipMemsetAsync(ovf, 0, sizeof(int), s));
  // initial factor = A (scaled): L = strict lower part, U = upper part
  XS_HIP(hipMemcpyAsync(F.rp_own, c->A.rp, sizeof(int) * (n + 1), hipMemcpyDeviceToDevice, s));
  XS_HIP(hipMemcpyAsync(F.ci_own, c->A.ci, sizeof(int) * nnz, hipMemcpyDeviceToDevice, s));
  XS_HIP(hipMemcpyAsync(F.lu, c->A.v, sizeof(double) * nnz, hipMemcpyDeviceToDevice, s));
  xs_pt_rowidx(s, n, F.rp_own, F.rowidx);
  xs_pt_scale_vec(s, n, c->A.rp, c->A.ci, c->A.v, F.scale);
  xs_pt_rescale(s, nnz, F.rowidx, F.ci_own, F.lu, F.scale, 0);
  xs_pt_dtol(s, n, c->A.rp, c->A.v, c->ilut_tol, dtol);
  xs_pt_dpos(s, n, F.rp_own, F.ci_own, F.dpos, ovf);
  // entries of A's lower part with the diagonal and of its upper part with
  // the diagonal: the bases of the keep counts (l0nnz / u0nnz,
  // hip/Precond/parilut.cpp:681-682)
  unsigned long long base[2] = {0, 0};
  xs_pt_count_parts(s, n, F.rp_own, F.dpos, cnt2);
  XS_HIP(hipMemcpyAsync(base, cnt2, sizeof(base), hipMemcpyDeviceToHost, s));
  XS_HIP(hipStreamSynchronize(s));
  const long nl0 = (long)base[0], nu0 = (long)base[1];
  int fnnz = nnz;
  const int steps = c->parilut_steps;
  const double fill = (double)c->ilut_fill;
  for (int step = 0; step < steps; step++) {
    // keep-rate ramp (hip/Precond/parilut.cpp:692) and keep counts (:1008-1013)
    const double rate = 1.0 + (fill - 1.0) * (double)(step + 1) / (double)steps;
    long keepL = (long)(rate * (double)nl0), keepU = (long)(rate * (double)nu0);
    // 1. candidate pattern
    xs_pt_symbolic(s, n, c->A.rp, c->A.ci, F.rp_own, F.ci_own, F.dpos, cnt, ovf, lrows, nlong);
    xs_scan_excl(s, n, cnt, crp, bsum);
    xs_pt_len_stats(s, n, cnt, lstat);
    int hdr[2] = {0, 0};
    unsigned cst[4];
    XS_HIP(hipMemcpyAsync(&hdr[0], crp + n, sizeof(int), hipMemcpyDeviceToHost, s));
    XS_HIP(hipMemcpyAsync(&hdr[1], ovf, sizeof(int), hipMemcpyDeviceToHost, s));
    XS_HIP(hipMemcpyAsync(cst, lstat, sizeof(cst), hipMemcpyDeviceToHost, s));
    XS_HIP(hipStreamSynchronize(s));
    const int crows = c->parilut_rows ? c->parilut_rows : xs_pt_pick_rows(n, cst);
    if (hdr[1] & 1) { snprintf(xs_errbuf, sizeof(xs_errbuf), "ParILUT: pattern overflow"); cleanup(); return XS_ERR_MEMORY; }
    if (hdr[1] & 2) { snprintf(xs_errbuf, sizeof(xs_errbuf), "ParILUT: a row without its diagonal"); cleanup(); return XS_ERR_PIVOT; }
    const int nc = hdr[0];
    {
      size_t c1 = cap_c, c2 = cap_c, c3 = cap_c, c4 = cap_c, c5 = cap_c;
      PT_TRY(grow((size_t)nc, c1, &cci, 2));
      PT_TRY(grow((size_t)nc, c2, &crow, 2));
      PT_TRY(grow((size_t)nc, c3, &cv, 2));
      PT_TRY(grow((size_t)nc, c4, &cv2, 2));
      PT_TRY(grow((size_t)nc, c5, &ckey, 2));
      cap_c = c1;
    }
    // 2-3. the candidates with the old values embedded, and a sweep over them
    xs_pt_fill_embed(s, n, c->A.rp, c->A.ci, c->A.v, F.scale, F.rp_own, F.ci_own, F.lu, F.dpos,
                     crp, cci, crow, cv, cdpos, ovf, crows, xs_pt_rows_over(crows, cst));
    xs_pt_sweep(s, n, crp, cci, cdpos, cv, cv2, c->A.rp, c->A.ci, c->A.v, F.scale, crows, ckey);
    std::swap(cv, cv2);
    // 4. thresholds (the part counts stay on the device); nothing is
    // dropped by rank while the candidates number at most round(rate nnz(A))
    // (hip/Precond/parilut.cpp:926-931)
    if ((long)nc - std::lround(rate * (double)nnz) <= 0) keepL = keepU = LONG_MAX;
    xs_pt_count_parts(s, n, crp, cdpos, cnt2);
    xs_pt_select(s, nc, ckey, n, cnt2, keepL, keepU, sel, hist, lcnt, list, thr);
    // compaction into the new factor
    xs_pt_keep(s, false, n, crp, cci, cdpos, cv, ckey, thr, dtol, cnt, nullptr, nullptr, nullptr,
               nullptr, nullptr, nullptr);
    xs_scan_excl(s, n, cnt, F.rp_own, bsum);
    xs_pt_len_stats(s, n, cnt, lstat);
    unsigned fst[4];
    XS_HIP(hipMemcpyAsync(&fnnz, F.rp_own + n, sizeof(int), hipMemcpyDeviceToHost, s));
    XS_HIP(hipMemcpyAsync(fst, lstat, sizeof(fst), hipMemcpyDeviceToHost, s));
    XS_HIP(hipStreamSynchronize(s));
    const int frows = c->parilut_rows ? c->parilut_rows : xs_pt_pick_rows(n, fst);
    // the old factor was last read by the embed (the stream is idle here)
    PT_TRY(grow((size_t)fnnz, cap_ci, &F.ci_own, 2));
    PT_TRY(grow((size_t)fnnz, cap_lu, &F.lu, 2));
    PT_TRY(grow((size_t)fnnz, cap_row, &F.rowidx, 2));
    PT_TRY(grow((size_t)fnnz, cap_v2, &fv2, 2));
    xs_pt_keep(s, true, n, crp, cci, cdpos, cv, ckey, thr, dtol, nullptr, F.rp_own, F.ci_own, F.lu,
               F.rowidx, F.dpos, ovf);
    // 5. sweep over the kept pattern
    xs_pt_sweep(s, n, F.rp_own, F.ci_own, F.dpos, F.lu, fv2, c->A.rp, c->A.ci, c->A.v, F.scale,
                frows, nullptr);
    std::swap(F.lu, fv2);
    std::swap(cap_lu, cap_v2);
  }
  xs_pt_rescale(s, fnnz, F.rowidx, F.ci_own, F.lu, F.scale, 1);
  F.n = n;
  F.nnz = fnnz;
  F.rp = F.rp_own;
  F.ci = F.ci_own;
  XS_HIP(hipMemsetAsync(c->d_int, 0x7f, sizeof(int), s));
  xs_ilu_dinv(s, F, c->d_int);
  int bad = 0, ov = 0;
  XS_HIP(hipMemcpyAsync(&bad, c->d_int, sizeof(int), hipMemcpyDeviceToHost, s));
  XS_HIP(hipMemcpyAsync(&ov, ovf, sizeof(int), hipMemcpyDeviceToHost, s));
  XS_HIP(hipStreamSynchronize(s));
  cleanup();
#undef PT_TRY
  if (ov & 1) return XS_ERR_MEMORY;
  if ((ov & 2) || bad != 0x7f7f7f7f) return XS_ERR_PIVOT;
  // row groups of the SpTRSV: built from the new pattern on the first exact solve
  F.sched_pending = true;
  return XS_OK;
}

extern "C" int xs_pc_setup(xs_ctx *c) {
  if (!c) return XS_ERR_ARG;
  if (!c->mat_ready) return XS_ERR_STATE;
  if (c->pc_ready) return XS_OK;
  int rc = XS_OK;
  XS_HIP(hipEventRecord(c->ev0, c->stream));
  switch (c->pc) {
    case XS_PC_NONE: break;
    case XS_PC_JACOBI: {
      if (!c->jac_dinv && (rc = dalloc(&c->jac_dinv, c->A.n))) return rc;
      XS_HIP(hipMemsetAsync(c->d_int, 0x7f, sizeof(int), c->stream));
      xs_jacobi_dinv(c->stream, c->A, c->jac_dinv, c->d_int);
      int bad = 0;
      XS_HIP(hipMemcpyAsync(&bad, c->d_int, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      XS_HIP(hipStreamSynchronize(c->stream));
      if (bad != 0x7f7f7f7f) rc = XS_ERR_PIVOT;
      break;
    }
    case XS_PC_ILU0:
      rc = factor_parilu(c);
      break;
    case XS_PC_IC0:
      rc = factor_paric(c);
      break;
    case XS_PC_ILUT:
      rc = xs_parilut_factor(c);
      break;
    default:
      return XS_ERR_STATE;
  }
  XS_HIP(hipEventRecord(c->ev1, c->stream));
  XS_HIP(hipEventSynchronize(c->ev1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->t_pc_ms = ms;
  if (xs_take_launch_guard()) rc = XS_ERR_STATE;
  if (rc == XS_OK) c->pc_ready = true;
  c->ilu.have_t = false;
  c->mixed_ready = false;   // FP32 copies of A and of the factors are rebuilt
  c->tell_ready = c->tell32_ready = false;   // and the sweep copies of the factors
  return rc;
}

// ---------------------------------------------------------------------------
// kernel timing: event pairs from a pool, summed after the solve
// ---------------------------------------------------------------------------
static const int XS_PROF_POOL = 16384;

void xs_prof_begin(xs_ctx *c, int kind) {
  if (!(c->prof_kind & (1 << kind))) return;
  if (c->prof_ev.empty()) {
    c->prof_ev.resize(XS_PROF_POOL);
    for (auto &e : c->prof_ev) (void)hipEventCreate(&e);
  }
  if (c->prof_used + 2 > XS_PROF_POOL) return;
  (void)hipEventRecord(c->prof_ev[c->prof_used], c->stream);
  c->prof_next_kind = kind;
}

void xs_prof_end(xs_ctx *c, int kind) {
  if (!(c->prof_kind & (1 << kind))) return;
  if (c->prof_used + 2 > XS_PROF_POOL) return;
  (void)hipEventRecord(c->prof_ev[c->prof_used + 1], c->stream);
  c->prof_kinds.push_back(kind);
  c->prof_used += 2;
}

void xs_prof_collect(xs_ctx *c) {
  for (int i = 0; i < c->prof_used; i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->prof_ev[i], c->prof_ev[i + 1]) == hipSuccess) {
      int k = c->prof_kinds[i / 2];
      c->prof_ms[k] += ms;
      c->prof_cnt[k] += 1;
    }
  }
  c->prof_used = 0;
  c->prof_kinds.clear();
}

// ---------------------------------------------------------------------------
// operators used by the solver loops
// ---------------------------------------------------------------------------
void xs_op_spmv(xs_ctx *c, const double *x, double *y, int epi, const double *b,
                const double *w, const xs_reduce *red) {
  xs_prof_begin(c, 1);
  if (c->use_t)
    xs_launch_spmv_csr(c->stream, c->At, c->planT, x, y, epi, b, w, red);
  else if (c->use_ell && c->ell.v)
    xs_launch_spmv_ell(c->stream, c->ell, x, y, epi, b, w, red);
  else
    xs_launch_spmv_csr(c->stream, c->A, c->plan, x, y, epi, b, w, red);
  xs_prof_end(c, 1);
}

// out = T^{-1} b approximated by k Jacobi sweeps from 0 (k = 0 and k = 1
// coincide: b for L, b / u_ii for U); ping-pong between out and w so that the
// last sweep lands in out.
static void ilu_sweeps(xs_ctx *c, bool upper, int k, const double *b, double *out, double *w,
                       const xs_state *st) {
  const int m = k < 1 ? 1 : k;
  double *cur = ((m - 1) % 2 == 0) ? out : w;
  xs_tri_sweep(c->stream, c->ilu, upper, b, nullptr, cur, st);
  for (int j = 2; j <= m; ++j) {
    double *nxt = (cur == out) ? w : out;
    xs_tri_sweep(c->stream, c->ilu, upper, b, cur, nxt, st);
    cur = nxt;
  }
}

// ELL copies of L and U for the sweep apply (pattern and values of the
// current factors); built before a solve so that no allocation happens inside
// a captured iteration batch.
static int ensure_tri_ell(xs_ctx *c) {
  if (c->tell_ready) return XS_OK;
  xs_ilu &F = c->ilu;
  int rc, w[2];
  if (!F.tell_shape) {   // widths once per factor pattern
    if (xs_tri_widths(c->stream, F, c->d_int, &w[0], &w[1])) return XS_ERR_DEVICE;
    for (int part = 0; part < 2; ++part) {
      xs_tell &E = part ? F.tu : F.tl;
      const int ld = (F.n + 1) & ~1;   // even: the FP32 sweeps load row pairs
      const size_t need = (size_t)w[part] * ld;
      if (need > E.cap) {
        dfree(E.ci); dfree(E.v); dfree(E.vlo); dfree(E.co); dfree(E.vs);
        E.vlo_bytes = 0;
        E.cap = 0;
        if ((rc = dalloc(&E.ci, need)) || (rc = dalloc(&E.v, need))) return rc;
        if (part == 1 && (rc = dalloc(&E.vs, need))) return rc;
        E.cap = need;
      }
      E.W = w[part];
      E.ld = ld;
    }
  }
  for (int part = 0; part < 2; ++part) xs_build_tell(c->stream, F, part == 1, part ? F.tu : F.tl);
  xs_tell_scale_cols(c->stream, F.tu, F.dinv);   // U's values for the fused first two sweeps
  if (!F.tell_shape) {   // offsets and group widths depend on the pattern only
    for (int part = 0; part < 2; ++part) {
      xs_tell &E = part ? F.tu : F.tl;
      if ((rc = make_offsets(c, F.n, E.ld, E.W, E.ci, &E.co))) return rc;
      dfree(E.ww);
      if ((rc = dalloc(&E.ww, ((size_t)F.n + XS_WAVE - 1) / XS_WAVE))) return rc;
      xs_tri_wave_widths(c->stream, F, part == 1, E.ww);
      // diagonal slots for the FP64 sweeps when the part has few offsets
      dfree(E.dv);
      dfree(E.dvs);
      E.nd = 0;
      if (c->sweep_dia && E.W > 0) {
        xs_dia_offs off{};
        const int nd = xs_tri_dia_offsets(c->stream, F, part == 1, c->d_int, off);
        if (nd < 0) return XS_ERR_DEVICE;
        if (nd > 0) {
          if ((rc = dalloc(&E.dv, (size_t)nd * E.ld))) return rc;
          if (part == 1 && (rc = dalloc(&E.dvs, (size_t)nd * E.ld))) return rc;
          E.nd = nd;
          E.doff = off;
        }
      }
    }
    F.tell_shape = true;
  }
  for (int part = 0; part < 2; ++part) {
    xs_tell &E = part ? F.tu : F.tl;
    if (E.nd > 0) xs_build_tri_dia(c->stream, F, part == 1, E);
  }
  XS_HIP(hipStreamSynchronize(c->stream));
  if (xs_take_launch_guard()) return XS_ERR_STATE;
  c->tell_ready = true;
  c->tell32_ready = false;
  c->mixed_ready = false;   // the FP32 copies are refreshed with the sweep values
  return XS_OK;
}

// k >= 2 sweeps over the ELL copy: sweeps 1 and 2 fused, then k - 2 more,
// ping-ponging so that the last lands in out (the results of ilu_sweeps, to
// rounding: the fused U pass multiplies by u_ij / u_jj where ilu_sweeps
// multiplies u_ij by (b_j / u_jj); L and the later passes are identical).
// With ru (L only): b is r, and the first pass forms s = r - gama u, writing
// it to sv, which the later passes then read as b.
static void ilu_sweeps_ell(xs_ctx *c, bool upper, int k, const double *b, double *out,
                           double *w, const xs_state *st, const double *ru = nullptr,
                           double *sv = nullptr) {
  double *cur = ((k - 2) % 2 == 0) ? out : w;
  if (ru) {
    xs_tri_sweep_ell_s(c->stream, c->ilu, b, ru, sv, cur, st);
    b = sv;
  } else {
    xs_tri_sweep_ell(c->stream, c->ilu, upper, true, b, nullptr, cur, st);
  }
  for (int j = 3; j <= k; ++j) {
    double *nxt = (cur == out) ? w : out;
    xs_tri_sweep_ell(c->stream, c->ilu, upper, false, b, cur, nxt, st);
    cur = nxt;
  }
}

static void apply_sweeps(xs_ctx *c, bool upper, int k, const double *b, double *out, double *w,
                         const xs_state *st) {
  if (k >= 2 && c->tell_ready) ilu_sweeps_ell(c, upper, k, b, out, w, st);
  else ilu_sweeps(c, upper, k, b, out, w, st);
}

// y = M^{-1} x on the device (tmp: scratch vector for the ILU forward solve)
void xs_op_pc(xs_ctx *c, const double *x, double *y, double *tmp, const xs_state *st) {
  switch (c->pc) {
    case XS_PC_JACOBI:
      xs_scale(c->stream, c->A.n, c->jac_dinv, x, y);
      break;
    case XS_PC_ILU0:
    case XS_PC_ILUT:
    case XS_PC_IC0:
      if (c->use_t) {
        xs_fill_sentinel(c->stream, tmp, c->A.n);
        xs_trsv_lower_t(c->stream, c->ilu, x, tmp, c->counters + 2, st);
        xs_fill_sentinel(c->stream, y, c->A.n);
        xs_trsv_upper_t(c->stream, c->ilu, tmp, y, c->counters + 4, st);
        break;
      }
      xs_prof_begin(c, 2);
      if (c->lsweeps >= 0) {
        apply_sweeps(c, false, c->lsweeps, x, tmp, xs_ctx_vec(c, 10), st);
      } else {
        xs_fill_sentinel(c->stream, tmp, c->A.n);
        xs_trsv_lower(c->stream, c->ilu, x, tmp, c->counters + 2, st);
      }
      xs_prof_end(c, 2);
      xs_prof_begin(c, 3);
      if (c->usweeps >= 0) {
        apply_sweeps(c, true, c->usweeps, tmp, y, xs_ctx_vec(c, 10), st);
      } else {
        xs_fill_sentinel(c->stream, y, c->A.n);
        xs_trsv_upper(c->stream, c->ilu, tmp, y, c->counters + 4, st);
      }
      xs_prof_end(c, 3);
      break;
    default:
      xs_copy(c->stream, c->A.n, x, y);
  }
}

bool xs_pc_fuse_l(const xs_ctx *c) {
  return xs_pc_factored(c->pc) && !c->use_t && c->lsweeps >= 2 && c->tell_ready;
}

// BiCGSTAB's p = r + omega (pin - alpha u) into pout, then y = M^{-1} p with
// the update formed inside the first L pass (one kernel less, same values).
void xs_op_pc_p(xs_ctx *c, const double *r, const double *u, const double *pin, double *pout,
                double *y, double *tmp, xs_state *st) {
  double *w = xs_ctx_vec(c, 10);
  const int k = c->lsweeps;
  double *cur = ((k - 2) % 2 == 0) ? tmp : w;
  xs_prof_begin(c, 2);
  xs_tri_sweep_ell_p(c->stream, c->ilu, r, u, pin, pout, cur, st);
  for (int j = 3; j <= k; ++j) {
    double *nxt = (cur == tmp) ? w : tmp;
    xs_tri_sweep_ell(c->stream, c->ilu, false, false, pout, cur, nxt, st);
    cur = nxt;
  }
  xs_prof_end(c, 2);
  xs_prof_begin(c, 3);
  if (c->usweeps >= 0) {
    apply_sweeps(c, true, c->usweeps, tmp, y, w, st);
  } else {
    xs_fill_sentinel(c->stream, y, c->A.n);
    xs_trsv_upper(c->stream, c->ilu, tmp, y, c->counters + 4, st);
  }
  xs_prof_end(c, 3);
}

// BiCGSTAB's s = r - gama u followed by y = M^{-1} s.  With an ILU applied
// by two or more L sweeps over the ELL copies, s is formed inside the first
// L pass (one kernel less, same values); otherwise the s update runs first.
void xs_op_pc_s(xs_ctx *c, const double *r, const double *u, double *sv, double *y, double *tmp,
                xs_state *st) {
  const bool fuse = xs_pc_fuse_l(c) && st != nullptr;
  if (!fuse) {
    xs_bicg_s(c->stream, c->A.n, r, u, sv, nullptr, nullptr, st);
    xs_op_pc(c, sv, y, tmp, st);
    return;
  }
  xs_prof_begin(c, 2);
  ilu_sweeps_ell(c, false, c->lsweeps, r, tmp, xs_ctx_vec(c, 10), st, u, sv);
  xs_prof_end(c, 2);
  xs_prof_begin(c, 3);
  if (c->usweeps >= 0) {
    apply_sweeps(c, true, c->usweeps, tmp, y, xs_ctx_vec(c, 10), st);
  } else {
    xs_fill_sentinel(c->stream, y, c->A.n);
    xs_trsv_upper(c->stream, c->ilu, tmp, y, c->counters + 4, st);
  }
  xs_prof_end(c, 3);
}

// ---------------------------------------------------------------------------
// solve
// ---------------------------------------------------------------------------
// Solve epilogue: stop the timer, collect kernel timings, and turn a refused
// launch or an expired triangular-solve wait into an error.  Returns rc
// (the solver's status) when nothing went wrong.
static int finish_solve(xs_ctx *c, int rc) {
  XS_HIP(hipEventRecord(c->ev1, c->stream));
  XS_HIP(hipEventSynchronize(c->ev1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->t_solve_ms = ms;
  xs_prof_collect(c);
  if (xs_take_launch_guard()) {
    snprintf(xs_errbuf, sizeof(xs_errbuf), "internal: a kernel launch found a null buffer");
    return XS_ERR_STATE;
  }
  // wait-expired flags: [0] grid reductions, [6]/[8] lower/upper triangular
  // solves, [10] tall-skinny reductions
  unsigned fl[12] = {0};
  XS_HIP(hipMemcpy(fl, c->counters, sizeof(fl), hipMemcpyDeviceToHost));
  if (fl[0] | fl[6] | fl[8] | fl[10]) {
    XS_HIP(hipMemset(c->counters, 0, sizeof(fl)));
    snprintf(xs_errbuf, sizeof(xs_errbuf), "%s wait expired",
             (fl[6] | fl[8]) ? "triangular solve dependency" : "grid reduction");
    return XS_ERR_DEVICE;
  }
  return rc;
}

// common checks and lazy set-up before a solve
static int prepare_solve(xs_ctx *c) {
  int rc;
  if (!c->mat_ready) return XS_ERR_STATE;
  if (c->ksp == XS_KSP_NONE) return XS_ERR_STATE;
  if (c->maxiter <= 0) return XS_ERR_STATE;
  if (!c->pc_ready && (rc = xs_pc_setup(c))) return rc;
  if ((rc = ensure_common(c))) return rc;
  if (xs_pc_factored(c->pc) && (c->lsweeps >= 2 || c->usweeps >= 2) &&
      (rc = ensure_tri_ell(c)))
    return rc;
  return ensure_trsv_sched(c);
}

static int run_solve(xs_ctx *c, const double *d_b, double *d_x) {
  int rc;
  if ((rc = prepare_solve(c))) return rc;
  if (c->resvec_cap < c->maxiter + 2) {
    dfree(c->resvec_d);
    if ((rc = dalloc(&c->resvec_d, (size_t)c->maxiter + 2))) return rc;
    c->resvec_cap = c->maxiter + 2;
  }
  XS_HIP(hipEventRecord(c->ev0, c->stream));
  if (c->precision == 1) {
    rc = xs_solve_mixed(c, d_b, d_x);
  } else switch (c->ksp) {
    case XS_KSP_CG: rc = xs_solve_cg(c, d_b, d_x); break;
    case XS_KSP_BICGSTAB: rc = xs_solve_bicgstab(c, d_b, d_x); break;
    case XS_KSP_GCR: rc = xs_solve_gcr(c, d_b, d_x); break;
    case XS_KSP_CG_LEFT: rc = xs_solve_cg_left(c, d_b, d_x, false); break;
    case XS_KSP_ICG: rc = xs_solve_cg_left(c, d_b, d_x, true); break;
    case XS_KSP_ICGM: rc = xs_solve_icgm(c, d_b, d_x); break;
    case XS_KSP_IBICGSTAB: rc = xs_solve_ibicgstab(c, d_b, d_x); break;
    case XS_KSP_GCR_LEFT: rc = xs_solve_gcr(c, d_b, d_x, true, false); break;
    case XS_KSP_IGCR_LEFT: rc = xs_solve_gcr(c, d_b, d_x, true, true); break;
    case XS_KSP_GMRES: rc = xs_solve_gmres(c, d_b, d_x, false); break;
    case XS_KSP_FGMRES: rc = xs_solve_gmres(c, d_b, d_x, true); break;
    default: return XS_ERR_STATE;
  }
  if ((rc = finish_solve(c, rc))) {
    if (rc < 0 && rc != XS_ERR_BREAKDOWN) return rc;
  }
  // results
  int nres = std::min(c->precision == 1 ? c->outer_iters + 1 : c->iters + 1, c->resvec_cap);
  c->resvec.resize(nres);
  XS_HIP(hipMemcpy(c->resvec.data(), c->resvec_d, sizeof(double) * nres, hipMemcpyDeviceToHost));
  return rc;
}

extern "C" int xs_solve_device(xs_ctx *c, const double *d_rhs, double *d_sol) {
  if (!c || !d_rhs || !d_sol) return XS_ERR_ARG;
  // the caller's producers of d_rhs/d_sol may run on any stream
  XS_HIP(hipDeviceSynchronize());
  return run_solve(c, d_rhs, d_sol);
}

// ---------------------------------------------------------------------------
// k right-hand sides
// ---------------------------------------------------------------------------
static int run_multi(xs_ctx *c, int k, const double *d_b, double *d_x, int layout) {
  int rc;
  if ((rc = prepare_solve(c))) return rc;
  if (c->precision != 0) {
    snprintf(xs_errbuf, sizeof(xs_errbuf), "block solve: FP64 only (precision 0)");
    return XS_ERR_UNSUPPORTED;
  }
  if (c->ksp != XS_KSP_BICGSTAB && c->ksp != XS_KSP_CG) {
    snprintf(xs_errbuf, sizeof(xs_errbuf), "block solve: BiCGSTAB or CG only");
    return XS_ERR_UNSUPPORTED;
  }
  if ((rc = xs_multi_reserve(c, k, 3))) return rc;   // 9-10: column-major copies, 11: sweeps
  const int n = c->A.n;
  const double *B = d_b;
  double *X = d_x;
  if (layout == XS_RHS_COLMAJOR) {
    double *mB = xs_multi_vec(c, k, 9), *mX = xs_multi_vec(c, k, 10);
    xs_mlayout(c->stream, n, k, d_b, mB, true);
    xs_mlayout(c->stream, n, k, d_x, mX, true);
    B = mB;
    X = mX;
  }
  XS_HIP(hipEventRecord(c->ev0, c->stream));
  rc = xs_solve_block(c, k, B, X);
  if (layout == XS_RHS_COLMAJOR && (rc >= 0 || rc == XS_ERR_BREAKDOWN))
    xs_mlayout(c->stream, n, k, X, d_x, false);
  return finish_solve(c, rc);
}

extern "C" int xs_solve_multi_device(xs_ctx *c, int nrhs, const double *d_rhs, double *d_sol,
                                     int layout) {
  if (!c || !d_rhs || !d_sol || nrhs < 1 || nrhs > XS_MAX_RHS) return XS_ERR_ARG;
  if (layout != XS_RHS_COLMAJOR && layout != XS_RHS_INTERLEAVED) return XS_ERR_ARG;
  XS_HIP(hipDeviceSynchronize());
  return run_multi(c, nrhs, d_rhs, d_sol, layout);
}

extern "C" int xs_solve_multi(xs_ctx *c, int nrhs, const double *rhs, double *sol, int layout) {
  if (!c || !rhs || !sol || nrhs < 1 || nrhs > XS_MAX_RHS) return XS_ERR_ARG;
  if (layout != XS_RHS_COLMAJOR && layout != XS_RHS_INTERLEAVED) return XS_ERR_ARG;
  if (!c->mat_ready) return XS_ERR_STATE;
  const size_t m = (size_t)c->A.n * nrhs;
  double *db = nullptr, *dx = nullptr;
  int rc = dalloc(&db, m);
  if (rc) return rc;
  if ((rc = dalloc(&dx, m))) { dfree(db); return rc; }
  if (hipMemcpy(db, rhs, sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dx, sol, sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess) {
    dfree(db); dfree(dx);
    return XS_ERR_DEVICE;
  }
  rc = run_multi(c, nrhs, db, dx, layout);
  if ((rc >= 0 || rc == XS_ERR_BREAKDOWN) &&
      hipMemcpy(sol, dx, sizeof(double) * m, hipMemcpyDeviceToHost) != hipSuccess)
    rc = XS_ERR_DEVICE;
  dfree(db);
  dfree(dx);
  return rc;
}

extern "C" int xs_get_multi_results(const xs_ctx *c, int *iters, double *res0, double *resl,
                                    int cap) {
  if (!c) return XS_ERR_ARG;
  const int k = std::min(c->mk, cap);
  for (int j = 0; j < k; ++j) {
    if (iters) iters[j] = c->m_iters[j];
    if (res0) res0[j] = c->m_res0[j];
    if (resl) resl[j] = c->m_resl[j];
  }
  return c->mk;
}

extern "C" int xs_get_multi_resvec(const xs_ctx *c, int col, double *out, int len) {
  if (!c || !out || col < 0 || col >= c->mk) return XS_ERR_ARG;
  const int cnt = std::min(len, c->m_iters[col] + 1);
  for (int i = 0; i < cnt; ++i) out[i] = c->m_resvec[(size_t)i * c->mk + col];
  return cnt;
}

extern "C" int xs_solve(xs_ctx *c, const double *rhs, double *sol) {
  if (!c || !rhs || !sol) return XS_ERR_ARG;
  if (!c->mat_ready) return XS_ERR_STATE;
  int rc;
  const int n = c->A.n;
  if (c->hb_n != n) {
    dfree(c->hb); dfree(c->hx);
    if ((rc = dalloc(&c->hb, n))) return rc;
    if ((rc = dalloc(&c->hx, n))) return rc;
    c->hb_n = n;
  }
  XS_HIP(hipMemcpy(c->hb, rhs, sizeof(double) * n, hipMemcpyHostToDevice));
  XS_HIP(hipMemcpy(c->hx, sol, sizeof(double) * n, hipMemcpyHostToDevice));
  rc = run_solve(c, c->hb, c->hx);
  if (rc < 0 && rc != XS_ERR_BREAKDOWN) return rc;
  XS_HIP(hipMemcpy(sol, c->hx, sizeof(double) * n, hipMemcpyDeviceToHost));
  return rc;
}

extern "C" int xs_get_iterations(const xs_ctx *c) { return c ? c->iters : -1; }
extern "C" double xs_get_residual(const xs_ctx *c) { return c ? c->resl : -1.0; }
extern "C" double xs_get_initial_residual(const xs_ctx *c) { return c ? c->res0 : -1.0; }
extern "C" int xs_get_resvec(const xs_ctx *c, double *out, int len) {
  if (!c || !out || len < 0) return XS_ERR_ARG;
  int m = std::min<int>(len, (int)c->resvec.size());
  for (int i = 0; i < m; i++) out[i] = c->resvec[i];
  return m;
}
extern "C" double xs_get_time_ms(const xs_ctx *c, const char *phase) {
  if (!c || !phase) return -1.0;
  if (!strcmp(phase, "pc_setup")) return c->t_pc_ms;
  if (!strcmp(phase, "solve")) return c->t_solve_ms;
  if (!strcmp(phase, "tune:chunk")) return c->tuned ? c->tune_chunk : -1.0;
  if (!strcmp(phase, "tune:ell")) return c->tuned ? c->use_ell : -1.0;
  if (!strcmp(phase, "tune:us")) return c->tuned ? c->tune_us : -1.0;
  if (!strcmp(phase, "tune:vec_lanes")) return c->tuned ? c->tune_L : -1.0;
  if (!strcmp(phase, "tune:long_thr")) return c->tuned ? c->tune_long_thr : -1.0;
  if (!strcmp(phase, "tune:spmm_cw")) return c->tuned ? c->spmm_cw : -1.0;
  if (!strcmp(phase, "tune:trsv_wg")) return c->tune_trsv_us > 0 ? c->trsv_wg : -1.0;
  if (!strcmp(phase, "tune:trsv_gather")) return c->tune_trsv_us > 0 ? c->trsv_gather : -1.0;
  if (!strcmp(phase, "tune:trsv_prog")) return c->tune_trsv_us > 0 ? c->trsv_prog : -1.0;
  if (!strcmp(phase, "tune:trsv_us")) return c->tune_trsv_us > 0 ? c->tune_trsv_us : -1.0;
  // SpMV kernel in use: 0 LDS row blocks, 1 vector CSR, 2 ELL
  if (!strcmp(phase, "spmv:kernel"))
    return !c->mat_ready ? -1.0 : (c->use_ell && c->ell.v) ? 2.0 : (double)c->plan.vec;
  // 1 when the ELL copy carries 16-bit column offsets
  if (!strcmp(phase, "spmv:idx16")) return c->ell.co ? 1.0 : 0.0;
  if (!strcmp(phase, "spmv:dia")) return (c->use_ell && c->ell.dv) ? (double)c->ell.nd : 0.0;
  if (!strcmp(phase, "pc:dia")) return c->tell_ready ? (double)c->ilu.tl.nd : -1.0;
  if (!strcmp(phase, "parilu:dia")) return (double)c->ilu.pnd;
  // 1 when both sweep copies of the factors carry 16-bit offsets (-1: not built)
  if (!strcmp(phase, "pc:idx16"))
    return !c->tell_ready ? -1.0
                          : ((c->ilu.tl.co || c->ilu.tl.W == 0) && (c->ilu.tu.co || c->ilu.tu.W == 0)) ? 1.0 : 0.0;
  static const char *kn[4] = {"", "spmv", "trsv_lower", "trsv_upper"};
  for (int k = 1; k < 4; k++) {
    std::string tot = std::string("total:") + kn[k], cnt = std::string("count:") + kn[k];
    if (tot == phase) return c->prof_ms[k];
    if (cnt == phase) return (double)c->prof_cnt[k];
  }
  return -1.0;
}

// ---------------------------------------------------------------------------
// transposes for the adjoint (no reference counterpart)
// ---------------------------------------------------------------------------
int xs_build_transpose(xs_ctx *c) {
  if (c->have_At) return XS_OK;
  const int n = c->A.n, nnz = c->A.nnz;
  int rc;
  dfree(c->At.rp); dfree(c->At.ci); dfree(c->At.v); dfree(c->perm_At); dfree(c->planT.blk_row);
  if ((rc = dalloc(&c->At.rp, n + 1)) || (rc = dalloc(&c->At.ci, nnz)) ||
      (rc = dalloc(&c->At.v, nnz)) || (rc = dalloc(&c->perm_At, nnz)))
    return rc;
  // the transpose on the device; only its row pointers come back (the
  // row-length bins of its SpMV plan)
  {
    int *row = nullptr;
    if ((rc = dalloc(&row, nnz))) return rc;
    xs_pt_rowidx(c->stream, n, c->A.rp, row);
    rc = device_transpose(c, n, nnz, c->A.ci, row, c->At.rp, c->At.ci, c->perm_At);
    dfree(row);
    if (rc) return rc;
  }
  std::vector<int> trp(n + 1);
  XS_HIP(hipMemcpy(trp.data(), c->At.rp, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  c->At.n = n;
  c->At.nnz = nnz;
  xs_permute_vals(c->stream, nnz, c->perm_At, c->A.v, c->At.v);
  if ((rc = make_plan(c, c->At, c->planT, xs_default_chunk(c->At)))) return rc;
  if ((rc = make_vec(c, trp, c->planT))) return rc;
  c->planT.vec = c->spmv_vec < 0 ? (regular_rows(trp, c->planT) ? 1 : 0) : c->spmv_vec;
  XS_HIP(hipStreamSynchronize(c->stream));
  if (xs_take_launch_guard()) return XS_ERR_STATE;
  c->have_At = true;
  return ensure_common(c);
}

int xs_build_factor_transpose(xs_ctx *c) {
  xs_ilu &F = c->ilu;
  if (F.have_t) return XS_OK;
  const int n = F.n, nnz = F.nnz;
  int rc;
  dfree(F.trp); dfree(F.tci); dfree(F.tdpos); dfree(F.tlu); dfree(F.tperm);
  dfree(F.tlgb); dfree(F.tugb); dfree(F.tlsched); dfree(F.tusched);
  F.tlgroups = F.tugroups = 0;
  if ((rc = dalloc(&F.trp, n + 1)) || (rc = dalloc(&F.tci, nnz)) || (rc = dalloc(&F.tdpos, n)) ||
      (rc = dalloc(&F.tlu, nnz)) || (rc = dalloc(&F.tperm, nnz)))
    return rc;
  // the transposed pattern and its pivot positions on the device (every
  // factor row holds its pivot, so every transposed row does)
  if ((rc = device_transpose(c, n, nnz, F.ci, F.rowidx, F.trp, F.tci, F.tperm))) return rc;
  XS_HIP(hipMemsetAsync(c->d_int, 0, sizeof(int), c->stream));
  xs_pt_dpos(c->stream, n, F.trp, F.tci, F.tdpos, c->d_int);
  // row groups and level order of the transposed triangular solves (host,
  // from a copy of the transposed pattern)
  std::vector<int> trp(n + 1), tci(nnz), tdpos(n);
  XS_HIP(hipStreamSynchronize(c->stream));
  XS_HIP(hipMemcpy(trp.data(), F.trp, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  XS_HIP(hipMemcpy(tci.data(), F.tci, sizeof(int) * nnz, hipMemcpyDeviceToHost));
  XS_HIP(hipMemcpy(tdpos.data(), F.tdpos, sizeof(int) * n, hipMemcpyDeviceToHost));
  {
    std::vector<int> gb, ord;
    for (int part = 0; part < 2; ++part) {
      const bool upper = part == 1;
      trsv_groups(n, trp.data(), tci.data(), tdpos.data(), upper, gb);
      const int G = (int)gb.size() - 1;
      int *&dgb = upper ? F.tugb : F.tlgb;
      if ((rc = dalloc(&dgb, (size_t)G + 1))) return rc;
      XS_HIP(hipMemcpy(dgb, gb.data(), sizeof(int) * (G + 1), hipMemcpyHostToDevice));
      (upper ? F.tugroups : F.tlgroups) = G;
      if (!c->trsv_schedule) continue;
      group_order(n, trp.data(), tci.data(), tdpos.data(), upper, gb, ord);
      int *&ds = upper ? F.tusched : F.tlsched;
      if ((rc = dalloc(&ds, (size_t)G))) return rc;
      XS_HIP(hipMemcpy(ds, ord.data(), sizeof(int) * G, hipMemcpyHostToDevice));
    }
  }
  xs_permute_vals(c->stream, nnz, F.tperm, F.lu, F.tlu);
  XS_HIP(hipStreamSynchronize(c->stream));
  if (xs_take_launch_guard()) return XS_ERR_STATE;
  F.have_t = true;
  return XS_OK;
}

// y = M^{-T} x: U^T (lower, 1/u_ii) then L^T (upper, unit diagonal)
int xs_op_pc_t(xs_ctx *c, const double *x, double *y, double *tmp, hipStream_t s) {
  if (xs_pc_factored(c->pc)) {
    int rc = xs_build_factor_transpose(c);
    if (rc) return rc;
    xs_fill_sentinel(s, tmp, c->A.n);
    xs_trsv_lower_t(s, c->ilu, x, tmp, c->counters + 2, nullptr);
    xs_fill_sentinel(s, y, c->A.n);
    xs_trsv_upper_t(s, c->ilu, tmp, y, c->counters + 4, nullptr);
  } else if (c->pc == XS_PC_JACOBI) {
    xs_scale(s, c->A.n, c->jac_dinv, x, y);
  } else {
    xs_copy(s, c->A.n, x, y);
  }
  XS_HIP(hipGetLastError());
  return xs_take_launch_guard() ? XS_ERR_STATE : XS_OK;
}

// ---------------------------------------------------------------------------
// kernel-level entry points
// ---------------------------------------------------------------------------
// Kernel-level entry points run on the caller's stream (0 = the default
// stream), so they order with the caller's own work on that stream.
static hipStream_t pick(xs_ctx *, void *s) { return (hipStream_t)s; }

extern "C" int xs_spmv(xs_ctx *c, const double *x, double *y, void *stream) {
  if (!c || !x || !y) return XS_ERR_ARG;
  if (!c->mat_ready) return XS_ERR_STATE;
  hipStream_t keep = c->stream;
  c->stream = pick(c, stream);
  xs_op_spmv(c, x, y, EPI_Y, nullptr, nullptr, nullptr);
  c->stream = keep;
  XS_HIP(hipGetLastError());
  return xs_take_launch_guard() ? XS_ERR_STATE : XS_OK;
}

extern "C" int xs_bmax(xs_ctx *c, const double *b, const double *x, double *y, void *stream) {
  if (!c || !b || !x || !y) return XS_ERR_ARG;
  if (!c->mat_ready) return XS_ERR_STATE;
  hipStream_t keep = c->stream;
  c->stream = pick(c, stream);
  xs_op_spmv(c, x, y, EPI_BMAX, b, nullptr, nullptr);
  c->stream = keep;
  XS_HIP(hipGetLastError());
  return xs_take_launch_guard() ? XS_ERR_STATE : XS_OK;
}

extern "C" int xs_spmv_t(xs_ctx *c, const double *x, double *y, void *stream) {
  if (!c || !x || !y) return XS_ERR_ARG;
  if (!c->mat_ready) return XS_ERR_STATE;
  int rc = xs_build_transpose(c);
  if (rc) return rc;
  xs_launch_spmv_csr(pick(c, stream), c->At, c->planT, x, y, EPI_Y, nullptr, nullptr, nullptr);
  XS_HIP(hipGetLastError());
  return xs_take_launch_guard() ? XS_ERR_STATE : XS_OK;
}

extern "C" int xs_pc_apply(xs_ctx *c, const double *x, double *y, void *stream) {
  if (!c || !x || !y) return XS_ERR_ARG;
  int rc;
  if (!c->pc_ready && (rc = xs_pc_setup(c))) return rc;
  if ((rc = xs_ctx_ensure_work(c, XS_NWORK))) return rc;
  if (xs_pc_factored(c->pc) && (c->lsweeps >= 2 || c->usweeps >= 2) &&
      (rc = ensure_tri_ell(c)))
    return rc;
  if ((rc = ensure_trsv_sched(c))) return rc;
  hipStream_t keep = c->stream;
  c->stream = pick(c, stream);
  xs_op_pc(c, x, y, xs_ctx_vec(c, XS_NWORK - 1), nullptr);
  c->stream = keep;
  XS_HIP(hipGetLastError());
  return xs_take_launch_guard() ? XS_ERR_STATE : XS_OK;
}

extern "C" int xs_pc_apply_t(xs_ctx *c, const double *x, double *y, void *stream) {
  if (!c || !x || !y) return XS_ERR_ARG;
  int rc;
  if (!c->pc_ready && (rc = xs_pc_setup(c))) return rc;
  if ((rc = xs_ctx_ensure_work(c, XS_NWORK))) return rc;
  return xs_op_pc_t(c, x, y, xs_ctx_vec(c, XS_NWORK - 1), pick(c, stream));
}

// Y = A X for k interleaved columns (n x k row-major device blocks).
extern "C" int xs_spmm(xs_ctx *c, int nrhs, const double *x, double *y, void *stream) {
  if (!c || !x || !y || nrhs < 1 || nrhs > XS_MAX_RHS) return XS_ERR_ARG;
  if (!c->mat_ready) return XS_ERR_STATE;
  const void *vecs[2] = {x, y};
  if (c->use_ell && xs_mspmm_ell_ok(c->ell, nrhs, vecs, 2))
    xs_mspmm_ell(pick(c, stream), c->ell, nrhs, x, y, EPI_Y, nullptr, nullptr, nullptr, nullptr,
                 c->spmm_cw);
  else
    xs_mspmm(pick(c, stream), c->A, nrhs, x, y, EPI_Y, nullptr, nullptr, nullptr, nullptr);
  XS_HIP(hipGetLastError());
  return xs_take_launch_guard() ? XS_ERR_STATE : XS_OK;
}

// Y = M^{-1} X for k interleaved columns.
extern "C" int xs_pc_apply_multi(xs_ctx *c, int nrhs, const double *x, double *y, void *stream) {
  if (!c || !x || !y || nrhs < 1 || nrhs > XS_MAX_RHS) return XS_ERR_ARG;
  if (!c->mat_ready) return XS_ERR_STATE;
  int rc;
  if (!c->pc_ready && (rc = xs_pc_setup(c))) return rc;
  if ((rc = ensure_common(c))) return rc;
  hipStream_t s = pick(c, stream);
  const int n = c->A.n;
  switch (c->pc) {
    case XS_PC_JACOBI:
      xs_mscale(s, n, nrhs, c->jac_dinv, x, y);
      break;
    case XS_PC_ILU0:
    case XS_PC_ILUT:
    case XS_PC_IC0: {
      if ((rc = xs_multi_reserve(c, nrhs, 1))) return rc;
      if ((c->lsweeps >= 2 || c->usweeps >= 2) && (rc = ensure_tri_ell(c))) return rc;
      if ((rc = ensure_trsv_sched(c))) return rc;
      xs_multi_ilu_apply(c, s, nrhs, x, y, xs_multi_vec(c, nrhs, 8), xs_multi_vec(c, nrhs, 9),
                         nullptr);
      break;
    }
    default:
      XS_HIP(hipMemcpyAsync(y, x, sizeof(double) * (size_t)n * nrhs, hipMemcpyDeviceToDevice, s));
  }
  XS_HIP(hipGetLastError());
  return xs_take_launch_guard() ? XS_ERR_STATE : XS_OK;
}

extern "C" int xs_trsv(xs_ctx *c, int which, const double *b, double *x, void *stream) {
  if (!c || !b || !x || (which != 0 && which != 1)) return XS_ERR_ARG;
  int rc;
  if (!xs_pc_factored(c->pc)) return XS_ERR_STATE;
  if (!c->pc_ready && (rc = xs_pc_setup(c))) return rc;
  if ((rc = ensure_trsv_sched(c, true))) return rc;   // an exact solve whatever the apply
  xs_fill_sentinel(pick(c, stream), x, c->A.n);
  if (which == 0) xs_trsv_lower(pick(c, stream), c->ilu, b, x, c->counters + 2, nullptr);
  else xs_trsv_upper(pick(c, stream), c->ilu, b, x, c->counters + 4, nullptr);
  XS_HIP(hipGetLastError());
  return xs_take_launch_guard() ? XS_ERR_STATE : XS_OK;
}

extern "C" int xs_pc_get_factors(xs_ctx *c, int *rowptr, int *colidx, double *val, int *diagpos,
                                 int cap) {
  if (!c) return XS_ERR_ARG;
  if (!xs_pc_factored(c->pc)) return XS_ERR_STATE;
  int rc;
  if (!c->pc_ready && (rc = xs_pc_setup(c))) return rc;
  const xs_ilu &F = c->ilu;
  if (cap < F.nnz) return F.nnz;
  XS_HIP(hipDeviceSynchronize());
  if (rowptr) XS_HIP(hipMemcpy(rowptr, F.rp, sizeof(int) * (F.n + 1), hipMemcpyDeviceToHost));
  if (colidx) XS_HIP(hipMemcpy(colidx, F.ci, sizeof(int) * F.nnz, hipMemcpyDeviceToHost));
  if (val) XS_HIP(hipMemcpy(val, F.lu, sizeof(double) * F.nnz, hipMemcpyDeviceToHost));
  if (diagpos) XS_HIP(hipMemcpy(diagpos, F.dpos, sizeof(int) * F.n, hipMemcpyDeviceToHost));
  return F.nnz;
}

// ---------------------------------------------------------------------------
// legacy singleton API (hip/xsolver_c.cpp)
// ---------------------------------------------------------------------------
static xs_ctx *g_ctx = nullptr;
static int g_err = XS_OK;
static int *g_usediter = nullptr;
static double *g_resvec = nullptr;

static int g_set(int rc) {
  g_err = rc;
  return rc;
}

extern "C" void xsolver_mat_setup(int n, int *rowptr, int *colidx, double *val, int *advanced_pam) {
  if (!g_ctx && g_set(xs_create(&g_ctx)) != XS_OK) return;
  g_set(xs_mat_setup(g_ctx, n, rowptr, colidx, val, advanced_pam));
}
extern "C" void xsolver_ksp_settype(char *fmt) {
  if (!g_ctx) { g_set(XS_ERR_STATE); return; }
  g_set(xs_ksp_settype(g_ctx, fmt));
}
extern "C" void xsolver_ksp_set_absolute_tol(double a) {
  if (!g_ctx) { g_set(XS_ERR_STATE); return; }
  g_set(xs_ksp_set_absolute_tol(g_ctx, a));
}
extern "C" void xsolver_ksp_setoption(double tol, int maxiter, int *usediter, double *resvec,
                                      int restart) {
  if (!g_ctx) { g_set(XS_ERR_STATE); return; }
  g_usediter = usediter;
  g_resvec = resvec;
  // IGCR restart = maxiter when 0 (hip/Solver/xsolver_igcr.h:36)
  g_set(xs_ksp_setoption(g_ctx, tol, maxiter, restart));
}
extern "C" void xsolver_pc_settype(char *fmt) {
  if (!g_ctx) { g_set(XS_ERR_STATE); return; }
  g_set(xs_pc_settype(g_ctx, fmt));
}
extern "C" void xsolver_pc_setilut(double tol, int max_fill) {
  if (!g_ctx) { g_set(XS_ERR_STATE); return; }
  g_set(xs_pc_setilut(g_ctx, tol, max_fill));
}
extern "C" void xsolver_pc_setict(double, int) { g_set(XS_ERR_UNSUPPORTED); }
extern "C" void xsolver_solve(double *rhs, double *sol) {
  if (!g_ctx) { g_set(XS_ERR_STATE); return; }
  int rc = xs_solve(g_ctx, rhs, sol);
  g_set(rc);
  if (rc >= 0 || rc == XS_ERR_BREAKDOWN) {
    if (g_usediter) *g_usediter = g_ctx->iters;
    if (g_resvec)
      for (size_t i = 0; i < g_ctx->resvec.size(); i++) g_resvec[i] = g_ctx->resvec[i];
  }
}
extern "C" void xsolver_free(void) {
  if (g_ctx) xs_destroy(g_ctx);
  g_ctx = nullptr;
  g_usediter = nullptr;
  g_resvec = nullptr;
}
extern "C" int xsolver_last_error(void) { return g_err; }
extern "C" int xsolver_get_iterations(void) { return g_ctx ? g_ctx->iters : -1; }
extern "C" double xsolver_get_residual(void) { return g_ctx ? g_ctx->resl : -1.0; }

// ---------------------------------------------------------------------------
// adjoint (no reference counterpart)
// ---------------------------------------------------------------------------
extern "C" int xs_solve_transpose_device(xs_ctx *c, const double *d_rhs, double *d_sol) {
  if (!c || !d_rhs || !d_sol) return XS_ERR_ARG;
  if (!c->mat_ready) return XS_ERR_STATE;
  if (c->precision != 0) return XS_ERR_UNSUPPORTED;
  XS_HIP(hipDeviceSynchronize());
  int rc;
  if (!c->pc_ready && (rc = xs_pc_setup(c))) return rc;
  if ((rc = xs_build_transpose(c))) return rc;
  if (xs_pc_factored(c->pc) && (rc = xs_build_factor_transpose(c))) return rc;
  c->use_t = true;
  rc = run_solve(c, d_rhs, d_sol);
  c->use_t = false;
  return rc;
}

extern "C" int xs_grad_values_device(xs_ctx *c, const double *d_lambda, const double *d_x,
                                     double *d_grad, void *stream) {
  if (!c || !d_lambda || !d_x || !d_grad) return XS_ERR_ARG;
  if (!c->mat_ready) return XS_ERR_STATE;
  xs_grad_vals(pick(c, stream), c->A, d_lambda, d_x, d_grad);
  XS_HIP(hipGetLastError());
  return xs_take_launch_guard() ? XS_ERR_STATE : XS_OK;
}
