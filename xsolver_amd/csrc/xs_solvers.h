// xs_solvers.h -- the context structure and the host-side solver drivers.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "xs_internal.h"

enum { XS_KSP_NONE = 0, XS_KSP_CG, XS_KSP_BICGSTAB, XS_KSP_GCR, XS_KSP_GMRES, XS_KSP_FGMRES,
       // the rest of the intel family (intel/mpisolver-C-v1.cpp)
       XS_KSP_CG_LEFT, XS_KSP_ICG, XS_KSP_ICGM, XS_KSP_IBICGSTAB, XS_KSP_GCR_LEFT,
       XS_KSP_IGCR_LEFT };
enum { XS_PC_NONE = 0, XS_PC_JACOBI, XS_PC_ILU0, XS_PC_ILUT, XS_PC_IC0 };
// preconditioners held as combined ILU factors (xs_ilu) and applied by the
// ILU solves: ParILU(0), ParILUT and ParIC(0) (L~ | D L~^T)
inline bool xs_pc_factored(int pc) {
  return pc == XS_PC_ILU0 || pc == XS_PC_ILUT || pc == XS_PC_IC0;
}

#define XS_NWORK 12   // work vectors of length n held by a context

struct xs_ctx {
  xs_ctx();
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // solver settings (defaults as the reference: tol/maxiter come from setoption)
  int ksp = XS_KSP_NONE;
  int pc = XS_PC_NONE;
  double tol = 1e-6, abstol = 0.0;
  int maxiter = 0, restart = 0;
  double ilut_tol = 1e-4;       // ParILUT relative drop tolerance
  int ilut_fill = 2;            // ParILUT final fill factor (nnz per factor / A's)
  int parilu_sweeps = 20;       // hip/Precond/precon_gpuilu0.cpp:5
  bool parilu_async = true;
  bool parilu_scale = true;
  bool pc_reuse = false;        // xs_mat_update_values keeps the preconditioner
  int parilut_steps = 10;       // hip/Precond/precon_gpuilut.cpp:5
  int parilut_rows = 0;         // ParILUT row table (128 / 256 / 512), 0 = from the row lengths
  bool parilut_small_sets = true;  // ParILUT candidate counts in 512-slot sets first
  // ILU application: -1 = exact sync-free triangular solve; k >= 0 = k Jacobi
  // sweeps per factor (the reference's GPU apply: 20/20 for ILU(0),
  // hip/Precond/precon_gpuilu0.cpp:17-18; 10/15 for ILUT, precon_gpuilut.cpp:19-20)
  int lsweeps = -1, usweeps = -1;
  int check_every = 8;
  bool tuned = false;           // SpMV format chosen by the autotuner
  int tune_chunk = 0;
  double tune_us = 0.0;
  int tune_L = 0, tune_long_thr = 0;  // vector-CSR lanes per row, long-row threshold
  double tune_trsv_us = 0.0;    // > 0 once the triangular solves are tuned
  bool fuse_p = false;          // BiCGSTAB forms p inside the first L pass (slower when launched directly)
  int use_graph = -1;           // replay batches as HIP graphs: 1 all solves, 0 none, -1 k-RHS blocks
  bool trsv_schedule = true;    // SpTRSV groups in dependency-level order
  int trsv_wg = 256;            // persistent workgroups of the SpTRSV
  int trsv_prog = -1;           // SpTRSV group mode (xs_tsched::prog), -1 = by factor width
  int trsv_gather = 0;          // SpTRSV gather width (4 or 16), 0 = by factor width
  bool spmv_dia = true;         // diagonal-slot copy of A for the FP64 SpMV when possible
  bool sweep_dia = true;        // diagonal-slot copies of L / U for the FP64 sweeps
  bool parilu_dia = true;       // ParILU(0) sweeps over diagonal slots when possible
  int spmm_cw = 0;              // k-RHS ELL SpMM columns per thread (2 or 4), 0 = 4 when k % 4 == 0
  int spmv_vec = -1;            // CSR kernel: -1 by row regularity, 0 LDS row blocks, 1 vector CSR
  int precision = 0;
  int use_ell = 0;
  bool ell_auto = false;        // use_ell chosen by the library (regular rows)

  // matrix
  bool mat_ready = false, own_matrix = false, sorted = true, has_diag = true;
  xs_csr A;
  xs_spmv_plan plan;
  xs_ell ell;
  std::vector<int> h_rp, h_ci;
  bool have_At = false;
  bool use_t = false;            // solve with A^T and M^T (adjoint)
  int *perm_At = nullptr;        // position in A^T of each entry of A
  xs_csr At;
  xs_spmv_plan planT;

  // preconditioner
  bool pc_ready = false;
  double *jac_dinv = nullptr;
  xs_ilu ilu;
  double *lu_tmp = nullptr;

  // device state and workspace
  xs_state *st = nullptr;
  xs_state *h_state = nullptr;   // pinned mirror
  unsigned *counters = nullptr;  // [0] reductions, [2..3] lower trsv, [4..5] upper trsv
  int *d_int = nullptr;
  double *partial = nullptr;
  size_t partial_blocks = 0;
  double *resvec_d = nullptr;
  int resvec_cap = 0;
  double *work = nullptr;
  size_t work_n = 0;
  int work_vecs = 0;
  double *hb = nullptr, *hx = nullptr;
  int hb_n = 0;
  // GMRES/GCR workspace (basis vectors, small dense data)
  double *basis = nullptr;
  size_t basis_len = 0;
  double *small_d = nullptr;
  size_t small_len = 0;
  double *ts_partial = nullptr;
  double *mdot_partial = nullptr;   // intel-family multi-dot partials
  size_t mdot_len = 0;
  size_t ts_partial_len = 0;

  // kernel timing (HIP events on the context stream around launches of one
  // kernel family while a solve runs): 0 off, 1 SpMV, 2 lower trsv, 3 upper trsv
  int prof_kind = 0;
  std::vector<hipEvent_t> prof_ev;
  int prof_used = 0;
  double prof_ms[4] = {0, 0, 0, 0};
  long prof_cnt[4] = {0, 0, 0, 0};
  int prof_next_kind = 0;
  std::vector<int> prof_kinds;

  // mixed precision (precision = 1): FP32 inner BiCGSTAB / CG inside an FP64
  // iterative-refinement loop; factors stored as FP32, BF16 or FP16
  int factor_type = 0;          // XS_F32 / XS_BF16 / XS_F16
  double inner_tol = 0.0;       // relative reduction asked of each inner solve;
                                // 0 = sqrt(tol), at least 1e-6 (two refinement steps)
  bool mixed_reliable = true;   // mixed BiCGSTAB: one FP32 recurrence with FP64
                                // residual replacement (else restarted refinement)
  bool mixed_ready = false;     // FP32 / low-precision copies of the factors
  bool mixed_mat_ready = false; // FP32 copy of A for the SpMV format in use
  bool tell_ready = false;      // ELL sweep copies of the factors match ilu.lu
  bool tell32_ready = false;    // and their FP32 values (mixed inner solver)
  float *A32 = nullptr, *ell32 = nullptr, *dinv32 = nullptr, *jac32 = nullptr;
  float *dia32 = nullptr;       // FP32 copy of A's diagonal-slot values (ell.dv)
  void *F32 = nullptr;
  size_t F32_bytes = 0;         // allocated bytes of F32
  float *work32 = nullptr;
  float *basis32 = nullptr;
  size_t basis32_len = 0;
  size_t work32_n = 0;
  xs_state *st_in = nullptr;
  int outer_iters = 0;

  // k right-hand sides solved together (interleaved n x k blocks)
  double *mwork = nullptr;
  size_t mwork_len = 0;
  void *ms = nullptr;            // xs_mstate on the device
  void *h_ms = nullptr;          // pinned mirror
  double *mpartial = nullptr;
  double *mresvec = nullptr;
  size_t mresvec_len = 0;
  int mk = 0;                    // columns of the last block solve
  std::vector<int> m_iters;
  std::vector<double> m_res0, m_resl, m_resvec;
  // results
  int iters = 0;
  double res0 = 0.0, resl = 0.0;
  std::vector<double> resvec;
  float t_pc_ms = 0.f, t_solve_ms = 0.f;
};

int xs_ctx_ensure_work(xs_ctx *c, int nvec);

// Replay of a batch of Krylov iterations as a HIP graph.  The first full
// batch (check_every iterations) of a solve is captured from the stream into
// a graph and every later full batch of that solve replays it: one graph
// launch instead of ~10 kernel launches per iteration.  A batch contains
// kernels only (no memset/memcpy nodes).  Off while kernel timing is on, and
// whenever capture fails (then the batch runs eagerly).
struct xs_graph_cache {
  hipGraphExec_t exec = nullptr;
  bool failed = false;
  ~xs_graph_cache() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

template <class F>
int xs_run_batch(xs_ctx *c, xs_graph_cache &gc, int count, F &iteration, bool block = false) {
  const bool graphs = c->use_graph > 0 || (c->use_graph < 0 && block);
  const bool want = graphs && c->prof_kind == 0 && !gc.failed && count == c->check_every;
  if (want && !gc.exec) {
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      for (int j = 0; j < count; j++) iteration();
      const hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (e != hipSuccess || !g || hipGraphInstantiate(&gc.exec, g, nullptr, nullptr, 0) != hipSuccess) {
        gc.exec = nullptr;
        gc.failed = true;
      }
      if (g) (void)hipGraphDestroy(g);
    } else {
      gc.failed = true;
    }
    (void)hipGetLastError();
  }
  if (want && gc.exec)
    return hipGraphLaunch(gc.exec, c->stream) == hipSuccess ? 0 : -4;
  for (int j = 0; j < count; j++) iteration();
  return 0;
}
// kernel timing helpers: bracket one launch of kernel family `kind`
void xs_prof_begin(xs_ctx *c, int kind);
void xs_prof_end(xs_ctx *c, int kind);
void xs_prof_collect(xs_ctx *c);
double *xs_ctx_vec(xs_ctx *c, int k);
xs_reduce xs_ctx_reduce(xs_ctx *c, int op, int slot);
void xs_op_spmv(xs_ctx *c, const double *x, double *y, int epi, const double *b,
                const double *w, const xs_reduce *red);
void xs_op_pc(xs_ctx *c, const double *x, double *y, double *tmp, const xs_state *st);
// s = r - gama u (gama in st), then y = M^{-1} s
void xs_op_pc_s(xs_ctx *c, const double *r, const double *u, double *sv, double *y, double *tmp,
                xs_state *st);
// true when the ILU is applied by two or more L sweeps over the ELL copies:
// BiCGSTAB's vector updates can then be formed inside the first L pass
bool xs_pc_fuse_l(const xs_ctx *c);
// BiCGSTAB's p = r + omega (pin - alpha u) into pout, then y = M^{-1} p, the
// update formed inside the first L pass (xs_pc_fuse_l must hold)
void xs_op_pc_p(xs_ctx *c, const double *r, const double *u, const double *pin, double *pout,
                double *y, double *tmp, xs_state *st);
int xs_op_pc_t(xs_ctx *c, const double *x, double *y, double *tmp, hipStream_t s);

int xs_solve_cg(xs_ctx *c, const double *b, double *x);
int xs_solve_bicgstab(xs_ctx *c, const double *b, double *x);
int xs_solve_gcr(xs_ctx *c, const double *b, double *x, bool left = false, bool mgs = false);
int xs_solve_cg_left(xs_ctx *c, const double *b, double *x, bool improved);
int xs_solve_icgm(xs_ctx *c, const double *b, double *x);
int xs_solve_ibicgstab(xs_ctx *c, const double *b, double *x);
int xs_solve_gmres(xs_ctx *c, const double *b, double *x, bool flexible);
int xs_parilut_factor(xs_ctx *c);
int xs_solve_mixed(xs_ctx *c, const double *b, double *x);
void xs_mixed_free(xs_ctx *c);
int xs_build_transpose(xs_ctx *c);
int xs_build_factor_transpose(xs_ctx *c);
void xs_solvers_free(xs_ctx *c);
// k right-hand sides (interleaved n x k device blocks; X = X0 in, solution out)
int xs_solve_block(xs_ctx *c, int k, const double *B, double *X);
int xs_multi_reserve(xs_ctx *c, int k, int extra_vecs);
double *xs_multi_vec(xs_ctx *c, int k, int idx);
void xs_multi_free(xs_ctx *c);
// k-column ILU apply (sweeps or sync-free solves; xs_multi_host.cpp)
struct xs_mstate;
void xs_multi_ilu_apply(xs_ctx *c, hipStream_t s, int k, const double *in, double *out,
                        double *tmp, double *w, const xs_mstate *ms);
