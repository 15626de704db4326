/*
 * xsolver_c.h -- C-ABI of the MI355X (gfx950) xsolver Krylov library.
 *
 * Drop-in for the reference API hip/xsolver_c.h:37-49 (implemented in
 * hip/xsolver_c.cpp:41-333).  The reference declares those functions with C++
 * linkage (the extern "C" is commented out), takes `int &usediter` and gives
 * `restart` a default.  Here the core is real C: plain pointers and sizes,
 * int status codes, no C++ or torch types.  For C++ callers the header also
 * provides inline overloads with the reference's exact signatures
 * (reference, default argument), so code written against the reference
 * header (e.g. hip/solve_base.h:4-24) recompiles unchanged.
 *
 * Everything runs on one GPU and one rank; the reference's MPI path
 * (advanced_pam[0] == 1, scatter_mpi/gather_mpi) is not provided.
 * The library is not thread-safe: the legacy entry points share one global
 * context, like the reference's singletons (hip/xsolver_c.cpp:12-21).
 */
#ifndef XSOLVER_C_H_
#define XSOLVER_C_H_

#include <stddef.h> /* NULL for callers of the Fortran-style surface */

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum xs_status {
  XS_OK = 0,
  XS_NOT_CONVERGED = 1,     /* maxiter reached (the reference's usediter = maxiter) */
  XS_ERR_ARG = -1,          /* invalid argument / unknown type string */
  XS_ERR_STATE = -2,        /* call out of order (e.g. solve before mat_setup) */
  XS_ERR_UNSUPPORTED = -3,  /* e.g. MPI distribution, CPU backend, IC0 on a nonsymmetric pattern */
  XS_ERR_DEVICE = -4,       /* HIP runtime error or no device */
  XS_ERR_BREAKDOWN = -5,    /* non-finite residual / Krylov breakdown */
  XS_ERR_PIVOT = -6,        /* missing diagonal entry / zero Jacobi pivot */
  XS_ERR_MEMORY = -7,
  XS_ERR_TOO_LARGE = -8     /* input beyond a fixed-size workspace's maximum */
};

/* ======================================================================== */
/* Legacy singleton API (reference hip/xsolver_c.h:37-49)                   */
/* ======================================================================== */

/* hip/xsolver_c.cpp:41-102.  CSR arrays are 0- or 1-based (rowptr[0] decides,
 * as in hip/HostMatrix/HostMatrixCSR.cpp:119-128); they are copied to the GPU
 * here.  advanced_pam[0] = 1 (MPI) -> XS_ERR_UNSUPPORTED; advanced_pam[1] = 1
 * selects ELL SpMV; advanced_pam[2] must be 1 (GPU). */
void xsolver_mat_setup(int n, int *rowptr, int *colidx, double *val, int *advanced_pam);
/* hip/xsolver_c.cpp:103-130: "Bicgstab|bicgstab|BICGSTAB", "Igcr|igcr|IGCR",
 * "Cg|cg|CG"; additionally "Gmres|gmres|GMRES", "Fgmres|fgmres|FGMRES" and the
 * rest of the intel family (intel/mpisolver-C-v1.cpp): "cgm" (= "cg"),
 * "cg_left" (CG on M^-1 A), "icg", "icgm", "ibicgstab", "gcr_left",
 * "igcr_left". */
void xsolver_ksp_settype(char *fmt);
/* hip/xsolver_c.cpp:131-133 */
void xsolver_ksp_set_absolute_tol(double absolute_tol);
/* hip/xsolver_c.cpp:134-149.  *usediter and resvec[0..usediter] (absolute
 * residual norms, resvec[0] = ||r0||) are written when xsolver_solve returns;
 * resvec may be NULL, otherwise it must hold maxiter + 1 doubles. */
void xsolver_ksp_setoption(double tol, int maxiter, int *usediter, double *resvec, int restart);
/* hip/xsolver_c.cpp:150-262: "Jacobi", "Ilu0", "Ilut", "Ilu0_mkl", "Ilut_mkl",
 * "Ic0" in the reference's spellings (on the GPU the *_mkl names map to
 * ParILU/ParILUT as in :196-211).  The reference runs IC0 on the host only
 * (:174-177 exits on the GPU); here "Ic0" is ParIC(0) on the GPU, whose fixed
 * point is the reference's IC(0) (hip/Precond/ic0.cpp), applied as
 * L~ D L~^T; it needs a structurally symmetric pattern (else
 * XS_ERR_UNSUPPORTED) and, like the reference, reads only the lower
 * triangle's values. */
void xsolver_pc_settype(char *fmt);
/* hip/xsolver_c.cpp:264-273 */
void xsolver_pc_setilut(double tol, int max_fill);
void xsolver_pc_setict(double tol, int max_fill);
/* hip/xsolver_c.cpp:274-292.  sol holds x0 on entry and the solution on exit. */
void xsolver_solve(double *rhs, double *sol);
/* hip/xsolver_c.cpp:308-333 */
void xsolver_free(void);

/* Result / iteration queries (the reference returns results only through the
 * usediter/resvec pointers captured by xsolver_ksp_setoption). */
int xsolver_last_error(void);
int xsolver_get_iterations(void);
double xsolver_get_residual(void);

/* ======================================================================== */
/* Handle API (same semantics, one context per system)                      */
/* ======================================================================== */
typedef struct xs_ctx xs_ctx;

/* Create a context on the current HIP device (no reference counterpart: the
 * reference keeps one global context, hip/xsolver_c.cpp:12-21). */
int xs_create(xs_ctx **ctx);
int xs_destroy(xs_ctx *ctx);
const char *xs_status_string(int status);

/* Matrix upload from host CSR (hip/xsolver_c.cpp:41-102). */
int xs_mat_setup(xs_ctx *ctx, int n, const int *rowptr, const int *colidx,
                 const double *val, const int *advanced_pam);
/* Matrix already resident on the GPU (0-based CSR device pointers, borrowed
 * until xs_destroy / the next setup; columns strictly ascending within rows,
 * else XS_ERR_ARG).  No reference counterpart: used when the caller owns
 * HBM-resident data. */
int xs_mat_setup_device(xs_ctx *ctx, int n, int nnz, const int *d_rowptr,
                        const int *d_colidx, const double *d_val, int use_ell);
/* Replace the values of the current pattern (host or device pointer), the
 * "*_mat_update" of the reference's Fortran-style API (intel/HostMatrix.h:197-201);
 * the preconditioner is re-factorised on the next solve. */
int xs_mat_update_values(xs_ctx *ctx, const double *val, int on_device);

int xs_ksp_settype(xs_ctx *ctx, const char *type);
int xs_ksp_set_absolute_tol(xs_ctx *ctx, double abstol);
int xs_ksp_setoption(xs_ctx *ctx, double tol, int maxiter, int restart);
int xs_pc_settype(xs_ctx *ctx, const char *type);
int xs_pc_setilut(xs_ctx *ctx, double tol, int max_fill);

/* Named options (no reference counterpart; the reference hard-codes them):
 *   "parilu_sweeps"  ParILU(0) fixed-point sweeps (reference: 20, precon_gpuilu0.cpp:5)
 *   "parilu_async"   1 = in-place asynchronous sweeps (reference behaviour),
 *                    0 = deterministic double-buffered sweeps
 *   "check_every"    iterations launched between host convergence checks
 *   "fuse_p"         1 = BiCGSTAB forms p inside the first L sweep of M^{-1} p
 *                    (sweep apply, even check_every; same iterates), 0
 *                    (default) = a separate update kernel
 *   "use_graph"      1 = replay each batch of check_every iterations as a
 *                    captured HIP graph, 0 = launch kernels one by one,
 *                    -1 (default) = graphs for k-RHS block solves only (a
 *                    single solve's short kernels ran faster launched one by
 *                    one, and its graph is captured anew in every solve)
 *   "precision"      0 = FP64, 1 = FP32 inner Krylov solve inside FP64
 *                    iterative refinement (BiCGSTAB/CG) or FP64 flexible
 *                    GMRES (GMRES/FGMRES)
 *   "factor_storage" ILU factors of the FP32 inner solve: 0 FP32, 1 BF16, 2 FP16
 *   "inner_tol"      relative residual reduction asked of each inner solve;
 *                    0 (default) = sqrt(tol), at least 1e-6
 *   "mixed_reliable" precision 1 with BiCGSTAB: 1 (default) = one FP32
 *                    BiCGSTAB recurrence whose residual is replaced by the FP64
 *                    true residual every check_every iterations (reliable
 *                    updates; inner_tol unused), 0 = restarted refinement
 *   "parilu_scale"   1 (default) = factorise S A S, S = |diag(A)|^-1/2, and
 *                    scale back (ParILU/ParILUT fixed-point stability)
 *   "parilut_steps"  ParILUT add/sweep/select/sweep steps (reference: 10)
 *   "parilut_rows"   ParILUT row table of the candidate and sweep kernels:
 *                    0 (default) = the smallest of 128 / 256 / 512 entries that
 *                    leaves at most n/1024 rows to the per-entry form; 128,
 *                    256 or 512 force it (longer rows take the per-entry form)
 *   "parilut_small_sets" 1 (default) = ParILUT candidate counts in 512-slot
 *                    sets first (rows that fill theirs counted again in
 *                    2048-slot sets; same counts), 0 = 2048-slot sets only
 *   "pc_reuse"       1 = xs_mat_update_values keeps the preconditioner built
 *                    from the earlier values (a preconditioner sets the rate
 *                    of convergence, not the converged solution); 0 (default)
 *                    = factorise again at the next solve
 *   "refactor"       any value: recompute the preconditioner at the next solve
 *   "profile"        bit mask of kernel families timed with HIP events
 *                    (2 SpMV, 4 lower solve, 8 upper solve); read back with
 *                    xs_get_time_ms("total:spmv"|"count:spmv"|...)
 *   "lsolve_sweeps", "usolve_sweeps"
 *                    -1 (default) = exact sync-free triangular solve; k >= 0 =
 *                    k Jacobi sweeps per factor, the reference's GPU ILU apply
 *                    (20/20 for ILU(0), precon_gpuilu0.cpp:17-18; 10/15 for
 *                    ILUT, precon_gpuilut.cpp:19-20)
 *   "trsv_schedule"  1 (default) = hand out the triangular-solve row groups in
 *                    dependency-level order, 0 = row order
 *   "trsv_progressive" -1 (default) = by factor width: factors with more than
 *                    six entries per row and triangle solve every 64-row group
 *                    lane by lane, publishing each row as soon as it is final;
 *                    narrower factors solve a group whose rows depend at most
 *                    on the row before with one affine scan; 0 / 1 force either
 *   "trsv_workgroups" persistent workgroups of a triangular solve (default 256):
 *                    bounds the number of polling waves
 *   "trsv_gather"    off-group entries a lane of the triangular solve loads
 *                    and polls together: 0 (default) = 16 for factors with
 *                    more than six entries per row and triangle, else 4
 *   "spmm_cw"        columns per thread of the k-RHS ELL SpMM: 0 (default) =
 *                    4 when k % 4 == 0, else 2
 *   "spmv_kernel"    SpMV kernel for CSR input: -1 (default) = automatic: an
 *                    internal ELL copy when the padding to the longest row is
 *                    under 20% of nnz, else vector CSR when the row lengths are
 *                    regular, else LDS row blocks; 0 = LDS-staged row blocks;
 *                    1 = row-binned vector CSR (lanes per row = mean row length
 *                    rounded up to a power of two, long rows one wavefront
 *                    each).  0 and 1 also drop the automatic ELL copy.
 *   "use_ell"        1 = ELL SpMV (same as advanced_pam[1] = 1)
 *   "spmv_dia"       1 (default) = when the ELL copy is built and the matrix
 *                    has at most 8 distinct column - row offsets (5- and
 *                    7-point stencils), the FP64 single-vector SpMV reads a
 *                    diagonal-slot copy (one value slab per offset, zeros
 *                    where a row lacks that diagonal, no per-entry index);
 *                    same entries in the same column order as the ELL rows.
 *                    Set before xs_mat_setup; 0 = ELL only
 *   "sweep_dia"      1 (default) = the same for the strictly lower and upper
 *                    factor parts in the FP64 Jacobi-sweep apply (at most 8
 *                    distinct offsets each: ILU(0) of a 7-point stencil has
 *                    3); 0 = ELL copies only
 *   "parilu_dia"     1 (default) = ParILU(0) sweeps over diagonal slots
 *                    (no index lists) when A has at most 8 offsets; the
 *                    deterministic sweeps give bit for bit the factors of
 *                    the list sweeps; 0 = the update lists */
int xs_set_option(xs_ctx *ctx, const char *name, double value);

/* Autotuner (no reference counterpart): times, on this matrix, the CSR
 * row-block kernel at nnz budgets per workgroup of 512..4096, the row-binned
 * vector-CSR kernel (reported as chunk 0) with lanes per row of half, one and
 * two times the mean row length rounded up to a power of two and long rows
 * past 4, 8 or 16 lane groups, and the ELL kernel (when its padding is under
 * 50%, with the k-RHS SpMM's 2 or 4 columns per thread timed at k = 16), and
 * keeps the fastest.  Called after xs_pc_setup with an ILU-type
 * preconditioner it also times the exact triangular solves over persistent
 * workgroups 128..1024, gather widths 4 / 16 and both group modes and keeps
 * the fastest.  With the environment variable XS_TUNE_CACHE naming a file,
 * results are read from and appended to it per matrix (and factor)
 * signature.  Choices and times are queried with xs_get_time_ms("tune:chunk"
 * | "tune:ell" | "tune:us" | "tune:vec_lanes" | "tune:long_thr" |
 * "tune:spmm_cw" | "tune:trsv_wg" | "tune:trsv_gather" | "tune:trsv_prog" |
 * "tune:trsv_us"; -1 when not tuned).  The SpMV kernel in
 * use: xs_get_time_ms("spmv:kernel") = 0 LDS row blocks, 1 vector CSR, 2 ELL;
 * xs_get_time_ms("spmv:idx16") = 1 when the ELL copy stores 16-bit column
 * offsets (every entry within 32767 of its row), 0 when it reads int32
 * columns; xs_get_time_ms("pc:idx16") the same for the ELL copies of L and U
 * used by the Jacobi-sweep apply (-1 before they are built);
 * xs_get_time_ms("spmv:dia") = the number of diagonal slots of the FP64
 * SpMV's copy (0: none), xs_get_time_ms("pc:dia") those of L's sweep copy
 * (-1 before the sweep copies are built), xs_get_time_ms("parilu:dia")
 * those of the ParILU(0) sweeps (0: update lists). */
int xs_autotune(xs_ctx *ctx);

/* Factorise / prepare the preconditioner now (otherwise done by the first
 * solve, like Xsolver::xsolver_init, hip/Solver/xsolver_bicgstab.cpp:3-19). */
int xs_pc_setup(xs_ctx *ctx);
/* Solve with host vectors (hip/xsolver_c.cpp:274-292): sol = x0 in, x out. */
int xs_solve(xs_ctx *ctx, const double *rhs, double *sol);
/* Solve with device vectors already in HBM (d_sol = x0 in, x out). */
int xs_solve_device(xs_ctx *ctx, const double *d_rhs, double *d_sol);

int xs_get_iterations(const xs_ctx *ctx);
double xs_get_residual(const xs_ctx *ctx);        /* last ||r_k|| (absolute) */
double xs_get_initial_residual(const xs_ctx *ctx);
/* Copies min(len, iterations+1) residual norms; returns the count copied. */
int xs_get_resvec(const xs_ctx *ctx, double *out, int len);
/* Milliseconds of the last pc setup and the last solve (HIP events). */
double xs_get_time_ms(const xs_ctx *ctx, const char *phase);

/* ======================================================================== */
/* Kernel-level entry points on device pointers (stream = hipStream_t or 0). */
/* They operate on the matrix/factors held by a context.                     */
/* ======================================================================== */
/* y = A x (CSR row-binned or ELL, per the context's format); reference
 * DeviceMatrixCSR::SpMV, hip/HostMatrix/DeviceMatrix/DeviceMatrixCSR.cpp:157-170,
 * and DeviceMatrixELL::SpMV, DeviceMatrixELL.cpp:147-162. */
int xs_spmv(xs_ctx *ctx, const double *d_x, double *d_y, void *stream);
/* y = b - A x; reference bmAxCSR/bmAxELL (DeviceMatrixCSR.cpp:72-84, 171-189). */
int xs_bmax(xs_ctx *ctx, const double *d_b, const double *d_x, double *d_y, void *stream);
/* y = A^T x (adjoint path; no reference counterpart). */
int xs_spmv_t(xs_ctx *ctx, const double *d_x, double *d_y, void *stream);
/* Preconditioner apply y = M^{-1} x; for ILU the sync-free L then U solves
 * (reference Lcsr_trsv_sync_free / Ucsc_trsv_sync_free,
 * hip/Precond/gpuparilu.cpp:257-391). */
int xs_pc_apply(xs_ctx *ctx, const double *d_x, double *d_y, void *stream);
/* Transposed preconditioner apply y = M^{-T} x (adjoint path). */
int xs_pc_apply_t(xs_ctx *ctx, const double *d_x, double *d_y, void *stream);
/* Single triangular solves with the ILU factors: which = 0 lower (unit),
 * 1 upper (with diagonal). */
int xs_trsv(xs_ctx *ctx, int which, const double *d_b, double *d_x, void *stream);
/* Adjoint solve A^T x = b with the transposed preconditioner M^{-T}
 * (U^T then L^T on the transposed factors), same Krylov method, device
 * vectors (d_sol = x0 in, x out).  No reference counterpart. */
int xs_solve_transpose_device(xs_ctx *ctx, const double *d_rhs, double *d_sol);
/* dL/dA_ij = -lambda_i x_j on the pattern of A (nnz values in CSR order).
 * No reference counterpart. */
int xs_grad_values_device(xs_ctx *ctx, const double *d_lambda, const double *d_x,
                          double *d_grad, void *stream);
/* Copy the ILU factors to the host in combined CSR form (A's pattern for
 * ILU(0)): nnz values, plus the position of each pivot.  Returns the nnz. */
int xs_pc_get_factors(xs_ctx *ctx, int *rowptr, int *colidx, double *val, int *diagpos, int cap);

/* ======================================================================== */
/* Block of right-hand sides: A X = B for k <= 64 columns at once             */
/* (no reference counterpart; the reference solves one right-hand side).      */
/* ======================================================================== */
enum xs_rhs_layout {
  XS_RHS_COLMAJOR = 0,     /* column j at rhs + j*n (k vectors back to back) */
  XS_RHS_INTERLEAVED = 1   /* row-major n x k: entry (i, j) at rhs[i*k + j] */
};
/* The context's Krylov method (BiCGSTAB or CG, FP64) and preconditioner run
 * on every column with its own scalars and convergence test, sharing one
 * SpMM and one k-column triangular solve per operator application.  sol =
 * X0 in, X out.  Returns XS_OK when every column converged, XS_NOT_CONVERGED
 * if some column reached maxiter, XS_ERR_BREAKDOWN if one broke down.
 * GMRES/GCR or precision 1 -> XS_ERR_UNSUPPORTED. */
int xs_solve_multi_device(xs_ctx *ctx, int nrhs, const double *d_rhs, double *d_sol,
                          int layout);
int xs_solve_multi(xs_ctx *ctx, int nrhs, const double *rhs, double *sol, int layout);
/* Per-column iterations, initial and final residual norms of the last block
 * solve (arrays of cap entries; any may be NULL).  Returns the column count. */
int xs_get_multi_results(const xs_ctx *ctx, int *iters, double *res0, double *resl, int cap);
/* Residual history of column col (absolute norms, entry 0 = ||r0||). */
int xs_get_multi_resvec(const xs_ctx *ctx, int col, double *out, int len);
/* Kernel level, k <= 64 interleaved columns (n x k row-major device blocks):
 * Y = A X (multi-RHS SpMM, the matrix read once for all columns) and
 * Y = M^{-1} X (k-column sync-free triangular solves for ILU). */
int xs_spmm(xs_ctx *ctx, int nrhs, const double *d_x, double *d_y, void *stream);
int xs_pc_apply_multi(xs_ctx *ctx, int nrhs, const double *d_x, double *d_y, void *stream);

/* ======================================================================== */
/* Batched solve: many independent systems in one persistent launch          */
/* (no reference counterpart; the reference solves one system per call).     */
/* ======================================================================== */
typedef struct xs_batch xs_batch;
/* Block-diagonal CSR (0-based): rows sys_ptr[s]..sys_ptr[s+1]-1 form system s
 * and its columns stay in that range.  Copied to the GPU. */
int xs_batch_create(xs_batch **bt, int nsys, const int *sys_ptr, const int *rowptr,
                    const int *colidx, const double *val);
/* CG + Jacobi on every system to ||b - A x|| < tol ||b - A x0|| (the
 * reference's Krylov convention, hip/Solver/xsolver_cg.cpp:16-74; FP64 check), each
 * system with its own iteration count; precision 0 = FP64, 1 = one FP32 CG
 * recurrence (FP32 matrix values and vectors) with FP64 reliable updates: at
 * every 1e-2 drop of the FP32 residual the FP64 residual b - A x replaces it
 * and decides convergence.  sol = x0 in, x out (host or device vectors). */
int xs_batch_solve(xs_batch *bt, const double *rhs, double *sol, double tol, int maxiter,
                   int precision);
int xs_batch_solve_device(xs_batch *bt, const double *d_rhs, double *d_sol, double tol,
                          int maxiter, int precision);
/* Allocation-free variant of xs_batch_solve_device (FP64 CG + Jacobi).  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  It takes no context: every argument is a device pointer
 * (the block-diagonal CSR as for xs_batch_create, rhs, sol = x0 in / x out,
 * per-system iters and FP64 relative residual ||b - A x|| / ||b - A x0|| out) except
 * the sizes, nsys systems with n rows and nnz entries in total.  Its workspace
 * (7 doubles per row) is a static device array for at most
 * XS_BATCH_NOALLOC_MAX_ROWS rows, held by each GPU it runs on until the
 * process ends; n above that -> XS_ERR_TOO_LARGE.  Calls no hipMalloc,
 * hipFree or allocating library routine and does not synchronise: it enqueues
 * kernels on `stream` (NULL = the null stream) and returns XS_OK; convergence
 * is read from relres (>= tol, or NaN, where a system did not converge; a
 * system whose rows or columns leave its block gets iters = -1, relres =
 * NaN).  Every call on one GPU shares that GPU's workspace: calls that
 * overlap on one GPU race. */
#define XS_BATCH_NOALLOC_MAX_ROWS 1048576
int xs_batch_solve_noalloc(int nsys, int n, int nnz, const int *d_sys_ptr, const int *d_rowptr,
                           const int *d_colidx, const double *d_val, const double *d_rhs,
                           double *d_sol, double tol, int maxiter, int *d_iters,
                           double *d_relres, void *stream);
/* Per-system iterations and final FP64 relative residual
 * ||b - A x|| / ||b - A x0||; returns nsys. */
int xs_batch_get_results(const xs_batch *bt, int *iters, double *relres);
double xs_batch_get_time_ms(const xs_batch *bt);
int xs_batch_destroy(xs_batch *bt);

/* ======================================================================== */
/* Matrix Market input (host only, no device needed)                          */
/* ======================================================================== */
/* Coordinate-format file -> 0-based CSR with sorted columns (reference reader
 * hip/Tool/ReadMatrix/csr_mmio_highlevel.h:8-233, mmio_info + mmio_data).
 * Call once with rowptr/colidx/val NULL to get *n and *nnz (after the
 * expansion of symmetric / skew-symmetric / Hermitian files), then with
 * arrays of those capacities (*n, *nnz in = capacities).  diag (optional, n
 * entries) receives the diagonal values as read (the reference's a_p).
 * Duplicate coordinates are summed; pattern files get value 1; complex and
 * non-square files -> XS_ERR_UNSUPPORTED.  The reference negates the values
 * of general files (:67, :178); XS_MTX_NEGATE_GENERAL reproduces that. */
#define XS_MTX_NEGATE_GENERAL 1
int xs_mtx_read(const char *path, int flags, int *n, int *nnz, int *rowptr, int *colidx,
                double *val, double *diag);

/* ======================================================================== */
/* Fortran-style MCSR surface of the reference's CUDA backend                */
/* (cuda/HostMatrix.h:162-193, cuda/DeviceMatrix.h:121-169,                  */
/* cuda/mpicudasolver-C-v1.cu:165-1530).  Every argument is passed by        */
/* address, as from Fortran (the reference's C++ references are pointers at */
/* the ABI).  One matrix at a time; not thread-safe; errors never exit:     */
/* gpu_last_error() returns the status of the last call.                    */
/* ======================================================================== */
/* Registers (borrows) the host matrix in MCSR form, 0-based: diag_val[n],
 * off-diagonals in CSR (offdiag_row_offset[n+1], offdiag_col_index /
 * offdiag_val[nSizes]), columns in any order.  nHalo != 0 (MPI halo) ->
 * XS_ERR_UNSUPPORTED. */
void cpu_mat_setup(int *nInterior, int *nHalo, int *nSizes, double *diag_val, double *offdiag_val,
                   int *offdiag_row_offset, int *offdiag_col_index);
/* new value arrays for the registered pattern (cuda/HostMatrix.h:170-173) */
void cpu_mat_update(double *diag_val, double *offdiag_val);
/* copies the registered matrix to the GPU (cuda/DeviceMatrix.h:121-126) */
void gpu_mat_setup(void);
/* "ELL": ELL SpMV; *num_nz (1..32) is written back as 16 or 32, as the
 * reference pads (:127-144); the ELL copy here has the true row width */
void gpu_mat_format(char *fmt, int *num_nz);
/* the values of the last cpu_mat_update to the GPU (:145-148) */
void gpu_mat_update(void);
void gpu_mat_free(void);
/* "Jacobi" (:153-158), or any preconditioner name xs_pc_settype accepts;
 * other strings: no preconditioner */
void gpu_preconditioner_setup(char *fmt);
/* refactorise on the current values (:159-165) */
void gpu_preconditioner_update(char *fmt);
void gpu_preconditioner_free(char *fmt);
/* The intel-family solves (cuda/mpicudasolver-C-v1.cu; intel/mpisolver-C-v1.cpp):
 * cg = CG on M^-1 A, icg = its one-reduction variant, cgm = preconditioned CG,
 * icgm = PCG with one reduction per iteration, bicgstab / ibicgstab =
 * right-preconditioned BiCGSTAB / improved BiCGSTAB, gcr / igcr =
 * left-preconditioned GCR(restart) with classical / modified Gram-Schmidt,
 * gmres / fgmres = GMRES(restart) left / flexible.  rhs, sol, sol_init and
 * resvec are host arrays; sol_init NULL = start from sol.  On return:
 * *usediter = 0-based index of the last iteration (maxiter - 1 at
 * exhaustion; maxiter for gcr / igcr, as those loops leave it), resvec[k] =
 * ||r_{k+1}|| / (||r_0|| + 1e-20) for the residual the method monitors. */
void gpu_cg_solve(double *rhs, double *sol, double *tol, int *maxiter, double *sol_init,
                  int *usediter, double *resvec);
void gpu_icg_solve(double *rhs, double *sol, double *tol, int *maxiter, double *sol_init,
                   int *usediter, double *resvec);
void gpu_cgm_solve(double *rhs, double *sol, double *tol, int *maxiter, double *sol_init,
                   int *usediter, double *resvec);
void gpu_icgm_solve(double *rhs, double *sol, double *tol, int *maxiter, double *sol_init,
                    int *usediter, double *resvec);
void gpu_bicgstab_solve(double *rhs, double *sol, double *tol, int *maxiter, double *sol_init,
                        int *usediter, double *resvec);
void gpu_ibicgstab_solve(double *rhs, double *sol, double *tol, int *maxiter, double *sol_init,
                         int *usediter, double *resvec);
void gpu_gcr_solve(double *rhs, double *sol, int *restart, double *tol, int *maxiter,
                   double *sol_init, int *usediter, double *resvec);
void gpu_igcr_solve(double *rhs, double *sol, int *restart, double *tol, int *maxiter,
                    double *sol_init, int *usediter, double *resvec);
void gpu_gmres_solve(double *rhs, double *sol, int *restart, double *tol, int *maxiter,
                     double *sol_init, int *usediter, double *resvec);
void gpu_fgmres_solve(double *rhs, double *sol, int *restart, double *tol, int *maxiter,
                      double *sol_init, int *usediter, double *resvec);
/* status of the last call of this surface (no reference counterpart) */
int gpu_last_error(void);

#ifdef __cplusplus
} /* extern "C" */

/* C++ overloads with the reference's exact signatures (hip/xsolver_c.h:40). */
inline void xsolver_ksp_setoption(double tol, int maxiter, int &usediter,
                                  double *resvec, int restart = 0) {
  xsolver_ksp_setoption(tol, maxiter, &usediter, resvec, restart);
}
/* and for the MCSR surface (cuda/HostMatrix.h:162, cuda/DeviceMatrix.h:127,
 * cuda/mpicudasolver-C-v1.cu): references in place of pointers */
inline void cpu_mat_setup(int &nInterior, int &nHalo, int &nSizes, double *diag_val,
                          double *offdiag_val, int *offdiag_row_offset, int *offdiag_col_index) {
  cpu_mat_setup(&nInterior, &nHalo, &nSizes, diag_val, offdiag_val, offdiag_row_offset,
                offdiag_col_index);
}
inline void gpu_mat_format(char *fmt, int &num_nz) { gpu_mat_format(fmt, &num_nz); }
#define XS_MCSR_SOLVE_REF(name)                                                            \
  inline void name(double *rhs, double *sol, double &tol, int &maxiter, double *sol_init,  \
                   int &usediter, double *resvec = nullptr) {                              \
    name(rhs, sol, &tol, &maxiter, sol_init, &usediter, resvec);                           \
  }
#define XS_MCSR_SOLVE_REF_R(name)                                                          \
  inline void name(double *rhs, double *sol, int &restart, double &tol, int &maxiter,      \
                   double *sol_init, int &usediter, double *resvec = nullptr) {            \
    name(rhs, sol, &restart, &tol, &maxiter, sol_init, &usediter, resvec);                 \
  }
XS_MCSR_SOLVE_REF(gpu_cg_solve)
XS_MCSR_SOLVE_REF(gpu_icg_solve)
XS_MCSR_SOLVE_REF(gpu_cgm_solve)
XS_MCSR_SOLVE_REF(gpu_icgm_solve)
XS_MCSR_SOLVE_REF(gpu_bicgstab_solve)
XS_MCSR_SOLVE_REF(gpu_ibicgstab_solve)
XS_MCSR_SOLVE_REF_R(gpu_gcr_solve)
XS_MCSR_SOLVE_REF_R(gpu_igcr_solve)
XS_MCSR_SOLVE_REF_R(gpu_gmres_solve)
XS_MCSR_SOLVE_REF_R(gpu_fgmres_solve)
#undef XS_MCSR_SOLVE_REF
#undef XS_MCSR_SOLVE_REF_R
#endif

#endif /* XSOLVER_C_H_ */
