/*
 * xs_oracle.h -- CPU restatement of the xsolver reference's preconditioned
 * Krylov path.  TEST INFRASTRUCTURE ONLY: this library is the checker that the
 * HIP path is compared against (tests/, __graft_entry__.smoke(), and the
 * cpu_baseline leg of bench.py).  The product library never links or calls it.
 *
 * Every function names the reference file:line it follows.  All matrices are
 * 0-based CSR with sorted column indices and the diagonal present.
 *
 * Conventions (HIP-framework conventions, hip/Solver/xsolver_*.cpp):
 *   resvec[0] = ||r0||, resvec[k] = ||r_k|| (absolute norms)
 *   *usediter = number of iterations performed
 * GMRES/FGMRES follow intel/mpisolver-C-v1.cpp:933-1247 with the iteration
 * count converted to "iterations performed" and with the final update using
 * all i+1 Krylov directions (see or_gmres).
 */
#ifndef XS_ORACLE_H
#define XS_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef void (*or_precond_fn)(void *ctx, const double *x, double *y);

/* ---- sparse kernels ---------------------------------------------------- */
void or_spmv(int n, const int *rp, const int *ci, const double *v,
             const double *x, double *y);
void or_bmax(int n, const int *rp, const int *ci, const double *v,
             const double *b, const double *x, double *y);
void or_spmv_t(int n, const int *rp, const int *ci, const double *v,
               const double *x, double *y);
void or_lower_solve(int n, const int *rp, const int *ci, const double *v,
                    const double *b, double *x);
void or_upper_solve(int n, const int *rp, const int *ci, const double *v,
                    const double *dinv, const double *b, double *x);

/* ---- preconditioners ---------------------------------------------------- */
int or_jacobi_setup(int n, const int *rp, const int *ci, const double *v,
                    double small, double *dinv);
void or_jacobi_apply(void *dinv, const double *x, double *y);

int or_ilu0(int n, const int *rp, const int *ci, const double *v,
            double *luval, int *uptr);
int or_ilut(int n, const int *rp, const int *ci, const double *v, int lfil,
            double droptol, int cap, double *luval, int *lucols, int *lurows,
            int *uptr, double *diag);

typedef struct {
  int n;
  const int *lrp, *lci; const double *lv;   /* strict lower, unit diag implied */
  const int *urp, *uci; const double *uv;   /* strict upper, rows pre-scaled by 1/u_ii */
  const double *dinv;                        /* 1/u_ii */
} or_lu_t;
void or_ilu_split(int n, const int *rp, const int *ci, const double *luval,
                  const int *uptr, int *lrp, int *lci, double *lv,
                  int *urp, int *uci, double *uv, double *dinv);
void or_lusolve(void *lu, const double *x, double *y);
/* kl / ku Jacobi sweeps on the combined factors (reference GPU apply) */
void or_ilu_sweeps(int n, const int *rp, const int *ci, const double *luval, const int *uptr,
                   int kl, int ku, const double *x, double *y);

/* ---- ParILU fixed-point sweep (synchronous / Jacobi form) --------------- */
void or_parilu_sweep(int n, const int *rp, const int *ci, const double *a,
                     const int *lrp, const int *lci, const double *lv_in,
                     const int *ucp, const int *uri, const double *uv_in,
                     double *lv_out, double *uv_out);

/* ---- Krylov solvers ------------------------------------------------------ */
int or_cg(int n, const int *rp, const int *ci, const double *v,
          const double *b, double *x, double tol, double abstol, int maxiter,
          or_precond_fn M, void *Mctx, double *resvec, int *usediter);
int or_bicgstab(int n, const int *rp, const int *ci, const double *v,
                const double *b, double *x, double tol, double abstol,
                int maxiter, or_precond_fn M, void *Mctx, double *resvec,
                int *usediter);
int or_gcr(int n, const int *rp, const int *ci, const double *v,
           const double *b, double *x, double tol, double abstol, int maxiter,
           int restart, or_precond_fn M, void *Mctx, double *resvec,
           int *usediter);
int or_gmres(int n, const int *rp, const int *ci, const double *v,
             const double *b, double *x, double tol, int maxiter, int restart,
             int flexible, or_precond_fn M, void *Mctx, double *resvec,
             int *usediter);

/* ---- independent systems (block-diagonal CSR), CG + Jacobi each --------- */
int or_batch_cg(int nsys, const int *sys_ptr, const int *rp, const int *ci, const double *v,
                const double *b, double *x, double tol, int maxiter, int *iters,
                double *relres);

/* ---- ParILUT (deterministic synchronous form) --------------------------- */
int or_parilut(int n, const int *arp, const int *aci, const double *av, int sweeps,
               double fill, double droptol, int cap, int *lrp, int *lci, double *lv,
               int *urp, int *uci, double *uv, int *trace);

/* ---- incomplete Cholesky (the reference's CPU IC0 / ICT) ----------------
 * Output L: lower CSR including the diagonal, stored last in each row as
 * 1/l_ii (ic0.cpp / ict_l.cpp layout).  Return 0, or i+1 for the first row
 * whose lower part does not end at the diagonal (IC0) / has no diagonal
 * (ICT).  or_ic0: cap >= nnz(lower(A)); or_ict: cap >= (maxfil + 1) n. */
int or_ic0(int n, const int *rp, const int *ci, const double *v, double smallnum,
           double droptol, int *lp, int *lc, double *lv);
int or_ict(int n, const int *rp, const int *ci, const double *v, double smallnum,
           double droptol, int maxfil, int *lp, int *lc, double *lv);
/* seqic0 / seqict's conversion to the LUsolve form (or_lu_t): L~ strict lower
 * (l_ij / l_jj), U = L~^T strict upper (CSR by transposition), dinv = 1/l_ii^2.
 * Arrays of capacity lp[n] - n (L~, U) and n (dinv, lrp/urp n + 1). */
void or_ic_split(int n, const int *lp, const int *lc, const double *lv, int *lrp, int *lci,
                 double *lvs, int *urp, int *uci, double *uvs, double *dinv);

/* ---- driver: pc in {"none","jacobi","ilu0","ilut","parilut","ic0","ict"},
 *      ksp in {"cg","bicgstab","gcr","gmres","fgmres"} ------------------- */
int or_solve(const char *ksp, const char *pc, int n, const int *rp,
             const int *ci, const double *v, const double *b, double *x,
             double tol, double abstol, int maxiter, int restart,
             double ilut_tol, int ilut_fill, int lsweeps, int usweeps,
             double *resvec, int *usediter);

#ifdef __cplusplus
}
#endif
#endif
