// xs_multi.h -- k right-hand sides solved together (SpMM, multi-column
// sync-free triangular solves, per-column Krylov vector kernels).
// Vectors are n x k row-major blocks ("interleaved": X[i*k + j]).
#pragma once
#include "xs_internal.h"

#define XS_MAX_RHS 64        // columns per block
#define XS_MRHS_TRSV 16      // columns one triangular-solve launch carries
#define XS_MGRID 2048        // workgroups of the reducing kernels

// per-column scalar slots
enum xs_mslot {
  MS_RES0 = 0, MS_RESL,
  MS_ALPHA, MS_BETA, MS_BETO, MS_GAMA, MS_OMEGA,   // BiCGSTAB (reference naming)
  MS_TAU, MS_CGALPHA, MS_CGBETA, MS_RR,            // CG
  MS_NSLOT
};

struct xs_mstate {
  double sc[MS_NSLOT][XS_MAX_RHS];
  int done[XS_MAX_RHS];   // 1 converged, 2 non-finite
  int iter[XS_MAX_RHS];
  int all;                // every column done
  int k;
};

// per-column finishing ops
enum xs_mfinish {
  MF_NONE = 0,
  MF_BICG_RES0,   // res0 = ||r0||, BiCGSTAB start scalars
  MF_BICG_GAMA,   // gama = beta / (rhat.u)
  MF_BICG_ALPHA,  // alpha = (s.v) / (v.v)
  MF_BICG_RES,    // ||r||, convergence, next beta/omega
  MF_CG_INIT,     // tau = r.z, res0 = ||r||
  MF_CG_ALPHA,    // alpha = tau / (p.Ap)
  MF_CG_RR,       // rr = r.r
  MF_CG_RES       // tau' = z.r, beta, convergence (rr from tot1 or the MS_RR slot)
};

struct xs_mred {
  double *partial;     // >= XS_MGRID * 2 * k doubles, XS_SENTINEL between reductions
  unsigned *counter;   // raised if a partial never arrives (bounded wait)
  xs_mstate *ms;
  double *resvec;      // [(maxiter + 1) * k] absolute residuals, may be null
  double tol, abstol;
  int op;
};

void xs_mspmm(hipStream_t s, const xs_csr &A, int k, const double *X, double *Y, int epi,
              const double *B, const double *W, const xs_mred *red, const xs_mstate *ms);
// the same over the ELL copy of A: k even and every vector 16-byte aligned
// (xs_mspmm_ell_ok); products in each row's entry order, as xs_mspmm.  cw:
// columns per thread, 2 or 4 (4 needs k % 4 == 0); 0 = 4 when k allows.
bool xs_mspmm_ell_ok(const xs_ell &E, int k, const void *const *vecs, int nvec);
void xs_mspmm_ell(hipStream_t s, const xs_ell &E, int k, const double *X, double *Y, int epi,
                  const double *B, const double *W, const xs_mred *red, const xs_mstate *ms,
                  int cw = 0);
void xs_mdot(hipStream_t s, int n, int k, const double *a, const double *b, const double *c,
             const double *d, const xs_mred &red);
void xs_mbicg_p(hipStream_t s, int n, int k, const double *R, const double *U, double *P,
                const double *dinv, double *Z, const xs_mstate *ms);
void xs_mbicg_s(hipStream_t s, int n, int k, const double *R, const double *U, double *Sv,
                const double *dinv, double *Z1, const xs_mstate *ms);
void xs_mbicg_xr(hipStream_t s, int n, int k, double *X, double *R, const double *Z,
                 const double *Z1, const double *Sv, const double *V, const double *Rh,
                 const xs_mred &red);
void xs_mcg_xr(hipStream_t s, int n, int k, double *X, double *R, const double *P,
               const double *AP, const double *dinv, double *Z, const xs_mred &red);
void xs_mcg_p(hipStream_t s, int n, int k, double *P, const double *Z, const xs_mstate *ms);
void xs_mscale(hipStream_t s, int n, int k, const double *d, const double *X, double *Y);
void xs_mtrsv(hipStream_t s, bool upper, int n, int k, const int *rp, const int *ci,
              const double *lu, const int *dpos, const double *dinv, const double *B, double *X,
              unsigned *tickets, const xs_tsched &S, const xs_mstate *ms);
// column-major (k vectors of length n) <-> interleaved n x k
void xs_mlayout(hipStream_t s, int n, int k, const double *in, double *out, bool to_interleaved);
void xs_mstate_init(hipStream_t s, xs_mstate *ms, int k);
// out = in over m doubles (a kernel, so it can sit in a captured graph)
void xs_mcopy(hipStream_t s, size_t m, const double *in, double *out);
// one k-column Jacobi sweep of a factor part over its ELL copy (fused: sweeps
// 1 and 2 from zero)
void xs_msweep_ell(hipStream_t s, const xs_tell &E, int n, int k, bool upper, bool fused,
                   const double *dinv, const double *B, const double *Yo, double *Yn,
                   const xs_mstate *ms);
