// xs_mixed.h -- FP32 inner-iteration kernels of the mixed-precision solve.
#pragma once
#include <hip/hip_runtime.h>

#include "xs_internal.h"

enum { XS_F32 = 0, XS_BF16 = 1, XS_F16 = 2 };   // factor storage types

void xs_launch_spmv_csr32(hipStream_t s, const xs_csr &A, const float *v32,
                          const xs_spmv_plan &P, const float *x, float *y, int epi,
                          const float *b, const float *w, const xs_reduce *red);
// FP32 SpMV over d32, the FP32 copy of E's diagonal-slot values (E.nd > 0)
void xs_launch_spmv_dia32(hipStream_t s, const xs_ell &E, const float *d32, const float *x,
                          float *y, int epi, const float *b, const float *w,
                          const xs_reduce *red);
void xs_launch_spmv_ell32(hipStream_t s, const xs_ell &E, const float *v32, const float *x,
                          float *y, int epi, const float *b, const float *w,
                          const xs_reduce *red);
// one Jacobi sweep of an FP32/BF16/FP16 factor (yo == null: yn = D b)
void xs_tri_sweep32(hipStream_t s, bool upper, int n, const int *rp, const int *ci, const void *lu,
                    int ftype, const int *dpos, const float *dinv, const float *b, const float *yo,
                    float *yn, const xs_state *st);
// FP32 sweep over the ELL copy of a factor part (values E.vlo stored as
// ftype: FP32 / BF16 / FP16); fused = sweeps 1+2
void xs_tri_sweep_ell32(hipStream_t s, bool upper, bool fused, int n, const xs_tell &E, int ftype,
                        const float *dinv, const float *b, const float *yo, float *yn,
                        const xs_state *st);
// the first L pass of M^{-1} p over the diagonal-slot copy with BiCGSTAB's
// p = r + beta (pin - omega u) formed in it and written to pout (pin != pout)
bool xs_tell32_dia(const xs_tell &E);
void xs_tri_sweep_ell32_p(hipStream_t s, int n, const xs_tell &E, int ftype, const float *r,
                          const float *u, const float *pin, float *pout, float *yn,
                          const xs_state *st);
// the fused first L pass with BiCGSTAB's s = r - gama u (gama in st) formed
// in the pass and written to sv
void xs_tri_sweep_ell32_s(hipStream_t s, int n, const xs_tell &E, int ftype, const float *r,
                          const float *u, float *sv, float *yn, const xs_state *st);
void xs_trsv32(hipStream_t s, bool upper, int n, const int *rp, const int *ci, const void *lu,
               int ftype, const int *dpos, const float *dinv, const float *b, float *x,
               unsigned *tickets, const xs_tsched &S, const xs_state *st);
void xs_cvt_factor(hipStream_t s, int nnz, const double *src, void *dst, int ftype);
void xs_d2f(hipStream_t s, int n, const double *a, float *b);
void xs_vec_d2f(hipStream_t s, int n, const double *d, float *f);
void xs_d2f_scaled(hipStream_t s, int n, const double *a, float *b, const xs_state *sref, int slot);
void xs_add_f2d(hipStream_t s, int n, const float *d, double *x, const xs_state *sref, int slot);
void xs_flush_f2d(hipStream_t s, int n, float *d, double *x, const xs_state *sref, int slot);
void xs_replace32(hipStream_t s, int n, const double *r64, float *r, const float *rh,
                  const xs_state *sref, int slot, const xs_reduce &red);
void xs_zero32(hipStream_t s, int n, float *a);
void xs_copy32(hipStream_t s, int n, const float *a, float *b);
void xs_scale32(hipStream_t s, int n, const float *d, const float *a, float *b);
void xs_bicg_p32(hipStream_t s, int n, const float *r, const float *u, float *p,
                 const float *dinv, float *z, xs_state *st);
void xs_bicg_s32(hipStream_t s, int n, const float *r, const float *u, float *sv,
                 const float *dinv, float *z1, xs_state *st);
void xs_bicg_xr32(hipStream_t s, int n, float *x, float *r, const float *z, const float *z1,
                  const float *sv, const float *v, const float *rh, const xs_reduce &red);
