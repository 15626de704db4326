// xs_krylov.h -- tall-skinny basis kernels and device-side Hessenberg work.
#pragma once
#include <hip/hip_runtime.h>

#include "xs_internal.h"

#define XS_TS_KMAX 64  // basis vectors per tall-skinny launch

enum { XS_G_NRM2 = 0, XS_G_SCALE = 1, XS_G_T0 = 2 };

// Small FP64 arrays of a GMRES(m) cycle, all in device memory.
struct xs_gmres_small {
  double *H;     // (m+1) x m, row-major
  double *cs, *sn, *s;  // m+1 each
  double *h1, *h2;      // m+1 each: the two CGS passes of the current column
  double *y;            // m+1
  double *misc;         // 8: [NRM2] ||w||^2, [SCALE] 1/(norm + 1e-20), [T0..] GCR t0,t1,||r||^2
};

int xs_ts_blocks(int n);
// T = double or float (instantiated in xs_krylov.hip); inner products in double
template <class T>
void xs_ts_dot(hipStream_t s, int n, int k, const T *V, size_t ld, const T *w,
               double *partial, unsigned *counter, double *out, const xs_state *st);
template <class T>
void xs_ts_update(hipStream_t s, int n, int k, const T *V, const T *U, size_t ld,
                  const double *h, T *w, T *wout, T *u, const xs_state *st);
template <class T>
void xs_ts_axpy(hipStream_t s, int n, int k, const T *V, size_t ld, const double *y, T *x);
template <class T>
void xs_ts_scale(hipStream_t s, int n, const T *src, T *dst, const double *scale,
                 const xs_state *st);
void xs_gmres_column(hipStream_t s, int i, int m, const xs_gmres_small &g, xs_state *st,
                     double *resvec);
void xs_gmres_cycle(hipStream_t s, int m, int first, const xs_gmres_small &g, xs_state *st,
                    double *resvec);
void xs_gmres_backsub(hipStream_t s, int k, int m, const xs_gmres_small &g);
void xs_gcr_apply(hipStream_t s, int n, double *u, double *c, double *x, double *r, double *t01,
                  double *partial, unsigned *counter, xs_state *st, double *resvec);
void xs_res0_from(hipStream_t s, const double *nrm2, xs_state *st, double *resvec);
