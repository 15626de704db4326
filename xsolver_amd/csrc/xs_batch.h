// xs_batch.h -- persistent batched CG kernels (xs_batch.hip).
#pragma once
#include <hip/hip_runtime.h>

void xs_batch_prep(hipStream_t s, int n, const int *rp, const int *ci, const double *v,
                   double *dinv, float *dinv32, float *v32, int *bad);
void xs_batch_prep64(hipStream_t s, int n, const int *rp, const int *ci, const double *v,
                     double *dinv, unsigned *queue);
// N rows and nnz entries in the whole batch (N is the stride of the six
// workspace vectors in w64 / w32); order == nullptr: systems are taken in
// index order.  A system whose rows or columns fall outside its bounds gets
// iters = -1 and relres = NaN.
void xs_batch_cg(hipStream_t s, int grid, int nsys, int N, int nnz, const int *order,
                 const int *sys_ptr, const int *rp, const int *ci, const double *v,
                 const float *v32, const double *dinv, const float *dinv32, const double *b,
                 double *x, double *w64, float *w32, double tol, int maxiter, int precision,
                 double inner_tol, unsigned *queue, int *iters, double *relres);
