// xs_batch_host.cpp -- C-ABI of the batched solve (no reference counterpart:
// the reference solves one system per xsolver_solve call).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/xsolver_c.h"
#include "xs_batch.h"

struct xs_batch {
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int nsys = 0, n = 0, nnz = 0;
  int *sys_ptr = nullptr, *order = nullptr, *rp = nullptr, *ci = nullptr;
  double *v = nullptr, *dinv = nullptr, *w64 = nullptr, *b = nullptr, *x = nullptr;
  double *relres = nullptr;
  float *v32 = nullptr, *dinv32 = nullptr, *w32 = nullptr;
  unsigned *queue = nullptr;
  int *iters = nullptr, *bad = nullptr;
  float ms = 0.f;
  std::vector<int> h_iters;
  std::vector<double> h_relres;
};

#define BT_CHK(call)                                   \
  do {                                                 \
    if ((call) != hipSuccess) { (void)hipGetLastError(); return XS_ERR_DEVICE; } \
  } while (0)

template <class T>
static int bt_alloc(T **p, size_t n) {
  if (hipMalloc((void **)p, sizeof(T) * (n ? n : 1)) != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return XS_ERR_MEMORY;
  }
  return XS_OK;
}

extern "C" int xs_batch_destroy(xs_batch *bt) {
  if (!bt) return XS_OK;
  if (bt->stream) (void)hipStreamSynchronize(bt->stream);
  auto f = [](auto *&p) { if (p) (void)hipFree((void *)p); p = nullptr; };
  f(bt->sys_ptr); f(bt->order); f(bt->rp); f(bt->ci); f(bt->v); f(bt->dinv); f(bt->w64);
  f(bt->b); f(bt->x); f(bt->relres); f(bt->v32); f(bt->dinv32); f(bt->w32); f(bt->queue);
  f(bt->iters); f(bt->bad);
  if (bt->e0) (void)hipEventDestroy(bt->e0);
  if (bt->e1) (void)hipEventDestroy(bt->e1);
  if (bt->stream) (void)hipStreamDestroy(bt->stream);
  delete bt;
  return XS_OK;
}

extern "C" int xs_batch_create(xs_batch **out, int nsys, const int *sys_ptr, const int *rowptr,
                               const int *colidx, const double *val) {
  if (!out || nsys <= 0 || !sys_ptr || !rowptr || !colidx || !val) return XS_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return XS_ERR_DEVICE;
  const int n = sys_ptr[nsys];
  if (sys_ptr[0] != 0 || n <= 0 || rowptr[0] != 0) return XS_ERR_ARG;
  const int nnz = rowptr[n];
  // block-diagonal check: every column inside its own system's row range
  for (int s = 0; s < nsys; s++) {
    if (sys_ptr[s + 1] <= sys_ptr[s]) return XS_ERR_ARG;
    for (int i = sys_ptr[s]; i < sys_ptr[s + 1]; i++)
      for (int e = rowptr[i]; e < rowptr[i + 1]; e++)
        if (colidx[e] < sys_ptr[s] || colidx[e] >= sys_ptr[s + 1]) return XS_ERR_ARG;
  }
  xs_batch *bt = new xs_batch();
  bt->nsys = nsys; bt->n = n; bt->nnz = nnz;
  int rc = XS_OK;
  if (hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking) != hipSuccess) {
    delete bt;
    return XS_ERR_DEVICE;
  }
  (void)hipEventCreate(&bt->e0);
  (void)hipEventCreate(&bt->e1);
  std::vector<int> order(nsys);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return sys_ptr[a + 1] - sys_ptr[a] > sys_ptr[b + 1] - sys_ptr[b];
  });
  if ((rc = bt_alloc(&bt->sys_ptr, nsys + 1)) || (rc = bt_alloc(&bt->order, nsys)) ||
      (rc = bt_alloc(&bt->rp, n + 1)) || (rc = bt_alloc(&bt->ci, nnz)) ||
      (rc = bt_alloc(&bt->v, nnz)) || (rc = bt_alloc(&bt->dinv, n)) ||
      (rc = bt_alloc(&bt->w64, 6 * (size_t)n)) || (rc = bt_alloc(&bt->b, n)) ||
      (rc = bt_alloc(&bt->x, n)) || (rc = bt_alloc(&bt->relres, nsys)) ||
      (rc = bt_alloc(&bt->v32, nnz)) || (rc = bt_alloc(&bt->dinv32, n)) ||
      (rc = bt_alloc(&bt->w32, 6 * (size_t)n)) || (rc = bt_alloc(&bt->queue, 4)) ||
      (rc = bt_alloc(&bt->iters, nsys)) || (rc = bt_alloc(&bt->bad, 1))) {
    xs_batch_destroy(bt);
    return rc;
  }
  BT_CHK(hipMemcpy(bt->sys_ptr, sys_ptr, sizeof(int) * (nsys + 1), hipMemcpyHostToDevice));
  BT_CHK(hipMemcpy(bt->order, order.data(), sizeof(int) * nsys, hipMemcpyHostToDevice));
  BT_CHK(hipMemcpy(bt->rp, rowptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice));
  BT_CHK(hipMemcpy(bt->ci, colidx, sizeof(int) * nnz, hipMemcpyHostToDevice));
  BT_CHK(hipMemcpy(bt->v, val, sizeof(double) * nnz, hipMemcpyHostToDevice));
  BT_CHK(hipMemset(bt->bad, 0x7f, sizeof(int)));
  xs_batch_prep(bt->stream, n, bt->rp, bt->ci, bt->v, bt->dinv, bt->dinv32, bt->v32, bt->bad);
  int bad = 0;
  BT_CHK(hipMemcpyAsync(&bad, bt->bad, sizeof(int), hipMemcpyDeviceToHost, bt->stream));
  BT_CHK(hipStreamSynchronize(bt->stream));
  if (bad != 0x7f7f7f7f) {
    xs_batch_destroy(bt);
    return XS_ERR_PIVOT;
  }
  *out = bt;
  return XS_OK;
}

static int batch_run(xs_batch *bt, const double *d_b, double *d_x, double tol, int maxiter,
                     int precision) {
  if (maxiter <= 0 || (precision != 0 && precision != 1)) return XS_ERR_ARG;
  hipStream_t s = bt->stream;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int grid = std::min(bt->nsys, 2 * cus);
  BT_CHK(hipMemsetAsync(bt->queue, 0, sizeof(unsigned), s));
  BT_CHK(hipEventRecord(bt->e0, s));
  xs_batch_cg(s, grid, bt->nsys, bt->n, bt->nnz, bt->order, bt->sys_ptr, bt->rp, bt->ci, bt->v, bt->v32,
              bt->dinv, bt->dinv32, d_b, d_x, bt->w64, bt->w32, tol, maxiter, precision, 1e-2,
              bt->queue, bt->iters, bt->relres);
  BT_CHK(hipGetLastError());
  BT_CHK(hipEventRecord(bt->e1, s));
  BT_CHK(hipEventSynchronize(bt->e1));
  (void)hipEventElapsedTime(&bt->ms, bt->e0, bt->e1);
  bt->h_iters.resize(bt->nsys);
  bt->h_relres.resize(bt->nsys);
  BT_CHK(hipMemcpy(bt->h_iters.data(), bt->iters, sizeof(int) * bt->nsys, hipMemcpyDeviceToHost));
  BT_CHK(hipMemcpy(bt->h_relres.data(), bt->relres, sizeof(double) * bt->nsys,
                   hipMemcpyDeviceToHost));
  for (int q = 0; q < bt->nsys; q++)
    if (!(bt->h_relres[q] < tol)) return XS_NOT_CONVERGED;
  return XS_OK;
}

extern "C" int xs_batch_solve(xs_batch *bt, const double *rhs, double *sol, double tol,
                              int maxiter, int precision) {
  if (!bt || !rhs || !sol) return XS_ERR_ARG;
  BT_CHK(hipMemcpy(bt->b, rhs, sizeof(double) * bt->n, hipMemcpyHostToDevice));
  BT_CHK(hipMemcpy(bt->x, sol, sizeof(double) * bt->n, hipMemcpyHostToDevice));
  int rc = batch_run(bt, bt->b, bt->x, tol, maxiter, precision);
  if (rc < 0) return rc;
  BT_CHK(hipMemcpy(sol, bt->x, sizeof(double) * bt->n, hipMemcpyDeviceToHost));
  return rc;
}

extern "C" int xs_batch_solve_device(xs_batch *bt, const double *d_rhs, double *d_sol,
                                     double tol, int maxiter, int precision) {
  if (!bt || !d_rhs || !d_sol) return XS_ERR_ARG;
  BT_CHK(hipDeviceSynchronize());
  return batch_run(bt, d_rhs, d_sol, tol, maxiter, precision);
}

extern "C" int xs_batch_get_results(const xs_batch *bt, int *iters, double *relres) {
  if (!bt) return XS_ERR_ARG;
  for (int q = 0; q < (int)bt->h_iters.size(); q++) {
    if (iters) iters[q] = bt->h_iters[q];
    if (relres) relres[q] = bt->h_relres[q];
  }
  return (int)bt->h_iters.size();
}

extern "C" double xs_batch_get_time_ms(const xs_batch *bt) { return bt ? bt->ms : -1.0; }

// ---------------------------------------------------------------------------
// Allocation-free batched solve (no reference counterpart).  For callers that
// cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  Its workspace -- the work-queue counter,
// 1/a_ii and the six CG vectors, 7 doubles per row -- lives in `scratch`, a
// static device array of fixed size that each GPU holds until the process
// ends.  Calls on one GPU share that GPU's array (the last call's contents
// stay in it): calls that overlap on one GPU, on any streams, race.
// ---------------------------------------------------------------------------
#define XS_NOALLOC_HEADER 256   // bytes before the vectors: the queue counter
static __device__ __align__(256) unsigned char
    scratch[XS_NOALLOC_HEADER + 7 * sizeof(double) * (size_t)XS_BATCH_NOALLOC_MAX_ROWS];

static std::mutex g_scratch_mu;
static std::vector<unsigned char *> g_scratch_ptr;   // device address per device

// the device address of `scratch` on the current device, looked up once per
// device with that device current (the host-side name is not a device address)
static int noalloc_scratch(unsigned char **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) { (void)hipGetLastError(); return XS_ERR_DEVICE; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return XS_ERR_DEVICE;
    }
    g_scratch_ptr[dev] = (unsigned char *)p;
  }
  *out = g_scratch_ptr[dev];
  return XS_OK;
}

extern "C" int xs_batch_solve_noalloc(int nsys, int n, int nnz, const int *d_sys_ptr,
                                      const int *d_rowptr, const int *d_colidx,
                                      const double *d_val, const double *d_rhs, double *d_sol,
                                      double tol, int maxiter, int *d_iters, double *d_relres,
                                      void *stream) {
  if (nsys <= 0 || n <= 0 || nnz <= 0 || maxiter <= 0 || !(tol >= 0.0) || !d_sys_ptr ||
      !d_rowptr || !d_colidx || !d_val || !d_rhs || !d_sol || !d_iters || !d_relres)
    return XS_ERR_ARG;
  if (n > XS_BATCH_NOALLOC_MAX_ROWS || nsys > n) return XS_ERR_TOO_LARGE;
  unsigned char *base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc) return rc;
  unsigned *queue = (unsigned *)base;
  double *dinv = (double *)(base + XS_NOALLOC_HEADER);
  double *w64 = dinv + n;
  hipStream_t s = (hipStream_t)stream;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int grid = std::min(nsys, 2 * cus);
  // one launch sets 1/a_ii and clears the queue counter (kernels only, so the
  // call can be captured into a graph)
  xs_batch_prep64(s, n, d_rowptr, d_colidx, d_val, dinv, queue);
  xs_batch_cg(s, grid, nsys, n, nnz, nullptr, d_sys_ptr, d_rowptr, d_colidx, d_val, nullptr,
              dinv, nullptr, d_rhs, d_sol, w64, nullptr, tol, maxiter, 0, 0.0, queue, d_iters,
              d_relres);
  if (hipGetLastError() != hipSuccess) return XS_ERR_DEVICE;
  return XS_OK;
}
