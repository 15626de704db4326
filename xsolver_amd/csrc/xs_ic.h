// xs_ic.h -- ParIC(0) kernels (xs_ic.hip).
#pragma once
#include "xs_internal.h"

// mirror positions of every entry (-1 when absent; *asym set to 1 then)
void xs_ic_mirror(hipStream_t s, const xs_csr &A, const int *rowidx, int *mpos, int *asym);
// update lists: count (pl == null: cnt and pd) or fill (pl, pu at offsets pp)
void xs_ic_pairs(hipStream_t s, const xs_csr &A, const xs_ilu &F, int *cnt, const int *pp,
                 int *pl, int *pu, int *pd);
// start values in sweep order into F.as, F.lc and lc2 (when given)
void xs_ic_start(hipStream_t s, const xs_csr &A, const xs_ilu &F, const double *sc, double *lc2);
// one sweep over the active entries (lc_in == lc_out: in place)
void xs_ic_sweep_pairs(hipStream_t s, const xs_ilu &F, const double *lc_in, double *lc_out);
// combined ILU factors on A's pattern into F.lu
void xs_ic_finish(hipStream_t s, const xs_csr &A, const xs_ilu &F, const int *mpos,
                  const double *sc, const double *lc);
