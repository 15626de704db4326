// C entry points onto the reference's own CPU ILU code, compiled in place from
// hip/Precond/ilut_diag_in_csr.cpp by oracle/Makefile (target "ref").  The
// result, oracle/_ref/libxsref.so, is used only to write and check golden
// vectors for the oracle's ILU(0) and ILUT restatements.
#include "ilut_diag_in_csr.h"

extern "C" void ref_ilu0_simple(double *luval, double *val, int *rows, int *cols,
                                int n, int *uptr) {
  ilu0_simple<double>(luval, val, rows, cols, n, uptr);
}

extern "C" void ref_ilutp_final(double *luval, double *val, int *rows, int *cols,
                                int n, double *diag, int lfil, int *lucols,
                                int *lurows, int *uptr, double droptol) {
  ilutp_final<double>(luval, val, rows, cols, n, diag, lfil, lucols, lurows,
                      uptr, 0.0, droptol, nullptr, nullptr);
}
