// C entry points onto the reference's own CPU ILU / IC code, compiled in
// place from hip/Precond/ilut_diag_in_csr.cpp, ic0.cpp and ict_l.cpp by
// oracle/Makefile (target "ref").  The result, oracle/_ref/libxsref.so, is
// used only to write and check golden vectors for the oracle's ILU(0), ILUT,
// IC(0) and ICT restatements.
#include "ilut_diag_in_csr.h"
#include "ic0.h"
#include "ict_l.h"

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

extern "C" void ref_ic0_csr_A(int n, int *rowptr, double *val, int *colidx, int *lptr,
                              double *lval, int *lcol, double smallnum, double droptol) {
  ic0_csr_A<double>(n, rowptr, val, colidx, lptr, lval, lcol, smallnum, droptol);
}

extern "C" void ref_ict_csr_A_l(int n, int *rowptr, double *val, int *colidx, int *lptr,
                                double *lval, int *lcol, double smallnum, double droptol,
                                int maxfil) {
  ict_csr_A_l<double>(n, rowptr, val, colidx, lptr, lval, lcol, smallnum, droptol, maxfil);
}
