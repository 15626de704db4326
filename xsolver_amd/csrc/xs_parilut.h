// xs_parilut.h -- ParILUT kernels (xs_parilut.hip) and shared scan/scale helpers.
#pragma once
#include <hip/hip_runtime.h>

void xs_scan_excl(hipStream_t s, int n, const int *in, int *out, int *bsum);
int xs_scan_bsum_len(int n);
// per-row candidate counts |C_i|, C_i = A_i u F_i u (union over k in L_i of
// the upper part of F_k with its pivot); *ovf |= 1 for a row beyond 1024
// longrows (n ints) and nlong (1 int): scratch of the two-tier count (512-slot
// sets, the rows that fill theirs counted again in 2048-slot sets); nullptr
// = every row in the 2048-slot set
void xs_pt_symbolic(hipStream_t s, int n, const int *arp, const int *aci, const int *frp,
                    const int *fci, const int *fdpos, int *cnt, int *ovf, int *longrows = nullptr,
                    int *nlong = nullptr);
// pivot positions of a sorted pattern (*bad |= 2 for a row without its diagonal)
void xs_pt_dpos(hipStream_t s, int n, const int *rp, const int *ci, int *dpos, int *bad);
// the candidate pattern C (crp from the counts; cci, crow, pivot positions
// cdpos) with the old factor F's values embedded (cv); one ParILUT sweep over
// a pattern (rp, ci, dpos).  maxr: the row table (128, 256 or 512 entries)
// from xs_pt_pick_rows; has_long: some rows are longer (they take the
// per-entry forms).  *ovf |= 2 for a row without its diagonal.
void xs_pt_fill_embed(hipStream_t s, int n, const int *arp, const int *aci, const double *av,
                      const double *sc, const int *frp, const int *fci, const double *fv,
                      const int *fdpos, const int *crp, int *cci, int *crow, double *cv,
                      int *cdpos, int *ovf, int maxr, bool has_long);
// keys (may be null): each entry's selection key (bit pattern of the
// unscaled magnitude, top bit set for U)
void xs_pt_sweep(hipStream_t s, int n, const int *rp, const int *ci, const int *dpos,
                 const double *vin, double *vout, const int *arp, const int *aci,
                 const double *av, const double *sc, int maxr, unsigned long long *keys);
// row-length statistics from per-row counts (4 words: longest row, rows
// longer than 128, 256, 512), the row table they call for and whether some
// rows exceed it
void xs_pt_len_stats(hipStream_t s, int n, const int *cnt, unsigned *stats);
int xs_pt_pick_rows(int n, const unsigned *stats);
bool xs_pt_rows_over(int maxr, const unsigned *stats);
// thresholds of both factors on the unscaled magnitudes from the entries'
// selection keys (xs_pt_sweep) and nunit unit diagonals of L: thr[0] (resp.
// thr[1]) = the value of rank cnt2[0] - keepL (resp. cnt2[1] - keepU) in
// ascending magnitude over L with its unit diagonal (U with its pivots), 0
// when the factor has no more than keep entries; sel: xs_pt_select_words()
// words, hist: 512, lcnt: 2, list: xs_pt_list_len()
void xs_pt_select(hipStream_t s, int nnz, const unsigned long long *keys, int nunit,
                  const unsigned long long *cnt2, long long keepL, long long keepU,
                  unsigned long long *sel, unsigned *hist, unsigned *lcnt,
                  unsigned long long *list, double *thr);
int xs_pt_select_words();
int xs_pt_list_len();
// entries with col <= row (L with its unit diagonal) -> cnt2[0], col >= row
// (U with its pivots) -> cnt2[1], from the pivot positions dpos of a sorted
// n-row pattern rp whose rows all hold their diagonal
void xs_pt_count_parts(hipStream_t s, int n, const int *rp, const int *dpos,
                       unsigned long long *cnt2);
// dtol[i] = droptol * ||a_i||_2 (unscaled A)
void xs_pt_dtol(hipStream_t s, int n, const int *arp, const double *av, double droptol,
                double *dtol);
// compaction of the candidates (count pass, then write pass) with the
// thresholds of xs_pt_select and the drop tolerance, on the selection keys;
// cdpos: the candidates' pivot positions (always kept).  The write pass also
// places the pivots (fdpos; *bad |= 2 for a row without)
void xs_pt_keep(hipStream_t s, bool write, int n, const int *crp, const int *cci,
                const int *cdpos, const double *cv, const unsigned long long *keys,
                const double *thr, const double *dtol, int *cnt, const int *frp, int *fci,
                double *fv, int *frow, int *fdpos, int *bad);
void xs_pt_scale_vec(hipStream_t s, int n, const int *arp, const int *aci, const double *av,
                     double *sc);
void xs_pt_rescale(hipStream_t s, int nnz, const int *row, const int *ci, double *v,
                   const double *sc, int unscale);
void xs_pt_rowidx(hipStream_t s, int n, const int *rp, int *row);

// Pattern analysis on the device (xs_split.hip).  xs_colsort: stable sort of
// m (key, val) pairs by key < ncol; returns 1 when the result is in key2 /
// val2, 0 when in key / val.  hist and off hold xs_colsort_hist_len(m) ints,
// bsum xs_scan_bsum_len(xs_colsort_hist_len(m)).
size_t xs_colsort_hist_len(int m);
int xs_colsort(hipStream_t s, int m, int ncol, unsigned *key, int *val, unsigned *key2, int *val2,
               int *hist, int *off, int *bsum);
// cp[j] = first sorted position with key >= j, j = 0..ncol
void xs_key_ptr(hipStream_t s, int m, int ncol, const unsigned *skey, int *cp);
// entries on and right of the pivot per row, and their (column, position)
// pairs in row-major order (uoff from the exclusive scan of the counts)
void xs_upper_count(hipStream_t s, int n, const int *rp, const int *dpos, int *cnt);
void xs_upper_pairs(hipStream_t s, int n, const int *rp, const int *ci, const int *dpos,
                    const int *uoff, unsigned *key, int *val);
void xs_all_pairs(hipStream_t s, int nnz, const int *ci, unsigned *key, int *val);
// out[q] = rowidx[pos[q]]; perm[pos[q]] = q when perm is given
void xs_sorted_rows(hipStream_t s, int m, const int *rowidx, const int *pos, int *out, int *perm);
// *flag |= 1: a column out of [0, n) or not strictly ascending in its row;
// |= 2: a row without its diagonal (rp must be monotone with rp[n] = nnz)
void xs_check_pattern(hipStream_t s, int n, const int *rp, const int *ci, int *flag);
