// xs_internal.h -- internal structures shared by the host driver and the
// gfx950 kernels.  Not part of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#define XS_BLOCK 256          // threads per workgroup for streaming kernels
#define XS_WAVE 64            // CDNA wavefront
#define XS_SPMV_LDS 4096      // products staged in LDS per CSR row block (32 KiB)
#define XS_MAX_REDUCE_BLOCKS 2048
#define XS_NSCAL 64

// Readiness marker for the sync-free triangular solves: a signalling-NaN bit
// pattern.  Arithmetic never produces a signalling NaN, and stored values that
// happen to carry it are replaced by the canonical quiet NaN.
#define XS_SENTINEL 0xFFF4DEADBEEF0001ull

// ---------------------------------------------------------------------------
// Device-resident solver state: scalars, flags, iteration counter.
// ---------------------------------------------------------------------------
enum xs_slot {
  S_RES0 = 0, S_RESL, S_TOL, S_ABSTOL,
  // BiCGSTAB (hip/Solver/xsolver_bicgstab.cpp naming)
  S_ALPHA, S_BETA, S_BETO, S_GAMA, S_OMEGA,
  // CG
  S_TAU, S_TAUO, S_CGALPHA, S_CGBETA,
  // GCR
  S_GCR_T0, S_GCR_T1,
  S_RHOO,           // BiCGSTAB: rhat.r of the iteration before the last
  S_SCRATCH0 = 32
};

struct xs_state {
  double sc[XS_NSCAL];
  int done;       // 1 = converged, 2 = breakdown / non-finite
  int iter;       // iterations completed
  int maxiter;
  int pad;
};

// Fused reduction ops: what the last workgroup does with the totals.
enum xs_finish_op {
  F_NONE = 0,
  F_STORE,          // sc[slot+k] = tot[k]
  F_RES0,           // res0 = sqrt(tot0); resvec[0] = res0; done if res0 == 0 or <= abstol
  F_CG_ALPHA,       // alpha = tau / tot0
  F_CG_RES,         // tau' = tot0 (z.r), rr = tot1; beta = tau'/tau; convergence
  F_CG_INIT,        // tau = tot0 (r.z), res0 = sqrt(tot1)
  F_BICG_GAMA,      // gama = beta / tot0
  F_BICG_ALPHA,     // alpha = tot0 / tot1
  F_BICG_RES,       // resl = sqrt(tot0), beta' = tot1, omega, convergence
  F_GCR_COEF,       // t0 = sqrt(tot0), t1 = tot1 -> scale and step
  F_BICG_REPL       // residual replaced by the true one: resl = sqrt(tot0), beta' = tot1
};

struct xs_reduce {
  double *partial;     // >= gridDim * nv slots, XS_SENTINEL between reductions
  unsigned *counter;   // raised if a partial never arrives (bounded wait)
  xs_state *st;
  double *resvec;      // device residual history (may be null)
  int op;
  int slot;
};

// ---------------------------------------------------------------------------
// Matrix on the device.
// ---------------------------------------------------------------------------
struct xs_csr {
  int n = 0, nnz = 0;
  int *rp = nullptr, *ci = nullptr;
  double *v = nullptr;
};

struct xs_spmv_plan {   // row blocks for the LDS-staged CSR kernel
  int nblk = 0;
  int chunk = 0;        // nnz budget per block
  int *blk_row = nullptr;  // [nblk + 1]
  // row-binned vector CSR (used when vec != 0): rows of length <= long_thr
  // by groups of vec_L lanes, the nlong longer rows one wavefront each
  int vec = 0, vec_L = 0, long_thr = 0, nlong = 0;
  int *long_rows = nullptr;
};

// Diagonal slots: at most this many distinct column - row offsets in a matrix
// (a 7-point stencil has 7) give its single-vector FP64 SpMV a copy without
// any per-entry index.
#define XS_DIA_MAX 8
struct xs_dia_offs {    // by value into the kernel: slot k holds column row + o[k]
  int o[XS_DIA_MAX];
};
// ParILU(0) over diagonal slots (xs_ilu::pnd > 0): for the slot of offset
// o = doff.o[s], the products l(i, i + o1) u(i + o1, i + o) as pairs of
// slots (pl, pu) with o1 = doff.o[pl] < min(o, 0) and doff.o[pu] = o - o1 > 0,
// ascending in o1 (the column of l); kind: 0 below the diagonal (divided by
// the pivot of its column), 1 the diagonal, 2 above with products, 3 above
// without any (never swept).
struct xs_dia_pairs {
  int np[XS_DIA_MAX];
  signed char pl[XS_DIA_MAX][XS_DIA_MAX], pu[XS_DIA_MAX][XS_DIA_MAX];
  signed char kind[XS_DIA_MAX];
  int diag;
};
struct xs_ell {         // column-major slabs: v[k * ld + row], width W
  int n = 0, ld = 0, W = 0;
  int *ci = nullptr;
  double *v = nullptr;
  short *co = nullptr;  // column - row as 16 bits (when every entry fits), else null
  // diagonal-slot copy (nd > 0): dv[k * ld + row] = a(row, row + doff.o[k])
  // or 0 where that entry is absent, the offsets ascending (the slots are in
  // column order, as the ELL rows); null when the matrix has more offsets
  int nd = 0;
  xs_dia_offs doff{};
  double *dv = nullptr;
};

// 16-bit column offsets: padding of the sweep copies (column -1)
#define XS_OFF16_PAD (-32768)
#define XS_OFF16_MAX 32767

// Combined ILU factors aligned with a CSR pattern: strict lower part of row i
// in [rp[i], dpos[i]), pivot at dpos[i], strict upper part in (dpos[i], rp[i+1]).
// ELL copy of the strictly lower (or strictly upper) part of the factors for
// the Jacobi-sweep apply: W column-major slabs of ld entries (entry k of row i
// at k * ld + i, in the row's column order), padding column -1 / value 0.
struct xs_tell {
  int W = 0, ld = 0;
  size_t cap = 0;            // allocated entries
  int *ci = nullptr;
  double *v = nullptr;
  void *vlo = nullptr;       // values for the mixed inner solver (factor storage type)
  size_t vlo_bytes = 0;      // allocated bytes of vlo
  short *co = nullptr;       // column - row as 16 bits, XS_OFF16_PAD = padding
  int *ww = nullptr;         // per 64-row group (rows 64g ..): its longest row, the
                             // slab count the sweeps walk for those rows
  double *vs = nullptr;      // U only: u_ij / u_jj, the values of the first two
                             // sweeps in one pass (no gather of 1/u_jj)
  // diagonal-slot copy for the FP64 sweeps (nd > 0): dv[k * ld + row] =
  // t(row, row + doff.o[k]) or 0, offsets ascending; dvs as vs (U only)
  int nd = 0;
  xs_dia_offs doff{};
  double *dv = nullptr, *dvs = nullptr;
  void *dvlo = nullptr;      // dv in the factor storage type (mixed inner solver)
  size_t dvlo_bytes = 0;
};

// Work order of a sync-free triangular solve: the row groups (gb: ngroups + 1
// boundaries in solve order -- ascending for L, descending for U; null =
// consecutive 64-row groups), the ticket -> group order (null = solve order)
// and the persistent workgroup count.
struct xs_tsched {
  const int *gb = nullptr;
  const int *order = nullptr;
  int ngroups = 0;
  int max_wg = 256;
  int gather = 4;   // off-group entries a lane loads and polls together (4 or 16)
  // 1: every group is solved lane by lane, each row published as soon as it
  // is final; 0: a group whose rows depend at most on the row before waits
  // for all its external values and is solved by one affine scan
  int prog = 0;
};

// Joint polling width of a factor's triangular solves: 16 off-group entries
// at a time when a triangle holds more than 6 entries per row on average
// (ParILUT fill), else 4 (wider gathers measured slower on 7-point factors).
inline int xs_trsv_gather_width(long nnz, int n) {
  return n > 0 && (nnz - (long)n) > 12L * n ? 16 : 4;
}

struct xs_ilu {
  int n = 0, nnz = 0;
  const int *rp = nullptr, *ci = nullptr;  // pattern (owned elsewhere or below)
  int *rp_own = nullptr, *ci_own = nullptr;
  double *lu = nullptr;     // factor values
  int *dpos = nullptr;      // pivot positions
  double *dinv = nullptr;   // 1 / u_ii
  double *scale = nullptr;  // |a_ii|^{-1/2} used while factorising
  // CSC of the upper triangle (incl. diagonal) for the ParILU sweep:
  int *ucp = nullptr, *uri = nullptr, *upos = nullptr;
  int *rowidx = nullptr;    // row of each nnz (COO)
  // ParILU(0) update lists (no reference counterpart): entry k's products
  // l_ik' u_k'j are lu[pl[e]] * lu[pu[e]] for e in [pp[k], pp[k+1]) in
  // ascending k'; pd[k] = position of u_jj below the diagonal, -2 on it, -1
  // above.
  int *pp = nullptr, *pl = nullptr, *pu = nullptr, *pd = nullptr;
  // entries a sweep can change (those with products, the pivots and the
  // entries below the diagonal); the others keep their start value
  int *act = nullptr;
  // per active entry t (entry act[t]): its products pl/pu[app[t] .. app[t+1])
  // and pivot code, so that a sweep reads them coalesced
  int *app = nullptr, *apd = nullptr;
  int nact = 0;
  // The sweeps run on the values in sweep order: entry k at cmap[k], the
  // active entries first (act[t] at t), the others after them in CSR order.
  // pl, pu and apd hold sweep-order positions; as = the (scaled) start
  // values and lc the iterate, both in sweep order.
  int *cmap = nullptr;
  double *as = nullptr, *lc = nullptr;
  // which update lists the pattern carries: 0 none, 1 ParILU(0), 2 ParIC(0)
  int lists = 0;
  // ParILU(0) over diagonal slots (A with at most XS_DIA_MAX offsets): pnd
  // slots of pld rows; pmask[i] bit s = row i has slot s; dmap[k] = the slot
  // position (s * pld + i) of CSR entry k; das the (scaled) start values and
  // dlc / dlc2 the iterates there (dlc2: the deterministic sweeps' second
  // buffer).  Positions without an entry hold 0 and are never written.
  int pnd = 0, pld = 0;
  xs_dia_offs poff{};
  xs_dia_pairs ppairs{};
  unsigned char *pmask = nullptr;
  int *dmap = nullptr;
  double *das = nullptr, *dlc = nullptr, *dlc2 = nullptr;
  // ParIC(0): position of each entry's mirror (ci[k], rowidx[k]) in A
  int *mpos = nullptr;
  // SpTRSV work order: the 64-row groups of L (resp. U, counted from the
  // bottom) sorted by dependency level; ticket t solves group sched[t]
  int *lsched = nullptr, *usched = nullptr;
  // row-group boundaries (at most 64 rows, cut where a row does not depend
  // on its neighbour in solve order) and their counts
  int *lgb = nullptr, *ugb = nullptr;
  int lgroups = 0, ugroups = 0;
  // the factor pattern changed on the device (ParILUT) and the groups and
  // their order are built from it on the first exact triangular solve
  bool sched_pending = false;
  int trsv_wg = 256;        // persistent workgroups of the triangular solves
  int trsv_prog = -1;       // xs_tsched::prog; -1 = by factor width (wide -> 1)
  int trsv_gather = 0;      // xs_tsched::gather (4 or 16); 0 = by factor width
  xs_tell tl, tu;           // sweep copies of L and U (built on demand)
  bool tell_shape = false;  // tl/tu widths and buffers match the pattern
  // transposed factors for the adjoint (built on demand), with the row
  // groups and level order of their triangular solves (as lgb / lsched)
  bool have_t = false;
  int *trp = nullptr, *tci = nullptr, *tdpos = nullptr, *tperm = nullptr;
  double *tlu = nullptr;
  int *tlgb = nullptr, *tugb = nullptr, *tlsched = nullptr, *tusched = nullptr;
  int tlgroups = 0, tugroups = 0;
};

// the gather width of a factor's triangular solves: the set one, else by width
inline int xs_ilu_gather(const xs_ilu &F) {
  return F.trsv_gather > 0 ? F.trsv_gather : xs_trsv_gather_width(F.nnz, F.n);
}

inline xs_tsched xs_ilu_tsched(const xs_ilu &F, bool upper) {
  xs_tsched t;
  t.gb = upper ? F.ugb : F.lgb;
  t.ngroups = upper ? F.ugroups : F.lgroups;
  t.order = upper ? F.usched : F.lsched;
  t.max_wg = F.trsv_wg;
  t.gather = xs_ilu_gather(F);
  t.prog = F.trsv_prog >= 0 ? F.trsv_prog : (t.gather == 16 ? 1 : 0);
  return t;
}

// ---------------------------------------------------------------------------
// Kernel launchers (xs_kernels.hip)
// ---------------------------------------------------------------------------
// SpMV epilogues: y = Ax | y = b - Ax | y = Ax, w.y | y = Ax, w.y, y.y | y = b - Ax, y.y
enum xs_spmv_epi { EPI_Y = 0, EPI_BMAX, EPI_DOT1, EPI_DOT2, EPI_BMAXN };

void xs_launch_spmv_csr(hipStream_t s, const xs_csr &A, const xs_spmv_plan &P,
                        const double *x, double *y, int epi, const double *b,
                        const double *w, const xs_reduce *red);
void xs_launch_spmv_ell(hipStream_t s, const xs_ell &E, const double *x, double *y,
                        int epi, const double *b, const double *w, const xs_reduce *red);
void xs_launch_spmv_t(hipStream_t s, const xs_csr &A, const double *x, double *y);
void xs_build_plan(hipStream_t s, const xs_csr &A, xs_spmv_plan &P, int *blk, int chunk);
int xs_default_chunk(const xs_csr &A);
void xs_build_ell(hipStream_t s, const xs_csr &A, int W, xs_ell &E);
// Distinct column - row offsets of A when there are at most XS_DIA_MAX of
// them (ascending, count returned), else 0; synchronous, -1 on failure.
// d_tmp: XS_DIA_MAX + 1 ints of device scratch.
int xs_dia_offsets(hipStream_t s, const xs_csr &A, int *d_tmp, xs_dia_offs &out);
// E.dv from A's values in the slots of E.doff (E.nd slots, E.ld rows)
void xs_build_dia(hipStream_t s, const xs_csr &A, xs_ell &E);
// The same for the strictly lower (upper) part of the factors: offsets, then
// E.dv (and for U E.dvs = u_ij / u_jj) from the part's current ELL copy.
int xs_tri_dia_offsets(hipStream_t s, const xs_ilu &F, bool upper, int *d_tmp, xs_dia_offs &out);
void xs_build_tri_dia(hipStream_t s, const xs_ilu &F, bool upper, xs_tell &E);
// ParILU(0) over diagonal slots: the slot map and row masks of A's pattern,
// the start values (and copies into lc2 when not null), one sweep (lc_in ==
// lc_out: asynchronous in place), and the factor back into F.lu in CSR order
void xs_parilu_dia_map(hipStream_t s, const xs_csr &A, xs_ilu &F);
void xs_parilu_dia_start(hipStream_t s, const xs_csr &A, const xs_ilu &F, const double *sc,
                         double *lc2);
void xs_parilu_dia_sweep(hipStream_t s, const xs_ilu &F, const double *lc_in, double *lc_out);
void xs_parilu_dia_finish(hipStream_t s, const xs_csr &A, const xs_ilu &F, const double *sc,
                          const double *lc);
// Largest |column - row| over the non-padding entries of ELL slabs (rows < n,
// column >= 0); synchronous, -1 on failure.  Then the 16-bit offsets: padding
// columns (-1) become XS_OFF16_PAD, rows >= n point at row n - 1.
int xs_ell_max_offset(hipStream_t s, int n, int ld, int W, const int *eci, int *d_tmp);
void xs_ell_build_offsets(hipStream_t s, int n, int ld, int W, const int *eci, short *eco);
int xs_max_row_len(hipStream_t s, const xs_csr &A, int *d_tmp);

// fills / vector ops
void xs_fill_sentinel(hipStream_t s, double *x, int n);
void xs_launch_dot(hipStream_t s, int n, const double *a, const double *b,
                   const double *c, const double *d, const xs_reduce &red);

// Krylov fused vector kernels
void xs_bicg_p(hipStream_t s, int n, const double *r, const double *u, double *p,
               const double *dinv, double *z, xs_state *st);
void xs_bicg_s(hipStream_t s, int n, const double *r, const double *u, double *sv,
               const double *dinv, double *z1, xs_state *st);
void xs_bicg_xr(hipStream_t s, int n, double *x, double *r, const double *z,
                const double *z1, const double *sv, const double *v,
                const double *rh, const xs_reduce &red);
void xs_cg_xr(hipStream_t s, int n, double *x, double *r, const double *p,
              const double *Ap, const double *dinv, double *z, const xs_reduce &red);
void xs_cg_zr_dot(hipStream_t s, int n, const double *z, const double *r, const xs_reduce &red);
void xs_cg_p(hipStream_t s, int n, double *p, const double *z, xs_state *st);
void xs_scale(hipStream_t s, int n, const double *d, const double *x, double *y);
void xs_copy(hipStream_t s, int n, const double *x, double *y);
void xs_init_state(hipStream_t s, xs_state *st, double tol, double abstol, int maxiter);

// ILU
void xs_ilu_copy_values(hipStream_t s, const xs_csr &A, xs_ilu &F);
// ParILU(0) update lists: count (pl == null: matches into cnt, pd filled) or
// fill (pl, pu at offsets pp)
void xs_parilu_pairs(hipStream_t s, const xs_csr &A, const xs_ilu &F, int *cnt, const int *pp,
                     int *pl, int *pu, int *pd);
// sweep order: cmap[k] = pos[k] for a flagged entry, nact + k - pos[k]
// otherwise (pos = exclusive scan of flag); idx[e] <- cmap[idx[e]]
void xs_parilu_cmap(hipStream_t s, int nnz, int nact, const int *flag, const int *pos, int *cmap);
void xs_remap_idx(hipStream_t s, int n, const int *cmap, int *idx);
// start values a_k (times s_i s_j with sc) in sweep order into F.as, F.lc
// and lc2 (when given)
void xs_parilu_start(hipStream_t s, const xs_csr &A, const xs_ilu &F, const double *sc,
                     double *lc2);
// F.lu[k] = lc[cmap[k]], scaled back with sc (as xs_pt_rescale with unscale)
void xs_parilu_finish(hipStream_t s, const xs_csr &A, const xs_ilu &F, const double *sc,
                      const double *lc);
// one ParILU sweep over the update lists in sweep order (same arithmetic as
// xs_parilu_sweep).  app[t] = pp[act[t]], apd[t] = cmap[pd[act[t]]] (codes
// -1 / -2 kept), app[nact] = pp[nnz] (the entries between two active ones
// have no products)
void xs_parilu_active_lists(hipStream_t s, int nact, int nnz, const int *act, const int *pp,
                            const int *pd, const int *cmap, int *app, int *apd);
void xs_parilu_sweep_pairs(hipStream_t s, const xs_ilu &F, const double *lc_in, double *lc_out);
// active-entry list: flag[k] = 1 for entries a sweep can change; compaction
// of the flagged indices at their scanned positions
void xs_parilu_active_flags(hipStream_t s, int nnz, const int *cnt, const int *pd, int *flag);
void xs_compact_flagged(hipStream_t s, int nnz, const int *flag, const int *pos, int *out);
void xs_parilu_sweep(hipStream_t s, const xs_csr &A, xs_ilu &F, double *lu_out, bool async,
                     const double *sc);
void xs_ilu_dinv(hipStream_t s, xs_ilu &F, int *d_bad);
void xs_trsv_lower(hipStream_t s, const xs_ilu &F, const double *b, double *x, unsigned *tickets,
                   const xs_state *st);
void xs_trsv_upper(hipStream_t s, const xs_ilu &F, const double *b, double *x, unsigned *tickets,
                   const xs_state *st);
void xs_trsv_lower_t(hipStream_t s, const xs_ilu &F, const double *b, double *x,
                     unsigned *tickets, const xs_state *st);
void xs_trsv_upper_t(hipStream_t s, const xs_ilu &F, const double *b, double *x,
                     unsigned *tickets, const xs_state *st);
// one Jacobi sweep of a factor: yn = D (b - T_off yo); yo == null: yn = D b
void xs_tri_sweep(hipStream_t s, const xs_ilu &F, bool upper, const double *b, const double *yo,
                  double *yn, const xs_state *st);
// Jacobi sweeps over the ELL copies: fused = sweeps 1 and 2 in one pass
// (yo implied: b for L, D b for U), else one sweep from yo.  xs_tri_widths
// returns the widths of L and U (synchronous; nonzero on failure).
int xs_tri_widths(hipStream_t s, const xs_ilu &F, int *d_tmp, int *wl, int *wu);
void xs_build_tell(hipStream_t s, const xs_ilu &F, bool upper, xs_tell &E);
// E.vs = E.v with every entry divided by the pivot of its column (dinv)
void xs_tell_scale_cols(hipStream_t s, const xs_tell &E, const double *dinv);
// E.ww[g] = the longest strictly lower (upper) row among rows 64g .. 64g + 63
void xs_tri_wave_widths(hipStream_t s, const xs_ilu &F, bool upper, int *ww);
void xs_tri_sweep_ell(hipStream_t s, const xs_ilu &F, bool upper, bool fused, const double *b,
                      const double *yo, double *yn, const xs_state *st);
// the fused first L pass with BiCGSTAB's s = r - gama u (gama in st) formed
// in the pass and written to sv
void xs_tri_sweep_ell_s(hipStream_t s, const xs_ilu &F, const double *r, const double *u,
                        double *sv, double *yn, const xs_state *st);
// the fused first L pass with BiCGSTAB's p = r + omega (pin - alpha u) formed
// in the pass and written to pout (pin != pout)
void xs_tri_sweep_ell_p(hipStream_t s, const xs_ilu &F, const double *r, const double *u,
                        const double *pin, double *pout, double *yn, const xs_state *st);
void xs_permute_vals(hipStream_t s, int nnz, const int *perm, const double *src, double *dst);
void xs_grad_vals(hipStream_t s, const xs_csr &A, const double *lam, const double *x, double *g);
void xs_jacobi_dinv(hipStream_t s, const xs_csr &A, double *dinv, int *d_bad);
// 1 if a launcher refused to launch since the last call (and clears it)
int xs_take_launch_guard();
void xs_raise_launch_guard();
// Launch guard: a launcher that finds a required pointer null (a set-up
// ordering bug) or a size out of range does not launch; it raises the flag,
// which the host driver turns into an error instead of letting a kernel touch
// address 0.
#define XS_GUARD(cond)            \
  do {                            \
    if (!(cond)) {                \
      xs_raise_launch_guard();    \
      return;                     \
    }                             \
  } while (0)
