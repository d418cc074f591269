// Host-side internal interfaces of the qexhip library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include "common.h"

namespace qh {

// Device-resident solver scalars.  Every per-solve parameter (mass, tolerances)
// lives here, so kernels and captured graphs do not depend on them.
struct Scal {
  double m2;        // operator mass^2: A = 4*m2 - (2D)(2D)
  double r2;        // latest residual^2 (one-step extrapolated, scaled units in mixed mode)
  double r2prev;    // exact residual^2 of the previous iterate (scaled units in mixed mode)
  double beta, alpha, pAp;
  double r2stop;    // absolute stopping threshold, unscaled
  double r2stop_s;  // r2stop in scaled units (mixed mode)
  double b2;
  double maxr2;     // mixed: r2 right after the last reliable update (scaled)
  double delta2;    // mixed: reliable update when r2 < delta2*maxr2
  double r2true;    // last FP64 true residual^2
  double sigma;     // mixed: sloppy vectors hold sigma*(true quantity)
  double pscale;    // mixed: extra factor on p in the next p update
  double tmp[4];    // generic reduction results
  int its, maxits, since, relmax;
  int done, need_rel, first, nrel;
  int mixed;
  int restart;       // mixed: next reliable update restarts the recursion (p = r)
  int rollback;      // mixed: the true residual grew; host returns x to the last good iterate
  int resinc;        // mixed: reliable updates whose true residual grew
  int improved;      // mixed: the last reliable update reached a new best residual
  int flushed;       // batched FP64 CG: the pending update of a converged system was applied
  double r2rel_last; // mixed: best true residual^2 at a reliable update
};

// mixed mode gives up after this many reliable updates that failed to reduce
// the true residual
#define QH_MAX_RES_INCREASE 20
// growth of the true residual^2 over the best one that counts as divergence
#define QH_RES_GROWTH 10.0

#define QH_MAXSHIFT 64

// Generic one-parity hop: out = ca*A + cb*B + chop*(2D)in on parity par.
struct HopArgs {
  Geom g;
  const char *links;  // link blocks [k][parity], LFmt layout
  long lkblk;         // bytes per (k, parity) block
  float qs[8];        // PQ16 links: value = int16 * qs[k]
  const void *in;     // other-parity input
  void *out;          // own-parity output
  const void *a;      // own-parity term A (precision PA)
  const void *b;      // own-parity term B (precision PA)
  double ca, cb, chop;
  int ca_m2;          // 1: ca is multiplied by 4*S->m2
  int par;            // output parity
  int pred;           // 0 none; 1 skip when done|need_rel
  int remap;          // XCD-aware block order
  Scal *S;
  RedWs red;
  long chunk;         // site walk (make_walk): t-march chunk, 0 = canonical order
  int tmarch;         // > 0 (plain links): t-march form, t-slices per thread (k_hop_tm)
  int two;            // 1: the stencil loop takes both links of a hop per trip (k_hop's TWO)
  int split;          // 2 or 4: that many threads per output site (k_hop's NS); else one
};

// Fused CG iteration (two kernels): see dslash.hip.
struct CgArgs {
  Geom g;
  const char *links;
  long lkblk;
  float qs[8];        // PQ16 links: value = int16 * qs[k]
  int par;            // solve parity; the first kernel writes the other parity
  int remap;
  void *x, *r, *p, *Ap, *t;
  Scal *S;
  RedWs red;
};

// Finish step of a reduction (reduce.hip): sum the block partials of the
// kernel just launched and apply one epilogue to the solver scalars.
enum Epi { EPI_TMP = 0, EPI_PAP = 1, EPI_RELIABLE = 2, EPI_CG2 = 3, EPI_FLUSH = 4 };
struct FinArgs {
  RedWs red;
  int nb;             // blocks of the producing kernel
  Scal *S;
  int slot;           // EPI_TMP: S->tmp[slot]
  double nreal;       // EPI_RELIABLE: real components of one parity vector
  int pred;           // 0 none; 1 skip when done|need_rel
};

// One system of a batched FP64 CG (batch.hip): blocks [blk0, blk0+nblk) of
// the batched grid belong to it.
struct BatchItem {
  CgArgs a;           // its CG arguments (a.red: the batch partials, slot = grid block)
  FinArgs fin;        // finish of the iteration dot products (its partials slice)
  FinArgs finf;       // finish of the flush
  long blk0, nblk;
};
#define QH_MAXBATCH 64

// One right-hand side of the batched multi-shift CG (batch.hip); shift 0 is
// the base system A = 4(m^2 - D_eo D_oe), shift k > 0 adds sg[k].
// The batched multi-shift solve advances every shifted x_k, p_k once per
// window of QH_MSB_WIN iterations (one graph chunk), from the residuals of the
// window kept in a ring of QH_MSB_RING vectors per right-hand side.
#define QH_MSB_WIN 16
#define QH_MSB_RING (QH_MSB_WIN + 1)
#define QH_MSB_NCOEF (2 + 2 * QH_MSB_WIN)
struct MsbScal {
  double r2i, r2stop, alpha, beta, alphaim1, betaim1;
  int its, maxits, continuing, nshift;
  int freeze;                  // stop updating shifts whose residual met r2stop
  int its0;                    // first iteration of the current window
  int winflow;                 // set if a window outgrew QH_MSB_WIN (never, with flushes every chunk)
  int pad;
  double sg[QH_MAXSHIFT], zi[QH_MAXSHIFT], zim1[QH_MAXSHIFT];
  double xc[QH_MAXSHIFT];      // x_k += xc[k] p_k (old p_k)
  double pc1[QH_MAXSHIFT];     // p_k = pc1[k] r_new + pc2[k] p_k
  double pc2[QH_MAXSHIFT];
  int active[QH_MAXSHIFT];
};
#define QH_MAXRHS 8
struct MsbPtrs {              // one group of nr right-hand sides
  double *p0[QH_MAXRHS];        // base search vectors (own parity), updated every iteration
  double *t[QH_MAXRHS];         // (2D) p0 on the other parity
  double *Ap[QH_MAXRHS];
  double *r[QH_MAXRHS];         // residual ring: QH_MSB_RING vectors of rstride doubles; r after
                                // iteration i in slot i % QH_MSB_RING
  double *p0s[QH_MAXRHS];       // p0 at the start of the window
  long rstride;
  double **x, **p;              // device arrays [nr*nshift]: shifted solutions and search vectors
  double *hist;                 // [nr][QH_MSB_WIN][nshift][3]: per iteration of the window and
                                // shift, (xc, pc1, pc2) of x_k += xc p_k, p_k = pc1 r + pc2 p_k
  double *coef;                 // [nr][nshift][QH_MSB_NCOEF]: window coefficients (k_msb_coef)
  int *cinfo;                   // [nr][2 + QH_MAXSHIFT]: window length, first slot, shift activity
};

struct MultiPtrs {     // one group of up to QH_MAXRHS fields of the multi-RHS stagD2
  const void *in[QH_MAXRHS];   // other-parity inputs
  const void *own[QH_MAXRHS];  // own-parity x
  void *out[QH_MAXRHS];        // own-parity r
};

struct HopKey {       // precision/format of one hop instantiation
  int pl, rc, pi, po, pa;
};

// reduce.hip
void launch_finish(int epi, const FinArgs &f, hipStream_t st);
// the CG finish into a second scalar buffer: Sout = Sin with EPI_CG2 applied
void launch_finish_cg2_to(const RedWs &red, int nb, const Scal *Sin, Scal *Sout, hipStream_t st);

// dslash.hip
void launch_hop(const HopKey &k, int nl, int mode, int redm, const HopArgs &a, hipStream_t st, int block, int cap);
// cap > 0: form 3's stencil kernels persistent on at most cap workgroups
void launch_cg_iter(int pl, int rc, int ps, int nl, const CgArgs &a, hipStream_t st, int block, int form,
                    int rx = P32, int block1 = 0, int cap = 0);
// niter iterations of CG form 2, 3 or 4 (4: scalars alternate between a.S and
// S1, the streaming kernel runs g0 workgroups); false if no instantiation
bool launch_cg_chunk(int pl, int rc, int ps, int nl, const CgArgs &a, hipStream_t st, int block, int form, int rx,
                     int block1, int niter, Scal *S1, int g0);
void launch_cg_flush(int ps, int rx, const CgArgs &a, hipStream_t st);
void launch_pack_links(int prec, int rc, const Geom &g, int nl, const double *canon, char *out, const float *qs,
                       hipStream_t st);
void launch_link_maxabs(const Geom &g, int nl, const double *canon, unsigned *maxbits, hipStream_t st);
void launch_recon_check(const Geom &g, const double *canon, unsigned *maxdev_bits, hipStream_t st);

// kernels.hip
void launch_pack_vec(int prec, const Geom &g, const double *canon, void *out, hipStream_t st);
void launch_unpack_vec(int prec, const Geom &g, const void *in, double *canon, hipStream_t st);
void launch_fill(void *p, size_t bytes, hipStream_t st);
void launch_stream_read(const void *in, long n16, double *out, int blocks, hipStream_t st);
void launch_fill_pattern(void *p, size_t bytes, hipStream_t st);
void launch_axpby64(const Geom &g, double a, const double *x, double b, double *y, hipStream_t st);
void launch_norm2(int prec, const Geom &g, const void *x, Scal *S, int slot, RedWs red, hipStream_t st);
void launch_redot(int prec, const Geom &g, const void *x, const void *y, Scal *S, int slot, RedWs red,
                  hipStream_t st);
void launch_rs_from_true(const Geom &g, void *rs, const double *rtrue, Scal *S, int pr, hipStream_t st);
void launch_xfold(const Geom &g, double *x, void *xs, Scal *S, int px, hipStream_t st);
// best = x when the last reliable update improved on the best residual and did not roll back
void launch_keep_best(const Geom &g, double *best, const double *x, const Scal *S, hipStream_t st);
void launch_scal_init(Scal *S, hipStream_t st, int mode);
// the multi-shift updates of the mixed solve's FP32 recursion (FP64 arithmetic);
// x, p: device arrays of nshift vector pointers
void launch_cgm_start32(const Geom &g, int nshift, const double *b, void *r, void *const *x, void *const *p,
                        hipStream_t st);
void launch_cgm_export32(const Geom &g, int nshift, void *const *x32, double *const *x64, hipStream_t st);
void launch_outer_acc(const Geom &g, const double *a_e, const double *a_o, const double *b_e, const double *b_o,
                      double ce, double co, double *F, int nl, hipStream_t st);
void launch_force(const Geom &g, const char *links64, long lkblk, const double *psi_e, const double *psi_o,
                  double *F, int outer_only, hipStream_t st, int nl = 4);

// batch.hip
void launch_bcg_iter(const BatchItem *T, int n, long nblocks, int nl, int block, hipStream_t st, int form);
void launch_bcg_flush(const BatchItem *T, int n, long nblocks, int block, hipStream_t st);
// false when no instantiation serves (prec, nl, nr, batches); chunk: site walk (make_walk)
bool launch_hop_multi(int prec, int nl, int nr, int batches, const Geom &g, const char *links, long lkblk, int par,
                      double ca, double cb, const MultiPtrs &P, int block, int cap, long chunk, hipStream_t st);
// workgroups of the multi-RHS multi-shift hops at this block size and site walk
unsigned mrhs_grid(const Geom &g, int block, long chunk);
// chunk: site walk of the two multi-RHS hops (make_walk; 0 = canonical order);
// batches: the hops load the right-hand sides' vectors of a link in 1, 2 or 4 batches;
// pv: vector storage and link precision (P64; P32 for a group of one right-hand side)
void launch_msb_iter(const Geom &g, int pv, const char *links, long lkblk, int nl, int par, int nr, int nshift,
                     double m2, const MsbPtrs &P, MsbScal *T, RedWs red, int block, long chunk, int batches,
                     hipStream_t st);
// end of a window: every x_k, p_k of the group advanced over the window's iterations
void launch_msb_flush(const Geom &g, int pv, int nr, int nshift, const MsbPtrs &P, MsbScal *T, hipStream_t st);

// gauge.hip
void launch_plaquette(const Geom &g, const double *U, Scal *S, int slot, RedWs red, hipStream_t st);
void launch_wilson_force(const Geom &g, const double *U, double beta, double *F, hipStream_t st);
void launch_random_momenta(const Geom &g, uint64_t seed, double *P, hipStream_t st);
void launch_link_update(const Geom &g, const double *P, double eps, double *U, hipStream_t st);

// smear.hip: HISQ link construction.  Workspace of 36 component-major link
// fields (input copy, 3-link staples, summed 5-link staples, first-level fat
// links, projected links), kept between calls on one lattice.
struct SmearWs {
  double2 *buf = nullptr;
  long V = 0;
  int nfields = 0;
  SmearWs() = default;
  SmearWs(const SmearWs &) = delete;
  SmearWs &operator=(const SmearWs &) = delete;
  ~SmearWs();
};
// events around k_fat3, k_fat5 and k_fatf of one fat7 level
struct SmearTimes {
  hipEvent_t ev[4];
};
void launch_make_imp_links(const Geom &g, SmearWs &w, double *fl, const double *gf, const double coef[5], double *ll,
                           const double *gl, double naik, int remap, hipStream_t st, SmearTimes *tm);
void launch_project_u(long n, double *u, const double *x, hipStream_t st);
void launch_hisq_smear(const Geom &g, SmearWs &w, const double first[5], const double second[5], double naik,
                       const double *gin, double *fl, double *ll, int remap, hipStream_t st, SmearTimes *tm);
// chain rule (derivatives in the pairing Re tr(A^+ B)); deriv/llderiv/dg accumulate
void launch_fat7l_deriv(const Geom &g, SmearWs &w, double *deriv, const double *gf, const double *mid,
                        const double coef[5], double *llderiv, const double *gl, const double *llmid, double naik,
                        int remap, hipStream_t st);
void launch_project_u_deriv(long n, double *r, const double *u, const double *x, const double *chain, hipStream_t st);
// dfl/dll: 4 fields each, element (mu, site) at base + (mu*mstride + site)*18 doubles
void launch_hisq_smear_deriv(const Geom &g, SmearWs &w, const double first[5], const double second[5], double naik,
                             const double *gin, const double *dfl, const double *dll, long mstride, double *dg,
                             int acc, int remap, hipStream_t st);
// 8 site-major link fields for the HISQ force outer products, kept in the
// smearing workspace (past the fields the chain rule uses)
double *smear_force_outer(SmearWs &w, long V);
// F_mu(x) = TAH(D_mu(x) U_mu(x)^+), in place on F = D allowed
void launch_tah_mul(long V, double *F, const double *D, const double *U, hipStream_t st);

// milcio.hip: MILC binary gauge files (single precision, site order)
struct MilcInfo {
  int dims[4];
  char stamp[65];
  int order, byteswapped;
  uint32_t sum29, sum31;
  int checksum_ok;
};
void milc_checksum(const uint32_t *w, size_t n, uint32_t *s29, uint32_t *s31);
void milc_read_header(const char *path, MilcInfo *info);
void milc_read(const Geom &g, const char *path, double *U, MilcInfo *info, hipStream_t st);
void milc_write(const Geom &g, const char *path, const double *U, const char *stamp, hipStream_t st);

// scidac.hip: SciDAC/ILDG (LIME) files; field data are canonical complex128
// arrays [nobj][site][nelem] on the device, nelem 9 (3x3 matrix), 3 (color
// vector) or 1 (complex number), nobj 1..4
struct ScidacInfo {
  int dims[4] = {0, 0, 0, 0};
  int nrecords = 0;
  std::string filemd, recmd, datatype, date;
  char prec = 0;
  int colors = 0, spins = 0, typesize = 0, datacount = 0;
  int ildg = 0, has_checksum = 0, checksum_ok = 0;
  uint32_t suma = 0, sumb = 0, calc_suma = 0, calc_sumb = 0;
};
void scidac_info(const char *path, int record, ScidacInfo *info);
void scidac_read(const Geom &g, const char *path, int record, double *dst, int nobj, int nelem, ScidacInfo *info,
                 hipStream_t st);
void scidac_create(const char *path, const int dims[4], const char *filemd);
void scidac_append(const Geom &g, const char *path, const char *recmd, const double *src, int nobj, int nelem,
                   char prec, int ildg, const char *date, ScidacInfo *info, hipStream_t st);
// qexlayout.hip: QEX's one-rank SIMD field storage (innerGeom lanes) <->
// canonical complex128 [nfields][site][ncomp]
void qex_layout_convert(const Geom &g, const int inner_geom[4], int from_qex, double *qex, int nfields, int ncomp,
                        double *canon, hipStream_t st);
// ms[0], ms[1]: device time per launch of the whole-lattice pack and unpack kernels
void scidac_time(const Geom &g, const double *src, int nobj, int nelem, char prec, int nrep, double ms[2],
                 hipStream_t st);

// gen.hip
void launch_gauge_random(const Geom &g, uint64_t seed, double *canon, hipStream_t st);
void launch_source_gauss(const Geom &g, uint64_t seed, int stream, int par, double *canon, hipStream_t st);
void launch_phase_bc(const Geom &g, double *canon, int nl, hipStream_t st);
void launch_hisq_links(const Geom &g, uint64_t seed, const double *Uph, double *g8, hipStream_t st);

}  // namespace qh
