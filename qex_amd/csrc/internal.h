// Host-side internal interfaces of the qexhip library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include "common.h"

namespace qh {

// Device-resident solver scalars.  Every per-solve parameter (mass, tolerances)
// lives here, so captured graphs do not depend on them.
struct Scal {
  double m2;        // operator mass^2: A = 4(m2 + shift) - (2D)(2D)
  double r2;        // current recursive residual^2 (scaled units in mixed mode)
  double r2old;     // r2 at the start of the current iteration
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
  int mixed, pad0, pad1, pad2;
  // multi-shift recurrences (cgm.nim:151-200), up to QH_MAXSHIFT shifts
};

#define QH_MAXSHIFT 64
struct ShiftScal {
  double sg[QH_MAXSHIFT];
  double zi[QH_MAXSHIFT], zim1[QH_MAXSHIFT];
  double xc[QH_MAXSHIFT];   // x_k += xc[k] * p_k      (alpha*zr)
  double pc1[QH_MAXSHIFT];  // p_k = pc1[k]*r + pc2[k]*p_k
  double pc2[QH_MAXSHIFT];
  double alpha, beta, alphaim1, betaim1, r2i, r2ip1;
  int nshift, continuing, its, maxits;
  double r2stop;
};

// Reduction workspace shared by consecutive kernels on one stream.
struct RedWs {
  double *partials;   // [4][maxblocks]
  unsigned *counter;
  int maxblocks;
};

struct HopArgs {
  Geom g;
  const void *links;  // [k][parity][ncomp][stride]
  long lkblk;         // ncomp*stride complex words per (k, parity)
  const void *in;     // other-parity input
  void *out;          // own-parity output
  const void *a;      // own-parity term A (precision PA)
  const void *b;      // own-parity term B (precision PA)
  double ca, cb, chop;
  int ca_m2;          // 1: ca is multiplied by 4*S->m2
  int par;            // output parity
  int pred;           // 0 none; 1 skip when done|need_rel; 2 run only when need_rel
  int fuse_x;         // reliable update: a := a + sigma^-1 * xs, xs := 0 (own parity)
  void *xs;           // sloppy solution (P32) for fuse_x
  Scal *S;
  RedWs red;
};

// launchers (kernels.hip)
void launch_hop(int PL, int PI, int PO, int PA, bool dbl, int nl, int mode, int redmode,
                const HopArgs &a, hipStream_t st, int block);
void launch_pack_vec(int prec, const Geom &g, const double *canon, void *out, hipStream_t st);
void launch_unpack_vec(int prec, const Geom &g, const void *in, double *canon, hipStream_t st);
void launch_pack_links(int prec, int recon, const Geom &g, int nl, const double *canon, void *out,
                       hipStream_t st);
void launch_fill(void *p, size_t bytes, hipStream_t st);
// y = a*x + b*y on one parity (FP64 internal layout); sc pointers may scale a
void launch_axpby64(const Geom &g, double a, const double *x, double b, double *y, hipStream_t st);
// reductions into S->tmp[slot] (FP64 accumulation)
void launch_norm2(int prec, const Geom &g, const void *x, Scal *S, int slot, RedWs red, hipStream_t st);
void launch_redot(int prec, const Geom &g, const void *x, const void *y, Scal *S, int slot, RedWs red,
                  hipStream_t st);
// CG vector kernels
void launch_pupdate(int PP, int PR, const Geom &g, void *p, const void *r, Scal *S, hipStream_t st);
void launch_xr(int PX, int PP, int PR, int PAp, const Geom &g, void *x, const void *p, void *r,
               const void *Ap, Scal *S, RedWs red, hipStream_t st);
void launch_rs_from_true(int PR, const Geom &g, void *rs, const double *rtrue, Scal *S, hipStream_t st);
void launch_xfold(const Geom &g, double *x, void *xs, Scal *S, hipStream_t st);
void launch_scal_init(Scal *S, hipStream_t st, int mode);
// multi-shift
void launch_cgm_xr(const Geom &g, int nshift, double *r, const double *Ap, double **x, double **p,
                   ShiftScal *T, Scal *S, RedWs red, hipStream_t st);
void launch_cgm_p(const Geom &g, int nshift, const double *r, double **p, ShiftScal *T, hipStream_t st);
void launch_cgm_alpha(ShiftScal *T, Scal *S, hipStream_t st);
// force
void launch_force(const Geom &g, const double *links64, long lkblk, const double *psi_e, const double *psi_o,
                  double *F, int outer_only, hipStream_t st);
// synthetic inputs (gen.hip)
void launch_gauge_random(const Geom &g, uint64_t seed, double *canon, hipStream_t st);
void launch_source_gauss(const Geom &g, uint64_t seed, int stream, int par, double *canon, hipStream_t st);
void launch_phase_bc(const Geom &g, double *canon, int nl, hipStream_t st);
void launch_hisq_links(const Geom &g, uint64_t seed, const double *Uph, double *g8, hipStream_t st);

}  // namespace qh
