// C-ABI entry points of the qexhip library and the host-side solver drivers.
// The drivers restate QEX's algorithms (cited per function); all vector work
// runs in the kernels of kernels.hip, on one library-owned HIP stream.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <chrono>
#include <stdexcept>
#include <string>
#include <vector>

#include "internal.h"
#include "qexhip.h"

using namespace qh;

namespace {

thread_local std::string g_err;
hipStream_t g_stream = nullptr;
int g_device = -1;
int g_block = 256;
int g_remap = 1;

#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess)                                                                          \
      throw std::runtime_error(std::string(#x) + " failed: " + hipGetErrorString(e_));             \
  } while (0)

#define API_BEGIN try {
#define API_END                                                                                    \
  }                                                                                                \
  catch (const std::exception &e_) {                                                               \
    g_err = e_.what();                                                                             \
    return 1;                                                                                      \
  }                                                                                                \
  return 0;

void require(bool c, const char *msg) {
  if (!c) throw std::runtime_error(msg);
}

void check_init() {
  if (!g_stream) throw std::runtime_error("qexhip_init has not been called");
}

void check_launch() { HIPCHK(hipGetLastError()); }

template <class T>
T *dalloc(size_t bytes) {
  void *p = nullptr;
  HIPCHK(hipMalloc(&p, bytes ? bytes : 16));
  return (T *)p;
}

size_t word_bytes(int prec) {
  switch (prec) {
    case P64: return 16;
    case P32: return 8;
    default: return 4;
  }
}

}  // namespace

// ------------------------------------------------------------------ objects
struct Work;

struct qexhip_lattice {
  Geom g;
  long V;
  RedWs red;
  Scal *S = nullptr;   // generic device scalars (API reductions)
  Scal *Sh = nullptr;  // pinned host mirror
  double *stage = nullptr;  // canonical staging buffer (V*6 doubles)
};

struct qexhip_field {
  qexhip_lattice *lat;
  int prec;
  char *data;  // [parity][3][stride] words
  size_t par_bytes;
  void *par(int p) const { return data + p * par_bytes; }
};

struct Work {
  // FP64 parity vectors
  double *y = nullptr, *rho = nullptr, *t64 = nullptr, *r64 = nullptr, *p64 = nullptr, *Ap64 = nullptr;
  // sloppy parity vectors (sized for FP32 words, which also fit FP16/BF16)
  void *xs = nullptr, *rs = nullptr, *Aps = nullptr, *ps = nullptr, *ts = nullptr;
  Scal *S = nullptr, *Sh = nullptr;
  // multi-shift
  int nms = 0;
  std::vector<double *> msp;
  double **d_x = nullptr, **d_p = nullptr;
  ShiftScal *T = nullptr, *Th = nullptr;
};

struct qexhip_links {
  qexhip_lattice *lat;
  int nl;
  double *canon = nullptr;             // device canonical FP64 copy, nl*V*18
  char *st[5][2] = {};                 // [precision][12-real?]
  int recon_ok = -1;                   // -1 unknown, 0 no, 1 phased SU(3)
  float qs[8] = {};                    // PQ16 scale/32767 per link field
  bool qs_ok = false;
  Work *ws = nullptr;
};

namespace {

size_t vec64_bytes(const qexhip_lattice *lat) { return (size_t)3 * lat->g.stride * 16; }

bool recon_allowed(qexhip_links *l) {
  if (l->nl != 4) return false;
  if (l->recon_ok < 0) {
    unsigned *d = dalloc<unsigned>(16);
    launch_fill(d, 16, g_stream);
    launch_recon_check(l->lat->g, l->canon, d, g_stream);
    check_launch();
    unsigned bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, d, 4, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    (void)hipFree(d);
    float dev;
    memcpy(&dev, &bits, 4);
    l->recon_ok = (dev < 1e-10f) ? 1 : 0;
  }
  return l->recon_ok == 1;
}

long lkblk_bytes(const qexhip_links *l, int prec, int rc) { return link_block_bytes(prec, rc, l->lat->g.stride); }

void link_scales(qexhip_links *l) {
  if (l->qs_ok) return;
  unsigned *d = dalloc<unsigned>(64);
  launch_fill(d, 64, g_stream);
  launch_link_maxabs(l->lat->g, l->nl, l->canon, d, g_stream);
  check_launch();
  unsigned bits[8] = {};
  HIPCHK(hipMemcpyAsync(bits, d, sizeof(unsigned) * l->nl, hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  (void)hipFree(d);
  for (int k = 0; k < l->nl; k++) {
    float m;
    memcpy(&m, &bits[k], 4);
    l->qs[k] = m / 32767.0f;
  }
  l->qs_ok = true;
}

const char *links_store(qexhip_links *l, int prec, int rc = 18) {
  require(rc == 18 || rc == 12, "recon must be 18 or 12");
  require(prec >= 0 && prec <= PQ16, "bad link precision");
  if (prec == PQ16) link_scales(l);
  if (rc == 12) require(recon_allowed(l), "12-real links need plain (nl=4) phased SU(3) links");
  const int ri = (rc == 12);
  if (!l->st[prec][ri]) {
    const size_t bytes = (size_t)l->nl * 2 * lkblk_bytes(l, prec, rc);
    l->st[prec][ri] = dalloc<char>(bytes);
    launch_fill(l->st[prec][ri], bytes, g_stream);
    launch_pack_links(prec, rc, l->lat->g, l->nl, l->canon, l->st[prec][ri], l->qs, g_stream);
    check_launch();
  }
  return l->st[prec][ri];
}

Work *work(qexhip_links *l) {
  if (!l->ws) {
    Work *w = new Work();
    const qexhip_lattice *lat = l->lat;
    size_t b64 = vec64_bytes(lat), b32 = (size_t)3 * lat->g.stride * 8;
    for (double **p : {&w->y, &w->rho, &w->t64, &w->r64, &w->p64, &w->Ap64}) {
      *p = dalloc<double>(b64);
      launch_fill(*p, b64, g_stream);
    }
    for (void **p : {&w->xs, &w->rs, &w->Aps, &w->ps, &w->ts}) {
      *p = dalloc<void>(b32);
      launch_fill(*p, b32, g_stream);
    }
    w->S = dalloc<Scal>(sizeof(Scal));
    HIPCHK(hipHostMalloc((void **)&w->Sh, sizeof(Scal), hipHostMallocDefault));
    memset(w->Sh, 0, sizeof(Scal));
    HIPCHK(hipMemcpyAsync(w->S, w->Sh, sizeof(Scal), hipMemcpyHostToDevice, g_stream));
    l->ws = w;
  }
  return l->ws;
}

void free_work(Work *w) {
  if (!w) return;
  for (void *p : {(void *)w->y, (void *)w->rho, (void *)w->t64, (void *)w->r64, (void *)w->p64, (void *)w->Ap64,
                  w->xs, w->rs, w->Aps, w->ps, w->ts, (void *)w->S, (void *)w->d_x, (void *)w->d_p,
                  (void *)w->T})
    if (p) (void)hipFree(p);
  for (double *p : w->msp) (void)hipFree(p);
  if (w->Sh) (void)hipHostFree(w->Sh);
  if (w->Th) (void)hipHostFree(w->Th);
  delete w;
}

void push(Scal *S, const Scal *Sh) { HIPCHK(hipMemcpyAsync(S, Sh, sizeof(Scal), hipMemcpyHostToDevice, g_stream)); }
void pull(Scal *Sh, const Scal *S) {
  HIPCHK(hipMemcpyAsync(Sh, S, sizeof(Scal), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
}

// One hop launch: out = ca*A + cb*B + chop*(2D)in on output parity par, with
// links and vectors in precision prec (arithmetic FP64 for FP64, else FP32).
void hop(qexhip_links *l, int prec, int rc, int par, int mode, int redm, const void *in, void *out, const void *a,
         const void *b, double ca, double cb, double chop, int ca_m2, int pred, Scal *S) {
  HopArgs h;
  h.g = l->lat->g;
  h.links = links_store(l, prec, rc);
  h.lkblk = lkblk_bytes(l, prec, rc);
  for (int k = 0; k < 8; k++) h.qs[k] = l->qs[k];
  h.in = in;
  h.out = out;
  h.a = a;
  h.b = b;
  h.ca = ca;
  h.cb = cb;
  h.chop = chop;
  h.ca_m2 = ca_m2;
  h.par = par;
  h.pred = pred;
  h.remap = g_remap;
  h.S = S;
  h.red = l->lat->red;
  const int pv = (prec == PQ16) ? P16 : prec;  // 16-bit fixed-point links go with FP16 vectors
  HopKey k{prec, rc, pv, pv, pv};
  launch_hop(k, l->nl, mode, redm, h, g_stream, g_block);
}

// A = 4(m2 - D_pq D_qp) applied to x (FP64, own parity par): out = cx*x + cb*b - sgn...
// Computes out = b - A x  (when with_b) with |out|^2 reduced per redm.
void resid64(qexhip_links *l, int par, const double *x, const double *b, double *out, Scal *S, int redm,
             double *t64) {
  hop(l, P64, 18, 1 - par, 0, 0, x, t64, nullptr, nullptr, 0, 0, 1.0, 0, 0, S);
  hop(l, P64, 18, par, 2, redm, t64, out, x, b, -1.0, 1.0, 1.0, 1, 0, S);
}

double flops_per_its(const qexhip_links *l) {
  return (double)(l->nl * 4 * 72 + 60) * (double)l->lat->g.Vh;
}

const int kChunk = 16;

// CG on A = 4(m^2 - D_pq D_qp), parity par (solvers/cg.nim:55-272, cpNone),
// two fused kernels per iteration (dslash.hip).  FP64 mode follows QEX's
// recurrence (alpha from exact norms; beta from a one-step extrapolation of
// |r|^2 computed in FP64, the stop test confirmed on the exact norm).  Mixed
// mode runs it in FP32 arithmetic on sloppy storage with FP64 reliable updates
// (true residual, FP64 solution accumulation).  x: FP64 in/out.
void cg_solve(qexhip_links *l, int par, double *x, const double *b, double m, double r2req, int maxits,
              int sloppy, int slinks, int recon, double delta, int relmax, int *its_out, int *nrel_out) {
  Work *w = work(l);
  qexhip_lattice *lat = l->lat;
  const Geom &g = lat->g;
  Scal *S = w->S, *Sh = w->Sh;
  const int rc = (recon == 12) ? 12 : 18;
  const bool mixed = (sloppy != P64);
  const int ps = sloppy, pl = mixed ? ((slinks < 0) ? sloppy : slinks) : P64;
  memset(Sh, 0, sizeof(Scal));
  Sh->m2 = m * m;
  Sh->maxits = maxits;
  Sh->mixed = mixed;
  Sh->delta2 = delta * delta;
  Sh->relmax = relmax > 0 ? relmax : (1 << 30);
  push(S, Sh);
  launch_scal_init(S, g_stream, 0);
  launch_norm2(P64, g, b, S, 1, lat->red, g_stream);
  check_launch();
  pull(Sh, S);
  const double b2 = Sh->tmp[1];
  *its_out = 0;
  *nrel_out = 0;
  if (b2 == 0.0) {  // cg.nim:138-143
    launch_fill(x, vec64_bytes(lat), g_stream);
    return;
  }
  Sh->b2 = b2;
  Sh->r2stop = r2req * b2;
  push(S, Sh);
  CgArgs a;
  a.g = g;
  a.links = links_store(l, pl, rc);
  a.lkblk = lkblk_bytes(l, pl, rc);
  for (int k = 0; k < 8; k++) a.qs[k] = l->qs[k];
  a.par = par;
  a.remap = g_remap;
  a.S = S;
  a.red = lat->red;
  if (!mixed) {
    a.x = x; a.r = w->r64; a.p = w->p64; a.Ap = w->Ap64; a.t = w->t64;
    launch_fill(w->Ap64, vec64_bytes(lat), g_stream);
    // r = b - A x ; r2 = |r|^2 (cg.nim:145-153)
    resid64(l, par, x, b, w->r64, S, 3, w->t64);
    pull(Sh, S);
    Sh->r2 = Sh->tmp[0];
    Sh->r2prev = Sh->r2;
    Sh->done = !(Sh->r2 > Sh->r2stop) || maxits <= 0;
    push(S, Sh);
  } else {
    a.x = w->xs; a.r = w->rs; a.p = w->ps; a.Ap = w->Aps; a.t = w->ts;
    const size_t b32 = (size_t)3 * g.stride * 8;
    launch_fill(w->xs, b32, g_stream);
    launch_fill(w->Aps, b32, g_stream);
    resid64(l, par, x, b, w->r64, S, 2, w->t64);  // epi_reliable: sigma, r2, done
    launch_rs_from_true(g, w->rs, w->r64, S, g_stream);
    // best FP64 iterate so far (the FP64-only p vector is free in mixed mode)
    HIPCHK(hipMemcpyAsync(w->p64, x, vec64_bytes(lat), hipMemcpyDeviceToDevice, g_stream));
    check_launch();
    pull(Sh, S);
  }
  while (!Sh->done) {
    for (int k = 0; k < kChunk; k++) launch_cg_iter(pl, rc, ps, l->nl, a, g_stream, g_block);
    check_launch();
    pull(Sh, S);
    if (Sh->done || Sh->need_rel) {
      launch_cg_flush(ps, a, g_stream);  // apply the pending x/r update; exact |r|^2
      if (mixed) {
        launch_xfold(g, x, w->xs, S, g_stream);
        resid64(l, par, x, b, w->r64, S, 2, w->t64);
        launch_rs_from_true(g, w->rs, w->r64, S, g_stream);
      }
      check_launch();
      pull(Sh, S);
      if (mixed && Sh->rollback) {
        // the sloppy recursion diverged: back to the best iterate, whose true
        // residual epi_reliable accepts again, restarting p = r
        HIPCHK(hipMemcpyAsync(x, w->p64, vec64_bytes(lat), hipMemcpyDeviceToDevice, g_stream));
        resid64(l, par, x, b, w->r64, S, 2, w->t64);
        launch_rs_from_true(g, w->rs, w->r64, S, g_stream);
        check_launch();
        pull(Sh, S);
      } else if (mixed && Sh->improved) {
        HIPCHK(hipMemcpyAsync(w->p64, x, vec64_bytes(lat), hipMemcpyDeviceToDevice, g_stream));
      }
    }
  }
  *its_out = Sh->its;
  *nrel_out = mixed ? Sh->nrel - 1 : 0;
}

// solveXX (src/physics/stagSolve.nim:330-421): outer defect-correction loop.
void solve_xx(qexhip_links *l, int par, double *r, const double *x, double m, qexhip_solver_params *sp) {
  auto t0 = std::chrono::steady_clock::now();
  Work *w = work(l);
  qexhip_lattice *lat = l->lat;
  const Geom &g = lat->g;
  Scal *S = w->S, *Sh = w->Sh;
  const size_t vb = vec64_bytes(lat);
  launch_fill(r, vb, g_stream);                                                  // r := 0
  HIPCHK(hipMemcpyAsync(w->rho, x, vb, hipMemcpyDeviceToDevice, g_stream));     // rho[par] := x
  memset(Sh, 0, sizeof(Scal));
  Sh->m2 = m * m;
  push(S, Sh);
  launch_norm2(P64, g, w->rho, S, 1, lat->red, g_stream);
  check_launch();
  pull(Sh, S);
  const double b2 = Sh->tmp[1];
  double r2 = b2, r2stop = sp->r2req * b2, r2prev = INFINITY;
  int its = 0, nrel = 0;
  while (r2 > r2stop && r2 < r2prev) {
    r2prev = r2;
    const int mx = sp->maxits - its;
    if (mx <= 0) break;
    launch_fill(w->y, vb, g_stream);
    int it = 0, nr = 0;
    cg_solve(l, par, w->y, w->rho, m, r2stop / r2, mx, sp->sloppy, sp->sloppy_links, sp->recon,
             sp->reliable_delta, sp->reliable_max_its, &it, &nr);
    its += it;
    nrel += nr;
    launch_axpby64(g, 1.0, w->y, 1.0, r, g_stream);  // r[par] += y
    // rho = x - A r ; r2 = |rho|^2
    Sh->m2 = m * m;
    push(S, Sh);
    resid64(l, par, r, x, w->rho, S, 3, w->t64);
    check_launch();
    pull(Sh, S);
    r2 = Sh->tmp[0];
  }
  sp->iterations = its;
  sp->reliable_updates = nrel;
  sp->r2 = (b2 > 0) ? r2 / b2 : 0.0;
  sp->flops = flops_per_its(l) * its;
  sp->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Multi-shift CG (src/solvers/cgm.nim:68-246), FP64, parity par.
void solve_ms(qexhip_links *l, int par, double *const *xs, const double *b, double m, const double *sigmas,
              int nshift, qexhip_solver_params *sp) {
  auto t0 = std::chrono::steady_clock::now();
  Work *w = work(l);
  qexhip_lattice *lat = l->lat;
  const Geom &g = lat->g;
  const size_t vb = vec64_bytes(lat);
  require(nshift >= 1 && nshift <= QH_MAXSHIFT, "multishift: nshift out of range");
  if (w->nms < nshift) {
    for (int k = w->nms; k < nshift; k++) w->msp.push_back(dalloc<double>(vb));
    if (w->d_x) (void)hipFree(w->d_x);
    if (w->d_p) (void)hipFree(w->d_p);
    w->d_x = dalloc<double *>(sizeof(double *) * nshift);
    w->d_p = dalloc<double *>(sizeof(double *) * nshift);
    w->nms = nshift;
  }
  if (!w->T) {
    w->T = dalloc<ShiftScal>(sizeof(ShiftScal));
    HIPCHK(hipHostMalloc((void **)&w->Th, sizeof(ShiftScal), hipHostMallocDefault));
  }
  std::vector<double *> xv(xs, xs + nshift);
  HIPCHK(hipMemcpyAsync(w->d_x, xv.data(), sizeof(double *) * nshift, hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(w->d_p, w->msp.data(), sizeof(double *) * nshift, hipMemcpyHostToDevice, g_stream));
  Scal *S = w->S, *Sh = w->Sh;
  ShiftScal *T = w->T, *Th = w->Th;
  // r := b ; xs := 0 ; b2 = |b|^2 (cgm.nim:122-127)
  HIPCHK(hipMemcpyAsync(w->r64, b, vb, hipMemcpyDeviceToDevice, g_stream));
  for (int k = 0; k < nshift; k++) launch_fill(xs[k], vb, g_stream);
  memset(Sh, 0, sizeof(Scal));
  Sh->m2 = m * m;
  push(S, Sh);
  launch_norm2(P64, g, b, S, 1, lat->red, g_stream);
  check_launch();
  pull(Sh, S);
  const double b2 = Sh->tmp[1];
  memset(Th, 0, sizeof(ShiftScal));
  Th->nshift = nshift;
  Th->maxits = sp->maxits;
  Th->r2stop = sp->r2req * b2;
  Th->r2i = b2;
  Th->alphaim1 = -1.0;
  Th->betaim1 = 0.0;
  for (int k = 0; k < nshift; k++) {
    Th->sg[k] = (k == 0) ? 0.0 : sigmas[k];
    Th->zi[k] = Th->zim1[k] = 1.0;
  }
  Th->continuing = (b2 > Th->r2stop) && sp->maxits > 0;
  HIPCHK(hipMemcpyAsync(T, Th, sizeof(ShiftScal), hipMemcpyHostToDevice, g_stream));
  Sh->done = !Th->continuing;
  push(S, Sh);
  for (int k = 0; k < nshift; k++)
    HIPCHK(hipMemcpyAsync(w->msp[k], w->r64, vb, hipMemcpyDeviceToDevice, g_stream));  // ps[m] := r
  while (Th->continuing) {
    for (int c = 0; c < kChunk; c++) {
      hop(l, P64, 18, 1 - par, 0, 0, w->msp[0], w->t64, nullptr, nullptr, 0, 0, 1.0, 0, 1, S);
      hop(l, P64, 18, par, 1, 1, w->t64, w->Ap64, w->msp[0], nullptr, 1.0, 0, -1.0, 1, 1, S);
      launch_cgm_alpha(T, S, g_stream);
      launch_cgm_xr(g, nshift, w->r64, w->Ap64, w->d_x, w->d_p, T, S, lat->red, g_stream);
      launch_cgm_p(g, nshift, w->r64, w->d_p, T, g_stream);
    }
    check_launch();
    HIPCHK(hipMemcpyAsync(Th, T, sizeof(ShiftScal), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    if (!Th->continuing) break;
  }
  sp->iterations = Th->its;
  sp->r2 = (b2 > 0) ? Th->r2i / b2 : 0.0;
  sp->flops = flops_per_its(l) * Th->its + (double)(4 * g.Vh + 15) * Th->its * nshift;
  sp->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void import_field(qexhip_field *f, int par, const double *canon, int on_device) {
  qexhip_lattice *lat = f->lat;
  const long Vh = lat->g.Vh;
  for (int p = 0; p < 2; p++) {
    if (par != 2 && par != p) continue;
    const double *src = canon + (size_t)p * Vh * 6;
    if (!on_device) {
      HIPCHK(hipMemcpyAsync(lat->stage, src, (size_t)Vh * 48, hipMemcpyHostToDevice, g_stream));
      src = lat->stage;
    }
    launch_pack_vec(f->prec, lat->g, src, f->par(p), g_stream);
    check_launch();
    if (!on_device) HIPCHK(hipStreamSynchronize(g_stream));
  }
}

void export_field(const qexhip_field *f, int par, double *canon, int on_device) {
  qexhip_lattice *lat = f->lat;
  const long Vh = lat->g.Vh;
  for (int p = 0; p < 2; p++) {
    if (par != 2 && par != p) continue;
    double *dst = canon + (size_t)p * Vh * 6;
    launch_unpack_vec(f->prec, lat->g, f->par(p), on_device ? dst : lat->stage, g_stream);
    check_launch();
    if (!on_device) HIPCHK(hipMemcpyAsync(dst, lat->stage, (size_t)Vh * 48, hipMemcpyDeviceToHost, g_stream));
  }
  HIPCHK(hipStreamSynchronize(g_stream));
}

void check_same_lat(const qexhip_links *l, const qexhip_field *f) {
  require(l && f && l->lat == f->lat, "field and links belong to different lattices");
}

}  // namespace

// ================================================================ C ABI
extern "C" {

const char *qexhip_last_error(void) { return g_err.c_str(); }

int qexhip_init(int device) {
  API_BEGIN
  if (g_stream) return 0;
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  require(device >= 0 && device < n, "qexhip_init: no such device");
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_device = device;
  if (const char *b = getenv("QEXHIP_BLOCK")) {
    int v = atoi(b);
    if (v >= 64 && v <= 256 && v % 64 == 0) g_block = v;
  }
  if (const char *r = getenv("QEXHIP_XCD_REMAP")) g_remap = atoi(r) != 0;
  API_END
}

int qexhip_finalize(void) {
  API_BEGIN
  if (g_stream) {
    HIPCHK(hipStreamSynchronize(g_stream));
    HIPCHK(hipStreamDestroy(g_stream));
    g_stream = nullptr;
  }
  API_END
}

int qexhip_sync(void) {
  API_BEGIN
  check_init();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

void *qexhip_stream(void) { return (void *)g_stream; }

int qexhip_lattice_create(const int dims[4], qexhip_lattice **out) {
  API_BEGIN
  check_init();
  for (int d = 0; d < 4; d++) require(dims[d] >= 2 && dims[d] % 2 == 0, "lattice extents must be even and >= 2");
  qexhip_lattice *lat = new qexhip_lattice();
  for (int d = 0; d < 4; d++) lat->g.L[d] = dims[d];
  lat->g.hx = dims[0] / 2;
  lat->V = (long)dims[0] * dims[1] * dims[2] * dims[3];
  lat->g.Vh = lat->V / 2;
  lat->g.stride = (lat->g.Vh + 255) / 256 * 256;
  lat->red.maxblocks = (int)((lat->g.stride + 63) / 64) + 64;
  lat->red.partials = dalloc<double>(sizeof(double) * 4 * lat->red.maxblocks);
  lat->S = dalloc<Scal>(sizeof(Scal));
  HIPCHK(hipHostMalloc((void **)&lat->Sh, sizeof(Scal), hipHostMallocDefault));
  memset(lat->Sh, 0, sizeof(Scal));
  HIPCHK(hipMemcpyAsync(lat->S, lat->Sh, sizeof(Scal), hipMemcpyHostToDevice, g_stream));
  lat->stage = dalloc<double>(sizeof(double) * 6 * lat->V);
  HIPCHK(hipStreamSynchronize(g_stream));
  *out = lat;
  API_END
}

int qexhip_lattice_destroy(qexhip_lattice *lat) {
  API_BEGIN
  if (!lat) return 0;
  HIPCHK(hipStreamSynchronize(g_stream));
  (void)hipFree(lat->red.partials);
  (void)hipFree(lat->S);
  (void)hipHostFree(lat->Sh);
  (void)hipFree(lat->stage);
  delete lat;
  API_END
}

int qexhip_links_create(qexhip_lattice *lat, int nl, const double *canon, int on_device, qexhip_links **out) {
  API_BEGIN
  check_init();
  require(lat != nullptr, "null lattice");
  require(nl == 4 || nl == 8, "nl must be 4 (plain) or 8 (HISQ fat+long)");
  if (nl == 8)
    for (int d = 0; d < 4; d++) require(lat->g.L[d] >= 4, "HISQ (3-hop) links need extents >= 4");
  qexhip_links *l = new qexhip_links();
  l->lat = lat;
  l->nl = nl;
  const size_t bytes = (size_t)nl * lat->V * 18 * sizeof(double);
  l->canon = dalloc<double>(bytes);
  HIPCHK(hipMemcpyAsync(l->canon, canon, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        g_stream));
  links_store(l, P64, 18);
  HIPCHK(hipStreamSynchronize(g_stream));
  *out = l;
  API_END
}

int qexhip_links_update(qexhip_links *l, const double *canon, int on_device) {
  API_BEGIN
  check_init();
  require(l != nullptr, "null links");
  const size_t bytes = (size_t)l->nl * l->lat->V * 18 * sizeof(double);
  HIPCHK(hipMemcpyAsync(l->canon, canon, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        g_stream));
  l->recon_ok = -1;
  l->qs_ok = false;
  if (l->st[PQ16][0] || l->st[PQ16][1]) link_scales(l);
  for (int p = 0; p <= PQ16; p++)
    for (int ri = 0; ri < 2; ri++) {
      if (!l->st[p][ri]) continue;
      if (ri == 1 && !recon_allowed(l)) {
        (void)hipFree(l->st[p][ri]);
        l->st[p][ri] = nullptr;
        continue;
      }
      launch_pack_links(p, ri ? 12 : 18, l->lat->g, l->nl, l->canon, l->st[p][ri], l->qs, g_stream);
    }
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_links_destroy(qexhip_links *l) {
  API_BEGIN
  if (!l) return 0;
  HIPCHK(hipStreamSynchronize(g_stream));
  (void)hipFree(l->canon);
  for (int p = 0; p <= PQ16; p++)
    for (int ri = 0; ri < 2; ri++)
      if (l->st[p][ri]) (void)hipFree(l->st[p][ri]);
  free_work(l->ws);
  delete l;
  API_END
}

int qexhip_field_create(qexhip_lattice *lat, int prec, qexhip_field **out) {
  API_BEGIN
  check_init();
  require(lat != nullptr, "null lattice");
  require(prec >= 0 && prec <= 3, "bad precision");
  qexhip_field *f = new qexhip_field();
  f->lat = lat;
  f->prec = prec;
  f->par_bytes = (size_t)3 * lat->g.stride * word_bytes(prec);
  f->data = dalloc<char>(2 * f->par_bytes);
  launch_fill(f->data, 2 * f->par_bytes, g_stream);
  HIPCHK(hipStreamSynchronize(g_stream));
  *out = f;
  API_END
}

int qexhip_field_destroy(qexhip_field *f) {
  API_BEGIN
  if (!f) return 0;
  HIPCHK(hipStreamSynchronize(g_stream));
  (void)hipFree(f->data);
  delete f;
  API_END
}

int qexhip_field_import(qexhip_field *f, int par, const double *canon, int on_device) {
  API_BEGIN
  check_init();
  require(f && canon, "null argument");
  require(par >= 0 && par <= 2, "bad parity");
  import_field(f, par, canon, on_device);
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_field_export(const qexhip_field *f, int par, double *canon, int on_device) {
  API_BEGIN
  check_init();
  require(f && canon, "null argument");
  require(par >= 0 && par <= 2, "bad parity");
  export_field(f, par, canon, on_device);
  API_END
}

int qexhip_field_norm2(const qexhip_field *f, int par, double *out) {
  API_BEGIN
  check_init();
  require(f && out, "null argument");
  qexhip_lattice *lat = f->lat;
  double s = 0;
  for (int p = 0; p < 2; p++) {
    if (par != 2 && par != p) continue;
    launch_norm2(f->prec, lat->g, f->par(p), lat->S, 0, lat->red, g_stream);
    check_launch();
    pull(lat->Sh, lat->S);
    s += lat->Sh->tmp[0];
  }
  *out = s;
  API_END
}

int qexhip_field_redot(const qexhip_field *x, const qexhip_field *y, int par, double *out) {
  API_BEGIN
  check_init();
  require(x && y && out && x->lat == y->lat && x->prec == y->prec, "bad arguments");
  qexhip_lattice *lat = x->lat;
  double s = 0;
  for (int p = 0; p < 2; p++) {
    if (par != 2 && par != p) continue;
    launch_redot(x->prec, lat->g, x->par(p), y->par(p), lat->S, 0, lat->red, g_stream);
    check_launch();
    pull(lat->Sh, lat->S);
    s += lat->Sh->tmp[0];
  }
  *out = s;
  API_END
}

int qexhip_field_axpby(const qexhip_field *x, qexhip_field *y, int par, double a, double b) {
  API_BEGIN
  check_init();
  require(x && y && x->lat == y->lat, "bad arguments");
  require(x->prec == P64 && y->prec == P64, "axpby fields must be FP64");
  require(par >= 0 && par <= 2, "bad parity");
  for (int p = 0; p < 2; p++) {
    if (par != 2 && par != p) continue;
    launch_axpby64(x->lat->g, a, (const double *)x->par(p), b, (double *)y->par(p), g_stream);
  }
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

// r = ca*r + cb*x + chop*(2D)x on parity par (general hop entry)
static void dslash_core(qexhip_links *l, int par, qexhip_field *r, const qexhip_field *x, double ca, double cb,
                        double chop) {
  check_same_lat(l, r);
  check_same_lat(l, x);
  require(r->prec == x->prec, "r and x must have the same precision");
  require(par >= 0 && par <= 2, "bad parity");
  const int pr = r->prec;
  for (int p = 0; p < 2; p++) {
    if (par != 2 && par != p) continue;
    if (ca == 0.0)
      hop(l, pr, 18, p, 1, 0, x->par(1 - p), r->par(p), x->par(p), nullptr, cb, 0, chop, 0, 0, nullptr);
    else
      hop(l, pr, 18, p, 2, 0, x->par(1 - p), r->par(p), r->par(p), x->par(p), ca, cb, chop, 0, 0, nullptr);
    check_launch();
  }
  HIPCHK(hipStreamSynchronize(g_stream));
}

int qexhip_stagD2(qexhip_links *l, int par, qexhip_field *r, const qexhip_field *x, double a, double b) {
  API_BEGIN
  check_init();
  dslash_core(l, par, r, x, a, b, 1.0);
  API_END
}

int qexhip_stagD(qexhip_links *l, int par, qexhip_field *r, const qexhip_field *x, double m, double sc, double a) {
  API_BEGIN
  check_init();
  // stagD2(a/(0.5sc), m/(0.5sc)) then r *= 0.5sc, folded into one pass
  dslash_core(l, par, r, x, a, m, 0.5 * sc);
  API_END
}

int qexhip_stagD2xx(qexhip_links *l, int par, qexhip_field *r, const qexhip_field *x, double m2) {
  API_BEGIN
  check_init();
  check_same_lat(l, r);
  check_same_lat(l, x);
  require(r->prec == x->prec, "r and x must have the same precision");
  require(par == 0 || par == 1, "stagD2xx parity must be 0 (ee) or 1 (oo)");
  Work *w = work(l);
  const int pr = r->prec;
  void *t = (pr == P64) ? (void *)w->t64 : w->ts;
  // t = (2D)_qp x_p  (stagDP, rir := 0) ; r = 4 m2 x - (2D)_pq t  (stagDM)
  hop(l, pr, 18, 1 - par, 0, 0, x->par(par), t, nullptr, nullptr, 0, 0, 1.0, 0, 0, nullptr);
  hop(l, pr, 18, par, 1, 0, t, r->par(par), x->par(par), nullptr, 4.0 * m2, 0, -1.0, 0, 0, nullptr);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_solve_xx(qexhip_links *l, int par, qexhip_field *r, const qexhip_field *x, double m,
                    qexhip_solver_params *sp) {
  API_BEGIN
  check_init();
  check_same_lat(l, r);
  check_same_lat(l, x);
  require(sp != nullptr, "null solver params");
  require(par == 0 || par == 1, "solve parity must be 0 (even) or 1 (odd)");
  require(r->prec == P64 && x->prec == P64, "solver fields must be FP64");
  require(sp->sloppy >= 0 && sp->sloppy <= 3, "bad sloppy precision");
  require(sp->sloppy_links >= -1 && sp->sloppy_links <= PQ16, "bad sloppy link precision");
  require(sp->recon == 0 || sp->recon == 18 || sp->recon == 12, "recon must be 0, 18 or 12");
  if (sp->sloppy != P64) {
    const int sl = sp->sloppy_links < 0 ? sp->sloppy : sp->sloppy_links;
    require(sl != P64, "mixed solve needs sloppy links below FP64");
    require(sl == sp->sloppy || sp->sloppy == P32 || (sl == PQ16 && sp->sloppy == P16),
            "sloppy link/spinor storage combination not supported");
  }
  if (sp->recon == 12) require(recon_allowed(l), "12-real links need plain (nl=4) phased SU(3) links");
  solve_xx(l, par, (double *)r->par(par), (const double *)x->par(par), m, sp);
  API_END
}

static int solve_host(qexhip_links *l, int par, double *r, const double *t, double m, qexhip_solver_params *sp) {
  API_BEGIN
  check_init();
  require(l && r && t && sp, "null argument");
  qexhip_field *fr = nullptr, *fx = nullptr;
  if (qexhip_field_create(l->lat, P64, &fr) || qexhip_field_create(l->lat, P64, &fx))
    throw std::runtime_error(g_err);
  import_field(fx, par, t, 0);
  int rc = qexhip_solve_xx(l, par, fr, fx, m, sp);
  if (rc == 0) export_field(fr, par, r, 0);
  std::string e = g_err;
  qexhip_field_destroy(fr);
  qexhip_field_destroy(fx);
  if (rc) throw std::runtime_error(e);
  API_END
}

int qexhip_solve_ee_host(qexhip_links *l, double *r, const double *t, double m, qexhip_solver_params *sp) {
  return solve_host(l, 0, r, t, m, sp);
}

int qexhip_solve_oo_host(qexhip_links *l, double *r, const double *t, double m, qexhip_solver_params *sp) {
  return solve_host(l, 1, r, t, m, sp);
}

int qexhip_solve_multishift(qexhip_links *l, int par, qexhip_field *const *xs, const qexhip_field *b, double m,
                            const double *sigmas, int nshift, qexhip_solver_params *sp) {
  API_BEGIN
  check_init();
  require(l && xs && b && sigmas && sp, "null argument");
  require(par == 0 || par == 1, "multishift parity must be 0 or 1");
  check_same_lat(l, b);
  require(b->prec == P64, "multishift fields must be FP64");
  std::vector<double *> xp(nshift);
  for (int k = 0; k < nshift; k++) {
    check_same_lat(l, xs[k]);
    require(xs[k]->prec == P64, "multishift fields must be FP64");
    xp[k] = (double *)xs[k]->par(par);
  }
  solve_ms(l, par, xp.data(), (const double *)b->par(par), m, sigmas, nshift, sp);
  API_END
}

int qexhip_force_solve(qexhip_links *l, qexhip_field *psi, const qexhip_field *phi, double m,
                       qexhip_solver_params *sp) {
  API_BEGIN
  check_init();
  check_same_lat(l, psi);
  check_same_lat(l, phi);
  require(psi->prec == P64 && phi->prec == P64, "force fields must be FP64");
  require(sp != nullptr, "null solver params");
  const Geom &g = l->lat->g;
  launch_fill(psi->data, 2 * psi->par_bytes, g_stream);
  // solveEE(varphi, phi) ; varphi.even *= 4 ; psi.odd = (2D)_oe varphi ; psi.even = varphi
  solve_xx(l, 0, (double *)psi->par(0), (const double *)phi->par(0), m, sp);
  launch_axpby64(g, 0.0, (const double *)psi->par(0), 4.0, (double *)psi->par(0), g_stream);
  hop(l, P64, 18, 1, 0, 0, psi->par(0), psi->par(1), nullptr, nullptr, 0, 0, 1.0, 0, 0, nullptr);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_force(qexhip_links *l, const qexhip_field *psi, double *F, int F_on_device, int outer_only) {
  API_BEGIN
  check_init();
  check_same_lat(l, psi);
  require(l->nl == 4, "qexhip_force: plain (nl=4) links required");
  require(psi->prec == P64, "force fields must be FP64");
  require(F != nullptr, "null output");
  qexhip_lattice *lat = l->lat;
  const size_t bytes = (size_t)4 * lat->V * 18 * sizeof(double);
  double *dF = F;
  if (!F_on_device) dF = dalloc<double>(bytes);
  launch_force(lat->g, links_store(l, P64, 18), lkblk_bytes(l, P64, 18), (const double *)psi->par(0),
               (const double *)psi->par(1), dF, outer_only, g_stream);
  check_launch();
  if (!F_on_device) {
    HIPCHK(hipMemcpyAsync(F, dF, bytes, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    (void)hipFree(dF);
  }
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_gen_gauge(qexhip_lattice *lat, uint64_t seed, double *U) {
  API_BEGIN
  check_init();
  launch_gauge_random(lat->g, seed, U, g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_gen_source(qexhip_lattice *lat, uint64_t seed, int stream, int par, double *x) {
  API_BEGIN
  check_init();
  require(par >= 0 && par <= 2, "bad parity");
  launch_source_gauss(lat->g, seed, stream, par, x, g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_gen_phase_bc(qexhip_lattice *lat, double *U) {
  API_BEGIN
  check_init();
  launch_phase_bc(lat->g, U, 4, g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_gen_hisq(qexhip_lattice *lat, uint64_t seed, const double *Uph, double *g8) {
  API_BEGIN
  check_init();
  for (int d = 0; d < 4; d++) require(lat->g.L[d] >= 4, "HISQ links need extents >= 4");
  launch_hisq_links(lat->g, seed, Uph, g8, g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_plaquette(qexhip_lattice *lat, const double *U, double *sum) {
  API_BEGIN
  check_init();
  require(lat && U && sum, "null argument");
  launch_plaquette(lat->g, U, lat->S, 0, lat->red, g_stream);
  check_launch();
  pull(lat->Sh, lat->S);
  *sum = lat->Sh->tmp[0];
  API_END
}

int qexhip_wilson_force(qexhip_lattice *lat, const double *U, double beta, double *F) {
  API_BEGIN
  check_init();
  require(lat && U && F, "null argument");
  launch_wilson_force(lat->g, U, beta, F, g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_random_momenta(qexhip_lattice *lat, uint64_t seed, double *P) {
  API_BEGIN
  check_init();
  require(lat && P, "null argument");
  launch_random_momenta(lat->g, seed, P, g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_link_update(qexhip_lattice *lat, const double *P, double eps, double *U) {
  API_BEGIN
  check_init();
  require(lat && P && U, "null argument");
  launch_link_update(lat->g, P, eps, U, g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_links_recon_ok(qexhip_links *l, int *ok) {
  API_BEGIN
  check_init();
  require(l && ok, "null argument");
  *ok = recon_allowed(l) ? 1 : 0;
  API_END
}

int qexhip_time_hop(qexhip_links *l, int prec, int recon, int nrep, double *ms) {
  API_BEGIN
  check_init();
  require(l && ms && nrep > 0, "bad arguments");
  require(prec >= 0 && prec <= PQ16, "bad precision");
  qexhip_lattice *lat = l->lat;
  qexhip_field *x = nullptr, *r = nullptr;
  const int pv = (prec == PQ16) ? P16 : prec;
  if (qexhip_field_create(lat, pv, &x) || qexhip_field_create(lat, pv, &r)) throw std::runtime_error(g_err);
  launch_source_gauss(lat->g, 99, 0, 2, lat->stage, g_stream);
  import_field(x, 2, lat->stage, 1);
  const int rc = (recon == 12) ? 12 : 18;
  for (int i = 0; i < 3; i++)
    hop(l, prec, rc, 0, 0, 0, x->par(1), r->par(0), nullptr, nullptr, 0, 0, 1.0, 0, 0, nullptr);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, g_stream));
  for (int i = 0; i < nrep; i++)
    hop(l, prec, rc, 0, 0, 0, x->par(1), r->par(0), nullptr, nullptr, 0, 0, 1.0, 0, 0, nullptr);
  HIPCHK(hipEventRecord(e1, g_stream));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  *ms = (double)t / nrep;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  qexhip_field_destroy(x);
  qexhip_field_destroy(r);
  API_END
}

}  // extern "C"
