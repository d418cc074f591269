// C-ABI entry points of the qexhip library and the host-side solver drivers.
// The drivers restate QEX's algorithms (cited per function); all vector work
// runs in the kernels of kernels.hip, on one library-owned HIP stream.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <chrono>
#include <algorithm>
#include <climits>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "internal.h"
#include "qexhip.h"

using namespace qh;

// Device scratch of qexhip_stagD2xx_noalloc: the odd-parity temporary of the
// two-hop operator, 128 MiB (an FP64 color vector of up to 2 796 032 sites per
// parity: 48^4 fits).  Each GPU the library runs on holds its own copy until
// the process ends.  Host code never passes the array's name to a kernel: the
// name is a host-side stand-in, and qexhip_init looks up this device's address
// of the array once (g_scratch below).
#define QEXHIP_SCRATCH_BYTES (128ul << 20)
static __device__ __align__(256) char scratch[QEXHIP_SCRATCH_BYTES];

namespace {

// device address of `scratch` per device, set by qexhip_init with that device current
constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};

thread_local std::string g_err;
hipStream_t g_stream = nullptr;
int g_device = -1;
int g_block = 256;
int g_remap = 1;
bool g_block_forced = false;  // QEXHIP_BLOCK set: no autotuning
bool g_autotune = true;
bool g_cg_graph = true;       // QEXHIP_CG_GRAPH=0: launch CG chunks kernel by kernel
bool g_cg_intervals = true;   // QEXHIP_CG_INTERVALS=0: mixed CG replays kChunk chunks, not whole intervals
bool g_smear_walk = true;     // QEXHIP_SMEAR_WALK=0: HISQ smearing kernels walk sites in canonical order
int g_multi_batches = 0;      // multi-RHS hop: vector-load batches per link pinned (1, 2) or 0 (tuner)
long g_multi_chunk = -1;      // multi-RHS hop: site-walk chunk pinned (0: canonical order) or -1 (tuner)
long g_hop_chunk = -1;        // one-parity hop: site-walk chunk pinned (0: canonical order) or -1 (tuner)
int g_hop_tmarch = -1;        // one-parity hop of plain links: t-march slices per thread pinned (0: off) or -1 (tuner)
int g_hop_two = -1;           // one-parity hop: both links of a hop per stencil trip pinned (0/1) or -1 (tuner)
int g_hop_split = -1;         // one-parity hop: threads per output site pinned (1, 2, 4) or -1 (tuner)
long g_msb_walk = -1;         // batched multi-shift hops: site-walk chunk pinned (0: canonical order) or -1 (tuner)
double g_ms_sloppy_r2 = 1e-12; // mixed multi-shift: FP32 phase ends at |r|^2 <= this |b|^2 (QEXHIP_MS_SLOPPY_R2)
int g_cg_block1 = 0;          // CG forms 3/4: other-parity stencil kernel's block size pinned, 0 (tuner)
int g_cg_cap = -1;            // CG forms 3/4: persistent workgroups pinned (0: form 3 one pass), -1 (tuner)
// Link bytes the stencils of a new lattice read with cache-allocating loads
// (Geom::lkcache), or -1: chosen per kernel family and link format by the
// tuner (link_cache_budgets), 0 where nothing is tuned.  QEXHIP_LINK_CACHE_MB
// at init pins a value for every lattice.
constexpr long kLinkCacheDefault = -1L;  // chosen by the tuner
long g_link_cache = kLinkCacheDefault;

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
  SmearWs smear;            // HISQ link construction workspace (allocated on first use)
};

struct qexhip_field {
  qexhip_lattice *lat;
  int prec;
  char *data;  // per parity: stride/64 blocks of [3 colors][64 sites] words
  size_t par_bytes;
  void *par(int p) const { return data + p * par_bytes; }
};

// An instantiated graph of cg_solve's launches, kept for the next solve on
// the same links whose launches would be the same: key holds every argument
// value the captured kernels bake in.
struct CachedGraph {
  std::string key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

struct Work {
  CachedGraph cg[2];  // cg_solve: [0] one reliable-update interval, [1] a chunk of iterations
  // FP64 parity vectors
  double *y = nullptr, *rho = nullptr, *t64 = nullptr, *r64 = nullptr, *p64 = nullptr, *Ap64 = nullptr;
  // sloppy parity vectors (sized for FP32 words, which also fit FP16/BF16)
  void *xs = nullptr, *rs = nullptr, *Aps = nullptr, *ps = nullptr, *ts = nullptr;
  Scal *S = nullptr, *Sh = nullptr;
  Scal *S1 = nullptr;  // second scalar buffer of the form-4 CG iteration
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
    w->S1 = dalloc<Scal>(sizeof(Scal));
    HIPCHK(hipHostMalloc((void **)&w->Sh, sizeof(Scal), hipHostMallocDefault));
    memset(w->Sh, 0, sizeof(Scal));
    HIPCHK(hipMemcpyAsync(w->S, w->Sh, sizeof(Scal), hipMemcpyHostToDevice, g_stream));
    l->ws = w;
  }
  return l->ws;
}

void free_work(Work *w) {
  if (!w) return;
  HIPCHK(hipStreamSynchronize(g_stream));
  for (CachedGraph &c : w->cg) {
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
  }
  for (void *p : {(void *)w->y, (void *)w->rho, (void *)w->t64, (void *)w->r64, (void *)w->p64, (void *)w->Ap64,
                  w->xs, w->rs, w->Aps, w->ps, w->ts, (void *)w->S, (void *)w->S1})
    if (p) (void)hipFree(p);
  if (w->Sh) (void)hipHostFree(w->Sh);
  delete w;
}

void push(Scal *S, const Scal *Sh) { HIPCHK(hipMemcpyAsync(S, Sh, sizeof(Scal), hipMemcpyHostToDevice, g_stream)); }
void pull(Scal *Sh, const Scal *S) {
  HIPCHK(hipMemcpyAsync(Sh, S, sizeof(Scal), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
}

// One hop launch: out = ca*A + cb*B + chop*(2D)in on output parity par, with
// links and vectors in precision prec (arithmetic FP64 for FP64, else FP32).
// ---------------------------------------------------------------- autotuner
// Launch shape (block size, XCD-aware block order) per kernel family, lattice
// and storage format, chosen by timing every candidate with HIP events on
// scratch data the first time a shape is needed, and kept for the process.
// Reductions sum block partials, so the chosen block size fixes the summation
// order: results are reproducible within a process and across processes that
// choose the same shape (set QEXHIP_BLOCK to pin it).
struct Tuned {
  int block, remap;
  double us;
  int form = 2;  // CG iteration: 2 or 3 kernels (dslash.hip)
  int cap = 0;   // hop: persistent form with at most cap workgroups (0: one pass)
  int block1 = 0;  // CG form 3: block size of the other-parity stencil kernel (0: block)
  int batches = 1;  // multi-RHS hop: vector-load batches per link
  long chunk = 0;   // site walk (make_walk): t-march chunk, 0 = canonical order
  int tmarch = 0;   // hop of plain links: t-slices per thread of the t-march form (0: off)
  int two = 0;      // hop: both links of a hop per stencil trip
  int split = 1;    // hop: threads per output site (1, 2 or 4)
  long lkcache = 0; // link bytes read with cache-allocating loads (Geom::lkcache), for lattices that leave it to the tuner
};
std::map<std::string, Tuned> g_tune;

// Candidate link-cache budgets (Geom::lkcache) of the stencil tuners.
// Measured at 32^4 (CG, us per iteration, budget 0 / best): Q16 12-real links
// 70.7 / 63.2 (128 MB: the whole 100 MB field cached), FP16 18-real 82.0 /
// 75.1, FP32 139.1 / 125.7 (192 MB: two of four directions), FP64 252.4 /
// 233.7 (192 MB: one direction, 151 MB); at 48^4, where the vectors alone
// nearly fill the 256 MB last-level cache, 0 was best (Q16 12-real 297.9 us,
// 2% slower at 128 MB; FP16 18-real 377.7 us, 6% slower at 192 MB).
const long kLinkCacheMB[] = {0, 64, 128, 160, 192, 224, 256};

// the link-cache budget a launch with tuned shape tu uses on lattice lat
long link_cache_of(const qexhip_lattice *lat, const Tuned &tu) {
  return lat->g.lkcache >= 0 ? lat->g.lkcache : tu.lkcache;
}

std::string geom_key(const Geom &g) {
  char b[96];
  snprintf(b, sizeof b, "%dx%dx%dx%d", g.L[0], g.L[1], g.L[2], g.L[3]);
  return b;
}

double time_us(const std::function<void()> &f, int warm, int reps) {
  for (int i = 0; i < warm; i++) f();
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, g_stream));
  for (int i = 0; i < reps; i++) f();
  HIPCHK(hipEventRecord(e1, g_stream));
  HIPCHK(hipEventSynchronize(e1));
  check_launch();
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 1e3 * (double)t / reps;
}

int g_cg_form = 0;  // QEXHIP_CG_FORM=2/3/4 pins the CG iteration form; 0: tuned

// Capture the launches of f on the library stream into an instantiated graph.
// A failure inside f ends the capture before rethrowing, so the stream is
// never left capturing.
void capture(const std::function<void()> &f, hipGraph_t *graph, hipGraphExec_t *exec) {
  HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
  try {
    f();
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(g_stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIPCHK(hipStreamEndCapture(g_stream, graph));
  HIPCHK(hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
}

// Bumped by every launch-shape setter of the C ABI: a graph cached before
// the call may hold launches of an earlier shape.
long g_cfg_gen = 0;

// c's graph when its key is key, else f captured into c (replacing the
// graph there).  The caller has synchronized the stream since c's last launch.
hipGraphExec_t cached_graph(CachedGraph &c, const std::string &key, const std::function<void()> &f) {
  if (c.exec && c.key == key) return c.exec;
  if (c.exec) (void)hipGraphExecDestroy(c.exec);
  if (c.graph) (void)hipGraphDestroy(c.graph);
  c.exec = nullptr;
  c.graph = nullptr;
  c.key.clear();
  capture(f, &c.graph, &c.exec);
  c.key = key;
  return c.exec;
}

// f(t) runs the kernel sequence once with shape t; the candidates are every
// block size x block order x listed form x listed workgroup cap
Tuned tune(const std::string &key, const std::function<void(const Tuned &)> &f,
           std::initializer_list<int> forms = {2}, std::initializer_list<int> caps = {0}) {
  auto it = g_tune.find(key);
  if (it != g_tune.end()) return it->second;
  const int fixed_form = g_cg_form ? g_cg_form : 2;
  if (g_block_forced || !g_autotune) return Tuned{g_block, g_remap, 0.0, fixed_form, 0};
  Tuned best{256, 1, 1e300, fixed_form, 0};
  bool pinned = false;  // a pinned CG form restricts only the lists that hold it
  for (int form : forms) pinned |= (g_cg_form != 0 && form == g_cg_form);
  for (int form : forms) {
    if (pinned && form != g_cg_form) continue;
    for (int cap : caps)
      for (int block : {64, 128, 256})
        for (int remap : {1, 0}) {
          const Tuned t{block, remap, 0.0, form, cap};
          const double us = time_us([&] { f(t); }, 2, 8);
          if (us < best.us) best = Tuned{block, remap, us, form, cap};
        }
  }
  g_tune[key] = best;
  return best;
}

Tuned tuned_hop(qexhip_links *l, int prec, int rc);

// One hop with links of storage prec packed in lk and launch shape tu.
void hop_shaped(qexhip_links *l, const char *lk, const Tuned &tu, int prec, int rc, int par, int mode, int redm,
                const void *in, void *out, const void *a, const void *b, double ca, double cb, double chop,
                int ca_m2, int pred, Scal *S) {
  HopArgs h;
  h.g = l->lat->g;
  h.links = lk;
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
  h.remap = tu.remap;
  h.S = S;
  h.red = l->lat->red;
  h.chunk = tu.chunk;
  h.tmarch = l->nl == 4 ? tu.tmarch : 0;
  h.two = tu.two;
  h.split = tu.split;
  h.g.lkcache = link_cache_of(l->lat, tu);
  const int pv = (prec == PQ16) ? P16 : prec;  // 16-bit fixed-point links go with FP16 vectors
  HopKey k{prec, rc, pv, pv, pv};
  launch_hop(k, l->nl, mode, redm, h, g_stream, tu.block, tu.cap);
}

void hop(qexhip_links *l, int prec, int rc, int par, int mode, int redm, const void *in, void *out, const void *a,
         const void *b, double ca, double cb, double chop, int ca_m2, int pred, Scal *S) {
  const char *lk = links_store(l, prec, rc);
  hop_shaped(l, lk, tuned_hop(l, prec, rc), prec, rc, par, mode, redm, in, out, a, b, ca, cb, chop, ca_m2, pred, S);
}

std::string hop_key(const qexhip_links *l, int prec, int rc) {
  char kb[96];
  if (g_hop_chunk >= 0 || g_hop_tmarch >= 0 || g_hop_two >= 0 || g_hop_split >= 0)
    snprintf(kb, sizeof kb, "hop nl=%d prec=%d rc=%d walk=%ld tmarch=%d two=%d split=%d ", l->nl, prec, rc,
             g_hop_chunk, g_hop_tmarch, g_hop_two, g_hop_split);
  else
    snprintf(kb, sizeof kb, "hop nl=%d prec=%d rc=%d ", l->nl, prec, rc);
  return kb + geom_key(l->lat->g);
}

// The hop's launch shape without tuning (tuning times candidates on buffers it
// allocates): the tuner's earlier choice for this shape, else the default.
Tuned cached_hop_shape(const qexhip_links *l, int prec, int rc) {
  auto it = g_tune.find(hop_key(l, prec, rc));
  if (it != g_tune.end()) return it->second;
  Tuned t{g_block, g_remap, 0.0, 2, 0};
  t.chunk = g_hop_chunk > 0 ? g_hop_chunk : 0;
  t.tmarch = g_hop_tmarch > 0 ? g_hop_tmarch : 0;
  t.two = g_hop_two > 0 ? 1 : 0;
  t.split = (g_hop_split == 2 || g_hop_split == 4) ? g_hop_split : 1;
  return t;
}

Tuned tuned_hop(qexhip_links *l, int prec, int rc) {
  const std::string key = hop_key(l, prec, rc);
  if (g_tune.count(key)) return g_tune[key];
  if (g_block_forced || !g_autotune) return cached_hop_shape(l, prec, rc);
  const size_t vb = vec64_bytes(l->lat);
  char *in = dalloc<char>(vb), *out = dalloc<char>(vb);
  launch_fill(in, vb, g_stream);
  HopArgs h{};
  h.g = l->lat->g;
  h.links = links_store(l, prec, rc);
  h.lkblk = lkblk_bytes(l, prec, rc);
  for (int k = 0; k < 8; k++) h.qs[k] = l->qs[k];
  h.in = in;
  h.out = out;
  h.chop = 1.0;
  h.red = l->lat->red;
  const int pv = (prec == PQ16) ? P16 : prec;
  const HopKey hk{prec, rc, pv, pv, pv};
  // block size x XCD order x workgroup cap (one pass or persistent) x site walk
  Tuned best{256, 1, 1e300};
  // (the walk trims the gathered vector's re-fetch but breaks the sequential
  // stream of the links, which carry most of the bytes: at 48^4 FP64 it read
  // 3.26 instead of 3.58 GB per hop in 582 instead of 545 us; a candidate still)
  std::vector<long> chunks = {0L, 2048L, 8192L};
  if (g_hop_chunk >= 0) chunks = {g_hop_chunk};
  const bool split_pinned = g_hop_split == 2 || g_hop_split == 4;
  if (g_hop_tmarch <= 0 && !split_pinned)
    for (long chunk : chunks) {
      if (chunk && g_hop_chunk < 0 && make_walk(h.g, 256, chunk).C == 0) continue;  // t-slices too small
      for (int cap : {0, 1024, 2048})
        for (int block : {64, 128, 256})
          for (int remap : {1, 0})
            for (int two : {0, 1}) {
              if (g_hop_two >= 0 && two != g_hop_two) continue;
              h.remap = remap;
              h.chunk = chunk;
              h.tmarch = 0;
              h.two = two;
              const double us = time_us([&] { launch_hop(hk, l->nl, 0, 0, h, g_stream, block, cap); }, 2, 8);
              if (us < best.us) {
                best = Tuned{block, remap, us, 2, cap};
                best.chunk = chunk;
                best.two = two;
              }
            }
    }
  // the t-march form (plain links): t-slices per thread x block size x XCD order
  std::vector<int> tms = {4, 8, 16};
  if (g_hop_tmarch >= 0) tms = {g_hop_tmarch};
  if (l->nl == 4 && g_hop_chunk <= 0 && !split_pinned)
    for (int tm : tms) {
      if (tm <= 0) continue;
      for (int block : {64, 128, 256})
        for (int remap : {1, 0}) {
          h.remap = remap;
          h.chunk = 0;
          h.tmarch = tm;
          h.two = 0;
          const double us = time_us([&] { launch_hop(hk, l->nl, 0, 0, h, g_stream, block, 0); }, 2, 8);
          if (us < best.us) {
            best = Tuned{block, remap, us, 2, 0};
            best.tmarch = tm;
          }
        }
    }
  // the split form: 2 or 4 threads per output site, each taking a share of
  // its links (for lattices too small to fill the device with one per site)
  std::vector<int> splits = {2, 4};
  if (g_hop_split >= 0) splits = {g_hop_split};
  if (g_hop_tmarch <= 0 && g_hop_chunk <= 0 && g_hop_two <= 0)
    for (int sp : splits) {
      if (sp != 2 && sp != 4) continue;
      for (int block : {64, 128, 256})
        for (int remap : {1, 0}) {
          h.remap = remap;
          h.chunk = 0;
          h.tmarch = 0;
          h.two = 0;
          h.split = sp;
          const double us = time_us([&] { launch_hop(hk, l->nl, 0, 0, h, g_stream, block, 0); }, 2, 8);
          if (us < best.us) {
            best = Tuned{block, remap, us, 2, 0};
            best.split = sp;
          }
        }
    }
  h.split = 1;
  // the link-cache budget, on the chosen shape (the t-march form and the
  // two-links trip read every link non-temporally; the split form caches the
  // links only when the whole field fits the budget)
  if (l->lat->g.lkcache < 0 && !best.tmarch && !best.two) {
    h.remap = best.remap;
    h.chunk = best.chunk;
    h.tmarch = 0;
    h.two = 0;
    h.split = best.split;
    double bus = 1e300;
    for (long mb : kLinkCacheMB) {
      h.g.lkcache = mb << 20;
      const double us = time_us([&] { launch_hop(hk, l->nl, 0, 0, h, g_stream, best.block, best.cap); }, 2, 8);
      if (us < bus) {
        bus = us;
        best.lkcache = mb << 20;
      }
    }
  }
  g_tune[key] = best;
  HIPCHK(hipStreamSynchronize(g_stream));
  (void)hipFree(in);
  (void)hipFree(out);
  return best;
}

// Fused CG iteration: the launch shape and iteration form, timed on chunks of
// a few iterations with scratch vectors and device scalars that never
// converge.  Forms 2 and 3 over block size x XCD order; form 3 also with its
// stencil kernels persistent (cap workgroups), form 4 over the workgroups of
// its persistent streaming kernel (cap); then, for forms 3 and 4, the block
// size of the other-parity stencil kernel (block1).  The us
// recorded is per iteration.
const int kTuneIters = 4;
Tuned tuned_cg(qexhip_links *l, int pl, int rc, int ps, int rx, const CgArgs &real) {
  char kb[80];
  snprintf(kb, sizeof kb, "cg nl=%d links=%d rc=%d vec=%d rx=%d ", l->nl, pl, rc, ps, rx);
  const std::string key = kb + geom_key(l->lat->g);
  if (g_tune.count(key) || g_block_forced || !g_autotune) return tune(key, nullptr);
  const size_t vb = vec64_bytes(l->lat);
  CgArgs a = real;
  void *v[5];
  for (int k = 0; k < 5; k++) {
    v[k] = dalloc<char>(vb);
    launch_fill(v[k], vb, g_stream);
  }
  a.x = v[0]; a.r = v[1]; a.p = v[2]; a.Ap = v[3]; a.t = v[4];
  Scal *S = dalloc<Scal>(2 * sizeof(Scal));
  Scal h;
  memset(&h, 0, sizeof h);
  h.r2stop = -1.0;
  h.r2stop_s = -1.0;
  h.maxits = INT_MAX;
  h.sigma = h.pscale = 1.0;
  HIPCHK(hipMemcpyAsync(S, &h, sizeof h, hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(S + 1, &h, sizeof h, hipMemcpyHostToDevice, g_stream));
  a.S = S;
  auto run = [&](const Tuned &c) {
    a.remap = c.remap;
    launch_cg_chunk(pl, rc, ps, l->nl, a, g_stream, c.block, c.form, rx, c.block1, kTuneIters, S + 1, c.cap);
  };
  Tuned best{256, 1, 1e300};
  for (int form : {2, 3, 4}) {
    if (g_cg_form && form != g_cg_form) continue;
    // form 4: its persistent streaming kernel's workgroups; form 3: one pass or
    // both stencil kernels persistent
    for (int cap : (form == 4 ? std::vector<int>{256, 512, 1024}
                              : form == 3 ? std::vector<int>{0, 1024, 2048} : std::vector<int>{0}))
      for (int block : {64, 128, 256})
        for (int remap : {1, 0}) {
          Tuned c{block, remap, 0.0, form, cap};
          const double us = time_us([&] { run(c); }, 2, 8);
          if (us < best.us) {
            best = c;
            best.us = us;
          }
        }
  }
  if (best.form >= 3) {
    // the other-parity stencil kernel reduces nothing, so its block size is free
    Tuned b = best;
    for (int b1 : {64, 128, 256}) {
      Tuned c = best;
      c.block1 = b1;
      const double us = time_us([&] { run(c); }, 2, 8);
      if (us < b.us) {
        b = c;
        b.us = us;
      }
    }
    best = b;
  }
  // the link-cache budget, on the chosen shape
  if (l->lat->g.lkcache < 0) {
    double bus = 1e300;
    for (long mb : kLinkCacheMB) {
      a.g.lkcache = mb << 20;
      const double us = time_us([&] { run(best); }, 2, 8);
      if (us < bus) {
        bus = us;
        best.lkcache = mb << 20;
      }
    }
    best.us = bus;
  }
  best.us /= kTuneIters;
  g_tune[key] = best;
  HIPCHK(hipStreamSynchronize(g_stream));
  for (int k = 0; k < 5; k++) (void)hipFree(v[k]);
  (void)hipFree(S);
  return best;
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
// longest reliable-update interval captured as one graph (4 kernels per iteration)
const int kMaxInterval = 256;

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
  // QEXHIP_FP16_ALL: FP16 search vectors and FP16 residual / solution increment
  const int rx = (sloppy == QEXHIP_FP16_ALL) ? P16 : P32;
  if (sloppy == QEXHIP_FP16_ALL) sloppy = P16;
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
  memset(&a, 0, sizeof a);  // padding included: the bytes of a are part of a graph key
  a.g = g;
  a.links = links_store(l, pl, rc);
  a.lkblk = lkblk_bytes(l, pl, rc);
  for (int k = 0; k < 8; k++) a.qs[k] = l->qs[k];
  a.par = par;
  a.S = S;
  a.red = lat->red;
  Tuned tu = tuned_cg(l, pl, rc, ps, rx, a);
  if (g_cg_form) tu.form = g_cg_form;
  if (g_cg_block1) tu.block1 = g_cg_block1;
  if (g_cg_cap >= 0) tu.cap = g_cg_cap;
  a.remap = tu.remap;
  a.g.lkcache = link_cache_of(lat, tu);
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
    launch_rs_from_true(g, w->rs, w->r64, S, rx, g_stream);
    // best FP64 iterate so far (the FP64-only p vector is free in mixed mode)
    HIPCHK(hipMemcpyAsync(w->p64, x, vec64_bytes(lat), hipMemcpyDeviceToDevice, g_stream));
    check_launch();
    pull(Sh, S);
  }
  // The chunk of iterations is captured once per solve and replayed: one graph
  // launch instead of ~4*kChunk kernel launches, so the host never limits the
  // short sloppy iterations (the kernels read every scalar from S).
  //
  // Mixed mode replays whole intervals instead.  A reliable update comes after
  // relmax iterations at the latest and leaves since = 0, so an interval of
  // relmax iterations always ends with need_rel set (relmax reached, or an
  // earlier residual drop, breakdown or maxits, after which the remaining
  // iterations do nothing), and the reliable update that follows it needs no
  // predicate.  One graph holds the interval, the update and the device-side
  // copy of the best iterate: one host round trip per interval instead of one
  // per kChunk iterations plus three per update.  An interval cut short by an
  // earlier trigger switches the rest of the solve to chunks, whose idle tail
  // after a trigger is at most kChunk - 1 iterations.
  //
  // The graphs stay with the links' workspace: a later solve whose launches
  // bake in the same values (kernel arguments, shapes, x and b, interval)
  // replays them instead of capturing again (0.5-1.6 ms per capture).
  hipGraphExec_t exec = nullptr, iexec = nullptr;
  auto graph_key = [&](int kind, int n) {
    const Tuned hop = tuned_hop(l, P64, 18);  // the FP64 hops of resid64
    const long v[] = {kind, n, tu.block, tu.remap, tu.form, tu.cap, tu.block1, tu.lkcache, pl, rc, ps, rx,
                      l->nl, mixed, (long)(uintptr_t)x, (long)(uintptr_t)b, hop.block, hop.remap, hop.cap,
                      hop.chunk, hop.tmarch, hop.two, hop.split, hop.lkcache, g_cfg_gen};
    std::string k((const char *)&a, sizeof a);
    k.append((const char *)v, sizeof v);
    return k;
  };
  auto chunk = [&](int n) {
    require(launch_cg_chunk(pl, rc, ps, l->nl, a, g_stream, tu.block, tu.form, rx, tu.block1, n, w->S1, tu.cap),
            "no CG kernels for this storage combination");
  };
  // the reliable update (FP64 mode: the flush of the pending update alone)
  auto update = [&] {
    launch_cg_flush(ps, rx, a, g_stream);  // apply the pending x/r update; exact |r|^2
    if (mixed) {
      launch_xfold(g, x, w->xs, S, rx, g_stream);
      resid64(l, par, x, b, w->r64, S, 2, w->t64);
      launch_rs_from_true(g, w->rs, w->r64, S, rx, g_stream);
    }
  };
  // the sloppy recursion diverged: back to the best iterate, whose true
  // residual epi_reliable accepts again, restarting p = r
  auto rollback = [&] {
    HIPCHK(hipMemcpyAsync(x, w->p64, vec64_bytes(lat), hipMemcpyDeviceToDevice, g_stream));
    resid64(l, par, x, b, w->r64, S, 2, w->t64);
    launch_rs_from_true(g, w->rs, w->r64, S, rx, g_stream);
    check_launch();
    pull(Sh, S);
  };
  const int nint = Sh->relmax;
  bool intervals = mixed && g_cg_graph && g_cg_intervals && nint <= kMaxInterval && !Sh->done;
  if (intervals) {
    iexec = cached_graph(w->cg[0], graph_key(0, nint), [&] {
      chunk(nint);
      update();
      launch_keep_best(g, w->p64, x, S, g_stream);
    });
  } else if (g_cg_graph && !Sh->done) {
    exec = cached_graph(w->cg[1], graph_key(1, kChunk), [&] { chunk(kChunk); });
  }
  double r2last = Sh->r2true;  // true residual^2 at the last reliable update
  while (!Sh->done) {
    if (intervals) {
      const int its0 = Sh->its;
      HIPCHK(hipGraphLaunch(iexec, g_stream));
      check_launch();
      pull(Sh, S);
      if (Sh->rollback) rollback();
      bool to_chunks = !Sh->done && Sh->its - its0 < nint;
      if (!Sh->done && !to_chunks && Sh->r2true > 0.0 && Sh->r2true < r2last) {
        // the true residual's decay over this interval predicts the
        // iterations left; with fewer than one interval to go, finish in
        // chunks, whose idle tail after convergence is under kChunk
        // iterations instead of under nint
        const double per_it = std::log(Sh->r2true / r2last) / (double)(Sh->its - its0);
        to_chunks = std::log(Sh->r2stop / Sh->r2true) / per_it < (double)nint;
      }
      r2last = Sh->r2true;
      if (to_chunks) {
        intervals = false;
        exec = cached_graph(w->cg[1], graph_key(1, kChunk), [&] { chunk(kChunk); });
      }
      continue;
    }
    if (exec) HIPCHK(hipGraphLaunch(exec, g_stream));
    else chunk(kChunk);
    check_launch();
    pull(Sh, S);
    if (Sh->done || Sh->need_rel) {
      update();
      check_launch();
      pull(Sh, S);
      if (mixed && Sh->rollback) rollback();
      else if (mixed && Sh->improved)
        HIPCHK(hipMemcpyAsync(w->p64, x, vec64_bytes(lat), hipMemcpyDeviceToDevice, g_stream));
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

// Defect correction of an FP64 solution x of 4(m2 - D_pq D_qp) x = b on parity
// par (the loop of solve_xx, started from x instead of 0): while the true
// residual exceeds r2stop and falls, solve for the correction with cg_solve
// (sp's sloppy storage, reliable updates) and add it.  Adds the iterations
// and reliable updates to *its, *nrel; returns the final |b - A x|^2.
double refine_xx(qexhip_links *l, int par, double *x, const double *b, double m2, double r2stop,
                 const qexhip_solver_params *sp, int *its, int *nrel) {
  Work *w = work(l);
  qexhip_lattice *lat = l->lat;
  const Geom &g = lat->g;
  Scal *S = w->S, *Sh = w->Sh;
  const size_t vb = vec64_bytes(lat);
  auto residual = [&] {
    memset(Sh, 0, sizeof(Scal));
    Sh->m2 = m2;
    push(S, Sh);
    resid64(l, par, x, b, w->rho, S, 3, w->t64);  // rho = b - A x
    check_launch();
    pull(Sh, S);
    return Sh->tmp[0];
  };
  double r2 = residual(), r2prev = INFINITY;
  int used = 0;
  while (r2 > r2stop && r2 < r2prev) {
    r2prev = r2;
    const int mx = sp->maxits - used;
    if (mx <= 0) break;
    launch_fill(w->y, vb, g_stream);
    int it = 0, nr = 0;
    cg_solve(l, par, w->y, w->rho, std::sqrt(m2), r2stop / r2, mx, sp->sloppy, sp->sloppy_links, sp->recon,
             sp->reliable_delta, sp->reliable_max_its, &it, &nr);
    used += it;
    *nrel += nr;
    launch_axpby64(g, 1.0, w->y, 1.0, x, g_stream);  // x += y
    r2 = residual();
  }
  *its += used;
  return r2;
}

// |b - (A + sigma) x|^2 in FP64 on parity par, A = 4(m^2 - D_pq D_qp): the
// true residual of one shifted system, which cgm.nim:208-239 computes for
// every shift at verbosity 3.  Uses the links' FP64 work vectors r64, t64.
double shifted_resid2(qexhip_links *l, int par, const double *x, const double *b, double m, double sigma) {
  Work *w = work(l);
  Scal *S = w->S, *Sh = w->Sh;
  memset(Sh, 0, sizeof(Scal));
  Sh->m2 = m * m + 0.25 * sigma;
  push(S, Sh);
  resid64(l, par, x, b, w->r64, S, 3, w->t64);
  check_launch();
  pull(Sh, S);
  return Sh->tmp[0];
}

// Fill the true-residual outputs of a multi-shift solve: sp->r2 is the base
// system's |b - A x_0|^2/|b|^2 and sp->shift_r2[k] (when given) each shifted
// system's, all recomputed in FP64 from the returned solutions.
void multishift_true_r2(qexhip_links *l, int par, const double *const *xs, const double *b, double b2, double m,
                        const double *sg, int nshift, qexhip_solver_params *sp) {
  for (int k = 0; k < nshift; k++) {
    if (k > 0 && !sp->shift_r2) break;
    const double r2 = (b2 > 0) ? shifted_resid2(l, par, xs[k], b, m, k == 0 ? 0.0 : sg[k]) / b2 : 0.0;
    if (k == 0) sp->r2 = r2;
    if (sp->shift_r2) sp->shift_r2[k] = r2;
  }
}

void msb_core(qexhip_links *l, int par, int nrhs, const double *const *b, double *const *x, double m,
              const double *sigmas, int nshift, const double *r2rel, const int *maxits, int flags, int pv, int *its,
              double *b2);

// Mixed-precision multi-shift CG, no QEX counterpart (cgm.nim runs in FP64
// only).  Two phases:
//  1. the multi-shift recursion (cgm.nim:68-246) on FP32 storage with FP32
//     links (FP32 stencil arithmetic; FP64 recurrences, vector updates and
//     reductions), the shifted x_k, p_k advanced once per window of
//     iterations from the window's residuals (msb_core), until the base
//     residual |r|^2 <= max(r2req, g_ms_sloppy_r2) |b|^2; a shift whose
//     residual zeta_k^2 |r|^2 met that bound stops being updated;
//  2. every shift refined from its phase-1 solution to r2req by the defect
//     correction of refine_xx on its own system 4(m^2 + sigma_k/4 - D D),
//     whose inner solves are the mixed CG of cg_solve on sp's sloppy storage.
// sp->iterations counts both phases; the true residuals are recomputed in FP64.
void solve_ms_mixed(qexhip_links *l, int par, double *const *xs, const double *b, double m, const double *sigmas,
                    int nshift, qexhip_solver_params *sp) {
  auto t0 = std::chrono::steady_clock::now();
  const Geom &g = l->lat->g;
  require(nshift >= 1 && nshift <= QH_MAXSHIFT, "multishift: nshift out of range");
  for (int k = 1; k < nshift; k++)
    require(m * m + 0.25 * sigmas[k] > 0.0, "multishift: a shifted system is not positive definite");
  // phase 1
  const double r2rel1 = std::max(sp->r2req, g_ms_sloppy_r2);
  const int flags = QEXHIP_BATCH_GRAPH | (g_autotune ? QEXHIP_BATCH_TUNE : 0);
  int its1 = 0;
  double b2 = 0.0;
  msb_core(l, par, 1, &b, xs, m, sigmas, nshift, &r2rel1, &sp->maxits, flags, P32, &its1, &b2);
  // phase 2
  int its = its1, nrel = 0;
  const double r2stop = sp->r2req * b2;
  for (int k = 0; k < nshift; k++)
    refine_xx(l, par, xs[k], b, m * m + (k == 0 ? 0.0 : 0.25 * sigmas[k]), r2stop, sp, &its, &nrel);
  sp->iterations = its;
  sp->reliable_updates = nrel;
  sp->flops = flops_per_its(l) * its + (double)(4 * g.Vh + 15) * its1 * nshift;
  multishift_true_r2(l, par, xs, b, b2, m, sigmas, nshift, sp);
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

void check_solver_params(qexhip_links *l, const qexhip_solver_params *sp) {
  require(sp != nullptr, "null solver params");
  require((sp->sloppy >= 0 && sp->sloppy <= 3) || sp->sloppy == QEXHIP_FP16_ALL, "bad sloppy precision");
  require(sp->sloppy_links >= -1 && sp->sloppy_links <= PQ16, "bad sloppy link precision");
  require(sp->recon == 0 || sp->recon == 18 || sp->recon == 12, "recon must be 0, 18 or 12");
  if (sp->sloppy != P64) {
    const int sv = (sp->sloppy == QEXHIP_FP16_ALL) ? P16 : sp->sloppy;
    const int sl = sp->sloppy_links < 0 ? sv : sp->sloppy_links;
    require(sl != P64, "mixed solve needs sloppy links below FP64");
    require(sl == sv || sv == P32 || (sl == PQ16 && sv == P16),
            "sloppy link/spinor storage combination not supported");
    if (sp->sloppy == QEXHIP_FP16_ALL)
      require(sl == P16 || sl == PQ16, "FP16_ALL needs FP16 or Q16 sloppy links");
  }
  if (sp->recon == 12) require(recon_allowed(l), "12-real links need plain (nl=4) phased SU(3) links");
}

// ------------------------------------------------------------ batched CG
// Independent FP64 CG solves (the fused iteration of cg_solve) on systems of
// different lattices, links and masses, one launch per stage for all of them
// (batch.hip).  The workspace (vectors, device scalars, block table, graph)
// is kept for the next call with the same links and parity.
struct BatchWs {
  std::vector<qexhip_links *> links;
  int par = 0, nl = 4, n = 0;
  std::vector<double *> x, r, p, Ap, t, b;
  Scal *S = nullptr, *Sh = nullptr, *Sx = nullptr;
  double *partials = nullptr;
  long maxblocks = 0, nblocks = 0;
  BatchItem *T = nullptr, *Tx = nullptr;  // device tables: solve, tuning scratch
  int block = 0, form = 2;
  long lkc = 0;  // link-cache bytes of the whole batch (batch_table)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~BatchWs() {
    for (auto *v : {&x, &r, &p, &Ap, &t, &b})
      for (double *q : *v) (void)hipFree(q);
    (void)hipFree(S);
    (void)hipHostFree(Sh);
    (void)hipFree(Sx);
    (void)hipFree(partials);
    (void)hipFree(T);
    (void)hipFree(Tx);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};
BatchWs *g_batch = nullptr;

void drop_batch() {
  if (g_batch) {
    HIPCHK(hipStreamSynchronize(g_stream));
    delete g_batch;
    g_batch = nullptr;
  }
}

// block table for block size `block`, scalars S (device array); returns the
// number of blocks of the batched grid.  lkc: link-cache bytes of the whole
// batch, shared out over the systems in proportion to their link bytes (a
// lattice with a pinned budget keeps its own).
long batch_table(BatchWs *w, int block, Scal *S, BatchItem *dst, long lkc = 0) {
  std::vector<BatchItem> T(w->n);
  double tot = 0.0;
  for (int k = 0; k < w->n; k++) tot += 8.0 * (double)lkblk_bytes(w->links[k], P64, 18);
  long b0 = 0;
  for (int k = 0; k < w->n; k++) {
    qexhip_links *l = w->links[k];
    const Geom &g = l->lat->g;
    BatchItem &it = T[k];
    memset(&it, 0, sizeof it);
    it.a.g = g;
    it.a.g.lkcache = g.lkcache >= 0 ? g.lkcache
                                    : (long)((double)lkc * 8.0 * (double)lkblk_bytes(l, P64, 18) / tot);
    it.a.links = links_store(l, P64, 18);
    it.a.lkblk = lkblk_bytes(l, P64, 18);
    it.a.par = w->par;
    it.a.remap = 0;
    it.a.x = w->x[k]; it.a.r = w->r[k]; it.a.p = w->p[k]; it.a.Ap = w->Ap[k]; it.a.t = w->t[k];
    it.a.S = S + k;
    it.a.red = RedWs{w->partials, (int)w->maxblocks};
    it.blk0 = b0;
    it.nblk = (g.Vh + block - 1) / block;
    const RedWs slice{w->partials + b0, (int)w->maxblocks};
    it.fin = FinArgs{slice, (int)it.nblk, S + k, 0, 0.0, 1};
    it.finf = FinArgs{slice, (int)it.nblk, S + k, 0, 0.0, 0};
    b0 += it.nblk;
  }
  HIPCHK(hipMemcpyAsync(dst, T.data(), sizeof(BatchItem) * w->n, hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  return b0;
}

BatchWs *get_batch(int n, qexhip_links *const *ls, int par) {
  if (g_batch && g_batch->n == n && g_batch->par == par &&
      std::equal(ls, ls + n, g_batch->links.begin()))
    return g_batch;
  drop_batch();
  BatchWs *w = new BatchWs();
  g_batch = w;
  w->n = n;
  w->par = par;
  w->links.assign(ls, ls + n);
  w->nl = ls[0]->nl;
  long mb = 0;
  for (int k = 0; k < n; k++) {
    const size_t vb = vec64_bytes(ls[k]->lat);
    for (auto *v : {&w->x, &w->r, &w->p, &w->Ap, &w->t, &w->b}) {
      v->push_back(dalloc<double>(vb));
      launch_fill(v->back(), vb, g_stream);
    }
    mb += (ls[k]->lat->g.Vh + 63) / 64;
  }
  w->maxblocks = mb;
  w->partials = dalloc<double>(sizeof(double) * 4 * mb);
  w->S = dalloc<Scal>(sizeof(Scal) * n);
  w->Sx = dalloc<Scal>(sizeof(Scal) * n);
  HIPCHK(hipHostMalloc((void **)&w->Sh, sizeof(Scal) * n, hipHostMallocDefault));
  w->T = dalloc<BatchItem>(sizeof(BatchItem) * n);
  w->Tx = dalloc<BatchItem>(sizeof(BatchItem) * n);
  return w;
}

void batch_chunk(BatchWs *w, const BatchItem *T) {
  for (int k = 0; k < kChunk; k++) launch_bcg_iter(T, w->n, w->nblocks, w->nl, w->block, g_stream, w->form);
  launch_bcg_flush(T, w->n, w->nblocks, w->block, g_stream);
}

void batch_shape(BatchWs *w, bool autotune) {
  // the batch has forms 2 and 3; a pinned form 4 (single systems only) means 3
  const int pinned = (g_cg_form == 4) ? 3 : g_cg_form;
  int block = 256, form = pinned ? pinned : 2;
  long lkc = 0;
  if (autotune && !g_block_forced) {
    std::string key = "batch-cg nl=" + std::to_string(w->nl);
    for (auto *l : w->links) key += " " + geom_key(l->lat->g);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      block = it->second.block;
      form = it->second.form;
      lkc = it->second.lkcache;
    } else {
      // scratch scalars that never converge
      std::vector<Scal> h(w->n);
      memset(h.data(), 0, sizeof(Scal) * w->n);
      for (auto &q : h) { q.r2stop = -1.0; q.maxits = INT_MAX; q.sigma = q.pscale = 1.0; q.flushed = 1; }
      HIPCHK(hipMemcpyAsync(w->Sx, h.data(), sizeof(Scal) * w->n, hipMemcpyHostToDevice, g_stream));
      Tuned best{256, 0, 1e300};
      for (int f : {2, 3}) {
        if (pinned && f != pinned) continue;
        for (int b : {64, 128, 256}) {
          const long nb = batch_table(w, b, w->Sx, w->Tx);
          const double us = time_us([&] { launch_bcg_iter(w->Tx, w->n, nb, w->nl, b, g_stream, f); }, 2, 8);
          if (us < best.us) best = Tuned{b, 0, us, f};
        }
      }
      // the batch's link-cache budget, on the chosen shape
      double bus = 1e300;
      for (long mbytes : kLinkCacheMB) {
        const long nb = batch_table(w, best.block, w->Sx, w->Tx, mbytes << 20);
        const double us =
            time_us([&] { launch_bcg_iter(w->Tx, w->n, nb, w->nl, best.block, g_stream, best.form); }, 2, 8);
        if (us < bus) {
          bus = us;
          best.lkcache = mbytes << 20;
        }
      }
      best.us = bus;
      g_tune[key] = best;
      block = best.block;
      form = best.form;
      lkc = best.lkcache;
    }
  } else if (g_block_forced) {
    block = g_block;
  }
  if (block != w->block || form != w->form || !w->exec) {
    w->block = block;
    w->form = form;
    if (w->exec) { (void)hipGraphExecDestroy(w->exec); w->exec = nullptr; }
    if (w->graph) { (void)hipGraphDestroy(w->graph); w->graph = nullptr; }
  }
  // the table every call: the link-cache budgets of pinned lattices may have
  // changed (the graph reads the table from device memory; its grid depends
  // only on the block size and the lattices)
  w->lkc = lkc;
  w->nblocks = batch_table(w, block, w->S, w->T, lkc);
}

void solve_batch(int n, qexhip_links *const *ls, int par, qexhip_field *const *xf, qexhip_field *const *bf,
                 const double *ms, qexhip_solver_params *sp, int flags) {
  auto t0 = std::chrono::steady_clock::now();
  BatchWs *w = get_batch(n, ls, par);
  batch_shape(w, (flags & QEXHIP_BATCH_TUNE) != 0);
  const bool graph = (flags & QEXHIP_BATCH_GRAPH) != 0;
  if (graph && !w->exec) {
    capture([&] { batch_chunk(w, w->T); }, &w->graph, &w->exec);
  }
  // x = 0, r = b, Ap = 0 ; scalars as cg_solve with x = 0 (cg.nim:138-153)
  for (int k = 0; k < n; k++) {
    const size_t vb = vec64_bytes(ls[k]->lat);
    HIPCHK(hipMemcpyAsync(w->b[k], bf[k]->par(par), vb, hipMemcpyDeviceToDevice, g_stream));
    HIPCHK(hipMemcpyAsync(w->r[k], w->b[k], vb, hipMemcpyDeviceToDevice, g_stream));
    launch_fill(w->x[k], vb, g_stream);
    launch_fill(w->Ap[k], vb, g_stream);
    Scal &h = w->Sh[k];
    memset(&h, 0, sizeof h);
    h.m2 = ms[k] * ms[k];
    h.maxits = sp[k].maxits;
    HIPCHK(hipMemcpyAsync(w->S + k, &h, sizeof h, hipMemcpyHostToDevice, g_stream));
    launch_scal_init(w->S + k, g_stream, 0);
    launch_norm2(P64, ls[k]->lat->g, w->b[k], w->S + k, 1, ls[k]->lat->red, g_stream);
  }
  check_launch();
  HIPCHK(hipMemcpyAsync(w->Sh, w->S, sizeof(Scal) * n, hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  std::vector<double> b2(n);
  for (int k = 0; k < n; k++) {
    Scal &h = w->Sh[k];
    b2[k] = h.tmp[1];
    h.b2 = b2[k];
    h.r2stop = sp[k].r2req * b2[k];
    h.r2 = h.r2prev = b2[k];
    h.done = !(b2[k] > h.r2stop) || sp[k].maxits <= 0;
    h.flushed = h.done;
  }
  HIPCHK(hipMemcpyAsync(w->S, w->Sh, sizeof(Scal) * n, hipMemcpyHostToDevice, g_stream));
  for (;;) {
    bool all = true;
    for (int k = 0; k < n; k++) all = all && w->Sh[k].done && w->Sh[k].flushed;
    if (all) break;
    if (graph) HIPCHK(hipGraphLaunch(w->exec, g_stream));
    else batch_chunk(w, w->T);
    check_launch();
    HIPCHK(hipMemcpyAsync(w->Sh, w->S, sizeof(Scal) * n, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
  }
  // FP64 true residuals, then the solutions out
  for (int k = 0; k < n; k++) {
    resid64(ls[k], par, w->x[k], w->b[k], w->r[k], w->S + k, 3, w->t[k]);
    HIPCHK(hipMemcpyAsync(xf[k]->par(par), w->x[k], vec64_bytes(ls[k]->lat), hipMemcpyDeviceToDevice, g_stream));
  }
  check_launch();
  HIPCHK(hipMemcpyAsync(w->Sh, w->S, sizeof(Scal) * n, hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int k = 0; k < n; k++) {
    sp[k].iterations = w->Sh[k].its;
    sp[k].reliable_updates = 0;
    sp[k].r2 = (b2[k] > 0) ? w->Sh[k].tmp[0] / b2[k] : 0.0;
    sp[k].flops = flops_per_its(ls[k]) * w->Sh[k].its;
    sp[k].seconds = sec;
  }
}

// ------------------------------------------------ batched multi-shift CG
// nrhs right-hand sides, one set of links (batch.hip (2)); the right-hand
// sides run in groups of 8/4/2/1 that share each link load.
struct MsbWs {
  qexhip_links *l = nullptr;
  int par = 0, nrhs = 0, nshift = 0;
  std::vector<double *> x, p, r, Ap, t, p0s;  // r: QH_MSB_RING vectors per right-hand side
  double **dx = nullptr, **dp = nullptr;
  double *hist = nullptr, *coef = nullptr;
  int *cinfo = nullptr;
  MsbScal *T = nullptr, *Th = nullptr, *Tx = nullptr;
  RedWs red{};
  std::vector<std::pair<int, int>> groups;  // (first rhs, size)
  int block = 0;
  long chunk = -1;   // site walk of the hops baked into the captured graph
  int batches = 0;   // vector-load batches of the hops baked into the captured graph
  long lkc = -2;     // link-cache bytes of the hops baked into the captured graph
  int pv = P64;      // vector storage and link precision of the captured graph (P64; P32: one source)
  double **dxo = nullptr;  // FP32 storage: device array of the FP64 output pointers
  double m2 = -1.0;  // mass^2 baked into the captured graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~MsbWs() {
    for (auto *v : {&x, &p, &r, &Ap, &t, &p0s})
      for (double *q : *v) (void)hipFree(q);
    (void)hipFree(dx);
    (void)hipFree(dp);
    (void)hipFree(dxo);
    (void)hipFree(hist);
    (void)hipFree(coef);
    (void)hipFree(cinfo);
    (void)hipFree(T);
    (void)hipHostFree(Th);
    (void)hipFree(Tx);
    (void)hipFree(red.partials);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};
MsbWs *g_msb = nullptr;

void drop_msb() {
  if (g_msb) {
    HIPCHK(hipStreamSynchronize(g_stream));
    delete g_msb;
    g_msb = nullptr;
  }
}

MsbWs *get_msb(qexhip_links *l, int par, int nrhs, int nshift) {
  if (g_msb && g_msb->l == l && g_msb->par == par && g_msb->nrhs == nrhs && g_msb->nshift == nshift) return g_msb;
  drop_msb();
  MsbWs *w = new MsbWs();
  g_msb = w;
  w->l = l;
  w->par = par;
  w->nrhs = nrhs;
  w->nshift = nshift;
  const size_t vb = vec64_bytes(l->lat);
  for (int k = 0; k < nrhs * nshift; k++) {
    w->x.push_back(dalloc<double>(vb));
    w->p.push_back(dalloc<double>(vb));
  }
  for (int j = 0; j < nrhs; j++) {
    for (auto *v : {&w->Ap, &w->t, &w->p0s}) {
      v->push_back(dalloc<double>(vb));
      launch_fill(v->back(), vb, g_stream);
    }
    w->r.push_back(dalloc<double>(vb * QH_MSB_RING));
    launch_fill(w->r.back(), vb * QH_MSB_RING, g_stream);
  }
  w->hist = dalloc<double>(sizeof(double) * nrhs * QH_MSB_WIN * nshift * 3);
  w->coef = dalloc<double>(sizeof(double) * nrhs * nshift * QH_MSB_NCOEF);
  w->cinfo = dalloc<int>(sizeof(int) * nrhs * (2 + QH_MAXSHIFT));
  launch_fill(w->hist, sizeof(double) * nrhs * QH_MSB_WIN * nshift * 3, g_stream);
  launch_fill(w->cinfo, sizeof(int) * nrhs * (2 + QH_MAXSHIFT), g_stream);
  w->dx = dalloc<double *>(sizeof(double *) * nrhs * nshift);
  w->dp = dalloc<double *>(sizeof(double *) * nrhs * nshift);
  HIPCHK(hipMemcpyAsync(w->dx, w->x.data(), sizeof(double *) * nrhs * nshift, hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(w->dp, w->p.data(), sizeof(double *) * nrhs * nshift, hipMemcpyHostToDevice, g_stream));
  w->T = dalloc<MsbScal>(sizeof(MsbScal) * nrhs);
  w->Tx = dalloc<MsbScal>(sizeof(MsbScal) * nrhs);
  HIPCHK(hipHostMalloc((void **)&w->Th, sizeof(MsbScal) * nrhs, hipHostMallocDefault));
  // one partial-sum slot per workgroup of the multi-RHS hops at any block size
  // and site walk (mrhs_grid: below 2 Vh/64 + 8)
  w->red.maxblocks = 2 * (int)((l->lat->g.Vh + 63) / 64) + 64;
  w->red.partials = dalloc<double>(sizeof(double) * 3 * QH_MAXRHS * w->red.maxblocks);
  for (int j = 0; j < nrhs;) {
    int g = 8;
    while (g > nrhs - j) g >>= 1;
    w->groups.push_back({j, g});
    j += g;
  }
  HIPCHK(hipStreamSynchronize(g_stream));
  return w;
}

MsbPtrs msb_ptrs(MsbWs *w, int j0, int nr) {
  MsbPtrs P;
  memset(&P, 0, sizeof P);
  for (int j = 0; j < nr; j++) {
    P.p0[j] = w->p[(j0 + j) * w->nshift];
    P.t[j] = w->t[j0 + j];
    P.Ap[j] = w->Ap[j0 + j];
    P.r[j] = w->r[j0 + j];
    P.p0s[j] = w->p0s[j0 + j];
  }
  P.rstride = (long)(vec64_bytes(w->l->lat) / sizeof(double));
  P.x = w->dx + j0 * w->nshift;
  P.p = w->dp + j0 * w->nshift;
  P.hist = w->hist + (size_t)j0 * QH_MSB_WIN * w->nshift * 3;
  P.coef = w->coef + (size_t)j0 * w->nshift * QH_MSB_NCOEF;
  P.cinfo = w->cinfo + j0 * (2 + QH_MAXSHIFT);
  return P;
}

// end of a window of iterations: every group's shifted x_k, p_k brought up to date
void msb_flush(MsbWs *w, MsbScal *T) {
  for (auto &gr : w->groups)
    launch_msb_flush(w->l->lat->g, w->pv, gr.second, w->nshift, msb_ptrs(w, gr.first, gr.second), T + gr.first,
                     g_stream);
}

void msb_iter(MsbWs *w, MsbScal *T, double m2, int block, long chunk, int batches, long lkc) {
  qexhip_links *l = w->l;
  require(mrhs_grid(l->lat->g, block, chunk) <= (unsigned)w->red.maxblocks, "multi-shift batch: hop grid too large");
  const char *links = links_store(l, w->pv, 18);
  const long lkblk = lkblk_bytes(l, w->pv, 18);
  Geom g = l->lat->g;
  g.lkcache = lkc;
  for (auto &gr : w->groups)
    launch_msb_iter(g, w->pv, links, lkblk, l->nl, w->par, gr.second, w->nshift, m2,
                    msb_ptrs(w, gr.first, gr.second), T + gr.first, w->red, block, chunk, batches, g_stream);
}

// The batched multi-shift CG (cgm.nim:68-246) with deferred shift updates,
// on storage pv: P64 (FP64 vectors, links and arithmetic), or P32 for one
// right-hand side (FP32 vectors and links, FP32 stencil arithmetic, FP64
// updates, recurrences and reductions: the first phase of the mixed solve).
// From b[j] (FP64, own parity, device) into x[j*nshift + k] (FP64, device),
// each right-hand side until |r|^2 <= r2rel[j] |b_j|^2 or maxits[j]
// iterations; its[j] and b2[j] = |b_j|^2 are returned.
void msb_core(qexhip_links *l, int par, int nrhs, const double *const *b, double *const *x, double m,
              const double *sigmas, int nshift, const double *r2rel, const int *maxits, int flags, int pv, int *its,
              double *b2) {
  require(pv == P64 || (pv == P32 && nrhs == 1), "multi-shift batch: FP64, or FP32 storage for one source");
  MsbWs *w = get_msb(l, par, nrhs, nshift);
  if (w->pv != pv) {  // the captured graph holds the storage precision's kernels
    w->pv = pv;
    if (w->exec) { (void)hipGraphExecDestroy(w->exec); w->exec = nullptr; }
    if (w->graph) { (void)hipGraphDestroy(w->graph); w->graph = nullptr; }
    w->block = 0;
  }
  links_store(l, pv, 18);  // packed before any capture (it may allocate)
  qexhip_lattice *lat = l->lat;
  const size_t vb = vec64_bytes(lat);
  const double m2 = m * m;
  // launch shape: block and site walk of the hops (a pinned walk fixes the latter)
  int block = g_block_forced ? g_block : 256;
  long chunk = g_msb_walk > 0 ? g_msb_walk : 0;
  int batches = 1;
  long lkc = 0;
  if ((flags & QEXHIP_BATCH_TUNE) && !g_block_forced) {
    const std::string key = "msb nl=" + std::to_string(l->nl) + " nrhs=" + std::to_string(nrhs) +
                            " nshift=" + std::to_string(nshift) + " walk=" + std::to_string(g_msb_walk) +
                            (pv == P32 ? " fp32 " : " ") + geom_key(lat->g);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      block = it->second.block;
      chunk = it->second.chunk;
      batches = it->second.batches;
      lkc = it->second.lkcache;
    } else {
      std::vector<MsbScal> h(nrhs);
      memset(h.data(), 0, sizeof(MsbScal) * nrhs);
      for (auto &q : h) {
        q.continuing = 1;
        q.maxits = INT_MAX;
        q.r2stop = -1.0;
        q.nshift = nshift;
        for (int k = 0; k < nshift; k++) { q.active[k] = 1; q.zi[k] = q.zim1[k] = 1.0; }
      }
      HIPCHK(hipMemcpyAsync(w->Tx, h.data(), sizeof(MsbScal) * nrhs, hipMemcpyHostToDevice, g_stream));
      for (int k = 0; k < nrhs * nshift; k++) {
        launch_fill(w->x[k], vb, g_stream);
        launch_fill(w->p[k], vb, g_stream);
      }
      Tuned best{256, 1, 1e300};
      const long S = lat->g.Vh / lat->g.L[3];
      for (int b : {64, 128, 256})
        for (long c : {0L, 1024L})
          for (int nb : {1, 2, 4}) {
            if (g_msb_walk >= 0 && c != 0) continue;  // a pinned walk: one candidate per block and batches
            const long cw = g_msb_walk >= 0 ? g_msb_walk : c;
            if (cw && (cw % b || cw >= S)) continue;
            const double us = time_us([&] { msb_iter(w, w->Tx, m2, b, cw, nb, 0); }, 1, 4);
            if (us < best.us) {
              best = Tuned{b, 1, us};
              best.chunk = cw;
              best.batches = nb;
            }
          }
      // the link-cache budget of the hops, on the chosen shape
      double bus = 1e300;
      for (long mbytes : kLinkCacheMB) {
        const double us =
            time_us([&] { msb_iter(w, w->Tx, m2, best.block, best.chunk, best.batches, mbytes << 20); }, 1, 4);
        if (us < bus) {
          bus = us;
          best.lkcache = mbytes << 20;
        }
      }
      best.us = bus;
      g_tune[key] = best;
      block = best.block;
      chunk = best.chunk;
      batches = best.batches;
      lkc = best.lkcache;
    }
  }
  if (lat->g.lkcache >= 0) lkc = lat->g.lkcache;  // a pinned lattice budget
  const bool graph = (flags & QEXHIP_BATCH_GRAPH) != 0;
  if (block != w->block || m2 != w->m2 || chunk != w->chunk || batches != w->batches || lkc != w->lkc) {
    w->block = block;
    w->chunk = chunk;
    w->batches = batches;
    w->lkc = lkc;
    w->m2 = m2;
    if (w->exec) { (void)hipGraphExecDestroy(w->exec); w->exec = nullptr; }
    if (w->graph) { (void)hipGraphDestroy(w->graph); w->graph = nullptr; }
  }
  // r := b ; p_k := b ; x_k := 0 (cgm.nim:122-158), in the storage precision
  for (int j = 0; j < nrhs; j++) {
    if (pv == P64) {
      HIPCHK(hipMemcpyAsync(w->r[j], b[j], vb, hipMemcpyDeviceToDevice, g_stream));
      for (int k = 0; k < nshift; k++) {
        HIPCHK(hipMemcpyAsync(w->p[j * nshift + k], w->r[j], vb, hipMemcpyDeviceToDevice, g_stream));
        launch_fill(w->x[j * nshift + k], vb, g_stream);
      }
    } else {
      launch_cgm_start32(lat->g, nshift, b[j], w->r[j], (void *const *)(w->dx + j * nshift),
                         (void *const *)(w->dp + j * nshift), g_stream);
    }
    launch_norm2(P64, lat->g, b[j], lat->S, 1, lat->red, g_stream);
    check_launch();
    pull(lat->Sh, lat->S);
    b2[j] = lat->Sh->tmp[1];
    MsbScal &h = w->Th[j];
    memset(&h, 0, sizeof h);
    h.r2i = b2[j];
    h.r2stop = r2rel[j] * b2[j];
    h.alphaim1 = -1.0;
    h.betaim1 = 0.0;
    h.maxits = maxits[j];
    h.nshift = nshift;
    h.freeze = (flags & QEXHIP_BATCH_NOFREEZE) ? 0 : 1;
    h.continuing = (b2[j] > h.r2stop) && maxits[j] > 0;
    for (int k = 0; k < nshift; k++) {
      h.sg[k] = (k == 0) ? 0.0 : sigmas[k];
      h.zi[k] = h.zim1[k] = 1.0;
      h.active[k] = h.continuing;
    }
  }
  HIPCHK(hipMemcpyAsync(w->T, w->Th, sizeof(MsbScal) * nrhs, hipMemcpyHostToDevice, g_stream));
  // one chunk = one window of the deferred shift updates: kChunk iterations, then the flush
  static_assert(kChunk <= QH_MSB_WIN, "a chunk must fit one window of the multi-shift updates");
  auto msb_chunk = [&] {
    for (int c = 0; c < kChunk; c++) msb_iter(w, w->T, m2, block, chunk, batches, lkc);
    msb_flush(w, w->T);
  };
  if (graph && !w->exec) capture(msb_chunk, &w->graph, &w->exec);
  for (;;) {
    bool any = false;
    for (int j = 0; j < nrhs; j++) any = any || w->Th[j].continuing;
    if (!any) break;
    if (graph) {
      HIPCHK(hipGraphLaunch(w->exec, g_stream));
    } else {
      msb_chunk();
    }
    check_launch();
    HIPCHK(hipMemcpyAsync(w->Th, w->T, sizeof(MsbScal) * nrhs, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
  }
  for (int j = 0; j < nrhs; j++) require(!w->Th[j].winflow, "multi-shift batch: update window overflow");
  if (pv == P64) {
    for (int j = 0; j < nrhs; j++)
      for (int k = 0; k < nshift; k++)
        HIPCHK(hipMemcpyAsync(x[j * nshift + k], w->x[j * nshift + k], vb, hipMemcpyDeviceToDevice, g_stream));
  } else {  // FP64 copies of the FP32 solutions
    if (!w->dxo) w->dxo = dalloc<double *>(sizeof(double *) * nrhs * nshift);
    HIPCHK(hipMemcpyAsync(w->dxo, x, sizeof(double *) * nrhs * nshift, hipMemcpyHostToDevice, g_stream));
    launch_cgm_export32(lat->g, nshift, (void *const *)w->dx, w->dxo, g_stream);
    check_launch();
  }
  HIPCHK(hipStreamSynchronize(g_stream));
  for (int j = 0; j < nrhs; j++) its[j] = w->Th[j].its;
}

void solve_msb(qexhip_links *l, int par, int nrhs, qexhip_field *const *xs, qexhip_field *const *bf, double m,
               const double *sigmas, int nshift, qexhip_solver_params *sp, int flags) {
  auto t0 = std::chrono::steady_clock::now();
  qexhip_lattice *lat = l->lat;
  std::vector<const double *> b(nrhs);
  std::vector<double *> x((size_t)nrhs * nshift);
  std::vector<double> r2rel(nrhs), b2(nrhs);
  std::vector<int> maxits(nrhs), its(nrhs);
  for (int j = 0; j < nrhs; j++) {
    b[j] = (const double *)bf[j]->par(par);
    r2rel[j] = sp[j].r2req;
    maxits[j] = sp[j].maxits;
    for (int k = 0; k < nshift; k++) x[j * nshift + k] = (double *)xs[j * nshift + k]->par(par);
  }
  msb_core(l, par, nrhs, b.data(), x.data(), m, sigmas, nshift, r2rel.data(), maxits.data(), flags, P64, its.data(),
           b2.data());
  for (int j = 0; j < nrhs; j++) {
    std::vector<const double *> xj(x.begin() + j * nshift, x.begin() + (j + 1) * nshift);
    multishift_true_r2(l, par, xj.data(), b[j], b2[j], m, sigmas, nshift, &sp[j]);
  }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int j = 0; j < nrhs; j++) {
    sp[j].iterations = its[j];
    sp[j].reliable_updates = 0;
    sp[j].flops = flops_per_its(l) * its[j] + (double)(4 * lat->g.Vh + 15) * its[j] * nshift;
    sp[j].seconds = sec;
  }
}

}  // namespace

// ================================================================ C ABI
extern "C" int qexhip_field_destroy(qexhip_field *f);

namespace {
struct FieldGuard {  // destroys a field created inside an entry point on every exit
  qexhip_field *&f;
  ~FieldGuard() {
    if (f) qexhip_field_destroy(f);
  }
};
}  // namespace

extern "C" {

const char *qexhip_last_error(void) { return g_err.c_str(); }

int qexhip_init(int device) {
  API_BEGIN
  if (g_stream) return 0;
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  require(device >= 0 && device < n, "qexhip_init: no such device");
  require(device < kMaxDevices, "qexhip_init: device index beyond the library's table");
  HIPCHK(hipSetDevice(device));
  if (!g_scratch[device]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "qexhip_init: no device address for the scratch buffer");
    g_scratch[device] = p;
  }
  HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_device = device;
  if (const char *b = getenv("QEXHIP_BLOCK")) {
    int v = atoi(b);
    if (v >= 64 && v <= 256 && v % 64 == 0) {
      g_block = v;
      g_block_forced = true;
    }
  }
  if (const char *t = getenv("QEXHIP_TUNE")) g_autotune = atoi(t) != 0;
  if (const char *c = getenv("QEXHIP_CG_GRAPH")) g_cg_graph = atoi(c) != 0;
  if (const char *c = getenv("QEXHIP_CG_INTERVALS")) g_cg_intervals = atoi(c) != 0;
  if (const char *f = getenv("QEXHIP_CG_FORM")) {
    const int v = atoi(f);
    if (v >= 2 && v <= 4) g_cg_form = v;
  }
  if (const char *r = getenv("QEXHIP_XCD_REMAP")) g_remap = atoi(r) != 0;
  if (const char *w = getenv("QEXHIP_SMEAR_WALK")) g_smear_walk = atoi(w) != 0;
  if (const char *w = getenv("QEXHIP_MS_SLOPPY_R2")) g_ms_sloppy_r2 = atof(w);
  if (const char *c = getenv("QEXHIP_LINK_CACHE_MB")) g_link_cache = atol(c) << 20;
  API_END
}

int qexhip_finalize(void) {
  API_BEGIN
  if (g_stream) {
    drop_batch();
    drop_msb();
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
  // the extents are checked before the device, so that a caller learns that a
  // lattice is refused without a GPU
  require(dims != nullptr && out != nullptr, "null argument");
  long vol = 1;
  for (int d = 0; d < 4; d++) {
    require(dims[d] >= 2 && dims[d] % 2 == 0, "lattice extents must be even and >= 2");
    require(dims[d] <= (1 << 20), "lattice extent too large");
    vol *= dims[d];
    // site indices and the generators' site counters are 32-bit
    require(vol <= (1L << 31), "lattice too large: more than 2^31 sites");
  }
  {
    // Kernels address one parity of a color-vector field (3 x stride complex
    // words), and a link block in the plane layout, with 32-bit byte offsets
    // from the field's base: the FP64 vector block, 48 B x stride, must stay
    // below 4 GiB (at most 89 478 400 sites per parity, 178 956 800 sites).
    // Blocked link storage (FP64/FP32) takes a 64-bit base per 64-site block.
    const long stride = (vol / 2 + 255) / 256 * 256;
    const long plane_link_bytes = link_blocked<P64, 8>() ? 36 : 144;  // FP16 18-real, or FP64 HISQ planes
    require(stride * 48 < (1L << 32) && stride * plane_link_bytes < (1L << 32),
            "lattice too large: one parity of an FP64 color vector would pass 4 GiB "
            "(at most 178956800 sites)");
  }
  check_init();
  qexhip_lattice *lat = new qexhip_lattice();
  for (int d = 0; d < 4; d++) lat->g.L[d] = dims[d];
  lat->g.hx = dims[0] / 2;
  lat->V = (long)dims[0] * dims[1] * dims[2] * dims[3];
  lat->g.Vh = lat->V / 2;
  lat->g.stride = (lat->g.Vh + 255) / 256 * 256;
  lat->g.lkcache = g_link_cache;
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
  if (g_batch && std::find(g_batch->links.begin(), g_batch->links.end(), l) != g_batch->links.end()) drop_batch();
  if (g_msb && g_msb->l == l) drop_msb();
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

// Launch shape of the multi-RHS hop for nr fields (block size, vector-load
// batches, site-walk chunk), timed (tuner on) with the group's inputs and
// scratch outputs of its own, one per field (a = b = 0), so the caller's
// fields are untouched and the candidates write as many bytes as the real call.
static Tuned tuned_hop_multi(qexhip_links *l, int prec, int nr, const MultiPtrs &P, const char *links, long lkblk,
                             int par) {
  std::vector<int> nbs;
  for (int nb : {1, 2})
    if ((nb == 1 || nr >= 2) && (!g_multi_batches || nb == g_multi_batches)) nbs.push_back(nb);
  if (nbs.empty()) nbs.push_back(1);  // a pin the group cannot take
  std::vector<long> chunks;
  if (g_multi_chunk >= 0) chunks.push_back(g_multi_chunk);
  else chunks = {0, 1024, 2048, 4096};
  char kb[96];
  snprintf(kb, sizeof kb, "hop_multi nl=%d prec=%d nr=%d pin=%d/%ld ", l->nl, prec, nr, g_multi_batches, g_multi_chunk);
  const std::string key = kb + geom_key(l->lat->g);
  auto it = g_tune.find(key);
  if (it != g_tune.end()) return it->second;
  Tuned best{256, 1, 1e300};
  best.batches = nbs[0];
  best.chunk = chunks[0];
  if (g_block_forced || !g_autotune) {
    best.block = g_block;
    best.us = 0.0;
    return best;
  }
  const size_t ob = (size_t)3 * l->lat->g.stride * word_bytes(prec);
  std::vector<char *> outs(nr);
  for (auto &o : outs) o = dalloc<char>(ob);
  MultiPtrs Q = P;
  for (int j = 0; j < nr; j++) Q.out[j] = outs[j];  // candidates only write these
  Geom g = l->lat->g;
  g.lkcache = 0;
  for (int nb : nbs)
    for (long ch : chunks)
      for (int block : {64, 128, 256}) {
        const double us = time_us(
            [&] { launch_hop_multi(prec, l->nl, nr, nb, g, links, lkblk, par, 0.0, 0.0, Q, block, 0, ch, g_stream); },
            2, 8);
        if (us < best.us) {
          best = Tuned{block, 1, us};
          best.batches = nb;
          best.chunk = ch;
        }
      }
  // the link-cache budget, on the chosen shape
  if (l->lat->g.lkcache < 0) {
    double bus = 1e300;
    for (long mbytes : kLinkCacheMB) {
      g.lkcache = mbytes << 20;
      const double us = time_us(
          [&] {
            launch_hop_multi(prec, l->nl, nr, best.batches, g, links, lkblk, par, 0.0, 0.0, Q, best.block, 0,
                             best.chunk, g_stream);
          },
          2, 8);
      if (us < bus) {
        bus = us;
        best.lkcache = mbytes << 20;
      }
    }
    best.us = bus;
  }
  HIPCHK(hipStreamSynchronize(g_stream));
  for (auto *o : outs) (void)hipFree(o);
  g_tune[key] = best;
  return best;
}

// r_j = a r_j + b x_j + (2D) x_j, j < n, links loaded once per group of up to 8
static void dslash_multi(qexhip_links *l, int par, int n, qexhip_field *const *r, const qexhip_field *const *x,
                         double ca, double cb) {
  require(n >= 1 && r && x, "stagD2_multi: bad arguments");
  const int pr = r[0]->prec;
  require(pr == P64 || pr == P32, "stagD2_multi: FP64 or FP32 fields");
  require(par >= 0 && par <= 2, "bad parity");
  for (int j = 0; j < n; j++) {
    require(r[j] && x[j], "null field");
    check_same_lat(l, r[j]);
    check_same_lat(l, x[j]);
    require(r[j]->prec == pr && x[j]->prec == pr, "stagD2_multi: all fields must share one precision");
  }
  const char *links = links_store(l, pr, 18);
  const long lkblk = lkblk_bytes(l, pr, 18);
  for (int p = 0; p < 2; p++) {
    if (par != 2 && par != p) continue;
    for (int j0 = 0; j0 < n;) {
      int gsz = QH_MAXRHS;
      while (gsz > n - j0) gsz >>= 1;
      if (gsz == 1) {  // one field: the one-parity hop of stagD2 (its own tuning)
        const qexhip_field *xj = x[j0];
        qexhip_field *rj = r[j0];
        if (ca == 0.0)
          hop(l, pr, 18, p, 1, 0, xj->par(1 - p), rj->par(p), xj->par(p), nullptr, cb, 0, 1.0, 0, 0, nullptr);
        else
          hop(l, pr, 18, p, 2, 0, xj->par(1 - p), rj->par(p), rj->par(p), xj->par(p), ca, cb, 1.0, 0, 0, nullptr);
        j0 += 1;
        continue;
      }
      MultiPtrs P;
      memset(&P, 0, sizeof P);
      for (int j = 0; j < gsz; j++) {
        P.in[j] = x[j0 + j]->par(1 - p);
        P.own[j] = x[j0 + j]->par(p);
        P.out[j] = r[j0 + j]->par(p);
      }
      const Tuned tu = tuned_hop_multi(l, pr, gsz, P, links, lkblk, p);
      Geom g = l->lat->g;
      g.lkcache = link_cache_of(l->lat, tu);
      require(launch_hop_multi(pr, l->nl, gsz, tu.batches, g, links, lkblk, p, ca, cb, P, tu.block, 0, tu.chunk,
                               g_stream),
              "stagD2_multi: no kernel for this shape");
      j0 += gsz;
    }
    check_launch();
  }
}

int qexhip_stagD2_multi(qexhip_links *l, int par, int n, qexhip_field *const *r, const qexhip_field *const *x,
                        double a, double b) {
  API_BEGIN
  check_init();
  dslash_multi(l, par, n, r, x, a, b);
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_stream_read(long mbytes, int nrep, double *gbs) {
  API_BEGIN
  check_init();
  require(mbytes >= 16 && mbytes <= (1L << 16) && nrep > 0 && gbs, "bad arguments");
  const long n16 = (mbytes << 20) / 16;
  const int blocks = 2048;
  char *in = dalloc<char>((size_t)n16 * 16);
  double *out = dalloc<double>(sizeof(double) * blocks * 256);
  launch_fill_pattern(in, (size_t)n16 * 16, g_stream);  // non-zero data, as real fields
  double best = 1e300;
  for (int rep = 0; rep < 3; rep++)
    best = std::min(best, time_us([&] { launch_stream_read(in, n16, out, blocks, g_stream); }, 1, nrep));
  HIPCHK(hipStreamSynchronize(g_stream));
  (void)hipFree(in);
  (void)hipFree(out);
  *gbs = (double)n16 * 16 / (best * 1e-6) / 1e9;
  API_END
}

int qexhip_time_hop_multi(qexhip_links *l, int prec, int nrhs, int nrep, double *ms) {
  API_BEGIN
  check_init();
  require(l && ms && nrep > 0 && nrhs >= 1 && nrhs <= 64, "bad arguments");
  require(prec == P64 || prec == P32, "multi-RHS hop: FP64 or FP32");
  qexhip_lattice *lat = l->lat;
  std::vector<qexhip_field *> x(nrhs, nullptr), r(nrhs, nullptr);
  struct Guard {
    std::vector<qexhip_field *> &a, &b;
    ~Guard() {
      for (auto *f : a) qexhip_field_destroy(f);
      for (auto *f : b) qexhip_field_destroy(f);
    }
  } guard{x, r};
  for (int j = 0; j < nrhs; j++) {
    if (qexhip_field_create(lat, prec, &x[j]) || qexhip_field_create(lat, prec, &r[j]))
      throw std::runtime_error(g_err);
    launch_source_gauss(lat->g, 99 + j, 0, 2, lat->stage, g_stream);
    import_field(x[j], 2, lat->stage, 1);
  }
  for (int i = 0; i < 3; i++) dslash_multi(l, 0, nrhs, r.data(), x.data(), 0.0, 0.0);
  const double us = time_us([&] { dslash_multi(l, 0, nrhs, r.data(), x.data(), 0.0, 0.0); }, 0, nrep);
  *ms = us * 1e-3;
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

int qexhip_stagD2xx_noalloc(qexhip_links *l, int par, qexhip_field *r, const qexhip_field *x, double m2) {
  API_BEGIN
  check_init();
  check_same_lat(l, r);
  check_same_lat(l, x);
  require(r->prec == x->prec, "r and x must have the same precision");
  require(r->prec == P64 || r->prec == P32, "stagD2xx_noalloc: fields must be FP64 or FP32");
  require(par == 0 || par == 1, "stagD2xx parity must be 0 (ee) or 1 (oo)");
  require(r != x, "stagD2xx_noalloc: r and x must be different fields");
  const int pr = r->prec;
  // no packing here: the FP64 copy exists from qexhip_links_create, the FP32
  // one from the first FP32 operator call or solve on these links
  const char *lk = l->st[pr][0];
  require(lk != nullptr,
          "stagD2xx_noalloc: the FP32 copy of these links is not packed yet (any FP32 operator call "
          "or solve on them packs it)");
  const size_t need = (size_t)3 * l->lat->g.stride * word_bytes(pr);
  if (need > QEXHIP_SCRATCH_BYTES) {
    char msg[200];
    snprintf(msg, sizeof msg,
             "stagD2xx_noalloc: the lattice needs %zu bytes of scratch, more than the %lu-byte static buffer "
             "(at most %lu sites per parity in FP64)",
             need, (unsigned long)QEXHIP_SCRATCH_BYTES, (unsigned long)(QEXHIP_SCRATCH_BYTES / 48 / 256 * 256));
    throw std::runtime_error(msg);
  }
  void *t = g_scratch[g_device];
  require(t != nullptr, "stagD2xx_noalloc: scratch buffer not set up");
  const Tuned tu = cached_hop_shape(l, pr, 18);
  // t = (2D)_qp x_p into the scratch ; r = 4 m2 x - (2D)_pq t.  The first hop
  // writes every site the second reads, so the scratch needs no clearing.
  hop_shaped(l, lk, tu, pr, 18, 1 - par, 0, 0, x->par(par), t, nullptr, nullptr, 0, 0, 1.0, 0, 0, nullptr);
  hop_shaped(l, lk, tu, pr, 18, par, 1, 0, t, r->par(par), x->par(par), nullptr, 4.0 * m2, 0, -1.0, 0, 0, nullptr);
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
  check_solver_params(l, sp);
  solve_xx(l, par, (double *)r->par(par), (const double *)x->par(par), m, sp);
  API_END
}

static int solve_host(qexhip_links *l, int par, double *r, const double *t, double m, qexhip_solver_params *sp) {
  API_BEGIN
  check_init();
  require(l && r && t && sp, "null argument");
  qexhip_field *fr = nullptr, *fx = nullptr;
  FieldGuard gr{fr}, gx{fx};
  if (qexhip_field_create(l->lat, P64, &fr) || qexhip_field_create(l->lat, P64, &fx))
    throw std::runtime_error(g_err);
  import_field(fx, par, t, 0);
  if (qexhip_solve_xx(l, par, fr, fx, m, sp)) throw std::runtime_error(g_err);
  export_field(fr, par, r, 0);
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
  require(nshift >= 1 && nshift <= QH_MAXSHIFT, "multishift: nshift out of range");
  if (sp->sloppy == P64) {
    // cgm.nim's FP64 recursion as the batched solve with one right-hand side:
    // the shift updates are deferred to one pass per window of iterations, and
    // every shift is updated until the base system converges, as cgm.nim does
    qexhip_field *bp = const_cast<qexhip_field *>(b);
    solve_msb(l, par, 1, xs, &bp, m, sigmas, nshift, sp, QEXHIP_BATCH_GRAPH | QEXHIP_BATCH_TUNE | QEXHIP_BATCH_NOFREEZE);
  } else {
    solve_ms_mixed(l, par, xp.data(), (const double *)b->par(par), m, sigmas, nshift, sp);
  }
  API_END
}

int qexhip_force_solve(qexhip_links *l, qexhip_field *psi, const qexhip_field *phi, double m,
                       qexhip_solver_params *sp) {
  API_BEGIN
  check_init();
  check_same_lat(l, psi);
  check_same_lat(l, phi);
  require(psi->prec == P64 && phi->prec == P64, "force fields must be FP64");
  check_solver_params(l, sp);
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

int qexhip_outer_acc(qexhip_links *l, const qexhip_field *a, const qexhip_field *b, double ce, double co, double *F) {
  API_BEGIN
  check_init();
  check_same_lat(l, a);
  check_same_lat(l, b);
  require(a->prec == P64 && b->prec == P64, "outer products need FP64 fields");
  require(F != nullptr, "null output");
  launch_outer_acc(l->lat->g, (const double *)a->par(0), (const double *)a->par(1), (const double *)b->par(0),
                   (const double *)b->par(1), ce, co, F, l->nl, g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_force(qexhip_links *l, const qexhip_field *psi, double *F, int F_on_device, int outer_only) {
  API_BEGIN
  check_init();
  check_same_lat(l, psi);
  require(l->nl == 4 || outer_only, "qexhip_force: HISQ (nl=8) links give outer products only (outer_only=1)");
  require(psi->prec == P64, "force fields must be FP64");
  require(F != nullptr, "null output");
  qexhip_lattice *lat = l->lat;
  const size_t bytes = (size_t)l->nl * lat->V * 18 * sizeof(double);
  double *dF = F;
  if (!F_on_device) dF = dalloc<double>(bytes);
  launch_force(lat->g, links_store(l, P64, 18), lkblk_bytes(l, P64, 18), (const double *)psi->par(0),
               (const double *)psi->par(1), dF, outer_only, g_stream, l->nl);
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

static void fat7_array(const qexhip_fat7_coefs *c, double out[5]) {
  out[0] = c->one_link;
  out[1] = c->three_staple;
  out[2] = c->five_staple;
  out[3] = c->seven_staple;
  out[4] = c->lepage;
}

static void hisq_fat7(qexhip_fat7_coefs *c, double f7lf, double naik) {
  c->one_link = (1.0 + 3.0 * f7lf + naik) / 8.0;
  c->three_staple = -1.0 / 16.0;
  c->five_staple = 1.0 / 64.0;
  c->seven_staple = -1.0 / 384.0;
  c->lepage = -f7lf / 16.0;
}

void qexhip_hisq_coefs_init(qexhip_hisq_coefs *c) {
  if (!c) return;
  hisq_fat7(&c->fat7first, 0.0, 0.0);
  hisq_fat7(&c->fat7second, 2.0, 1.0);
  c->naik = -1.0 / 24.0;
}

int qexhip_make_imp_links(qexhip_lattice *lat, double *fl, const double *gf, const qexhip_fat7_coefs *coef,
                          double *ll, const double *gf_long, double naik) {
  API_BEGIN
  check_init();
  require(lat && fl && gf && coef, "null argument");
  require(fl != gf && (!ll || (ll != gf && ll != fl)), "outputs must not alias the input links");
  if (ll && naik != 0.0)
    for (int d = 0; d < 4; d++) require(lat->g.L[d] >= 3, "long links need extents >= 3");
  double c[5];
  fat7_array(coef, c);
  launch_make_imp_links(lat->g, lat->smear, fl, gf, c, ll, gf_long, naik, g_remap | (g_smear_walk ? 2 : 0), g_stream, nullptr);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_project_u(long n, double *u, const double *x) {
  API_BEGIN
  check_init();
  require(u && x && n >= 0, "bad arguments");
  if (n > 0) launch_project_u(n, u, x, g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_hisq_smear(qexhip_lattice *lat, const qexhip_hisq_coefs *c, const double *g, double *fl, double *ll) {
  API_BEGIN
  check_init();
  require(lat && g && fl && ll, "null argument");
  require(fl != g && ll != g && fl != ll, "outputs must not alias the input links");
  for (int d = 0; d < 4; d++) require(lat->g.L[d] >= 3, "HISQ links need extents >= 3");
  qexhip_hisq_coefs def;
  if (!c) {
    qexhip_hisq_coefs_init(&def);
    c = &def;
  }
  double c1[5], c2[5];
  fat7_array(&c->fat7first, c1);
  fat7_array(&c->fat7second, c2);
  launch_hisq_smear(lat->g, lat->smear, c1, c2, c->naik, g, fl, ll, g_remap | (g_smear_walk ? 2 : 0), g_stream, nullptr);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_fat7l_deriv(qexhip_lattice *lat, double *deriv, const double *gauge, const double *mid,
                       const qexhip_fat7_coefs *coef, double *llderiv, const double *llgauge, const double *llmid,
                       double naik) {
  API_BEGIN
  check_init();
  require(lat && deriv && gauge && mid && coef, "null argument");
  require(deriv != gauge && deriv != mid, "deriv must not alias gauge or mid");
  if (llderiv && llmid && naik != 0.0)
    for (int d = 0; d < 4; d++) require(lat->g.L[d] >= 3, "long links need extents >= 3");
  double c[5];
  fat7_array(coef, c);
  launch_fat7l_deriv(lat->g, lat->smear, deriv, gauge, mid, c, llderiv, llgauge, llmid, naik, g_remap | (g_smear_walk ? 2 : 0), g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_project_u_deriv(long n, double *r, const double *u, const double *x, const double *chain) {
  API_BEGIN
  check_init();
  require(r && u && x && chain && n >= 0, "bad arguments");
  if (n > 0) launch_project_u_deriv(n, r, u, x, chain, g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

static void hisq_arrays(const qexhip_hisq_coefs *c, double c1[5], double c2[5], double *naik) {
  qexhip_hisq_coefs def;
  if (!c) {
    qexhip_hisq_coefs_init(&def);
    c = &def;
  }
  fat7_array(&c->fat7first, c1);
  fat7_array(&c->fat7second, c2);
  *naik = c->naik;
}

int qexhip_hisq_smear_deriv(qexhip_lattice *lat, const qexhip_hisq_coefs *c, const double *g, const double *dfl,
                            const double *dll, double *dg) {
  API_BEGIN
  check_init();
  require(lat && g && dfl && dll && dg, "null argument");
  require(dg != g && dg != dfl && dg != dll, "dg must not alias the inputs");
  for (int d = 0; d < 4; d++) require(lat->g.L[d] >= 3, "HISQ links need extents >= 3");
  double c1[5], c2[5], naik;
  hisq_arrays(c, c1, c2, &naik);
  launch_hisq_smear_deriv(lat->g, lat->smear, c1, c2, naik, g, dfl, dll, lat->V, dg, 1, g_remap | (g_smear_walk ? 2 : 0), g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_hisq_force(qexhip_links *l, const qexhip_hisq_coefs *c, const double *g, const qexhip_field *psi,
                      double *F) {
  API_BEGIN
  check_init();
  check_same_lat(l, psi);
  require(l->nl == 8, "qexhip_hisq_force: HISQ (nl=8) links required");
  require(psi->prec == P64, "force fields must be FP64");
  require(g && F && F != g, "bad arguments");
  qexhip_lattice *lat = l->lat;
  double c1[5], c2[5], naik;
  hisq_arrays(c, c1, c2, &naik);
  // eps psi psi(x+mu)^+ at [2mu], eps psi psi(x+3mu)^+ at [2mu+1]: twice dS/d(fat, long)
  double *outer = smear_force_outer(lat->smear, lat->V);
  launch_force(lat->g, links_store(l, P64, 18), lkblk_bytes(l, P64, 18), (const double *)psi->par(0),
               (const double *)psi->par(1), outer, 1, g_stream, 8);
  // F = 2 dS/dg, through both smearing levels; then F <- TAH(F g^+)
  launch_hisq_smear_deriv(lat->g, lat->smear, c1, c2, naik, g, outer, outer + (size_t)lat->V * 18, 2 * lat->V, F, 0,
                          g_remap | (g_smear_walk ? 2 : 0), g_stream);
  launch_tah_mul(lat->V, F, F, g, g_stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_time_hisq_smear(qexhip_lattice *lat, const double *g, double *fl, double *ll, int nrep, double *ms) {
  API_BEGIN
  check_init();
  require(lat && g && fl && ll && ms && nrep > 0, "bad arguments");
  qexhip_hisq_coefs c;
  qexhip_hisq_coefs_init(&c);
  double c1[5], c2[5];
  fat7_array(&c.fat7first, c1);
  fat7_array(&c.fat7second, c2);
  SmearTimes tm[2];
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  for (auto &t : tm)
    for (auto &e : t.ev) HIPCHK(hipEventCreate(&e));
  launch_hisq_smear(lat->g, lat->smear, c1, c2, c.naik, g, fl, ll, g_remap | (g_smear_walk ? 2 : 0), g_stream, nullptr);
  HIPCHK(hipEventRecord(e0, g_stream));
  for (int i = 0; i < nrep; i++)
    launch_hisq_smear(lat->g, lat->smear, c1, c2, c.naik, g, fl, ll, g_remap | (g_smear_walk ? 2 : 0), g_stream, nullptr);
  HIPCHK(hipEventRecord(e1, g_stream));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  ms[0] = (double)t / nrep;
  for (int k = 1; k < 7; k++) ms[k] = 0.0;
  for (int i = 0; i < nrep; i++) {
    launch_hisq_smear(lat->g, lat->smear, c1, c2, c.naik, g, fl, ll, g_remap | (g_smear_walk ? 2 : 0), g_stream, tm);
    HIPCHK(hipStreamSynchronize(g_stream));
    for (int lv = 0; lv < 2; lv++)
      for (int k = 0; k < 3; k++) {
        HIPCHK(hipEventElapsedTime(&t, tm[lv].ev[k], tm[lv].ev[k + 1]));
        ms[1 + 3 * lv + k] += (double)t / nrep;
      }
  }
  check_launch();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (auto &tt : tm)
    for (auto &e : tt.ev) (void)hipEventDestroy(e);
  API_END
}

static void milc_info_out(const MilcInfo &m, qexhip_milc_info *info) {
  if (!info) return;
  for (int d = 0; d < 4; d++) info->dims[d] = m.dims[d];
  memcpy(info->timestamp, m.stamp, sizeof info->timestamp);
  info->order = m.order;
  info->byteswapped = m.byteswapped;
  info->sum29 = m.sum29;
  info->sum31 = m.sum31;
  info->checksum_ok = m.checksum_ok;
}

int qexhip_milc_read_info(const char *path, qexhip_milc_info *info) {
  API_BEGIN
  require(path && info, "null argument");
  MilcInfo m{};
  milc_read_header(path, &m);
  m.checksum_ok = 0;
  milc_info_out(m, info);
  API_END
}

int qexhip_milc_read(qexhip_lattice *lat, const char *path, double *U, qexhip_milc_info *info) {
  MilcInfo m{};
  API_BEGIN
  check_init();
  require(lat && path && U, "null argument");
  try {
    milc_read(lat->g, path, U, &m, g_stream);
  } catch (...) {
    milc_info_out(m, info);
    throw;
  }
  milc_info_out(m, info);
  API_END
}

int qexhip_milc_write(qexhip_lattice *lat, const char *path, const double *U, const char *timestamp) {
  API_BEGIN
  check_init();
  require(lat && path && U, "null argument");
  milc_write(lat->g, path, U, timestamp, g_stream);
  API_END
}

static void copy_str(char *dst, size_t cap, const std::string &s) {
  if (!dst || cap == 0) return;
  const size_t n = s.size() < cap - 1 ? s.size() : cap - 1;
  memcpy(dst, s.data(), n);
  dst[n] = 0;
}

static void scidac_info_out(const ScidacInfo &m, qexhip_scidac_info *info) {
  if (!info) return;
  memset(info, 0, sizeof *info);
  for (int d = 0; d < 4; d++) info->dims[d] = m.dims[d];
  info->nrecords = m.nrecords;
  copy_str(info->datatype, sizeof info->datatype, m.datatype);
  copy_str(info->date, sizeof info->date, m.date);
  info->precision = m.prec;
  info->colors = m.colors;
  info->spins = m.spins;
  info->typesize = m.typesize;
  info->datacount = m.datacount;
  info->ildg = m.ildg;
  info->has_checksum = m.has_checksum;
  info->suma = m.suma;
  info->sumb = m.sumb;
  info->calc_suma = m.calc_suma;
  info->calc_sumb = m.calc_sumb;
  info->checksum_ok = m.checksum_ok;
  info->filemd_len = (long long)m.filemd.size();
  info->recmd_len = (long long)m.recmd.size();
}

int qexhip_scidac_read_info(const char *path, int record, qexhip_scidac_info *info, char *filemd,
                            long long filemd_cap, char *recmd, long long recmd_cap) {
  API_BEGIN
  require(path && info, "null argument");
  require(filemd_cap >= 0 && recmd_cap >= 0, "negative buffer capacity");
  ScidacInfo m;
  scidac_info(path, record, &m);
  scidac_info_out(m, info);
  copy_str(filemd, (size_t)filemd_cap, m.filemd);
  copy_str(recmd, (size_t)recmd_cap, m.recmd);
  API_END
}

int qexhip_scidac_read(qexhip_lattice *lat, const char *path, int record, double *data, int nobj, int nelem,
                       qexhip_scidac_info *info) {
  ScidacInfo m;
  API_BEGIN
  check_init();
  require(lat && path && data, "null argument");
  try {
    scidac_read(lat->g, path, record, data, nobj, nelem, &m, g_stream);
  } catch (...) {
    scidac_info_out(m, info);
    throw;
  }
  scidac_info_out(m, info);
  API_END
}

int qexhip_scidac_create(const char *path, const int dims[4], const char *filemd) {
  API_BEGIN
  require(path && dims, "null argument");
  scidac_create(path, dims, filemd);
  API_END
}

int qexhip_scidac_append(qexhip_lattice *lat, const char *path, const char *recmd, const double *data, int nobj,
                         int nelem, char precision, int ildg, const char *date, qexhip_scidac_info *info) {
  API_BEGIN
  check_init();
  require(lat && path && data, "null argument");
  ScidacInfo m;
  scidac_append(lat->g, path, recmd, data, nobj, nelem, precision, ildg, date, &m, g_stream);
  scidac_info_out(m, info);
  API_END
}

int qexhip_time_scidac(qexhip_lattice *lat, const double *data, int nobj, int nelem, char precision, int nrep,
                       double *ms) {
  API_BEGIN
  check_init();
  require(lat && data && ms, "null argument");
  scidac_time(lat->g, data, nobj, nelem, precision, nrep, ms, g_stream);
  API_END
}

int qexhip_from_qex_layout(qexhip_lattice *lat, const int inner_geom[4], const double *qex_dev, int nfields,
                           int ncomp, double *canon_dev) {
  API_BEGIN
  check_init();
  require(lat && inner_geom && qex_dev && canon_dev, "null argument");
  qex_layout_convert(lat->g, inner_geom, 1, const_cast<double *>(qex_dev), nfields, ncomp, canon_dev, g_stream);
  HIPCHK(hipStreamSynchronize(g_stream));
  API_END
}

int qexhip_to_qex_layout(qexhip_lattice *lat, const int inner_geom[4], const double *canon_dev, int nfields,
                         int ncomp, double *qex_dev) {
  API_BEGIN
  check_init();
  require(lat && inner_geom && qex_dev && canon_dev, "null argument");
  qex_layout_convert(lat->g, inner_geom, 0, qex_dev, nfields, ncomp, const_cast<double *>(canon_dev), g_stream);
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

int qexhip_solve_batch(int n, qexhip_links *const *links, int par, qexhip_field *const *x,
                       qexhip_field *const *b, const double *masses, qexhip_solver_params *sp, int flags) {
  API_BEGIN
  check_init();
  require(n >= 1 && n <= QH_MAXBATCH, "solve_batch: n out of range");
  require(links && x && b && masses && sp, "null argument");
  require(par == 0 || par == 1, "solve_batch: parity must be 0 or 1");
  for (int k = 0; k < n; k++) {
    require(links[k] && links[k]->nl == links[0]->nl, "solve_batch: all systems need the same link count");
    check_same_lat(links[k], x[k]);
    check_same_lat(links[k], b[k]);
    require(x[k]->prec == P64 && b[k]->prec == P64, "solve_batch fields must be FP64");
  }
  solve_batch(n, links, par, x, b, masses, sp, flags);
  API_END
}

int qexhip_solve_multishift_batch(qexhip_links *l, int par, int nrhs, qexhip_field *const *xs,
                                  qexhip_field *const *b, double m, const double *sigmas, int nshift,
                                  qexhip_solver_params *sp, int flags) {
  API_BEGIN
  check_init();
  require(l && xs && b && sigmas && sp, "null argument");
  require(par == 0 || par == 1, "multishift parity must be 0 or 1");
  require(nrhs >= 1 && nrhs <= 64, "multishift batch: nrhs out of range");
  require(nshift >= 1 && nshift <= QH_MAXSHIFT, "multishift: nshift out of range");
  for (int j = 0; j < nrhs; j++) {
    check_same_lat(l, b[j]);
    require(b[j]->prec == P64, "multishift fields must be FP64");
  }
  for (int k = 0; k < nrhs * nshift; k++) {
    check_same_lat(l, xs[k]);
    require(xs[k]->prec == P64, "multishift fields must be FP64");
  }
  solve_msb(l, par, nrhs, xs, b, m, sigmas, nshift, sp, flags);
  API_END
}

int qexhip_set_autotune(int on) {
  g_cfg_gen++;
  g_autotune = on != 0;
  return 0;
}

int qexhip_set_cg_block1(int block1) {
  API_BEGIN
  g_cfg_gen++;
  require(block1 == 0 || block1 == 64 || block1 == 128 || block1 == 256, "block1 must be 0 (tuner), 64, 128 or 256");
  g_cg_block1 = block1;
  API_END
}

int qexhip_set_cg_cap(int cap) {
  API_BEGIN
  g_cfg_gen++;
  require(cap >= -1 && cap <= (1 << 20), "cg cap must be -1 (tuner), 0 or a workgroup count up to 2^20");
  g_cg_cap = cap;
  API_END
}

int qexhip_set_link_cache(qexhip_lattice *lat, long bytes) {
  API_BEGIN
  g_cfg_gen++;
  require(bytes >= -1, "link cache bytes must be -1 (default) or >= 0");
  if (lat) {
    lat->g.lkcache = bytes < 0 ? g_link_cache : bytes;
  } else {
    g_link_cache = bytes < 0 ? kLinkCacheDefault : bytes;
  }
  API_END
}

int qexhip_set_hop_walk(long chunk) {
  API_BEGIN
  g_cfg_gen++;
  require(chunk == -1 || (chunk >= 0 && chunk % 256 == 0), "site-walk chunk must be -1 (tuner), 0 or a multiple of 256");
  g_hop_chunk = chunk;
  API_END
}

int qexhip_set_hop_split(int split) {
  API_BEGIN
  g_cfg_gen++;
  require(split == -1 || split == 0 || split == 1 || split == 2 || split == 4,
          "hop threads per site must be -1 (tuner), 0 or 1 (one), 2 or 4");
  g_hop_split = split;
  API_END
}

int qexhip_set_hop_two(int two) {
  API_BEGIN
  g_cfg_gen++;
  require(two >= -1 && two <= 1, "hop stencil trip must be -1 (tuner), 0 (one link) or 1 (both links of a hop)");
  g_hop_two = two;
  API_END
}

int qexhip_set_hop_tmarch(int slices) {
  API_BEGIN
  g_cfg_gen++;
  require(slices >= -1 && slices <= 1024, "t-march slices per thread must be -1 (tuner), 0 (off) or 1..1024");
  g_hop_tmarch = slices;
  API_END
}

int qexhip_set_msb_walk(long chunk) {
  API_BEGIN
  g_cfg_gen++;
  require(chunk == -1 || (chunk >= 0 && chunk % 256 == 0), "site-walk chunk must be -1 (tuner), 0 or a multiple of 256");
  g_msb_walk = chunk;
  API_END
}

int qexhip_set_hop_multi_shape(int batches, long chunk) {
  API_BEGIN
  g_cfg_gen++;
  require(batches >= 0 && batches <= 2, "vector batches must be 0 (tuner), 1 or 2");
  require(chunk == -1 || (chunk >= 0 && chunk % 256 == 0), "site-walk chunk must be -1 (tuner), 0 or a multiple of 256");
  g_multi_batches = batches;
  g_multi_chunk = chunk;
  API_END
}

int qexhip_set_cg_graph(int on) {
  API_BEGIN
  g_cfg_gen++;
  require(on >= 0 && on <= 2, "cg graph mode must be 0, 1 or 2");
  g_cg_graph = on != 0;
  g_cg_intervals = on == 1;
  API_END
}

int qexhip_set_cg_form(int form) {
  API_BEGIN
  g_cfg_gen++;
  require(form == 0 || (form >= 2 && form <= 4), "cg form must be 0 (tuned), 2, 3 or 4");
  g_cg_form = form;
  API_END
}

int qexhip_autotune_report(char *buf, int len) {
  API_BEGIN
  require(buf && len > 0, "bad buffer");
  std::string out;
  for (auto &kv : g_tune) {
    // the key can be long (a batch lists every lattice): it is appended whole
    char line[256];
    snprintf(line, sizeof line,
             ": block=%d remap=%d form=%d cap=%d block1=%d batches=%d chunk=%ld tmarch=%d two=%d split=%d "
             "lkcache=%ld us=%.2f\n",
             kv.second.block, kv.second.remap, kv.second.form, kv.second.cap, kv.second.block1, kv.second.batches,
             kv.second.chunk, kv.second.tmarch, kv.second.two, kv.second.split, kv.second.lkcache >> 20,
             kv.second.us);
    out += kv.first;
    out += line;
  }
  if (out.size() >= (size_t)len)
    throw std::runtime_error("autotune report buffer too small: " + std::to_string(out.size() + 1) + " bytes needed");
  snprintf(buf, (size_t)len, "%s", out.c_str());
  API_END
}

int qexhip_time_hop(qexhip_links *l, int prec, int recon, int nrep, double *ms) {
  API_BEGIN
  check_init();
  require(l && ms && nrep > 0, "bad arguments");
  require(prec >= 0 && prec <= PQ16, "bad precision");
  qexhip_lattice *lat = l->lat;
  qexhip_field *x = nullptr, *r = nullptr;
  FieldGuard gx{x}, gr{r};
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
  API_END
}

}  // extern "C"
