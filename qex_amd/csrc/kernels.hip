// Device kernels of the qexhip library besides the stencil: layout conversion,
// vector kernels with fused FP64 reductions, the reliable-update helpers, the
// multi-shift CG updates and the fermion-force outer product.
#include <type_traits>
#include <algorithm>
#include <stdio.h>
#include "internal.h"

namespace qh {

// ---------------------------------------------------------- layout packing
template <int PO>
__global__ void k_pack_vec(Geom g, const double2 *canon, void *out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  using A = typename std::conditional<PO == P64, double2, float2>::type;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double2 d = canon[3 * i + c];
    A v;
    v.x = d.x; v.y = d.y;
    stc<PO, A>(out, vix(g, c, i), v);
  }
}

template <int PI>
__global__ void k_unpack_vec(Geom g, const void *in, double2 *canon) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (PI == P64) {
      canon[3 * i + c] = ldc<P64, double2>(in, vix(g, c, i));
    } else {
      float2 f = ldc<PI, float2>(in, vix(g, c, i));
      canon[3 * i + c] = make_double2(f.x, f.y);
    }
  }
}

static inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_pack_vec(int prec, const Geom &g, const double *canon, void *out, hipStream_t st) {
  const int b = 256;
  const double2 *c = (const double2 *)canon;
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_pack_vec<P64>, dim3(nblk(g.Vh, b)), dim3(b), 0, st, g, c, out); break;
    case P32: hipLaunchKernelGGL(k_pack_vec<P32>, dim3(nblk(g.Vh, b)), dim3(b), 0, st, g, c, out); break;
    case P16: hipLaunchKernelGGL(k_pack_vec<P16>, dim3(nblk(g.Vh, b)), dim3(b), 0, st, g, c, out); break;
    case PBF: hipLaunchKernelGGL(k_pack_vec<PBF>, dim3(nblk(g.Vh, b)), dim3(b), 0, st, g, c, out); break;
  }
}

void launch_unpack_vec(int prec, const Geom &g, const void *in, double *canon, hipStream_t st) {
  const int b = 256;
  double2 *c = (double2 *)canon;
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_unpack_vec<P64>, dim3(nblk(g.Vh, b)), dim3(b), 0, st, g, in, c); break;
    case P32: hipLaunchKernelGGL(k_unpack_vec<P32>, dim3(nblk(g.Vh, b)), dim3(b), 0, st, g, in, c); break;
    case P16: hipLaunchKernelGGL(k_unpack_vec<P16>, dim3(nblk(g.Vh, b)), dim3(b), 0, st, g, in, c); break;
    case PBF: hipLaunchKernelGGL(k_unpack_vec<PBF>, dim3(nblk(g.Vh, b)), dim3(b), 0, st, g, in, c); break;
  }
}

// -------------------------------------------------------------- fill / blas
__global__ void k_fill(uint4 *p, long n16) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (long)gridDim.x * blockDim.x)
    p[i] = make_uint4(0, 0, 0, 0);
}

void launch_fill(void *p, size_t bytes, hipStream_t st) {
  // all library allocations are multiples of 16 bytes
  long n16 = (long)(bytes / 16);
  if (n16 == 0) return;
  unsigned nb = (unsigned)std::min<long>(nblk(n16, 256), 8192);
  hipLaunchKernelGGL(k_fill, dim3(nb), dim3(256), 0, st, (uint4 *)p, n16);
}

// Streaming-read reference: grid-stride non-temporal 16-byte loads, one store
// per thread (out: gridDim.x*blockDim.x doubles).  Measures the read bandwidth
// a kernel can reach on this device, the practical ceiling of the hop.
__global__ void __launch_bounds__(256) k_stream_read(const double2 *in, long n16, double *out) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  const d2v *p = (const d2v *)in;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  d2v acc = {0.0, 0.0};
  for (long i = t; i < n16; i += (long)gridDim.x * blockDim.x) acc += __builtin_nontemporal_load(p + i);
  out[t] = acc.x + acc.y;
}

// non-trivial bit patterns (a 64-bit multiplicative hash of the index)
__global__ void k_fill_pattern(uint64_t *p, long n8) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n8; i += (long)gridDim.x * blockDim.x)
    p[i] = ((uint64_t)i + 1) * 0x9E3779B97F4A7C15ull;
}

void launch_fill_pattern(void *p, size_t bytes, hipStream_t st) {
  const long n8 = (long)(bytes / 8);
  if (n8 == 0) return;
  hipLaunchKernelGGL(k_fill_pattern, dim3(4096), dim3(256), 0, st, (uint64_t *)p, n8);
}

void launch_stream_read(const void *in, long n16, double *out, int blocks, hipStream_t st) {
  hipLaunchKernelGGL(k_stream_read, dim3(blocks), dim3(256), 0, st, (const double2 *)in, n16, out);
}

__global__ void k_axpby64(Geom g, double a, const double2 *x, double b, double2 *y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double2 xv = x[vix(g, c, i)], yv = y[vix(g, c, i)];
    y[vix(g, c, i)] = make_double2(a * xv.x + b * yv.x, a * xv.y + b * yv.y);
  }
}

void launch_axpby64(const Geom &g, double a, const double *x, double b, double *y, hipStream_t st) {
  hipLaunchKernelGGL(k_axpby64, dim3(nblk(g.Vh, 256)), dim3(256), 0, st, g, a, (const double2 *)x, b,
                     (double2 *)y);
}

template <int P, int MODE>  // MODE 0: |x|^2, 1: Re<x,y>
__global__ void __launch_bounds__(256) k_reduce(Geom g, const void *x, const void *y, Scal *S, int slot,
                                                RedWs red) {
  __shared__ double lds[16];
  using A = typename std::conditional<P == P64, double2, float2>::type;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < g.Vh) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      A a = ldc<P, A>(x, vix(g, c, i));
      if (MODE == 0) {
        s += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
      } else {
        A b = ldc<P, A>(y, vix(g, c, i));
        s += (double)a.x * (double)b.x + (double)a.y * (double)b.y;
      }
    }
  }
  double v[1] = {s};
  block_partials<1>(v, lds, red);
}

static void finish_tmp(unsigned nb, Scal *S, int slot, RedWs red, hipStream_t st) {
  FinArgs f{red, (int)nb, S, slot, 0.0, 0};
  launch_finish(EPI_TMP, f, st);
}

void launch_norm2(int prec, const Geom &g, const void *x, Scal *S, int slot, RedWs red, hipStream_t st) {
  dim3 grid(nblk(g.Vh, 256));
  switch (prec) {
    case P64: hipLaunchKernelGGL((k_reduce<P64, 0>), grid, dim3(256), 0, st, g, x, x, S, slot, red); break;
    case P32: hipLaunchKernelGGL((k_reduce<P32, 0>), grid, dim3(256), 0, st, g, x, x, S, slot, red); break;
    case P16: hipLaunchKernelGGL((k_reduce<P16, 0>), grid, dim3(256), 0, st, g, x, x, S, slot, red); break;
    case PBF: hipLaunchKernelGGL((k_reduce<PBF, 0>), grid, dim3(256), 0, st, g, x, x, S, slot, red); break;
  }
  finish_tmp(grid.x, S, slot, red, st);
}

void launch_redot(int prec, const Geom &g, const void *x, const void *y, Scal *S, int slot, RedWs red,
                  hipStream_t st) {
  dim3 grid(nblk(g.Vh, 256));
  switch (prec) {
    case P64: hipLaunchKernelGGL((k_reduce<P64, 1>), grid, dim3(256), 0, st, g, x, y, S, slot, red); break;
    case P32: hipLaunchKernelGGL((k_reduce<P32, 1>), grid, dim3(256), 0, st, g, x, y, S, slot, red); break;
    case P16: hipLaunchKernelGGL((k_reduce<P16, 1>), grid, dim3(256), 0, st, g, x, y, S, slot, red); break;
    case PBF: hipLaunchKernelGGL((k_reduce<PBF, 1>), grid, dim3(256), 0, st, g, x, y, S, slot, red); break;
  }
  finish_tmp(grid.x, S, slot, red, st);
}

// reliable update, part 1: x += xs/sigma ; xs = 0   (FP64 solution accumulation;
// xs stored in PX = FP32 or FP16)
template <int PX>
__global__ void k_xfold(Geom g, double2 *x, void *xs, Scal *S) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  const double is = 1.0 / S->sigma;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const long o = vix(g, c, i);
    const float2 s = ldc<PX, float2>(xs, o);
    double2 v = x[o];
    v.x += is * (double)s.x;
    v.y += is * (double)s.y;
    x[o] = v;
    stc<PX, float2>(xs, o, make_float2(0.f, 0.f));
  }
}

// reliable update, last part: rs = sigma * r_true (sloppy copy of the true residual,
// stored in PR = FP32 or FP16)
template <int PR>
__global__ void k_rs_from_true(Geom g, void *rs, const double2 *rt, Scal *S) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  const double sg = S->sigma;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const long o = vix(g, c, i);
    double2 v = rt[o];
    stc<PR, float2>(rs, o, make_float2((float)(sg * v.x), (float)(sg * v.y)));
  }
}

void launch_rs_from_true(const Geom &g, void *rs, const double *rtrue, Scal *S, int pr, hipStream_t st) {
  if (pr == P16)
    hipLaunchKernelGGL(k_rs_from_true<P16>, dim3(nblk(g.Vh, 256)), dim3(256), 0, st, g, rs, (const double2 *)rtrue,
                       S);
  else
    hipLaunchKernelGGL(k_rs_from_true<P32>, dim3(nblk(g.Vh, 256)), dim3(256), 0, st, g, rs, (const double2 *)rtrue,
                       S);
}

void launch_xfold(const Geom &g, double *x, void *xs, Scal *S, int px, hipStream_t st) {
  if (px == P16)
    hipLaunchKernelGGL(k_xfold<P16>, dim3(nblk(g.Vh, 256)), dim3(256), 0, st, g, (double2 *)x, xs, S);
  else
    hipLaunchKernelGGL(k_xfold<P32>, dim3(nblk(g.Vh, 256)), dim3(256), 0, st, g, (double2 *)x, xs, S);
}

// reliable update, after the true residual: best = x when epi_reliable found
// the residual no larger than the best one so far (and did not roll back);
// the device-side form of the mixed CG's best-iterate copy, for a reliable
// update that runs inside a captured graph
__global__ void k_keep_best(Geom g, double2 *best, const double2 *x, const Scal *S) {
  if (S->rollback || !S->improved) return;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const long o = vix(g, c, i);
    best[o] = x[o];
  }
}

void launch_keep_best(const Geom &g, double *best, const double *x, const Scal *S, hipStream_t st) {
  hipLaunchKernelGGL(k_keep_best, dim3(nblk(g.Vh, 256)), dim3(256), 0, st, g, (double2 *)best, (const double2 *)x,
                     S);
}

__global__ void k_scal_init(Scal *S, int mode) {
  if (threadIdx.x != 0) return;
  if (mode == 0) {  // start of a CG solve: fresh state (cg.nim:21-27)
    S->its = 0; S->since = 0; S->done = 0; S->need_rel = 0; S->first = 1; S->nrel = 0;
    S->r2prev = 1.0; S->beta = 0.0; S->alpha = 0.0; S->pscale = 1.0; S->sigma = 1.0;
    S->restart = 0; S->rollback = 0; S->resinc = 0; S->improved = 0; S->r2rel_last = __builtin_inf();
  } else if (mode == 1) {  // force a reliable update now
    S->need_rel = 1;
  }
}

void launch_scal_init(Scal *S, hipStream_t st, int mode) {
  hipLaunchKernelGGL(k_scal_init, dim3(1), dim3(64), 0, st, S, mode);
}

// ------------------------------------------------------------- multi-shift
// Loads and stores of one complex entry of an FP64 (double2) or FP32
// (float2) parity vector, computing in FP64.
__device__ __forceinline__ double2 ld2d(const double2 *p, long o) { return p[o]; }
__device__ __forceinline__ double2 ld2d(const float2 *p, long o) {
  const float2 v = p[o];
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st2d(double2 *p, long o, double2 v) { p[o] = v; }
__device__ __forceinline__ void st2d(float2 *p, long o, double2 v) { p[o] = make_float2((float)v.x, (float)v.y); }

// Start of the multi-shift recursion on FP32 storage: r = b, p_k = b, x_k = 0
// (cgm.nim:122-158)
__global__ void k_cgm_start32(Geom g, int nshift, const double2 *b, float2 *r, float2 *const *x, float2 *const *p) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const long o = vix(g, c, i);
    const double2 bv = b[o];
    st2d(r, o, bv);
    for (int k = 0; k < nshift; k++) {
      st2d(p[k], o, bv);
      x[k][o] = make_float2(0.f, 0.f);
    }
  }
}

// FP64 copies of the FP32 shifted solutions
__global__ void k_cgm_export32(Geom g, int nshift, float2 *const *x32, double2 *const *x64) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const long o = vix(g, c, i);
    for (int k = 0; k < nshift; k++) x64[k][o] = ld2d(x32[k], o);
  }
}

void launch_cgm_start32(const Geom &g, int nshift, const double *b, void *r, void *const *x, void *const *p,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_cgm_start32, dim3(nblk(g.Vh, 256)), dim3(256), 0, st, g, nshift, (const double2 *)b,
                     (float2 *)r, (float2 *const *)x, (float2 *const *)p);
}

void launch_cgm_export32(const Geom &g, int nshift, void *const *x32, double *const *x64, hipStream_t st) {
  hipLaunchKernelGGL(k_cgm_export32, dim3(nblk(g.Vh, 256)), dim3(256), 0, st, g, nshift, (float2 *const *)x32,
                     (double2 *const *)x64);
}

// ------------------------------------------------------------------- force
// F_mu(x) = TAH(eps(x) psi(x) psi(x+mu)^+ U_mu(x)^+)   (hmc/hmcAction.nim:735-790,
// unsmeared; TAH = maths/matrixFunctions.nim:414-419).  One thread per
// (canonical site, mu); output in QEX canonical order [mu][site][3][3].
// nl = 8 (HISQ links, outer products only): output k = 2mu holds the 1-hop
// outer product, k = 2mu+1 the 3-hop one (both link x to a site of the other parity)
__global__ void __launch_bounds__(256) k_force(Geom g, const char *links, long lkblk, const double2 *pe,
                                               const double2 *po, double2 *F, int outer_only, int nl) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  const int mu = nl == 8 ? k >> 1 : k;
  if (idx >= V) return;
  const int par = idx >= g.Vh;
  const long i = idx - par * g.Vh;
  const Site s = decode(g, i, par);
  const long j = nbr(g, s, mu, (nl == 8 && (k & 1)) ? 3 : 1);
  const double2 *ps = par ? po : pe, *pn = par ? pe : po;
  const double eps = par ? -1.0 : 1.0;
  double2 a[3], b[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    a[c] = ps[vix(g, c, i)];
    b[c] = pn[vix(g, c, j)];
    a[c].x *= eps; a[c].y *= eps;
  }
  double2 G[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++)  // a_r conj(b_c)
      G[3 * r + c] = make_double2(a[r].x * b[c].x + a[r].y * b[c].y, a[r].y * b[c].x - a[r].x * b[c].y);
  double2 *out = F + ((long)k * V + idx) * 9;
  if (outer_only) {
#pragma unroll
    for (int e = 0; e < 9; e++) out[e] = G[e];
    return;
  }
  double2 U[9];
  ld_link<P64, 18, false, double2, link_blocked<P64, 4>()>(links + (long)(2 * mu + par) * lkblk, g.stride, i, U, 1.0);
  double2 T[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {  // (G U^+)_{rc} = sum_k G_rk conj(U_ck)
      double2 acc = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double2 gv = G[3 * r + k], uv = U[3 * c + k];
        acc.x += gv.x * uv.x + gv.y * uv.y;
        acc.y += gv.y * uv.x - gv.x * uv.y;
      }
      T[3 * r + c] = acc;
    }
  double2 Rm[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double2 x = T[3 * r + c], y = T[3 * c + r];
      Rm[3 * r + c] = make_double2(0.5 * (x.x - y.x), 0.5 * (x.y + y.y));
    }
  const double tri = (Rm[0].y + Rm[4].y + Rm[8].y) / 3.0;
  const double trr = (Rm[0].x + Rm[4].x + Rm[8].x) / 3.0;
  Rm[0].x -= trr; Rm[4].x -= trr; Rm[8].x -= trr;
  Rm[0].y -= tri; Rm[4].y -= tri; Rm[8].y -= tri;
#pragma unroll
  for (int e = 0; e < 9; e++) out[e] = Rm[e];
}

// F_k(y) += c(y) a(y) b(y + d_k)^+ on every site y, c = ce on even and co on
// odd sites, d_k the hop of link k (one site in direction mu; three for
// k = 2mu+1 when nl = 8): the outer-product accumulations of stagD2deriv and
// stagDeriv (physics/stagD.nim:640-715, peqOuter/meqOuter of
// maths/matrixConcept.nim:718-725)
__global__ void __launch_bounds__(256) k_outer_acc(Geom g, const double2 *ae, const double2 *ao, const double2 *be,
                                                   const double2 *bo, double ce, double co, double2 *F, int nl) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  const int mu = nl == 8 ? k >> 1 : k;
  if (idx >= V) return;
  const int par = idx >= g.Vh;
  const long i = idx - par * g.Vh;
  const Site s = decode(g, i, par);
  const int len = (nl == 8 && (k & 1)) ? 3 : 1;
  const long j = nbr(g, s, mu, len);
  // a 3-hop neighbour has the other parity too (odd displacement)
  const double2 *as = par ? ao : ae, *bn = par ? be : bo;
  const double c = par ? co : ce;
  double2 a[3], b[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    a[q] = as[vix(g, q, i)];
    b[q] = bn[vix(g, q, j)];
    a[q].x *= c;
    a[q].y *= c;
  }
  double2 *out = F + ((long)k * V + idx) * 9;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int q = 0; q < 3; q++) {  // a_r conj(b_q)
      double2 o = out[3 * r + q];
      o.x += a[r].x * b[q].x + a[r].y * b[q].y;
      o.y += a[r].y * b[q].x - a[r].x * b[q].y;
      out[3 * r + q] = o;
    }
}

void launch_outer_acc(const Geom &g, const double *a_e, const double *a_o, const double *b_e, const double *b_o,
                      double ce, double co, double *F, int nl, hipStream_t st) {
  dim3 grid(nblk(2 * g.Vh, 256), nl);
  hipLaunchKernelGGL(k_outer_acc, grid, dim3(256), 0, st, g, (const double2 *)a_e, (const double2 *)a_o,
                     (const double2 *)b_e, (const double2 *)b_o, ce, co, (double2 *)F, nl);
}

void launch_force(const Geom &g, const char *links64, long lkblk, const double *psi_e, const double *psi_o,
                  double *F, int outer_only, hipStream_t st, int nl) {
  dim3 grid(nblk(2 * g.Vh, 256), nl);
  hipLaunchKernelGGL(k_force, grid, dim3(256), 0, st, g, links64, lkblk, (const double2 *)psi_e,
                     (const double2 *)psi_o, (double2 *)F, outer_only, nl);
}

}  // namespace qh
