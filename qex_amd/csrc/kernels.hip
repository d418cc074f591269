// Device kernels of the qexhip library: the one-parity staggered hop (plain and
// HISQ-form), layout conversion, CG/CGM vector updates with fused FP64
// reductions, and the fermion-force outer product.
//
// The hop computes, for every site x of the output parity p,
//   H(x) = sum_k [ U_k(x) in(x + l_k mu_k) - U_k(x - l_k mu_k)^+ in(x - l_k mu_k) ]
// i.e. QEX's (2D) stencil (physics/stagD.nim:383-448; nl=4: l_k=1, mu_k=k;
// nl=8: k=2mu fat with l=1, k=2mu+1 long with l=3, stagD.nim:40-51), and
//   out = ca*A + cb*B + chop*H
// with optional FP64 reductions of the result fused into the epilogue.
#include <type_traits>
#include <algorithm>
#include <stdio.h>
#include "internal.h"

namespace qh {

// -------------------------------------------------------------- reductions
__device__ __forceinline__ double ld_coherent(const double *p) {
  unsigned long long u = __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)u);
}

// Block partials -> global partials; the last block to arrive sums all partials
// in a fixed order (deterministic) and returns true (result valid in thread 0).
template <int NV>
__device__ __forceinline__ bool grid_finish(double (&v)[NV], double *lds, const RedWs &red) {
  __shared__ int last;
  const int nb = gridDim.x * gridDim.y;
  const int bid = blockIdx.y * gridDim.x + blockIdx.x;
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) red.partials[(long)k * red.maxblocks + bid] = v[k];
    __threadfence();
    unsigned t = atomicAdd(red.counter, 1u);
    last = (t == (unsigned)(nb - 1));
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) s[k] = 0.0;
  for (int j = threadIdx.x; j < nb; j += blockDim.x)
#pragma unroll
    for (int k = 0; k < NV; k++) s[k] += ld_coherent(red.partials + (long)k * red.maxblocks + j);
  block_sum<NV>(s, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = s[k];
    atomicExch(red.counter, 0u);
  }
  return true;
}

// ---------------------------------------------------------------- epilogues
// after p.Ap (solvers/cg.nim:206-208): alpha = rz/pAp
__device__ void epi_alpha(Scal *S, double pAp) {
  S->pAp = pAp;
  S->alpha = S->r2 / pAp;
  S->first = 0;
  S->pscale = 1.0;
}

// after r2 = |r|^2 (cg.nim:209-213, loop test :174)
__device__ void epi_r2(Scal *S, double r2new) {
  S->r2old = S->r2;
  S->r2 = r2new;
  S->its += 1;
  S->since += 1;
  S->beta = r2new / S->r2old;
  if (S->mixed) {
    S->need_rel = (r2new < S->delta2 * S->maxr2) || (S->since >= S->relmax) ||
                  (r2new <= S->r2stop_s) || (S->its >= S->maxits);
  } else {
    S->done = (S->its >= S->maxits) || (r2new <= S->r2stop);
  }
}

// after the FP64 true residual r = b - A x (reliable update)
__device__ void epi_reliable(Scal *S, double s, double nreal) {
  S->r2true = s;
  const double sig_old = S->sigma;
  const double sig_new = (s > 0.0) ? sqrt(nreal / s) : 1.0;
  const double r2old_unscaled = S->r2old / (sig_old * sig_old);
  S->beta = (r2old_unscaled > 0.0) ? s / r2old_unscaled : 0.0;
  S->pscale = sig_new / sig_old;
  S->sigma = sig_new;
  S->r2 = s * sig_new * sig_new;
  S->r2stop_s = S->r2stop * sig_new * sig_new;
  S->maxr2 = S->r2;
  S->since = 0;
  S->need_rel = 0;
  S->nrel += 1;
  S->done = (s <= S->r2stop) || (S->its >= S->maxits);
}

__device__ __forceinline__ bool skip_pred(int pred, const Scal *S) {
  if (pred == 1) return (S->done | S->need_rel) != 0;
  if (pred == 2) return S->need_rel == 0;
  return false;
}

// -------------------------------------------------------------------- hop
template <int PL, class A>
__device__ __forceinline__ void load_link(const void *base, long off, long stride, A (&U)[9]) {
#pragma unroll
  for (int c = 0; c < 9; c++) U[c] = ldc<PL, A>(base, off + c * stride);
}

template <int PL, int PI, int PO, int PA, class A, int NL, int MODE, int RED>
__global__ void __launch_bounds__(256) k_hop(HopArgs P) {
  using R = typename cx<A>::R;
  __shared__ double lds[16 * 2];
  if (P.S && skip_pred(P.pred, P.S)) return;
  const Geom &g = P.g;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double red = 0.0;
  if (i < g.Vh) {
    const Site s = decode(g, i, P.par);
    A af[3], ab[3];
#pragma unroll
    for (int c = 0; c < 3; c++) { af[c] = mk<A>(0, 0); ab[c] = mk<A>(0, 0); }
#pragma unroll
    for (int k = 0; k < NL; k++) {
      const int mu = (NL == 4) ? k : (k >> 1);
      const int len = (NL == 8 && (k & 1)) ? 3 : 1;
      const long jf = nbr(g, s, mu, len);
      const long jb = nbr(g, s, mu, -len);
      A U[9], v[3];
      load_link<PL, A>(P.links, (long)(2 * k + P.par) * P.lkblk + i, g.stride, U);
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = ldc<PI, A>(P.in, c * g.stride + jf);
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) cmac(af[r], U[3 * r + c], v[c]);
      load_link<PL, A>(P.links, (long)(2 * k + 1 - P.par) * P.lkblk + jb, g.stride, U);
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = ldc<PI, A>(P.in, c * g.stride + jb);
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) cjmac(ab[r], U[3 * c + r], v[c]);
    }
    const R chop = (R)P.chop;
    A res[3];
#pragma unroll
    for (int c = 0; c < 3; c++) res[c] = mk<A>(chop * (af[c].x - ab[c].x), chop * (af[c].y - ab[c].y));
    A av[3];
    if (MODE >= 1) {
      double ca = P.ca;
      if (P.ca_m2) ca *= 4.0 * P.S->m2;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        av[c] = ldc<PA, A>(P.a, c * g.stride + i);
        res[c].x += (R)ca * av[c].x;
        res[c].y += (R)ca * av[c].y;
      }
    }
    if (MODE == 2) {
      const R cb = (R)P.cb;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        A bv = ldc<PA, A>(P.b, c * g.stride + i);
        res[c].x += cb * bv.x;
        res[c].y += cb * bv.y;
      }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) stc<PO, A>(P.out, c * g.stride + i, res[c]);
    if (RED == 1) {
#pragma unroll
      for (int c = 0; c < 3; c++) red += (double)av[c].x * (double)res[c].x + (double)av[c].y * (double)res[c].y;
    } else if (RED >= 2) {
#pragma unroll
      for (int c = 0; c < 3; c++) red += (double)res[c].x * (double)res[c].x + (double)res[c].y * (double)res[c].y;
    }
  }
  if (RED) {
    double v[1] = {red};
    if (grid_finish<1>(v, lds, P.red) && threadIdx.x == 0) {
      if (RED == 1) epi_alpha(P.S, v[0]);
      else if (RED == 2) epi_reliable(P.S, v[0], 6.0 * (double)g.Vh);
      else P.S->tmp[0] = v[0];
    }
  }
}

template <int PL, int PI, int PO, int PA, class A, int NL>
static void hop_nl(int mode, int redm, const HopArgs &a, hipStream_t st, int block) {
  dim3 grid((unsigned)((a.g.Vh + block - 1) / block));
#define QH_HOP(M, R) \
  if (mode == M && redm == R) { hipLaunchKernelGGL((k_hop<PL, PI, PO, PA, A, NL, M, R>), grid, dim3(block), 0, st, a); return; }
  QH_HOP(0, 0) QH_HOP(1, 0) QH_HOP(1, 1) QH_HOP(2, 0) QH_HOP(2, 2) QH_HOP(2, 3)
#undef QH_HOP
  fprintf(stderr, "qexhip: unsupported hop mode %d/%d\n", mode, redm);
}

template <int PL, int PI, int PO, int PA, class A>
static void hop_t(int nl, int mode, int redm, const HopArgs &a, hipStream_t st, int block) {
  if (nl == 4) hop_nl<PL, PI, PO, PA, A, 4>(mode, redm, a, st, block);
  else hop_nl<PL, PI, PO, PA, A, 8>(mode, redm, a, st, block);
}

void launch_hop(int PL, int PI, int PO, int PA, bool dbl, int nl, int mode, int redm, const HopArgs &a,
                hipStream_t st, int block) {
  const int key = (PL << 12) | (PI << 8) | (PO << 4) | PA;
#define K4(a_, b_, c_, d_) (((a_) << 12) | ((b_) << 8) | ((c_) << 4) | (d_))
  if (dbl) {
    if (key == K4(P64, P64, P64, P64)) return hop_t<P64, P64, P64, P64, double2>(nl, mode, redm, a, st, block);
  } else {
    switch (key) {
      case K4(P32, P32, P32, P32): return hop_t<P32, P32, P32, P32, float2>(nl, mode, redm, a, st, block);
      case K4(P16, P16, P16, P16): return hop_t<P16, P16, P16, P16, float2>(nl, mode, redm, a, st, block);
      case K4(P16, P16, P32, P16): return hop_t<P16, P16, P32, P16, float2>(nl, mode, redm, a, st, block);
      case K4(PBF, PBF, PBF, PBF): return hop_t<PBF, PBF, PBF, PBF, float2>(nl, mode, redm, a, st, block);
      case K4(PBF, PBF, P32, PBF): return hop_t<PBF, PBF, P32, PBF, float2>(nl, mode, redm, a, st, block);
      case K4(P16, P32, P32, P32): return hop_t<P16, P32, P32, P32, float2>(nl, mode, redm, a, st, block);
      case K4(PBF, P32, P32, P32): return hop_t<PBF, P32, P32, P32, float2>(nl, mode, redm, a, st, block);
      default: break;
    }
  }
#undef K4
  fprintf(stderr, "qexhip: unsupported hop precision combination %x (dbl=%d)\n", key, (int)dbl);
}

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
    stc<PO, A>(out, c * g.stride + i, v);
  }
}

template <int PI>
__global__ void k_unpack_vec(Geom g, const void *in, double2 *canon) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (PI == P64) {
      canon[3 * i + c] = ldc<P64, double2>(in, c * g.stride + i);
    } else {
      float2 f = ldc<PI, float2>(in, c * g.stride + i);
      canon[3 * i + c] = make_double2(f.x, f.y);
    }
  }
}

template <int PO>
__global__ void k_pack_links(Geom g, int nl, const double2 *canon, void *out) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (idx >= V) return;
  const int par = idx >= g.Vh;
  const long i = idx - par * g.Vh;
  using A = typename std::conditional<PO == P64, double2, float2>::type;
  const long base = (long)(2 * k + par) * 9 * g.stride + i;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    double2 d = canon[((long)k * V + idx) * 9 + e];
    A v;
    v.x = d.x; v.y = d.y;
    stc<PO, A>(out, base + e * g.stride, v);
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

void launch_pack_links(int prec, int recon, const Geom &g, int nl, const double *canon, void *out,
                       hipStream_t st) {
  (void)recon;
  const int b = 256;
  dim3 grid(nblk(2 * g.Vh, b), nl);
  const double2 *c = (const double2 *)canon;
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_pack_links<P64>, grid, dim3(b), 0, st, g, nl, c, out); break;
    case P32: hipLaunchKernelGGL(k_pack_links<P32>, grid, dim3(b), 0, st, g, nl, c, out); break;
    case P16: hipLaunchKernelGGL(k_pack_links<P16>, grid, dim3(b), 0, st, g, nl, c, out); break;
    case PBF: hipLaunchKernelGGL(k_pack_links<PBF>, grid, dim3(b), 0, st, g, nl, c, out); break;
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

__global__ void k_axpby64(Geom g, double a, const double2 *x, double b, double2 *y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double2 xv = x[c * g.stride + i], yv = y[c * g.stride + i];
    y[c * g.stride + i] = make_double2(a * xv.x + b * yv.x, a * xv.y + b * yv.y);
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
      A a = ldc<P, A>(x, c * g.stride + i);
      if (MODE == 0) {
        s += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
      } else {
        A b = ldc<P, A>(y, c * g.stride + i);
        s += (double)a.x * (double)b.x + (double)a.y * (double)b.y;
      }
    }
  }
  double v[1] = {s};
  if (grid_finish<1>(v, lds, red) && threadIdx.x == 0) S->tmp[slot] = v[0];
}

void launch_norm2(int prec, const Geom &g, const void *x, Scal *S, int slot, RedWs red, hipStream_t st) {
  dim3 grid(nblk(g.Vh, 256));
  switch (prec) {
    case P64: hipLaunchKernelGGL((k_reduce<P64, 0>), grid, dim3(256), 0, st, g, x, x, S, slot, red); break;
    case P32: hipLaunchKernelGGL((k_reduce<P32, 0>), grid, dim3(256), 0, st, g, x, x, S, slot, red); break;
    case P16: hipLaunchKernelGGL((k_reduce<P16, 0>), grid, dim3(256), 0, st, g, x, x, S, slot, red); break;
    case PBF: hipLaunchKernelGGL((k_reduce<PBF, 0>), grid, dim3(256), 0, st, g, x, x, S, slot, red); break;
  }
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
}

// p = r + beta*p (cg.nim:186-193; first iteration p = r)
template <int PP, int PR, class A>
__global__ void k_pupdate(Geom g, void *p, const void *r, Scal *S) {
  if ((S->done | S->need_rel) != 0) return;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  using R = typename cx<A>::R;
  if (S->first) {
#pragma unroll
    for (int c = 0; c < 3; c++) stc<PP, A>(p, c * g.stride + i, ldc<PR, A>(r, c * g.stride + i));
    return;
  }
  const R beta = (R)(S->beta * S->pscale);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    A rv = ldc<PR, A>(r, c * g.stride + i), pv = ldc<PP, A>(p, c * g.stride + i);
    pv.x = rv.x + beta * pv.x;
    pv.y = rv.y + beta * pv.y;
    stc<PP, A>(p, c * g.stride + i, pv);
  }
}

// x += alpha p ; r -= alpha Ap ; r2 = |r|^2  (cg.nim:208-213)
template <int PX, int PP, int PR, int PAp, class A>
__global__ void __launch_bounds__(256) k_xr(Geom g, void *x, const void *p, void *r, const void *Ap, Scal *S,
                                            RedWs red) {
  __shared__ double lds[16];
  if ((S->done | S->need_rel) != 0) return;
  using R = typename cx<A>::R;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < g.Vh) {
    const R alpha = (R)S->alpha;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long o = c * g.stride + i;
      A pv = ldc<PP, A>(p, o), xv = ldc<PX, A>(x, o), rv = ldc<PR, A>(r, o), av = ldc<PAp, A>(Ap, o);
      xv.x += alpha * pv.x; xv.y += alpha * pv.y;
      rv.x -= alpha * av.x; rv.y -= alpha * av.y;
      stc<PX, A>(x, o, xv);
      stc<PR, A>(r, o, rv);
      s += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
    }
  }
  double v[1] = {s};
  if (grid_finish<1>(v, lds, red) && threadIdx.x == 0) epi_r2(S, v[0]);
}

void launch_pupdate(int PP, int PR, const Geom &g, void *p, const void *r, Scal *S, hipStream_t st) {
  dim3 grid(nblk(g.Vh, 256)), blk(256);
  if (PP == P64 && PR == P64) { hipLaunchKernelGGL((k_pupdate<P64, P64, double2>), grid, blk, 0, st, g, p, r, S); return; }
  if (PR != P32) { fprintf(stderr, "qexhip: pupdate r precision\n"); return; }
  switch (PP) {
    case P32: hipLaunchKernelGGL((k_pupdate<P32, P32, float2>), grid, blk, 0, st, g, p, r, S); break;
    case P16: hipLaunchKernelGGL((k_pupdate<P16, P32, float2>), grid, blk, 0, st, g, p, r, S); break;
    case PBF: hipLaunchKernelGGL((k_pupdate<PBF, P32, float2>), grid, blk, 0, st, g, p, r, S); break;
  }
}

void launch_xr(int PX, int PP, int PR, int PAp, const Geom &g, void *x, const void *p, void *r, const void *Ap,
               Scal *S, RedWs red, hipStream_t st) {
  dim3 grid(nblk(g.Vh, 256)), blk(256);
  if (PX == P64 && PP == P64) {
    hipLaunchKernelGGL((k_xr<P64, P64, P64, P64, double2>), grid, blk, 0, st, g, x, p, r, Ap, S, red);
    return;
  }
  (void)PX; (void)PR; (void)PAp;
  switch (PP) {
    case P32: hipLaunchKernelGGL((k_xr<P32, P32, P32, P32, float2>), grid, blk, 0, st, g, x, p, r, Ap, S, red); break;
    case P16: hipLaunchKernelGGL((k_xr<P32, P16, P32, P32, float2>), grid, blk, 0, st, g, x, p, r, Ap, S, red); break;
    case PBF: hipLaunchKernelGGL((k_xr<P32, PBF, P32, P32, float2>), grid, blk, 0, st, g, x, p, r, Ap, S, red); break;
  }
}

// reliable update, part 1: x += xs/sigma ; xs = 0   (FP64 solution accumulation)
__global__ void k_xfold(Geom g, double2 *x, float2 *xs, Scal *S) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  const double is = 1.0 / S->sigma;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const long o = c * g.stride + i;
    float2 s = xs[o];
    double2 v = x[o];
    v.x += is * (double)s.x;
    v.y += is * (double)s.y;
    x[o] = v;
    xs[o] = make_float2(0.f, 0.f);
  }
}

// reliable update, last part: rs = sigma * r_true (sloppy copy of the true residual)
__global__ void k_rs_from_true(Geom g, float2 *rs, const double2 *rt, Scal *S) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  const double sg = S->sigma;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const long o = c * g.stride + i;
    double2 v = rt[o];
    rs[o] = make_float2((float)(sg * v.x), (float)(sg * v.y));
  }
}

void launch_rs_from_true(int PR, const Geom &g, void *rs, const double *rtrue, Scal *S, hipStream_t st) {
  (void)PR;
  hipLaunchKernelGGL(k_rs_from_true, dim3(nblk(g.Vh, 256)), dim3(256), 0, st, g, (float2 *)rs,
                     (const double2 *)rtrue, S);
}

void launch_xfold(const Geom &g, double *x, void *xs, Scal *S, hipStream_t st) {
  hipLaunchKernelGGL(k_xfold, dim3(nblk(g.Vh, 256)), dim3(256), 0, st, g, (double2 *)x, (float2 *)xs, S);
}

__global__ void k_scal_init(Scal *S, int mode) {
  if (threadIdx.x != 0) return;
  if (mode == 0) {  // start of a CG solve: fresh state (cg.nim:21-27)
    S->its = 0; S->since = 0; S->done = 0; S->need_rel = 0; S->first = 1; S->nrel = 0;
    S->r2old = 1.0; S->beta = 0.0; S->alpha = 0.0; S->pscale = 1.0; S->sigma = 1.0;
  } else if (mode == 1) {  // force a reliable update now
    S->need_rel = 1;
  }
}

void launch_scal_init(Scal *S, hipStream_t st, int mode) {
  hipLaunchKernelGGL(k_scal_init, dim3(1), dim3(64), 0, st, S, mode);
}

// ------------------------------------------------------------- multi-shift
// Multi-shift CG (solvers/cgm.nim:163-200).  After Ap and p.Ap:
//   alpha = r2i/pAp ; for each shift k>0 the zeta recurrence gives
//   xc[k] = alpha*zr (x_k update) and the p_k update coefficients.
__global__ void k_cgm_alpha(ShiftScal *T, Scal *S) {
  if (threadIdx.x != 0) return;
  if (!T->continuing) return;
  const double pAp = S->pAp;
  const double alpha = (pAp != 0.0) ? T->r2i / pAp : 0.0;
  T->alpha = alpha;
  T->xc[0] = alpha;
  for (int k = 1; k < T->nshift; k++) {
    double zip1d = alpha * T->betaim1 * (T->zim1[k] - T->zi[k]);
    zip1d += T->zim1[k] * T->alphaim1 * (1.0 + T->sg[k] * alpha);
    const double zip1 = (zip1d != 0.0) ? T->zi[k] * T->zim1[k] * T->alphaim1 / zip1d : 0.0;
    const double zr = (T->zi[k] != 0.0) ? zip1 / T->zi[k] : 0.0;
    T->xc[k] = alpha * zr;
    T->pc1[k] = zip1;
    T->pc2[k] = zr * zr;  // times beta, known after |r|^2
    T->zim1[k] = T->zi[k];
    T->zi[k] = zip1;
  }
}

void launch_cgm_alpha(ShiftScal *T, Scal *S, hipStream_t st) {
  hipLaunchKernelGGL(k_cgm_alpha, dim3(1), dim3(64), 0, st, T, S);
}

// r -= alpha Ap ; x_k += xc[k] p_k (all k) ; r2 = |r|^2
__global__ void __launch_bounds__(256) k_cgm_xr(Geom g, int nshift, double2 *r, const double2 *Ap,
                                                double2 *const *x, double2 *const *p, ShiftScal *T,
                                                Scal *S, RedWs red) {
  __shared__ double lds[16];
  if (!T->continuing) return;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < g.Vh) {
    const double alpha = T->alpha;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long o = c * g.stride + i;
      double2 rv = r[o], av = Ap[o];
      rv.x -= alpha * av.x; rv.y -= alpha * av.y;
      r[o] = rv;
      s += rv.x * rv.x + rv.y * rv.y;
    }
    for (int k = 0; k < nshift; k++) {
      const double xc = T->xc[k];
      double2 *xk = x[k];
      const double2 *pk = p[k];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const long o = c * g.stride + i;
        double2 xv = xk[o], pv = pk[o];
        xv.x += xc * pv.x; xv.y += xc * pv.y;
        xk[o] = xv;
      }
    }
  }
  double v[1] = {s};
  if (grid_finish<1>(v, lds, red) && threadIdx.x == 0) {
    const double r2ip1 = v[0];
    T->r2ip1 = r2ip1;
    T->its += 1;
    const double beta = (T->r2i != 0.0) ? r2ip1 / T->r2i : 0.0;
    T->beta = beta;
    T->continuing = (T->its < T->maxits) && (r2ip1 > T->r2stop);
    S->done = !T->continuing;
    T->pc1[0] = 1.0;
    T->pc2[0] = beta;
    for (int k = 1; k < T->nshift; k++) T->pc2[k] *= beta;
    T->alphaim1 = T->alpha;
    T->betaim1 = beta;
    T->r2i = r2ip1;
  }
}

// p_k = pc1[k] r + pc2[k] p_k (only while continuing, cgm.nim:186,198-199)
__global__ void k_cgm_p(Geom g, int nshift, const double2 *r, double2 *const *p, ShiftScal *T) {
  if (!T->continuing) return;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  double2 rv[3];
#pragma unroll
  for (int c = 0; c < 3; c++) rv[c] = r[c * g.stride + i];
  for (int k = 0; k < nshift; k++) {
    const double a = T->pc1[k], b = T->pc2[k];
    double2 *pk = p[k];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long o = c * g.stride + i;
      double2 pv = pk[o];
      pk[o] = make_double2(a * rv[c].x + b * pv.x, a * rv[c].y + b * pv.y);
    }
  }
}

void launch_cgm_xr(const Geom &g, int nshift, double *r, const double *Ap, double **x, double **p, ShiftScal *T,
                   Scal *S, RedWs red, hipStream_t st) {
  hipLaunchKernelGGL(k_cgm_xr, dim3(nblk(g.Vh, 256)), dim3(256), 0, st, g, nshift, (double2 *)r,
                     (const double2 *)Ap, (double2 *const *)x, (double2 *const *)p, T, S, red);
}

void launch_cgm_p(const Geom &g, int nshift, const double *r, double **p, ShiftScal *T, hipStream_t st) {
  hipLaunchKernelGGL(k_cgm_p, dim3(nblk(g.Vh, 256)), dim3(256), 0, st, g, nshift, (const double2 *)r,
                     (double2 *const *)p, T);
}

// ------------------------------------------------------------------- force
// F_mu(x) = TAH(eps(x) psi(x) psi(x+mu)^+ U_mu(x)^+)   (hmc/hmcAction.nim:735-790,
// unsmeared; TAH = maths/matrixFunctions.nim:414-419).  One thread per
// (canonical site, mu); output in QEX canonical order [mu][site][3][3].
__global__ void __launch_bounds__(256) k_force(Geom g, const double2 *links, long lkblk, const double2 *pe,
                                               const double2 *po, double2 *F, int outer_only) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (idx >= V) return;
  const int par = idx >= g.Vh;
  const long i = idx - par * g.Vh;
  const Site s = decode(g, i, par);
  const long j = nbr(g, s, mu, 1);
  const double2 *ps = par ? po : pe, *pn = par ? pe : po;
  const double eps = par ? -1.0 : 1.0;
  double2 a[3], b[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    a[c] = ps[c * g.stride + i];
    b[c] = pn[c * g.stride + j];
    a[c].x *= eps; a[c].y *= eps;
  }
  double2 G[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++)  // a_r conj(b_c)
      G[3 * r + c] = make_double2(a[r].x * b[c].x + a[r].y * b[c].y, a[r].y * b[c].x - a[r].x * b[c].y);
  double2 *out = F + ((long)mu * V + idx) * 9;
  if (outer_only) {
#pragma unroll
    for (int e = 0; e < 9; e++) out[e] = G[e];
    return;
  }
  double2 U[9];
  load_link<P64, double2>(links, (long)(2 * mu + par) * lkblk + i, g.stride, U);
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

void launch_force(const Geom &g, const double *links64, long lkblk, const double *psi_e, const double *psi_o,
                  double *F, int outer_only, hipStream_t st) {
  dim3 grid(nblk(2 * g.Vh, 256), 4);
  hipLaunchKernelGGL(k_force, grid, dim3(256), 0, st, g, (const double2 *)links64, lkblk, (const double2 *)psi_e,
                     (const double2 *)psi_o, (double2 *)F, outer_only);
}

}  // namespace qh
