// The staggered stencil on gfx950: the one-parity hop (plain and HISQ form,
// FP64/FP32/FP16/BF16 storage, 18- or 12-real links) and the fused two-kernel
// CG iteration built on it.
//
// Hop, for every site x of the output parity p (QEX's (2D) stencil,
// physics/stagD.nim:383-448; nl=4: k=mu, 1 hop; nl=8: k=2mu fat 1 hop,
// k=2mu+1 long 3 hops, stagD.nim:40-51):
//   H(x) = sum_k [ U_k(x) in(x + l_k mu_k) - U_k(x - l_k mu_k)^+ in(x - l_k mu_k) ]
// One thread per output site; links are streamed with non-temporal 16-byte
// loads (each link is read once per hop), the other-parity color vectors are
// gathered through L2 with an XCD-aware block order.
#include <stdio.h>
#include <type_traits>
#include "internal.h"

namespace qh {

template <class A> using RT = typename cx<A>::R;

// One link of the stencil: j = 2k (forward U_k(x) at this site) or 2k+1
// (backward U_k(x - l mu)^+ at the neighbour), with its gathered vector.
template <class A> struct Leg {
  A U[9];
  A v[3];
};

template <int PL, int RC, int NL, class A, class LD>
__device__ __forceinline__ void leg_load(const Geom &g, const char *links, long lkblk, const float *qs, int par,
                                         long i, const Site &s, LD &ld, int j, Leg<A> &L) {
  const int k = j >> 1, fwd = !(j & 1);
  const int mu = (NL == 4) ? k : (k >> 1);
  const int len = (NL == 8 && (k & 1)) ? 3 : 1;
  const long n = nbr(g, s, mu, fwd ? len : -len);
  ld_link<PL, RC, true, A>(links + (long)(2 * k + (fwd ? par : 1 - par)) * lkblk, g.stride, fwd ? i : n, L.U,
                           (RT<A>)(RC == 12 ? link_sign(g, s, mu, fwd) : 1.0f), PL == PQ16 ? qs[k] : 0.0f);
  ld(n, L.v);
}

template <class A>
__device__ __forceinline__ void leg_fma(int j, const Leg<A> &L, A (&af)[3], A (&ab)[3]) {
  if (!(j & 1)) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) cmac(af[r], L.U[3 * r + c], L.v[c]);
  } else {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) cjmac(ab[r], L.U[3 * c + r], L.v[c]);
  }
}

// Compiler-level memory barrier: keeps later loads from being hoisted above
// it, which bounds the links a wave holds in registers.
__device__ __forceinline__ void mem_fence_compiler() { asm volatile("" ::: "memory"); }

// res = sum_k [U_k(x) v(x+l mu) - U_k(x-l mu)^+ v(x-l mu)], v from the loader.
// QH_PIPE >= 0: the 2*NL links are unrolled and software-pipelined QH_PIPE
// legs deep (loads of leg j+QH_PIPE issued before the FMAs of leg j, a
// compiler barrier after each leg).  QH_PIPE = -1 / -2: a rolled loop over
// hops / links, which bounds the registers a wave holds.
template <int PL, int RC, int NL, class A, class LD>
__device__ __forceinline__ void stencil(const Geom &g, const char *links, long lkblk, const float *qs, int par,
                                        long i, const Site &s, LD &&ld, A (&res)[3]) {
  constexpr int NJ = 2 * NL;
  constexpr int D = QH_PIPE;
  A af[3], ab[3];
#pragma unroll
  for (int c = 0; c < 3; c++) { af[c] = mk<A>(0, 0); ab[c] = mk<A>(0, 0); }
  if constexpr (D == -1) {  // rolled loop, one hop (two links) per trip
#pragma unroll 1
    for (int j = 0; j < NJ; j += 2) {
      Leg<A> L0, L1;
      leg_load<PL, RC, NL, A>(g, links, lkblk, qs, par, i, s, ld, j, L0);
      leg_load<PL, RC, NL, A>(g, links, lkblk, qs, par, i, s, ld, j + 1, L1);
      leg_fma<A>(j, L0, af, ab);
      leg_fma<A>(j + 1, L1, af, ab);
    }
  } else if constexpr (D < 0) {  // rolled loop, one link per trip
#pragma unroll 1
    for (int j = 0; j < NJ; j++) {
      Leg<A> L0;
      leg_load<PL, RC, NL, A>(g, links, lkblk, qs, par, i, s, ld, j, L0);
      leg_fma<A>(j, L0, af, ab);
    }
  } else {
    constexpr int DD = (D < 0) ? 0 : D;
    Leg<A> L[DD + 1];
#pragma unroll
    for (int j = 0; j < DD; j++) leg_load<PL, RC, NL, A>(g, links, lkblk, qs, par, i, s, ld, j, L[j]);
#pragma unroll
    for (int j = 0; j < NJ; j++) {
      if (j + DD < NJ) leg_load<PL, RC, NL, A>(g, links, lkblk, qs, par, i, s, ld, j + DD, L[(j + DD) % (DD + 1)]);
      leg_fma<A>(j, L[j % (DD + 1)], af, ab);
      mem_fence_compiler();
    }
  }
#pragma unroll
  for (int c = 0; c < 3; c++) res[c] = mk<A>(af[c].x - ab[c].x, af[c].y - ab[c].y);
}

// value after a round trip through storage precision P (exact for P32/P64)
template <int P, class A> __device__ __forceinline__ A rt_store(A v) {
  if (P == P16) {
    float2 f = __half22float2(__float22half2_rn(make_float2((float)v.x, (float)v.y)));
    return mk<A>(f.x, f.y);
  } else if (P == PBF) {
    const uint32_t a = f_to_bf16((float)v.x), b = f_to_bf16((float)v.y);
    return mk<A>(__uint_as_float(a << 16), __uint_as_float(b << 16));
  }
  return v;
}

// ------------------------------------------------------------ generic hop
template <int PL, int RC, int PI, int PO, int PA, class A, int NL, int MODE, int RED>
__global__ void __launch_bounds__(256) k_hop(HopArgs P) {
  using R = RT<A>;
  __shared__ double lds[16];
  if (P.pred && (P.S->done | P.S->need_rel)) return;
  const Geom &g = P.g;
  const long i = xcd_block(P.remap) * blockDim.x + threadIdx.x;
  double red = 0.0;
  if (i < g.Vh) {
    const Site s = decode(g, i, P.par);
    A res[3];
    auto ld = [&](long j, A(&v)[3]) {
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = ldc<PI, A>(P.in, c * g.stride + j);
    };
    stencil<PL, RC, NL, A>(g, P.links, P.lkblk, P.qs, P.par, i, s, ld, res);
    const R chop = (R)P.chop;
#pragma unroll
    for (int c = 0; c < 3; c++) res[c] = mk<A>(chop * res[c].x, chop * res[c].y);
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
        const A bv = ldc<PA, A>(P.b, c * g.stride + i);
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
    block_partials<1>(v, lds, P.red);
  }
}

template <int PL, int RC, int PI, int PO, int PA, class A, int NL>
static bool hop_modes(int mode, int redm, const HopArgs &a, hipStream_t st, int block) {
  dim3 grid((unsigned)((a.g.Vh + block - 1) / block));
#define QH_HOP(M, R)                                                                                 \
  if (mode == M && redm == R) {                                                                      \
    hipLaunchKernelGGL((k_hop<PL, RC, PI, PO, PA, A, NL, M, R>), grid, dim3(block), 0, st, a);       \
    if (R) {                                                                                         \
      FinArgs f{a.red, (int)grid.x, a.S, nullptr, 0, 6.0 * (double)a.g.Vh, a.pred};                  \
      launch_finish(R == 1 ? EPI_PAP : R == 2 ? EPI_RELIABLE : EPI_TMP, f, st);                      \
    }                                                                                                \
    return true;                                                                                     \
  }
  QH_HOP(0, 0) QH_HOP(1, 0) QH_HOP(2, 0)
  if constexpr (PL == P64) { QH_HOP(1, 1) QH_HOP(2, 2) QH_HOP(2, 3) }
#undef QH_HOP
  return false;
}

template <int PL, class A>
static bool hop_prec(int rc, int nl, int mode, int redm, const HopArgs &a, hipStream_t st, int block) {
  if (rc == 18 && nl == 4) return hop_modes<PL, 18, PL, PL, PL, A, 4>(mode, redm, a, st, block);
  if (rc == 18 && nl == 8) return hop_modes<PL, 18, PL, PL, PL, A, 8>(mode, redm, a, st, block);
  if (rc == 12 && nl == 4) return hop_modes<PL, 12, PL, PL, PL, A, 4>(mode, redm, a, st, block);
  return false;
}

template <class A>
static bool hop_q16(int rc, int nl, int mode, int redm, const HopArgs &a, hipStream_t st, int block) {
  if (rc == 18 && nl == 4) return hop_modes<PQ16, 18, P16, P16, P16, A, 4>(mode, redm, a, st, block);
  if (rc == 18 && nl == 8) return hop_modes<PQ16, 18, P16, P16, P16, A, 8>(mode, redm, a, st, block);
  if (rc == 12 && nl == 4) return hop_modes<PQ16, 12, P16, P16, P16, A, 4>(mode, redm, a, st, block);
  return false;
}

void launch_hop(const HopKey &k, int nl, int mode, int redm, const HopArgs &a, hipStream_t st, int block) {
  bool ok = false;
  if (k.pl == PQ16 && k.pi == P16 && k.po == P16 && k.pa == P16) ok = hop_q16<float2>(k.rc, nl, mode, redm, a, st, block);
  if (k.pi == k.pl && k.po == k.pl && k.pa == k.pl) {
    switch (k.pl) {
      case P64: ok = hop_prec<P64, double2>(k.rc, nl, mode, redm, a, st, block); break;
      case P32: ok = hop_prec<P32, float2>(k.rc, nl, mode, redm, a, st, block); break;
      case P16: ok = hop_prec<P16, float2>(k.rc, nl, mode, redm, a, st, block); break;
      case PBF: ok = hop_prec<PBF, float2>(k.rc, nl, mode, redm, a, st, block); break;
    }
  }
  if (!ok)
    fprintf(stderr, "qexhip: no hop instantiation for pl=%d rc=%d pi=%d po=%d pa=%d nl=%d mode=%d red=%d\n", k.pl,
            k.rc, k.pi, k.po, k.pa, nl, mode, redm);
}

// ---------------------------------------------------------- fused CG
// Iteration n in two kernels.  On entry the own-parity vectors hold
// x_{n-2}, r_{n-2}, p_{n-1}, Ap_{n-1} and S holds alpha_{n-1} (pending) and
// beta_n.  Kernel 1 (other parity) forms p_n = (r_{n-2} - alpha Ap_{n-1}) +
// beta p_{n-1} on the fly at each neighbour and writes t = (2D) p_n.
// Kernel 2 (own parity) applies the pending update, forms p_n again, writes
// x_{n-1}, r_{n-1}, p_n, Ap_n = 4 m^2 p_n - (2D) t and reduces the four FP64
// dot products epi_cg2 needs.  A final flush applies the last pending update.
template <int PS, int PR, class A>
__device__ __forceinline__ void p_next(const CgArgs &P, long off, RT<A> alpha, RT<A> beta, bool first, A &rn,
                                       A &pn, A &pv, A &av) {
  const A rv = ldc<PR, A>(P.r, off);
  av = ldc<PS, A>(P.Ap, off);
  rn = mk<A>(rv.x - alpha * av.x, rv.y - alpha * av.y);
  if (first) {
    pv = mk<A>(0, 0);
    pn = rt_store<PS, A>(rn);
  } else {
    pv = ldc<PS, A>(P.p, off);
    pn = rt_store<PS, A>(mk<A>(rn.x + beta * pv.x, rn.y + beta * pv.y));
  }
}

template <int PL, int RC, int PS, int PR, int PX, class A, int NL>
__global__ void __launch_bounds__(256) k_cg1(CgArgs P) {
  using R = RT<A>;
  const Scal *S = P.S;
  if (S->done | S->need_rel) return;
  const Geom &g = P.g;
  const long i = xcd_block(P.remap) * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  const int q = 1 - P.par;
  const bool first = S->first != 0;
  const R alpha = (R)S->alpha, beta = (R)(S->beta * S->pscale);
  const Site s = decode(g, i, q);
  auto ld = [&](long j, A(&v)[3]) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      A rn, pv, av;
      p_next<PS, PR, A>(P, c * g.stride + j, alpha, beta, first, rn, v[c], pv, av);
    }
  };
  A res[3];
  stencil<PL, RC, NL, A>(g, P.links, P.lkblk, P.qs, q, i, s, ld, res);
#pragma unroll
  for (int c = 0; c < 3; c++) stc<PS, A>(P.t, c * g.stride + i, res[c]);
}

template <int PL, int RC, int PS, int PR, int PX, class A, int NL>
__global__ void __launch_bounds__(256) k_cg2(CgArgs P) {
  using R = RT<A>;
  __shared__ double lds[16 * 4];
  Scal *S = P.S;
  if (S->done | S->need_rel) return;
  const Geom &g = P.g;
  const long i = xcd_block(P.remap) * blockDim.x + threadIdx.x;
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < g.Vh) {
    const bool first = S->first != 0;
    const R alpha = (R)S->alpha, beta = (R)(S->beta * S->pscale), m4 = (R)(4.0 * S->m2);
    // own-site vectors first: their loads overlap the stencil's
    A rv[3], av[3], pv[3], xv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long off = c * g.stride + i;
      rv[c] = ldc<PR, A>(P.r, off);
      av[c] = ldc<PS, A>(P.Ap, off);
      pv[c] = first ? mk<A>(0, 0) : ldc<PS, A>(P.p, off);
      xv[c] = ldc<PX, A>(P.x, off);
    }
    const Site s = decode(g, i, P.par);
    auto ld = [&](long j, A(&v)[3]) {
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = ldc<PS, A>(P.t, c * g.stride + j);
    };
    A H[3];
    stencil<PL, RC, NL, A>(g, P.links, P.lkblk, P.qs, P.par, i, s, ld, H);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long off = c * g.stride + i;
      const A rn = mk<A>(rv[c].x - alpha * av[c].x, rv[c].y - alpha * av[c].y);
      const A pn = rt_store<PS, A>(first ? rn : mk<A>(rn.x + beta * pv[c].x, rn.y + beta * pv[c].y));
      const A xn = mk<A>(xv[c].x + alpha * pv[c].x, xv[c].y + alpha * pv[c].y);
      const A apn = rt_store<PS, A>(mk<A>(m4 * pn.x - H[c].x, m4 * pn.y - H[c].y));
      stc<PX, A>(P.x, off, xn);
      stc<PR, A>(P.r, off, rn);
      stc<PS, A>(P.p, off, pn);
      stc<PS, A>(P.Ap, off, apn);
      red[0] += (double)pn.x * (double)apn.x + (double)pn.y * (double)apn.y;
      red[1] += (double)rn.x * (double)apn.x + (double)rn.y * (double)apn.y;
      red[2] += (double)apn.x * (double)apn.x + (double)apn.y * (double)apn.y;
      red[3] += (double)rn.x * (double)rn.x + (double)rn.y * (double)rn.y;
    }
  }
  block_partials<4>(red, lds, P.red);
}

// x += alpha p ; r -= alpha Ap ; |r|^2 (the pending update of the last iteration)
template <int PS, int PR, int PX, class A>
__global__ void __launch_bounds__(256) k_cg_flush(CgArgs P) {
  using R = RT<A>;
  __shared__ double lds[16];
  Scal *S = P.S;
  const Geom &g = P.g;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0;
  if (i < g.Vh) {
    const R alpha = (R)S->alpha;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long off = c * g.stride + i;
      A rv = ldc<PR, A>(P.r, off), av = ldc<PS, A>(P.Ap, off), pv = ldc<PS, A>(P.p, off), xv = ldc<PX, A>(P.x, off);
      rv.x -= alpha * av.x; rv.y -= alpha * av.y;
      xv.x += alpha * pv.x; xv.y += alpha * pv.y;
      stc<PR, A>(P.r, off, rv);
      stc<PX, A>(P.x, off, xv);
      rr += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
    }
  }
  double v[1] = {rr};
  block_partials<1>(v, lds, P.red);
}

template <int PL, int RC, int PS, int NL>
static void cg_iter_t(const CgArgs &a, hipStream_t st, int block) {
  dim3 grid((unsigned)((a.g.Vh + block - 1) / block)), blk(block);
  if constexpr (PL == P64) {
    hipLaunchKernelGGL((k_cg1<PL, RC, P64, P64, P64, double2, NL>), grid, blk, 0, st, a);
    hipLaunchKernelGGL((k_cg2<PL, RC, P64, P64, P64, double2, NL>), grid, blk, 0, st, a);
  } else {
    hipLaunchKernelGGL((k_cg1<PL, RC, PS, P32, P32, float2, NL>), grid, blk, 0, st, a);
    hipLaunchKernelGGL((k_cg2<PL, RC, PS, P32, P32, float2, NL>), grid, blk, 0, st, a);
  }
  FinArgs f{a.red, (int)grid.x, a.S, nullptr, 0, 0.0, 1};
  launch_finish(EPI_CG2, f, st);
}

template <int PL, int PS>
static bool cg_iter_rc(int rc, int nl, const CgArgs &a, hipStream_t st, int block) {
  if (rc == 18 && nl == 4) { cg_iter_t<PL, 18, PS, 4>(a, st, block); return true; }
  if (rc == 18 && nl == 8) { cg_iter_t<PL, 18, PS, 8>(a, st, block); return true; }
  if (rc == 12 && nl == 4) { cg_iter_t<PL, 12, PS, 4>(a, st, block); return true; }
  return false;
}

void launch_cg_iter(int pl, int rc, int ps, int nl, const CgArgs &a, hipStream_t st, int block) {
  bool ok = false;
  if (pl == P64 && ps == P64) ok = cg_iter_rc<P64, P64>(rc, nl, a, st, block);
  else if (pl == P32 && ps == P32) ok = cg_iter_rc<P32, P32>(rc, nl, a, st, block);
  else if (pl == P16 && ps == P16) ok = cg_iter_rc<P16, P16>(rc, nl, a, st, block);
  else if (pl == P16 && ps == P32) ok = cg_iter_rc<P16, P32>(rc, nl, a, st, block);
  else if (pl == PBF && ps == PBF) ok = cg_iter_rc<PBF, PBF>(rc, nl, a, st, block);
  else if (pl == PBF && ps == P32) ok = cg_iter_rc<PBF, P32>(rc, nl, a, st, block);
  else if (pl == PQ16 && ps == P16) ok = cg_iter_rc<PQ16, P16>(rc, nl, a, st, block);
  else if (pl == PQ16 && ps == P32) ok = cg_iter_rc<PQ16, P32>(rc, nl, a, st, block);
  if (!ok) fprintf(stderr, "qexhip: no CG instantiation for pl=%d rc=%d ps=%d nl=%d\n", pl, rc, ps, nl);
}

void launch_cg_flush(int ps, const CgArgs &a, hipStream_t st) {
  dim3 grid((unsigned)((a.g.Vh + 255) / 256)), blk(256);
  switch (ps) {
    case P64: hipLaunchKernelGGL((k_cg_flush<P64, P64, P64, double2>), grid, blk, 0, st, a); break;
    case P32: hipLaunchKernelGGL((k_cg_flush<P32, P32, P32, float2>), grid, blk, 0, st, a); break;
    case P16: hipLaunchKernelGGL((k_cg_flush<P16, P32, P32, float2>), grid, blk, 0, st, a); break;
    case PBF: hipLaunchKernelGGL((k_cg_flush<PBF, P32, P32, float2>), grid, blk, 0, st, a); break;
  }
  FinArgs f{a.red, (int)grid.x, a.S, nullptr, 0, 0.0, 0};
  launch_finish(EPI_FLUSH, f, st);
}

// ------------------------------------------------------------ link packing
struct QScale { float inv[8]; };  // PQ16: 32767/scale per link field

template <int PL, int RC>
__global__ void k_pack_links(Geom g, int nl, const double2 *canon, char *out, long lkblk, QScale q) {
  using F = LFmt<PL, RC>;
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (idx >= V) return;
  const int par = idx >= g.Vh;
  const long i = idx - par * g.Vh;
  uint32_t w[F::words];
#pragma unroll
  for (int c = 0; c < RC / 2; c++) {
    const double2 d = canon[((long)k * V + idx) * 9 + c];
    if (PL == P64) {
      w[4 * c] = (uint32_t)__double2loint(d.x); w[4 * c + 1] = (uint32_t)__double2hiint(d.x);
      w[4 * c + 2] = (uint32_t)__double2loint(d.y); w[4 * c + 3] = (uint32_t)__double2hiint(d.y);
    } else if (PL == P32) {
      w[2 * c] = __float_as_uint((float)d.x); w[2 * c + 1] = __float_as_uint((float)d.y);
    } else if (PL == P16) {
      __half2 h = __floats2half2_rn((float)d.x, (float)d.y);
      __builtin_memcpy(&w[c], &h, 4);
    } else if (PL == PQ16) {
      const double f = (double)q.inv[k];
      const int re = (int)fmax(-32767.0, fmin(32767.0, rint(d.x * f)));
      const int im = (int)fmax(-32767.0, fmin(32767.0, rint(d.y * f)));
      w[c] = ((uint32_t)re & 0xffffu) | ((uint32_t)im << 16);
    } else {
      w[c] = f_to_bf16((float)d.x) | (f_to_bf16((float)d.y) << 16);
    }
  }
  st_link_words<PL, RC>(out + (long)(2 * k + par) * lkblk, g.stride, i, w);
}

template <int PL>
static void pack_rc(int rc, const Geom &g, int nl, const double *canon, char *out, const QScale &q, hipStream_t st) {
  dim3 grid((unsigned)((2 * g.Vh + 255) / 256), nl);
  const long blk = link_block_bytes(PL, rc, g.stride);
  if (rc == 12)
    hipLaunchKernelGGL((k_pack_links<PL, 12>), grid, dim3(256), 0, st, g, nl, (const double2 *)canon, out, blk, q);
  else
    hipLaunchKernelGGL((k_pack_links<PL, 18>), grid, dim3(256), 0, st, g, nl, (const double2 *)canon, out, blk, q);
}

void launch_pack_links(int prec, int rc, const Geom &g, int nl, const double *canon, char *out, const float *qs,
                       hipStream_t st) {
  QScale q;
  for (int k = 0; k < 8; k++) q.inv[k] = (qs && k < nl && qs[k] > 0.0f) ? 1.0f / qs[k] : 0.0f;
  switch (prec) {
    case P64: pack_rc<P64>(rc, g, nl, canon, out, q, st); break;
    case P32: pack_rc<P32>(rc, g, nl, canon, out, q, st); break;
    case P16: pack_rc<P16>(rc, g, nl, canon, out, q, st); break;
    case PBF: pack_rc<PBF>(rc, g, nl, canon, out, q, st); break;
    case PQ16: pack_rc<PQ16>(rc, g, nl, canon, out, q, st); break;
  }
}

// largest |re|, |im| of each link field (for the PQ16 scale)
__global__ void k_link_maxabs(Geom g, int nl, const double2 *canon, unsigned *maxbits) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  float m = 0.0f;
  if (idx < V) {
    const double2 *u = canon + ((long)k * V + idx) * 9;
    for (int e = 0; e < 9; e++) m = fmaxf(m, fmaxf(fabsf((float)u[e].x), fabsf((float)u[e].y)));
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(maxbits + k, __float_as_uint(m));
}

void launch_link_maxabs(const Geom &g, int nl, const double *canon, unsigned *maxbits, hipStream_t st) {
  dim3 grid((unsigned)((2 * g.Vh + 255) / 256), nl);
  hipLaunchKernelGGL(k_link_maxabs, grid, dim3(256), 0, st, g, nl, (const double2 *)canon, maxbits);
}

// Largest |row2 - sign*conj(row0 x row1)| over all plain links: 12-real
// compression is only used when the links are phased SU(3) matrices.
__global__ void k_recon_check(Geom g, const double2 *canon, unsigned *maxdev_bits) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (idx >= V) return;
  const int par = idx >= g.Vh;
  const Site s = decode(g, idx - par * g.Vh, par);
  const double sgn = link_sign(g, s, mu, 1);
  const double2 *u = canon + ((long)mu * V + idx) * 9;
  double dev = 0.0;
  for (int k = 0; k < 3; k++) {
    const double2 a1 = u[(k + 1) % 3], a2 = u[(k + 2) % 3], b1 = u[3 + (k + 1) % 3], b2 = u[3 + (k + 2) % 3];
    const double re = (a1.x * b2.x - a1.y * b2.y) - (a2.x * b1.x - a2.y * b1.y);
    const double im = (a1.x * b2.y + a1.y * b2.x) - (a2.x * b1.y + a2.y * b1.x);
    dev = fmax(dev, fabs(sgn * re - u[6 + k].x) + fabs(-sgn * im - u[6 + k].y));
  }
  if (!(dev <= 1.0)) dev = 1.0;  // NaN or gross mismatch
  atomicMax(maxdev_bits, __float_as_uint((float)dev));
}

void launch_recon_check(const Geom &g, const double *canon, unsigned *maxdev_bits, hipStream_t st) {
  dim3 grid((unsigned)((2 * g.Vh + 255) / 256), 4);
  hipLaunchKernelGGL(k_recon_check, grid, dim3(256), 0, st, g, (const double2 *)canon, maxdev_bits);
}

}  // namespace qh
