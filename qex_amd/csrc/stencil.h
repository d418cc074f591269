// Device building blocks of the staggered stencil shared by the single-system
// kernels (dslash.hip) and the batched ones (batch.hip): the per-link loads,
// the rolled or pipelined one-parity stencil sum, storage round trips, and the
// bodies of the fused two-kernel CG iteration.
#pragma once
#include "internal.h"
#include "finish.h"

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
  ld_link<PL, RC, true, A, link_blocked<PL, NL>()>(links + (long)(2 * k + (fwd ? par : 1 - par)) * lkblk, g.stride, fwd ? i : n, L.U,
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
// compiler barrier after each leg).  QH_PIPE = -1 / -2 / -3: a rolled loop over
// hops / links / links with a single accumulator, which bounds the registers
// a wave holds.
template <int PL> constexpr int stencil_form() {
  return (PL == P16 || PL == PBF || PL == PQ16) ? QH_PIPE16 : QH_PIPE;
}

// stencil in form D (the forms above); stencil() takes the build's form for PL
template <int D, int PL, int RC, int NL, class A, class LD>
__device__ __forceinline__ void stencil_in(const Geom &g, const char *links, long lkblk, const float *qs, int par,
                                           long i, const Site &s, LD &&ld, A (&res)[3]) {
  constexpr int NJ = 2 * NL;
  A af[3], ab[3];
#pragma unroll
  for (int c = 0; c < 3; c++) { af[c] = mk<A>(0, 0); ab[c] = mk<A>(0, 0); }
  if constexpr (D == -4) {  // rolled loop, one accumulator, next link loaded ahead
    Leg<A> L0;
    leg_load<PL, RC, NL, A>(g, links, lkblk, qs, par, i, s, ld, 0, L0);
#pragma unroll 1
    for (int j = 0; j < NJ; j++) {
      Leg<A> L1;
      if (j + 1 < NJ) leg_load<PL, RC, NL, A>(g, links, lkblk, qs, par, i, s, ld, j + 1, L1);
      if (!(j & 1)) {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) cmac(af[r], L0.U[3 * r + c], L0.v[c]);
      } else {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) cjmsub(af[r], L0.U[3 * c + r], L0.v[c]);
      }
      L0 = L1;
    }
  } else if constexpr (D == -3) {  // rolled loop, one link per trip, one accumulator
#pragma unroll 1
    for (int j = 0; j < NJ; j++) {
      Leg<A> L0;
      leg_load<PL, RC, NL, A>(g, links, lkblk, qs, par, i, s, ld, j, L0);
      if (!(j & 1)) {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) cmac(af[r], L0.U[3 * r + c], L0.v[c]);
      } else {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) cjmsub(af[r], L0.U[3 * c + r], L0.v[c]);
      }
    }
  } else if constexpr (D == -1) {  // rolled loop, one hop (two links) per trip
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

template <int PL, int RC, int NL, class A, class LD>
__device__ __forceinline__ void stencil(const Geom &g, const char *links, long lkblk, const float *qs, int par,
                                        long i, const Site &s, LD &&ld, A (&res)[3]) {
  stencil_in<stencil_form<PL>(), PL, RC, NL, A>(g, links, lkblk, qs, par, i, s, ld, res);
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
__device__ __forceinline__ void cg1_body(const CgArgs &P, long blk) {
  using R = RT<A>;
  const Scal *S = P.S;
  if (S->done | S->need_rel) return;
  const Geom &g = P.g;
  const long i = blk * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  const int q = 1 - P.par;
  const bool first = S->first != 0;
  const R alpha = (R)S->alpha, beta = (R)(S->beta * S->pscale);
  const Site s = decode(g, i, q);
  auto ld = [&](long j, A(&v)[3]) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      A rn, pv, av;
      p_next<PS, PR, A>(P, vix(g, c, j), alpha, beta, first, rn, v[c], pv, av);
    }
  };
  A res[3];
  stencil<PL, RC, NL, A>(g, P.links, P.lkblk, P.qs, q, i, s, ld, res);
#pragma unroll
  for (int c = 0; c < 3; c++) stc<PS, A>(P.t, vix(g, c, i), res[c]);
}

// lds: 16*4 doubles; the block's four partial sums go to slot pidx
template <int PL, int RC, int PS, int PR, int PX, class A, int NL>
__device__ __forceinline__ void cg2_body(const CgArgs &P, long blk, double *lds, long pidx) {
  using R = RT<A>;
  Scal *S = P.S;
  if (S->done | S->need_rel) return;
  const Geom &g = P.g;
  const long i = blk * blockDim.x + threadIdx.x;
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < g.Vh) {
    const bool first = S->first != 0;
    const R alpha = (R)S->alpha, beta = (R)(S->beta * S->pscale), m4 = (R)(4.0 * S->m2);
    // own-site vectors first: their loads overlap the stencil's
    A rv[3], av[3], pv[3], xv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long off = vix(g, c, i);
      rv[c] = ldc<PR, A>(P.r, off);
      av[c] = ldc<PS, A>(P.Ap, off);
      pv[c] = first ? mk<A>(0, 0) : ldc<PS, A>(P.p, off);
      xv[c] = ldc<PX, A>(P.x, off);
    }
    const Site s = decode(g, i, P.par);
    auto ld = [&](long j, A(&v)[3]) {
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = ldc<PS, A>(P.t, vix(g, c, j));
    };
    A H[3];
    stencil<PL, RC, NL, A>(g, P.links, P.lkblk, P.qs, P.par, i, s, ld, H);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long off = vix(g, c, i);
      const A rn = rt_store<PR, A>(mk<A>(rv[c].x - alpha * av[c].x, rv[c].y - alpha * av[c].y));
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
  block_partials_at<4>(red, lds, P.red, pidx);
}

// Three-kernel form of the same iteration: (0) own parity, streaming: apply
// the pending update and store p_n = r_{n-1} + beta p_{n-1}; (1) other parity:
// t = (2D) p_n gathering p_n alone; (2) own parity: Ap_n = 4 m^2 p_n - (2D) t
// and the four dot products.  Same algorithmic bytes as the two-kernel form,
// but the stencil kernels gather one vector instead of forming p_n from three
// at every neighbour.
template <int PS, int PR, int PX, class A>
__device__ __forceinline__ void cg0_body(const CgArgs &P, long blk) {
  using R = RT<A>;
  const Scal *S = P.S;
  if (S->done | S->need_rel) return;
  const Geom &g = P.g;
  const long i = blk * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  const bool first = S->first != 0;
  const R alpha = (R)S->alpha, beta = (R)(S->beta * S->pscale);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const long off = vix(g, c, i);
    const A rv = ldc<PR, A>(P.r, off), av = ldc<PS, A>(P.Ap, off);
    const A pv = first ? mk<A>(0, 0) : ldc<PS, A>(P.p, off);
    const A xv = ldc<PX, A>(P.x, off);
    const A rn = mk<A>(rv.x - alpha * av.x, rv.y - alpha * av.y);
    const A pn = rt_store<PS, A>(first ? rn : mk<A>(rn.x + beta * pv.x, rn.y + beta * pv.y));
    stc<PX, A>(P.x, off, mk<A>(xv.x + alpha * pv.x, xv.y + alpha * pv.y));
    stc<PR, A>(P.r, off, rn);
    stc<PS, A>(P.p, off, pn);
  }
}

// Form 4: the streaming kernel of form 3 with the previous iteration's
// dot-product finish folded in.  Every workgroup adds the nb block partials
// of the last cg2b launch (sum_partials: the same order, hence the same bits,
// as the one-block finish of form 3 when blockDim.x = QH_FIN_THREADS),
// applies the CG epilogue to its own copy of the scalars read from Sr, and
// uses that copy; workgroup 0 stores it to Sw, which the iteration's other
// kernels read.  Sr and Sw alternate between two buffers from one iteration
// to the next, so no workgroup reads scalars that another one is writing.
// The workgroup's vector loads are issued before the partial sums, whose
// latency they overlap; form 4 runs cg2b on a persistent grid so that nb
// (its workgroups) stays small.  lds: 16*4 doubles.
template <int PS, int PR, int PX, class A>
__device__ __forceinline__ void cg0f_body(const CgArgs &P, const Scal *Sr, Scal *Sw, int nb, double *lds,
                                          Scal *sc) {
  using R = RT<A>;
  const Geom &g = P.g;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < g.Vh;
  A rv[3], av[3], pv[3], xv[3];
  if (in)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long off = vix(g, c, i);
      rv[c] = ldc<PR, A>(P.r, off);
      av[c] = ldc<PS, A>(P.Ap, off);
      pv[c] = ldc<PS, A>(P.p, off);
      xv[c] = ldc<PX, A>(P.x, off);
    }
  double tot[4];
  sum_partials<4>(P.red, nb, tot, lds);
  if (threadIdx.x == 0) {
    Scal v = *Sr;
    if (!(v.done | v.need_rel)) epi_cg2(&v, tot[0], tot[1], tot[2], tot[3]);
    *sc = v;
    if (blockIdx.x == 0) *Sw = v;
  }
  __syncthreads();
  if ((sc->done | sc->need_rel) || !in) return;
  const bool first = sc->first != 0;
  const R alpha = (R)sc->alpha, beta = (R)(sc->beta * sc->pscale);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const long off = vix(g, c, i);
    const A p0 = first ? mk<A>(0, 0) : pv[c];
    const A rn = mk<A>(rv[c].x - alpha * av[c].x, rv[c].y - alpha * av[c].y);
    const A pn = rt_store<PS, A>(first ? rn : mk<A>(rn.x + beta * p0.x, rn.y + beta * p0.y));
    stc<PX, A>(P.x, off, mk<A>(xv[c].x + alpha * p0.x, xv[c].y + alpha * p0.y));
    stc<PR, A>(P.r, off, rn);
    stc<PS, A>(P.p, off, pn);
  }
}

template <int PL, int RC, int PS, class A, int NL>
__device__ __forceinline__ void cg1b_body(const CgArgs &P, long blk) {
  const Scal *S = P.S;
  if (S->done | S->need_rel) return;
  const Geom &g = P.g;
  const long i = blk * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  const int q = 1 - P.par;
  const Site s = decode(g, i, q);
  auto ld = [&](long j, A(&v)[3]) {
#pragma unroll
    for (int c = 0; c < 3; c++) v[c] = ldc<PS, A>(P.p, vix(g, c, j));
  };
  A res[3];
  stencil<PL, RC, NL, A>(g, P.links, P.lkblk, P.qs, q, i, s, ld, res);
#pragma unroll
  for (int c = 0; c < 3; c++) stc_out<PS, A>(P.t, vix(g, c, i), res[c]);
}

// cg2b's work on logical block blk, its four dot products added to red
template <int PL, int RC, int PS, int PR, class A, int NL>
__device__ __forceinline__ void cg2b_accum(const CgArgs &P, long blk, double (&red)[4]) {
  using R = RT<A>;
  Scal *S = P.S;
  const Geom &g = P.g;
  const long i = blk * blockDim.x + threadIdx.x;
  if (i < g.Vh) {
    const R m4 = (R)(4.0 * S->m2);
    A rv[3], pv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long off = vix(g, c, i);
      rv[c] = ldc<PR, A>(P.r, off);
      pv[c] = ldc<PS, A>(P.p, off);
    }
    const Site s = decode(g, i, P.par);
    auto ld = [&](long j, A(&v)[3]) {
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = ldc<PS, A>(P.t, vix(g, c, j));
    };
    A H[3];
    stencil<PL, RC, NL, A>(g, P.links, P.lkblk, P.qs, P.par, i, s, ld, H);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const A apn = rt_store<PS, A>(mk<A>(m4 * pv[c].x - H[c].x, m4 * pv[c].y - H[c].y));
      stc_out<PS, A>(P.Ap, vix(g, c, i), apn);
      red[0] += (double)pv[c].x * (double)apn.x + (double)pv[c].y * (double)apn.y;
      red[1] += (double)rv[c].x * (double)apn.x + (double)rv[c].y * (double)apn.y;
      red[2] += (double)apn.x * (double)apn.x + (double)apn.y * (double)apn.y;
      red[3] += (double)rv[c].x * (double)rv[c].x + (double)rv[c].y * (double)rv[c].y;
    }
  }
}

template <int PL, int RC, int PS, int PR, class A, int NL>
__device__ __forceinline__ void cg2b_body(const CgArgs &P, long blk, double *lds, long pidx) {
  if (P.S->done | P.S->need_rel) return;
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  cg2b_accum<PL, RC, PS, PR, A, NL>(P, blk, red);
  block_partials_at<4>(red, lds, P.red, pidx);
}

// cg2b on a persistent grid (form 4): each workgroup walks its logical blocks
// (block_iter) and writes one set of partials, gridDim.x in all
template <int PL, int RC, int PS, int PR, class A, int NL>
__device__ __forceinline__ void cg2b_persist_body(const CgArgs &P, double *lds) {
  if (P.S->done | P.S->need_rel) return;
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  const long nblk = (P.g.Vh + blockDim.x - 1) / blockDim.x;
  const BlockIter it = block_iter(nblk, P.remap);
#pragma unroll 1
  for (long lb = it.cur; lb < it.end; lb += it.step) cg2b_accum<PL, RC, PS, PR, A, NL>(P, lb, red);
  block_partials_at<4>(red, lds, P.red, blockIdx.x);
}

// x += alpha p ; r -= alpha Ap ; |r|^2 (the pending update of the last iteration)
template <int PS, int PR, int PX, class A>
__device__ __forceinline__ void cg_flush_body(const CgArgs &P, long blk, double *lds, long pidx) {
  using R = RT<A>;
  Scal *S = P.S;
  const Geom &g = P.g;
  const long i = blk * blockDim.x + threadIdx.x;
  double rr = 0.0;
  if (i < g.Vh) {
    const R alpha = (R)S->alpha;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long off = vix(g, c, i);
      A rv = ldc<PR, A>(P.r, off), av = ldc<PS, A>(P.Ap, off), pv = ldc<PS, A>(P.p, off), xv = ldc<PX, A>(P.x, off);
      rv.x -= alpha * av.x; rv.y -= alpha * av.y;
      xv.x += alpha * pv.x; xv.y += alpha * pv.y;
      stc<PR, A>(P.r, off, rv);
      stc<PX, A>(P.x, off, xv);
      rr += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
    }
  }
  double v[1] = {rr};
  block_partials_at<1>(v, lds, P.red, pidx);
}

}  // namespace qh
