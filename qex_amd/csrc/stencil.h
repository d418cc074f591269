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

// Cache policy of the stencils' link loads.  A sweep reads every link once,
// so by default links are loaded with the non-temporal hint, which keeps the
// gathered color vectors resident in L2 and the last-level cache.  But a
// solver sweeps the same links again and again, and a link field that fits in
// the 256 MB last-level cache (16-bit 12-real links: 100 MB at 32^4) can stay
// there from one sweep to the next if it is read with plain loads.  Geom's
// lkcache gives the link bytes to read that way: the first link_cached_blocks
// link blocks (block b = 2k + parity, lkblk bytes each), the rest
// non-temporally.  A hop of either parity reads every block once, so a cached
// block serves both.  Direction k's forward link comes from block 2k + par and
// its backward one from block 2k + 1 - par: the directions below mb/2 have
// both blocks cached, and for odd mb direction mb/2 has block mb - 1, which
// is its forward link's block when par is even and its backward one's when
// par is odd.  Each range has a loop of its own, so the two kinds of load
// never meet in one instruction.
template <int NL> __device__ __forceinline__ int link_cached_blocks(const Geom &g, long lkblk) {
  int m = 0;
  while (m < 2 * NL && (long)(m + 1) * lkblk <= g.lkcache) m++;
  return m;
}

template <int PL, int RC, int NL, class A, class LD, bool NT = true>
__device__ __forceinline__ void leg_load(const Geom &g, const char *links, long lkblk, const float *qs, int par,
                                         long i, const Site &s, LD &ld, int j, Leg<A> &L) {
  const int k = j >> 1, fwd = !(j & 1);
  const int mu = (NL == 4) ? k : (k >> 1);
  const int len = (NL == 8 && (k & 1)) ? 3 : 1;
  const long n = nbr(g, s, mu, fwd ? len : -len);
  ld_link<PL, RC, NT, A, link_blocked<PL, NL>()>(links + (long)(2 * k + (fwd ? par : 1 - par)) * lkblk, g.stride, fwd ? i : n, L.U,
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

// ---- form -5: lean loop for compact links.  One direction per trip, its
// forward and then its backward link, each link's loads and arithmetic in one
// basic block (no branch on the link's orientation), so that 16-bit values
// enter the fused multiply-adds unconverted (FP16 operands of an FP32 FMA).
// Neighbour indices come from 32-bit per-site bits and wave-uniform steps
// instead of a decode per link, the staggered signs likewise, and the PQ16
// scale of a direction multiplies that direction's sum once instead of every
// link real (the integers enter the products as they are; the rebuilt row 2,
// quadratic in them, takes one factor of the scale with its sign).

// Per-site bits of the lean form.  wrap: bit 2k / 2k+1 set when the forward /
// backward neighbour of direction k (NL = 4: k = mu; NL = 8: k = 2 mu for the
// one-link hop, 2 mu + 1 for the three-link hop) lies across the lattice
// edge.  sgn: bit 2 mu / 2 mu + 1 set when link_sign(g, s, mu, 1 / 0) is -1.
template <int NL>
__device__ __forceinline__ void lean_bits(const Geom &g, const Site &s, uint32_t &wrap, uint32_t &sgn) {
  wrap = 0u;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const int mu = (NL == 4) ? k : (k >> 1);
    const int len = (NL == 8 && (k & 1)) ? 3 : 1;
    const int c = mu == 0 ? s.x : mu == 1 ? s.y : mu == 2 ? s.z : s.t;
    wrap |= (uint32_t)(c + len >= g.L[mu]) << (2 * k);
    wrap |= (uint32_t)(c < len) << (2 * k + 1);
  }
  const uint32_t e0 = (uint32_t)s.t, e1 = (uint32_t)(s.x + s.t), e2 = (uint32_t)(s.x + s.y + s.t);
  sgn = (e0 & 1u) * 3u | ((e1 & 1u) * 3u) << 2 | ((e2 & 1u) * 3u) << 4 |
        (uint32_t)(s.t == g.L[3] - 1) << 6 | (uint32_t)(s.t == 0) << 7;
}

// sign bit (bit 31) of link_sign for orientation bit b = 2 mu + (backward)
__device__ __forceinline__ uint32_t lean_signbit(uint32_t sgn, int b) { return (sgn << (31 - b)) & 0x80000000u; }

// One link of the lean form: row 2 of a 12-real link rebuilt as
// f * conj(row0 x row1), f = the link's sign (times the PQ16 scale: rows 0-1
// are then unscaled, and so is the rebuilt row relative to the true one).
// The products start from fma(a, b, 0) so that 16-bit operands stay
// unconverted in every one of them.
template <int PL, int RC, class A, bool BLK, bool NT>
__device__ __forceinline__ void lean_link(const char *blk, long stride, long i, RT<A> f, A (&U)[9]) {
  if (link_wide<PL, RC, BLK>(stride))
    ld_link_planes<PL, RC, NT, A, BLK, true, false>(blk, stride, i, U, 0.0f);
  else
    ld_link_planes<PL, RC, NT, A, BLK, false, false>(blk, stride, i, U, 0.0f);
  if constexpr (RC == 12) {
    using R = RT<A>;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const A a1 = U[(k + 1) % 3], a2 = U[(k + 2) % 3], b1 = U[3 + (k + 1) % 3], b2 = U[3 + (k + 2) % 3];
      R re = fma(a1.x, b2.x, (R)0), im = fma(a1.x, b2.y, (R)0);
      re = fma(-a1.y, b2.y, re);
      re = fma(-a2.x, b1.x, re);
      re = fma(a2.y, b1.y, re);
      im = fma(a1.y, b2.x, im);
      im = fma(-a2.x, b1.y, im);
      im = fma(-a2.y, b1.x, im);
      U[6 + k] = mk<A>(f * re, -f * im);
    }
  }
}

// A scheduling barrier between the two links of a trip and after the trip:
// without it the compiler issues the backward link's loads with the forward
// one's, and the Q16 12-real cg2b kernel takes 71 VGPRs instead of 58 (7
// waves per SIMD instead of 8).  QH_LEAN_FENCE=0 drops it.
#ifndef QH_LEAN_FENCE
#define QH_LEAN_FENCE 1
#endif
__device__ __forceinline__ void lean_fence() {
  if (QH_LEAN_FENCE) {
    mem_fence_compiler();
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int PL, int RC, int NL, class A, class LD>
__device__ __forceinline__ void stencil_lean(const Geom &g, const char *links, long lkblk, const float *qs, int par,
                                             long i, const Site &s, LD &&ld, A (&acc)[3]) {
  using R = RT<A>;
  constexpr bool BLK = link_blocked<PL, NL>();
  constexpr bool Q = (PL == PQ16);
  uint32_t wrap, sgn;
  lean_bits<NL>(g, s, wrap, sgn);
  const uint32_t i32 = (uint32_t)i, ix = i32 + (uint32_t)(s.x & 1);
  const uint32_t hx = (uint32_t)g.hx, L1 = (uint32_t)g.L[1], L12 = (uint32_t)g.L[1] * (uint32_t)g.L[2];
#pragma unroll
  for (int c = 0; c < 3; c++) acc[c] = mk<A>(0, 0);
  // one trip: direction k, its forward / backward link loaded non-temporally
  // when ntf / ntb is true
  auto trip = [&](int k, auto ntf, auto ntb) {
    constexpr bool NTF = decltype(ntf)::value, NTB = decltype(ntb)::value;
    const int mu = (NL == 4) ? k : (k >> 1);
    const uint32_t len = (NL == 8 && (k & 1)) ? 3u : 1u;
    // wave-uniform: index step of one hop in mu, and its edge correction
    const uint32_t Lmu = (uint32_t)(mu == 1 ? g.L[1] : mu == 2 ? g.L[2] : g.L[3]);
    const uint32_t step = mu == 1 ? hx : mu == 2 ? hx * L1 : hx * L12;
    const uint32_t dF = mu == 0 ? (len >> 1) : len * step;
    const uint32_t dB = mu == 0 ? 0u - ((len + 1u) >> 1) : 0u - dF;
    const uint32_t wF = mu == 0 ? 0u - hx : 0u - step * Lmu;
    const uint32_t base = mu == 0 ? ix : i32;
    const uint32_t nF = base + dF + (((wrap >> (2 * k)) & 1u) ? wF : 0u);
    const uint32_t nB = base + dB + (((wrap >> (2 * k + 1)) & 1u) ? 0u - wF : 0u);
    const R fs = Q ? (R)qs[k] : (R)1;
    A T[3];
#pragma unroll
    for (int c = 0; c < 3; c++) T[c] = Q ? mk<A>(0, 0) : acc[c];
    {
      A U[9], v[3];
      const R f = (R)__uint_as_float(__float_as_uint((float)fs) ^ lean_signbit(sgn, 2 * mu));
      lean_link<PL, RC, A, BLK, NTF>(links + (long)(2 * k + par) * lkblk, g.stride, (long)i32, f, U);
      ld((long)nF, v);
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) cmac(T[r], U[3 * r + c], v[c]);
    }
    lean_fence();
    {
      A U[9], v[3];
      const R f = (R)__uint_as_float(__float_as_uint((float)fs) ^ lean_signbit(sgn, 2 * mu + 1));
      lean_link<PL, RC, A, BLK, NTB>(links + (long)(2 * k + 1 - par) * lkblk, g.stride, (long)nB, f, U);
      ld((long)nB, v);
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) cjmsub(T[r], U[3 * c + r], v[c]);
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
      acc[c] = Q ? mk<A>(fma(fs, T[c].x, acc[c].x), fma(fs, T[c].y, acc[c].y)) : T[c];
    lean_fence();
  };
  // the first link_cached_blocks blocks with cache-allocating loads
  const int mb = link_cached_blocks<NL>(g, lkblk), kf = mb >> 1;
#pragma unroll 1
  for (int k = 0; k < kf; k++) trip(k, std::false_type(), std::false_type());
  int k1 = kf;
  if (mb & 1) {  // direction kf: only block mb - 1 = 2 kf cached
    if (par & 1)
      trip(kf, std::true_type(), std::false_type());
    else
      trip(kf, std::false_type(), std::true_type());
    k1++;
  }
#pragma unroll 1
  for (int k = k1; k < NL; k++) trip(k, std::true_type(), std::true_type());
}

// The links j in [sub NJ/NS, (sub+1) NJ/NS) of the stencil sum for one of the
// NS adjacent threads of a site (split form of k_hop), then their sum over the
// NS threads (lane shuffles within the group): every thread of the group ends
// with the whole sum.  The links are read with cache-allocating loads when the
// whole link field fits the link-cache budget (a per-launch choice, the same
// for every thread), else non-temporally.
template <int PL, int RC, int NL, class A, int NS, class LD>
__device__ __forceinline__ void stencil_share(const Geom &g, const char *links, long lkblk, const float *qs, int par,
                                              long i, const Site &s, LD &&ld, int sub, A (&res)[3]) {
  constexpr int NJ = 2 * NL, PER = NJ / NS;
  static_assert(NJ % NS == 0, "NS must divide the links of a site");
  A af[3];
#pragma unroll
  for (int c = 0; c < 3; c++) af[c] = mk<A>(0, 0);
  auto trip = [&](int j, auto ntc) {
    Leg<A> L0;
    leg_load<PL, RC, NL, A, LD, decltype(ntc)::value>(g, links, lkblk, qs, par, i, s, ld, j, L0);
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
  };
  const int j0 = sub * PER;
  if (link_cached_blocks<NL>(g, lkblk) == NJ) {
#pragma unroll 1
    for (int j = j0; j < j0 + PER; j++) trip(j, std::false_type());
  } else {
#pragma unroll 1
    for (int j = j0; j < j0 + PER; j++) trip(j, std::true_type());
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    auto x = af[c].x, y = af[c].y;
#pragma unroll
    for (int o = 1; o < NS; o <<= 1) {
      x += __shfl_xor(x, o, 64);
      y += __shfl_xor(y, o, 64);
    }
    res[c] = mk<A>(x, y);
  }
}

// stencil in form D (the forms above); stencil() takes the build's form for PL
template <int D, int PL, int RC, int NL, class A, class LD>
__device__ __forceinline__ void stencil_in(const Geom &g, const char *links, long lkblk, const float *qs, int par,
                                           long i, const Site &s, LD &&ld, A (&res)[3]) {
  if constexpr (D == -5) {
    stencil_lean<PL, RC, NL, A>(g, links, lkblk, qs, par, i, s, ld, res);
    return;
  }
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
    // the links of the first link_cached_blocks blocks with cache-allocating loads
    auto trip = [&](int j, auto ntc) {
      Leg<A> L0;
      leg_load<PL, RC, NL, A, LD, decltype(ntc)::value>(g, links, lkblk, qs, par, i, s, ld, j, L0);
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
    };
    const int mb = link_cached_blocks<NL>(g, lkblk), kf = mb >> 1;
#pragma unroll 1
    for (int j = 0; j < 2 * kf; j++) trip(j, std::false_type());
    int j1 = 2 * kf;
    if (mb & 1) {  // direction kf: only block mb - 1 = 2 kf cached
      if (par & 1) {
        trip(j1, std::true_type());
        trip(j1 + 1, std::false_type());
      } else {
        trip(j1, std::false_type());
        trip(j1 + 1, std::true_type());
      }
      j1 += 2;
    }
#pragma unroll 1
    for (int j = j1; j < NJ; j++) trip(j, std::true_type());
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
