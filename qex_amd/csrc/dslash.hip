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
// gathered through L2 with an XCD-aware block order and, where the tuner
// picks it, the t-first site walk (make_walk), which keeps a chunk's t +- 1
// neighbours in L2 between their uses.
#include <stdio.h>
#include <type_traits>
#include "stencil.h"

namespace qh {


// ------------------------------------------------------------ generic hop
#ifndef QH_HOP_WAVES
#define QH_HOP_WAVES 1  // minimum waves per SIMD asked of the register allocator
#endif
// TWO: the stencil loop takes the two links of a hop per trip (form -1: twice
// the loads in flight per wave, 10-20 more registers): 17-27% faster at 16^4,
// where too few waves fill the device to hide latency otherwise, and 1% at 48^4,
// but 1% slower at 32^4; a tuned launch dimension.
// NS > 1 (split form): NS adjacent threads per output site, each taking
// 2 NL / NS of its links, their sums added across the group by lane shuffles;
// the group's first thread finishes the site.  More waves and loads in flight
// for lattices too small to fill the device with one thread per site.
template <int PL, int RC, int PI, int PO, int PA, class A, int NL, int MODE, int RED, int TWO = 0, int NS = 1>
__global__ void __launch_bounds__(256, QH_HOP_WAVES) k_hop(HopArgs P) {
  using R = RT<A>;
  __shared__ double lds[16];
  if (P.pred && (P.S->done | P.S->need_rel)) return;
  const Geom &g = P.g;
  double red = 0.0;
  const long spb = blockDim.x / NS;  // sites per workgroup
  const int sub = (int)(threadIdx.x % NS);
  const Walk wk = make_walk(g, spb, P.chunk);
  const BlockIter it = block_iter(wk.nblk, P.remap);
#pragma unroll 1
  for (long lb = it.cur; lb < it.end; lb += it.step) {
    // the NS threads of a site share i, so they skip together
    const long i = walk_site(wk, lb, spb, threadIdx.x / NS);
    if (i < 0 || i >= g.Vh) continue;
    const Site s = decode(g, i, P.par);
    A res[3];
    auto ld = [&](long j, A(&v)[3]) {
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = ldc<PI, A>(P.in, vix(g, c, j));
    };
    if constexpr (NS == 1) {
      stencil_in<TWO ? -1 : stencil_form<PL>(), PL, RC, NL, A>(g, P.links, P.lkblk, P.qs, P.par, i, s, ld, res);
    } else {
      stencil_share<PL, RC, NL, A, NS>(g, P.links, P.lkblk, P.qs, P.par, i, s, ld, sub, res);
      if (sub != 0) continue;
    }
    const R chop = (R)P.chop;
#pragma unroll
    for (int c = 0; c < 3; c++) res[c] = mk<A>(chop * res[c].x, chop * res[c].y);
    A av[3];
    if (MODE >= 1) {
      double ca = P.ca;
      if (P.ca_m2) ca *= 4.0 * P.S->m2;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        av[c] = ldc<PA, A>(P.a, vix(g, c, i));
        res[c].x += (R)ca * av[c].x;
        res[c].y += (R)ca * av[c].y;
      }
    }
    if (MODE == 2) {
      const R cb = (R)P.cb;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const A bv = ldc<PA, A>(P.b, vix(g, c, i));
        res[c].x += cb * bv.x;
        res[c].y += cb * bv.y;
      }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) stc_out<PO, A>(P.out, vix(g, c, i), res[c]);
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

// ------------------------------------------------------- t-march hop
// The hop of plain links (nl = 4) with each thread walking a column of
// t-slices: thread l of logical workgroup (seg, blk) serves the half-site
// index is = blk*blockDim.x + l of the output-parity slices t = seg*T ..
// seg*T + T - 1.  A t-hop keeps x/2, y and z, so the t-neighbours of output
// site (t, is) are the input sites (t -/+ 1, is), and the input vector loaded
// as the forward neighbour of slice t is the backward neighbour of slice t+2:
// it stays in registers (two slots, one per slice parity), and a column's
// input vectors are each loaded once per segment, plus two at its start.
// Workgroups are segment-major, so an XCD's contiguous range holds
// neighbouring blocks of the same segment, which march through t together and
// share their x, y, z neighbours through its L2; the t-neighbours, which the
// canonical order reads again a slice later (from memory once two slices
// outgrow the L2, as at 48^4), no longer come through the caches at all.
template <int PL, int RC, int PI, int PO, int PA, class A, int MODE, int RED>
__global__ void __launch_bounds__(256, QH_HOP_WAVES) k_hop_tm(HopArgs P) {
  using R = RT<A>;
  __shared__ double lds[16];
  if (P.pred && (P.S->done | P.S->need_rel)) return;
  const Geom &g = P.g;
  const int Lt = g.L[3], T = P.tmarch;
  const long S = g.Vh / Lt;
  const long nblk = (S + blockDim.x - 1) / blockDim.x;
  const long G = xcd_block(P.remap);
  const long seg = G / nblk, is = (G % nblk) * blockDim.x + threadIdx.x;
  const int t0 = (int)seg * T;
  double red = 0.0;
  if (is < S && t0 < Lt) {
    const int t1 = t0 + T < Lt ? t0 + T : Lt;
    auto vin = [&](int t, A(&v)[3]) {
      const long j = (long)wrap(t, Lt) * S + is;
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = ldc<PI, A>(P.in, vix(g, c, j));
    };
    // W[0]: backward t-neighbour of the even steps t0 + 2k, W[1] of the odd ones
    A W[2][3];
    vin(t0 - 1, W[0]);
    vin(t0, W[1]);
    // one output slice: the legs of the rolled stencil loop, with the backward
    // t-neighbour from w and the forward one, loaded, left in w for slice t + 2
    auto step = [&](int t, A(&w)[3]) {
      const long i = (long)t * S + is;
      const Site s = decode(g, i, P.par);
      A af[3], nf[3];
#pragma unroll
      for (int c = 0; c < 3; c++) af[c] = mk<A>(0, 0);
#pragma unroll 1
      for (int j = 0; j < 8; j++) {
        const int mu = j >> 1, fwd = !(j & 1);
        const long n = nbr(g, s, mu, fwd ? 1 : -1);
        A U[9], v[3];
        ld_link<PL, RC, true, A, link_blocked<PL, 4>()>(P.links + (long)(2 * mu + (fwd ? P.par : 1 - P.par)) * P.lkblk,
                                                        g.stride, fwd ? i : n, U,
                                                        (R)(RC == 12 ? link_sign(g, s, mu, fwd) : 1.0f),
                                                        PL == PQ16 ? P.qs[mu] : 0.0f);
        if (j == 7) {
#pragma unroll
          for (int c = 0; c < 3; c++) v[c] = w[c];
        } else {
#pragma unroll
          for (int c = 0; c < 3; c++) v[c] = ldc<PI, A>(P.in, vix(g, c, n));
          if (j == 6)
#pragma unroll
            for (int c = 0; c < 3; c++) nf[c] = v[c];
        }
        if (fwd) {
#pragma unroll
          for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) cmac(af[r], U[3 * r + c], v[c]);
        } else {
#pragma unroll
          for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) cjmsub(af[r], U[3 * c + r], v[c]);
        }
      }
#pragma unroll
      for (int c = 0; c < 3; c++) w[c] = nf[c];
      const R chop = (R)P.chop;
      A res[3];
#pragma unroll
      for (int c = 0; c < 3; c++) res[c] = mk<A>(chop * af[c].x, chop * af[c].y);
      A av[3];
      if (MODE >= 1) {
        double ca = P.ca;
        if (P.ca_m2) ca *= 4.0 * P.S->m2;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          av[c] = ldc<PA, A>(P.a, vix(g, c, i));
          res[c].x += (R)ca * av[c].x;
          res[c].y += (R)ca * av[c].y;
        }
      }
      if (MODE == 2) {
        const R cb = (R)P.cb;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const A bv = ldc<PA, A>(P.b, vix(g, c, i));
          res[c].x += cb * bv.x;
          res[c].y += cb * bv.y;
        }
      }
#pragma unroll
      for (int c = 0; c < 3; c++) stc_out<PO, A>(P.out, vix(g, c, i), res[c]);
      if (RED == 1) {
#pragma unroll
        for (int c = 0; c < 3; c++) red += (double)av[c].x * (double)res[c].x + (double)av[c].y * (double)res[c].y;
      } else if (RED >= 2) {
#pragma unroll
        for (int c = 0; c < 3; c++) red += (double)res[c].x * (double)res[c].x + (double)res[c].y * (double)res[c].y;
      }
    };
#pragma unroll 1
    for (int t = t0; t < t1; t += 2) {
      step(t, W[0]);
      if (t + 1 < t1) step(t + 1, W[1]);
    }
  }
  if (RED) {
    double v[1] = {red};
    block_partials<1>(v, lds, P.red);
  }
}

// workgroups of the t-march hop: (slices / tmarch, rounded up) x blocks per
// slice, rounded up to a multiple of 8 for the XCD-aware order
inline unsigned tmarch_grid(const Geom &g, int block, int tmarch, int remap) {
  const long S = g.Vh / g.L[3];
  const long nseg = (g.L[3] + tmarch - 1) / tmarch;
  return grid_blocks(nseg * ((S + block - 1) / block), remap, 0);
}

template <int PL, int RC, int PI, int PO, int PA, class A, int NL>
static bool hop_modes(int mode, int redm, const HopArgs &a, hipStream_t st, int block, int cap) {
  const bool tm = NL == 4 && a.tmarch > 0;
  const int spb = (a.split == 2 || a.split == 4) ? block / a.split : block;  // sites per workgroup
  dim3 grid(tm ? tmarch_grid(a.g, block, a.tmarch, a.remap) : grid_blocks(make_walk(a.g, spb, a.chunk).nblk, a.remap, cap));
#define QH_HOP(M, R)                                                                                 \
  if (mode == M && redm == R) {                                                                      \
    if (tm) hipLaunchKernelGGL((k_hop_tm<PL, RC, PI, PO, PA, A, M, R>), grid, dim3(block), 0, st, a);   \
    else if (a.split == 2) hipLaunchKernelGGL((k_hop<PL, RC, PI, PO, PA, A, NL, M, R, 0, 2>), grid, dim3(block), 0, st, a); \
    else if (a.split == 4) hipLaunchKernelGGL((k_hop<PL, RC, PI, PO, PA, A, NL, M, R, 0, 4>), grid, dim3(block), 0, st, a); \
    else if (a.two) hipLaunchKernelGGL((k_hop<PL, RC, PI, PO, PA, A, NL, M, R, 1>), grid, dim3(block), 0, st, a); \
    else hipLaunchKernelGGL((k_hop<PL, RC, PI, PO, PA, A, NL, M, R>), grid, dim3(block), 0, st, a);  \
    if (R) {                                                                                         \
      FinArgs f{a.red, (int)grid.x, a.S, 0, 6.0 * (double)a.g.Vh, a.pred};                  \
      launch_finish(R == 1 ? EPI_PAP : R == 2 ? EPI_RELIABLE : EPI_TMP, f, st);                      \
    }                                                                                                \
    return true;                                                                                     \
  }
  QH_HOP(0, 0) QH_HOP(1, 0) QH_HOP(2, 0)
  if constexpr (PL == P64) { QH_HOP(1, 1) QH_HOP(2, 2) QH_HOP(2, 3) }
  if constexpr (PL == P32) { QH_HOP(1, 1) }
#undef QH_HOP
  return false;
}

template <int PL, class A>
static bool hop_prec(int rc, int nl, int mode, int redm, const HopArgs &a, hipStream_t st, int block, int cap) {
  if (rc == 18 && nl == 4) return hop_modes<PL, 18, PL, PL, PL, A, 4>(mode, redm, a, st, block, cap);
  if (rc == 18 && nl == 8) return hop_modes<PL, 18, PL, PL, PL, A, 8>(mode, redm, a, st, block, cap);
  if (rc == 12 && nl == 4) return hop_modes<PL, 12, PL, PL, PL, A, 4>(mode, redm, a, st, block, cap);
  return false;
}

template <class A>
static bool hop_q16(int rc, int nl, int mode, int redm, const HopArgs &a, hipStream_t st, int block, int cap) {
  if (rc == 18 && nl == 4) return hop_modes<PQ16, 18, P16, P16, P16, A, 4>(mode, redm, a, st, block, cap);
  if (rc == 18 && nl == 8) return hop_modes<PQ16, 18, P16, P16, P16, A, 8>(mode, redm, a, st, block, cap);
  if (rc == 12 && nl == 4) return hop_modes<PQ16, 12, P16, P16, P16, A, 4>(mode, redm, a, st, block, cap);
  return false;
}

// cap > 0: persistent form with at most cap workgroups (block_iter)
void launch_hop(const HopKey &k, int nl, int mode, int redm, const HopArgs &a, hipStream_t st, int block, int cap) {
  bool ok = false;
  if (k.pl == PQ16 && k.pi == P16 && k.po == P16 && k.pa == P16)
    ok = hop_q16<float2>(k.rc, nl, mode, redm, a, st, block, cap);
  if (k.pi == k.pl && k.po == k.pl && k.pa == k.pl) {
    switch (k.pl) {
      case P64: ok = hop_prec<P64, double2>(k.rc, nl, mode, redm, a, st, block, cap); break;
      case P32: ok = hop_prec<P32, float2>(k.rc, nl, mode, redm, a, st, block, cap); break;
      case P16: ok = hop_prec<P16, float2>(k.rc, nl, mode, redm, a, st, block, cap); break;
      case PBF: ok = hop_prec<PBF, float2>(k.rc, nl, mode, redm, a, st, block, cap); break;
    }
  }
  if (!ok)
    fprintf(stderr, "qexhip: no hop instantiation for pl=%d rc=%d pi=%d po=%d pa=%d nl=%d mode=%d red=%d\n", k.pl,
            k.rc, k.pi, k.po, k.pa, nl, mode, redm);
}

// ---------------------------------------------------------- fused CG
// (iteration bodies in stencil.h)
template <int PL, int RC, int PS, int PR, int PX, class A, int NL>
__global__ void __launch_bounds__(256) k_cg1(CgArgs P) {
  cg1_body<PL, RC, PS, PR, PX, A, NL>(P, xcd_block(P.remap));
}

template <int PL, int RC, int PS, int PR, int PX, class A, int NL>
__global__ void __launch_bounds__(256) k_cg2(CgArgs P) {
  __shared__ double lds[16 * 4];
  cg2_body<PL, RC, PS, PR, PX, A, NL>(P, xcd_block(P.remap), lds, blockIdx.x);
}

template <int PS, int PR, int PX, class A>
__global__ void __launch_bounds__(256) k_cg0(CgArgs P) {
  cg0_body<PS, PR, PX, A>(P, blockIdx.x);
}

template <int PS, int PR, int PX, class A>
__global__ void __launch_bounds__(QH_FIN_THREADS) k_cg0f(CgArgs P, const Scal *Sr, Scal *Sw, int nb) {
  __shared__ double lds[16 * 4];
  __shared__ Scal sc;
  cg0f_body<PS, PR, PX, A>(P, Sr, Sw, nb, lds, &sc);
}

template <int PL, int RC, int PS, int PR, class A, int NL>
__global__ void __launch_bounds__(256) k_cg2bp(CgArgs P) {
  __shared__ double lds[16 * 4];
  cg2b_persist_body<PL, RC, PS, PR, A, NL>(P, lds);
}

template <int PL, int RC, int PS, class A, int NL>
__global__ void __launch_bounds__(256) k_cg1b(CgArgs P) {
  cg1b_body<PL, RC, PS, A, NL>(P, xcd_block(P.remap));
}

// the same on a persistent grid: each workgroup serves the logical blocks
// block_iter gives it (an XCD's workgroups walk its contiguous range together)
template <int PL, int RC, int PS, class A, int NL>
__global__ void __launch_bounds__(256) k_cg1bp(CgArgs P) {
  const long nblk = (P.g.Vh + blockDim.x - 1) / blockDim.x;
  const BlockIter it = block_iter(nblk, P.remap);
#pragma unroll 1
  for (long lb = it.cur; lb < it.end; lb += it.step) cg1b_body<PL, RC, PS, A, NL>(P, lb);
}

template <int PL, int RC, int PS, int PR, class A, int NL>
__global__ void __launch_bounds__(256) k_cg2b(CgArgs P) {
  __shared__ double lds[16 * 4];
  cg2b_body<PL, RC, PS, PR, A, NL>(P, xcd_block(P.remap), lds, blockIdx.x);
}

template <int PS, int PR, int PX, class A>
__global__ void __launch_bounds__(256) k_cg_flush(CgArgs P) {
  __shared__ double lds[16];
  cg_flush_body<PS, PR, PX, A>(P, blockIdx.x, lds, blockIdx.x);
}

// cap > 0 (form 3): both stencil kernels persistent on at most cap workgroups
// (k_cg1bp, k_cg2bp; the finish then adds cap partial sums)
template <int PL, int RC, int PS, int PRX, int NL>
static void cg_iter_t(const CgArgs &a, hipStream_t st, int block, int form, int block1, int cap) {
  dim3 grid((unsigned)((a.g.Vh + block - 1) / block)), blk(block);
  // the other-parity stencil kernel of form 3 may run its own block size
  const int b1 = block1 > 0 ? block1 : block;
  dim3 grid1((unsigned)((a.g.Vh + b1 - 1) / b1)), blk1(b1);
  const bool pers = form == 3 && cap > 0;
  const dim3 grid2p(grid_blocks(grid.x, a.remap, cap)), grid1p(grid_blocks(grid1.x, a.remap, cap));
  using A = typename std::conditional<PL == P64, double2, float2>::type;
  constexpr int PSA = (PL == P64) ? P64 : PS;
  constexpr int PRA = (PL == P64) ? P64 : PRX;
  if (form == 3) {
    hipLaunchKernelGGL((k_cg0<PSA, PRA, PRA, A>), grid, blk, 0, st, a);
    if (pers) {
      hipLaunchKernelGGL((k_cg1bp<PL, RC, PSA, A, NL>), grid1p, blk1, 0, st, a);
      hipLaunchKernelGGL((k_cg2bp<PL, RC, PSA, PRA, A, NL>), grid2p, blk, 0, st, a);
    } else {
      hipLaunchKernelGGL((k_cg1b<PL, RC, PSA, A, NL>), grid1, blk1, 0, st, a);
      hipLaunchKernelGGL((k_cg2b<PL, RC, PSA, PRA, A, NL>), grid, blk, 0, st, a);
    }
  } else {
    hipLaunchKernelGGL((k_cg1<PL, RC, PSA, PRA, PRA, A, NL>), grid, blk, 0, st, a);
    hipLaunchKernelGGL((k_cg2<PL, RC, PSA, PRA, PRA, A, NL>), grid, blk, 0, st, a);
  }
  FinArgs f{a.red, (int)(pers ? grid2p.x : grid.x), a.S, 0, 0.0, 1};
  launch_finish(EPI_CG2, f, st);
}

// Form 4 chunk of niter iterations: the first iteration as form 3 without
// its finish; in each following one the streaming kernel k_cg0f finishes the
// previous iteration's dot products itself, reading the scalars from one
// buffer and writing the other (a.S, S1 alternately); a one-block finish
// leaves the chunk's last result in a.S for the host.  The own-parity stencil
// kernel runs persistent on cap2 workgroups (k_cg2bp), so it leaves cap2
// partials for k_cg0f to re-add instead of one per block.
template <int PL, int RC, int PS, int PRX, int NL>
static void cg_chunk4_t(const CgArgs &a, Scal *S1, hipStream_t st, int block, int block1, int niter, int cap2) {
  constexpr int PX = (PL == P64) ? P64 : PRX;
  constexpr int PR = PX;
  using A = typename std::conditional<PL == P64, double2, float2>::type;
  const long nblk = (a.g.Vh + block - 1) / block;
  dim3 grid((unsigned)nblk), blk(block);
  // cap2 = 0 (no tuned shape): 512 workgroups, two per CU
  dim3 grid2(grid_blocks(nblk, a.remap, cap2 > 0 ? cap2 : 512)), blk2(block);
  const int b1 = block1 > 0 ? block1 : block;
  dim3 grid1((unsigned)((a.g.Vh + b1 - 1) / b1)), blk1(b1);
  dim3 grid0((unsigned)((a.g.Vh + QH_FIN_THREADS - 1) / QH_FIN_THREADS)), blk0(QH_FIN_THREADS);
  Scal *S[2] = {a.S, S1};
  const int nb = (int)grid2.x;  // partials of k_cg2bp
  for (int k = 0; k < niter; k++) {
    CgArgs ak = a;
    ak.S = S[k & 1];
    if (k == 0) hipLaunchKernelGGL((k_cg0<PS, PR, PX, A>), grid, blk, 0, st, ak);
    else hipLaunchKernelGGL((k_cg0f<PS, PR, PX, A>), grid0, blk0, 0, st, a, S[(k - 1) & 1], S[k & 1], nb);
    hipLaunchKernelGGL((k_cg1b<PL, RC, PS, A, NL>), grid1, blk1, 0, st, ak);
    hipLaunchKernelGGL((k_cg2bp<PL, RC, PS, PR, A, NL>), grid2, blk2, 0, st, ak);
  }
  launch_finish_cg2_to(a.red, nb, S[(niter - 1) & 1], a.S, st);
}

template <int PL, int PS, int PRX = P32>
static bool cg_chunk4_rc(int rc, int nl, const CgArgs &a, Scal *S1, hipStream_t st, int block, int b1, int niter,
                         int g0) {
  if (rc == 18 && nl == 4) { cg_chunk4_t<PL, 18, PS, PRX, 4>(a, S1, st, block, b1, niter, g0); return true; }
  if (rc == 18 && nl == 8) { cg_chunk4_t<PL, 18, PS, PRX, 8>(a, S1, st, block, b1, niter, g0); return true; }
  if (rc == 12 && nl == 4) { cg_chunk4_t<PL, 12, PS, PRX, 4>(a, S1, st, block, b1, niter, g0); return true; }
  return false;
}

template <int PL, int PS, int PRX = P32>
static bool cg_iter_rc(int rc, int nl, const CgArgs &a, hipStream_t st, int block, int form, int b1, int cap) {
  if (rc == 18 && nl == 4) { cg_iter_t<PL, 18, PS, PRX, 4>(a, st, block, form, b1, cap); return true; }
  if (rc == 18 && nl == 8) { cg_iter_t<PL, 18, PS, PRX, 8>(a, st, block, form, b1, cap); return true; }
  if (rc == 12 && nl == 4) { cg_iter_t<PL, 12, PS, PRX, 4>(a, st, block, form, b1, cap); return true; }
  return false;
}

// form 2: two stencil kernels (p formed on the fly); form 3: a streaming update
// kernel plus two stencil kernels that gather one vector each.  rx: storage of
// the sloppy residual and solution increment (P32, or P16 with FP16 vectors).
void launch_cg_iter(int pl, int rc, int ps, int nl, const CgArgs &a, hipStream_t st, int block, int form, int rx,
                    int block1, int cap) {
  bool ok = false;
  if (rx == P16 && ps == P16) {
    if (pl == P16) ok = cg_iter_rc<P16, P16, P16>(rc, nl, a, st, block, form, block1, cap);
    else if (pl == PQ16) ok = cg_iter_rc<PQ16, P16, P16>(rc, nl, a, st, block, form, block1, cap);
  } else if (pl == P64 && ps == P64) ok = cg_iter_rc<P64, P64>(rc, nl, a, st, block, form, block1, cap);
  else if (pl == P32 && ps == P32) ok = cg_iter_rc<P32, P32>(rc, nl, a, st, block, form, block1, cap);
  else if (pl == P16 && ps == P16) ok = cg_iter_rc<P16, P16>(rc, nl, a, st, block, form, block1, cap);
  else if (pl == P16 && ps == P32) ok = cg_iter_rc<P16, P32>(rc, nl, a, st, block, form, block1, cap);
  else if (pl == PBF && ps == PBF) ok = cg_iter_rc<PBF, PBF>(rc, nl, a, st, block, form, block1, cap);
  else if (pl == PBF && ps == P32) ok = cg_iter_rc<PBF, P32>(rc, nl, a, st, block, form, block1, cap);
  else if (pl == PQ16 && ps == P16) ok = cg_iter_rc<PQ16, P16>(rc, nl, a, st, block, form, block1, cap);
  else if (pl == PQ16 && ps == P32) ok = cg_iter_rc<PQ16, P32>(rc, nl, a, st, block, form, block1, cap);
  if (!ok)
    fprintf(stderr, "qexhip: no CG instantiation for pl=%d rc=%d ps=%d rx=%d nl=%d\n", pl, rc, ps, rx, nl);
}

// niter CG iterations of form 2, 3 (launch_cg_iter each) or 4 (cg_chunk4_t,
// whose second scalar buffer is S1; g0: its streaming kernel's workgroups)
bool launch_cg_chunk(int pl, int rc, int ps, int nl, const CgArgs &a, hipStream_t st, int block, int form, int rx,
                     int block1, int niter, Scal *S1, int g0) {
  if (form != 4) {
    for (int k = 0; k < niter; k++) launch_cg_iter(pl, rc, ps, nl, a, st, block, form, rx, block1, g0);
    return true;
  }
  if (rx == P16 && ps == P16) {
    if (pl == P16) return cg_chunk4_rc<P16, P16, P16>(rc, nl, a, S1, st, block, block1, niter, g0);
    if (pl == PQ16) return cg_chunk4_rc<PQ16, P16, P16>(rc, nl, a, S1, st, block, block1, niter, g0);
    return false;
  }
  if (pl == P64 && ps == P64) return cg_chunk4_rc<P64, P64>(rc, nl, a, S1, st, block, block1, niter, g0);
  if (pl == P32 && ps == P32) return cg_chunk4_rc<P32, P32>(rc, nl, a, S1, st, block, block1, niter, g0);
  if (pl == P16 && ps == P16) return cg_chunk4_rc<P16, P16>(rc, nl, a, S1, st, block, block1, niter, g0);
  if (pl == P16 && ps == P32) return cg_chunk4_rc<P16, P32>(rc, nl, a, S1, st, block, block1, niter, g0);
  if (pl == PBF && ps == PBF) return cg_chunk4_rc<PBF, PBF>(rc, nl, a, S1, st, block, block1, niter, g0);
  if (pl == PBF && ps == P32) return cg_chunk4_rc<PBF, P32>(rc, nl, a, S1, st, block, block1, niter, g0);
  if (pl == PQ16 && ps == P16) return cg_chunk4_rc<PQ16, P16>(rc, nl, a, S1, st, block, block1, niter, g0);
  if (pl == PQ16 && ps == P32) return cg_chunk4_rc<PQ16, P32>(rc, nl, a, S1, st, block, block1, niter, g0);
  return false;
}

void launch_cg_flush(int ps, int rx, const CgArgs &a, hipStream_t st) {
  dim3 grid((unsigned)((a.g.Vh + 255) / 256)), blk(256);
  if (rx == P16 && ps == P16) {
    hipLaunchKernelGGL((k_cg_flush<P16, P16, P16, float2>), grid, blk, 0, st, a);
  } else {
    switch (ps) {
      case P64: hipLaunchKernelGGL((k_cg_flush<P64, P64, P64, double2>), grid, blk, 0, st, a); break;
      case P32: hipLaunchKernelGGL((k_cg_flush<P32, P32, P32, float2>), grid, blk, 0, st, a); break;
      case P16: hipLaunchKernelGGL((k_cg_flush<P16, P32, P32, float2>), grid, blk, 0, st, a); break;
      case PBF: hipLaunchKernelGGL((k_cg_flush<PBF, P32, P32, float2>), grid, blk, 0, st, a); break;
    }
  }
  FinArgs f{a.red, (int)grid.x, a.S, 0, 0.0, 0};
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
  // the layout each reader expects: blocked for FP64 and FP32 links
  if (nl == 4 ? link_blocked<PL, 4>() : link_blocked<PL, 8>())
    st_link_words<PL, RC, true>(out + (long)(2 * k + par) * lkblk, g.stride, i, w);
  else
    st_link_words<PL, RC, false>(out + (long)(2 * k + par) * lkblk, g.stride, i, w);
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
