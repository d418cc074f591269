// Batched solves on one GPU, FP64.
//
// (1) Independent CG solves (solvers/cg.nim:55-272 each) on lattices of
//     different sizes, each with its own links, mass and source, advanced
//     together: every CG stage is ONE launch over all systems, whose blocks
//     are dealt to the systems through a table (BatchItem), and the
//     per-system dot products finish in one launch with one block per system.
//     Converged systems drop out block by block; the host captures a chunk of
//     iterations in a hipGraph and replays it.
//
// (2) Multi-shift CG (solvers/cgm.nim:68-246) for up to QH_MAXRHS right-hand
//     sides on one set of links.  Per iteration: t_j = (2D)_oe p_j and
//     Ap_j = 4 m^2 p_j - (2D)_eo t_j for all j in two multi-RHS hops that load
//     each link once for every right-hand side; alpha_j and the zeta recurrence
//     of every shift from the fused dot products <p,Ap>, <r,Ap>, |Ap|^2 (the
//     new |r|^2 extrapolated in one step, as in the fused CG); then ONE pass
//     updates r_j, every x_jk and every p_jk (x_k += alpha_k p_k and
//     p_k = zeta r + beta_k p_k share the read of p_k), reducing the exact
//     |r_j|^2 for the next alpha and the stopping test.  Shifts whose residual
//     zeta_k^2 |r|^2 met the request stop being updated (optional).
#include <stdio.h>
#include "stencil.h"
#include "finish.h"

namespace qh {

// ------------------------------------------------------- (1) CG batch
__device__ __forceinline__ int batch_item(const BatchItem *T, int n, long b) {
  int k = 0;
  while (k + 1 < n && T[k + 1].blk0 <= b) k++;
  return k;
}

template <int NL>
__global__ void __launch_bounds__(256) k_bcg1(const BatchItem *T, int n) {
  const long b = blockIdx.x;
  const BatchItem &it = T[batch_item(T, n, b)];
  cg1_body<P64, 18, P64, P64, P64, double2, NL>(it.a, b - it.blk0);
}

template <int NL>
__global__ void __launch_bounds__(256) k_bcg2(const BatchItem *T, int n) {
  __shared__ double lds[16 * 4];
  const long b = blockIdx.x;
  const BatchItem &it = T[batch_item(T, n, b)];
  cg2_body<P64, 18, P64, P64, P64, double2, NL>(it.a, b - it.blk0, lds, b);
}

// three-kernel form (stencil.h): streaming update, then one-vector gathers
__global__ void __launch_bounds__(256) k_bcg0(const BatchItem *T, int n) {
  const long b = blockIdx.x;
  const BatchItem &it = T[batch_item(T, n, b)];
  cg0_body<P64, P64, P64, double2>(it.a, b - it.blk0);
}

template <int NL>
__global__ void __launch_bounds__(256) k_bcg1b(const BatchItem *T, int n) {
  const long b = blockIdx.x;
  const BatchItem &it = T[batch_item(T, n, b)];
  cg1b_body<P64, 18, P64, double2, NL>(it.a, b - it.blk0);
}

template <int NL>
__global__ void __launch_bounds__(256) k_bcg2b(const BatchItem *T, int n) {
  __shared__ double lds[16 * 4];
  const long b = blockIdx.x;
  const BatchItem &it = T[batch_item(T, n, b)];
  cg2b_body<P64, 18, P64, P64, double2, NL>(it.a, b - it.blk0, lds, b);
}

__global__ void __launch_bounds__(1024) k_bfin_cg2(const BatchItem *T) {
  __shared__ double lds[16 * 4];
  finish_body<4, EPI_CG2>(T[blockIdx.x].fin, lds);
}

// apply the pending update of systems that just converged (once each)
__global__ void __launch_bounds__(256) k_bflush(const BatchItem *T, int n) {
  __shared__ double lds[16];
  const long b = blockIdx.x;
  const BatchItem &it = T[batch_item(T, n, b)];
  const Scal *S = it.a.S;
  if (!S->done || S->flushed) return;
  cg_flush_body<P64, P64, P64, double2>(it.a, b - it.blk0, lds, b);
}

__global__ void __launch_bounds__(1024) k_bfin_flush(const BatchItem *T) {
  __shared__ double lds[16];
  const BatchItem &it = T[blockIdx.x];
  Scal *S = it.a.S;
  const bool run = S->done && !S->flushed;  // read by every thread before any write
  __syncthreads();
  if (!run) return;
  finish_body<1, EPI_FLUSH>(it.finf, lds);
  if (threadIdx.x == 0) S->flushed = 1;
}

void launch_bcg_iter(const BatchItem *T, int n, long nblocks, int nl, int block, hipStream_t st, int form) {
  const dim3 grid((unsigned)nblocks), blk(block);
  if (form == 3) {
    hipLaunchKernelGGL(k_bcg0, grid, blk, 0, st, T, n);
    if (nl == 8) {
      hipLaunchKernelGGL(k_bcg1b<8>, grid, blk, 0, st, T, n);
      hipLaunchKernelGGL(k_bcg2b<8>, grid, blk, 0, st, T, n);
    } else {
      hipLaunchKernelGGL(k_bcg1b<4>, grid, blk, 0, st, T, n);
      hipLaunchKernelGGL(k_bcg2b<4>, grid, blk, 0, st, T, n);
    }
  } else if (nl == 8) {
    hipLaunchKernelGGL(k_bcg1<8>, grid, blk, 0, st, T, n);
    hipLaunchKernelGGL(k_bcg2<8>, grid, blk, 0, st, T, n);
  } else {
    hipLaunchKernelGGL(k_bcg1<4>, grid, blk, 0, st, T, n);
    hipLaunchKernelGGL(k_bcg2<4>, grid, blk, 0, st, T, n);
  }
  hipLaunchKernelGGL(k_bfin_cg2, dim3(n), dim3(1024), 0, st, T);
}

void launch_bcg_flush(const BatchItem *T, int n, long nblocks, int block, hipStream_t st) {
  hipLaunchKernelGGL(k_bflush, dim3((unsigned)nblocks), dim3(block), 0, st, T, n);
  hipLaunchKernelGGL(k_bfin_flush, dim3(n), dim3(1024), 0, st, T);
}

// --------------------------------------------- (2) multi-shift, multi-RHS

// MODE 0: t_j = (2D) p0_j on parity q (x = site of parity q).
// MODE 1: Ap_j = m4 p0_j - (2D) t_j on parity q with the partial sums of
//         <p0_j, Ap_j>, <r_j, Ap_j>, |Ap_j|^2 (slots 3j, 3j+1, 3j+2).
// Each link is loaded once and applied to the NR right-hand sides.  The
// vector loads of every right-hand side are issued before the link's
// arithmetic (as in k_hop_multi): loads and arithmetic run for all NR, and a
// right-hand side that has converged is left out only at the stores and sums
// (its vectors stay allocated in the workspace, so the loads stay in bounds).
// Skipping it inside the loop made every right-hand side's loads wait for the
// previous one's.  Sites are taken in the walk of make_walk (chunk 0: the
// canonical order; else each XCD runs a chunk of a t-slice through all t, so
// the t +-1, +-3 neighbours of the 8 gathered vectors are read while the
// same chunk's nearby t-slices are in flight); a workgroup may serve several
// logical blocks and writes one set of partial sums.
template <int PV, class A, int NL, int NR, int MODE, int NB>
__global__ void __launch_bounds__(256) k_mrhs_hop(Geom g, const char *links, long lkblk, int q, double m4,
                                                 MsbPtrs P, const MsbScal *T, RedWs red, long chunk) {
  static_assert(NR % NB == 0, "NB must divide NR");
  using R = RT<A>;
  __shared__ double lds[16 * 3 * NR];
  const long spb = blockDim.x;
  const Walk wk = make_walk(g, spb, chunk);
  // one logical block per workgroup (the launch has a partial-sum slot for each)
  const BlockIter it = block_iter(wk.nblk, 1);
  const long i = it.cur < it.end ? walk_site(wk, it.cur, spb, threadIdx.x) : -1;
  A acc[NR][3];
  if (i >= 0 && i < g.Vh) {
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) acc[r][c] = mk<A>(0, 0);
    const Site s = decode(g, i, q);
    // one link j, loaded non-temporally when ntc is true (link_cached_blocks)
    auto trip = [&](int j, auto ntc) {
      const int k = j >> 1, fwd = !(j & 1);
      const int mu = (NL == 4) ? k : (k >> 1);
      const int len = (NL == 8 && (k & 1)) ? 3 : 1;
      const long n = nbr(g, s, mu, fwd ? len : -len);
      A U[9];
      ld_link<PV, 18, decltype(ntc)::value, A, link_blocked<PV, NL>()>(
          links + (long)(2 * k + (fwd ? q : 1 - q)) * lkblk, g.stride, fwd ? i : n, U, (R)1.0);
      // the right-hand sides' vectors in batches of NB, each batch's loads
      // issued before its arithmetic
#pragma unroll
      for (int r0 = 0; r0 < NR; r0 += NB) {
        A v[NB][3];
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
          for (int c = 0; c < 3; c++)
            v[b][c] = ldc<PV, A>(MODE == 0 ? (const void *)P.p0[r0 + b] : P.t[r0 + b], vix(g, c, n));
        if (fwd) {
#pragma unroll
          for (int b = 0; b < NB; b++)
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
              for (int c = 0; c < 3; c++) cmac(acc[r0 + b][a], U[3 * a + c], v[b][c]);
        } else {
#pragma unroll
          for (int b = 0; b < NB; b++)
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
              for (int c = 0; c < 3; c++) cjmsub(acc[r0 + b][a], U[3 * c + a], v[b][c]);
        }
      }
    };
    // the first link_cached_blocks blocks with cache-allocating loads (block
    // 2k + q holds direction k's forward link, 2k + 1 - q its backward one)
    const int mb = link_cached_blocks<NL>(g, lkblk), kf = mb >> 1;
#pragma unroll 1
    for (int j = 0; j < 2 * kf; j++) trip(j, std::false_type());
    int j1 = 2 * kf;
    if (mb & 1) {
      if (q & 1) {
        trip(j1, std::true_type());
        trip(j1 + 1, std::false_type());
      } else {
        trip(j1, std::false_type());
        trip(j1 + 1, std::true_type());
      }
      j1 += 2;
    }
#pragma unroll 1
    for (int j = j1; j < 2 * NL; j++) trip(j, std::true_type());
  }
  if constexpr (MODE == 0) {
    if (i >= 0 && i < g.Vh) {
#pragma unroll
      for (int r = 0; r < NR; r++) {
        if (!T[r].continuing) continue;
#pragma unroll
        for (int c = 0; c < 3; c++) stc<PV, A>(P.t[r], vix(g, c, i), acc[r][c]);
      }
    }
  } else {
    double sums[3 * NR];
#pragma unroll
    for (int k = 0; k < 3 * NR; k++) sums[k] = 0.0;
    if (i >= 0 && i < g.Vh) {
#pragma unroll
      for (int r = 0; r < NR; r++) {
        if (!T[r].continuing) continue;
        const double *rr = P.r[r] + (long)(T[r].its % QH_MSB_RING) * P.rstride;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const long o = vix(g, c, i);
          const A pv = ldc<PV, A>(P.p0[r], o), rv = ldc<PV, A>(rr, o);
          const A av = rt_store<PV, A>(mk<A>((R)m4 * pv.x - acc[r][c].x, (R)m4 * pv.y - acc[r][c].y));
          stc<PV, A>(P.Ap[r], o, av);
          sums[3 * r] += (double)pv.x * (double)av.x + (double)pv.y * (double)av.y;
          sums[3 * r + 1] += (double)rv.x * (double)av.x + (double)rv.y * (double)av.y;
          sums[3 * r + 2] += (double)av.x * (double)av.x + (double)av.y * (double)av.y;
        }
      }
    }
    block_partials<3 * NR>(sums, lds, red);
  }
}

// Multi-RHS stagD2 (QEX's stagDPN/stagDN, stagD.nim:132-199,451-456): on parity
// par, r_j = a r_j + b x_j + (2D) x_j for NR fields, each link loaded once for
// all of them.  FP64 (A = double2, P64 links) or FP32 (float2, P32 links).
// One thread per site.  The vectors of NB right-hand sides are loaded
// together, before their arithmetic: NB = NR puts every load of a link in
// flight at once, NB = NR/2 halves the vector registers.  (Loading one
// right-hand side at a time and waiting for it made the 8-field FP64 hop
// latency-bound: 41 us per field at 32^4, 29.4 us with all loads in flight.)
// The sites follow the t-march walk of chunk (0: canonical order): with 8
// fields the inputs otherwise come back from memory 3.3 times each at 32^4.
template <int PL, class A, int NL, int NR, int NB>
__global__ void __launch_bounds__(256) k_hop_multi(Geom g, const char *links, long lkblk, int par, double ca,
                                                  double cb, MultiPtrs P, long chunk) {
  using R = RT<A>;
  static_assert(NR % NB == 0, "NB must divide NR");
  const long spb = blockDim.x;
  const Walk wk = make_walk(g, spb, chunk);
  const BlockIter it = block_iter(wk.nblk, 1);
#pragma unroll 1
  for (long lb = it.cur; lb < it.end; lb += it.step) {
    const long i = walk_site(wk, lb, spb, threadIdx.x);
    if (i < 0 || i >= g.Vh) continue;
    const Site s = decode(g, i, par);
    A acc[NR][3];
#pragma unroll
    for (int q = 0; q < NR; q++)
#pragma unroll
      for (int c = 0; c < 3; c++) acc[q][c] = mk<A>(0, 0);
    // one link j, loaded non-temporally when ntc is true (link_cached_blocks)
    auto trip = [&](int j, auto ntc) {
      const int k = j >> 1;
      const bool fwd = !(j & 1);
      const int mu = (NL == 4) ? k : (k >> 1);
      const int len = (NL == 8 && (k & 1)) ? 3 : 1;
      const long n = nbr(g, s, mu, fwd ? len : -len);
      A U[9];
      ld_link<PL, 18, decltype(ntc)::value, A, link_blocked<PL, NL>()>(
          links + (long)(2 * k + (fwd ? par : 1 - par)) * lkblk, g.stride, fwd ? i : n, U, (R)1.0);
#pragma unroll
      for (int q0 = 0; q0 < NR; q0 += NB) {
        A v[NB][3];
#pragma unroll
        for (int q = 0; q < NB; q++)
#pragma unroll
          for (int c = 0; c < 3; c++) v[q][c] = ldc<PL, A>(P.in[q0 + q], vix(g, c, n));
        if (fwd) {
#pragma unroll
          for (int q = 0; q < NB; q++)
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
              for (int c = 0; c < 3; c++) cmac(acc[q0 + q][a], U[3 * a + c], v[q][c]);
        } else {
#pragma unroll
          for (int q = 0; q < NB; q++)
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
              for (int c = 0; c < 3; c++) cjmsub(acc[q0 + q][a], U[3 * c + a], v[q][c]);
        }
      }
    };
    // the first link_cached_blocks blocks with cache-allocating loads
    const int mb = link_cached_blocks<NL>(g, lkblk), kf = mb >> 1;
#pragma unroll 1
    for (int j = 0; j < 2 * kf; j++) trip(j, std::false_type());
    int j1 = 2 * kf;
    if (mb & 1) {
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
    for (int j = j1; j < 2 * NL; j++) trip(j, std::true_type());
#pragma unroll
    for (int q = 0; q < NR; q++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const long o = vix(g, c, i);
        A v = acc[q][c];
        if (cb != 0.0) {
          const A x = ldc<PL, A>(P.own[q], o);
          v.x += (R)cb * x.x;
          v.y += (R)cb * x.y;
        }
        if (ca != 0.0) {
          const A y = ldc<PL, A>(P.out[q], o);
          v.x += (R)ca * y.x;
          v.y += (R)ca * y.y;
        }
        stc<PL, A>(P.out[q], o, v);
      }
  }
}

// batches: the vectors of all (1) or half (2) of the right-hand sides loaded
// together per link
template <int PL, class A, int NL, int NR>
static bool hop_multi_nb(int batches, const Geom &g, const char *links, long lkblk, int par, double ca, double cb,
                          const MultiPtrs &P, int block, int cap, long chunk, hipStream_t st) {
  const Walk wk = make_walk(g, block, chunk);
  const dim3 grid(grid_blocks(wk.nblk, 1, cap)), blk(block);
  if (batches == 2) {
    if constexpr (NR >= 2) {
      hipLaunchKernelGGL((k_hop_multi<PL, A, NL, NR, NR / 2>), grid, blk, 0, st, g, links, lkblk, par, ca, cb, P, chunk);
      return true;
    }
    return false;
  }
  hipLaunchKernelGGL((k_hop_multi<PL, A, NL, NR, NR>), grid, blk, 0, st, g, links, lkblk, par, ca, cb, P, chunk);
  return true;
}

template <int PL, class A, int NL>
static bool hop_multi_nr(int nr, int batches, const Geom &g, const char *links, long lkblk, int par, double ca,
                         double cb, const MultiPtrs &P, int block, int cap, long chunk, hipStream_t st) {
  switch (nr) {
    case 1: return hop_multi_nb<PL, A, NL, 1>(batches, g, links, lkblk, par, ca, cb, P, block, cap, chunk, st);
    case 2: return hop_multi_nb<PL, A, NL, 2>(batches, g, links, lkblk, par, ca, cb, P, block, cap, chunk, st);
    case 4: return hop_multi_nb<PL, A, NL, 4>(batches, g, links, lkblk, par, ca, cb, P, block, cap, chunk, st);
    case 8: return hop_multi_nb<PL, A, NL, 8>(batches, g, links, lkblk, par, ca, cb, P, block, cap, chunk, st);
  }
  return false;
}

bool launch_hop_multi(int prec, int nl, int nr, int batches, const Geom &g, const char *links, long lkblk, int par,
                      double ca, double cb, const MultiPtrs &P, int block, int cap, long chunk, hipStream_t st) {
  bool ok = false;
  if (prec == P64)
    ok = (nl == 8) ? hop_multi_nr<P64, double2, 8>(nr, batches, g, links, lkblk, par, ca, cb, P, block, cap, chunk, st)
                   : hop_multi_nr<P64, double2, 4>(nr, batches, g, links, lkblk, par, ca, cb, P, block, cap, chunk, st);
  else if (prec == P32)
    ok = (nl == 8) ? hop_multi_nr<P32, float2, 8>(nr, batches, g, links, lkblk, par, ca, cb, P, block, cap, chunk, st)
                   : hop_multi_nr<P32, float2, 4>(nr, batches, g, links, lkblk, par, ca, cb, P, block, cap, chunk, st);
  return ok;
}

// After the MODE-1 hop: alpha_j, the extrapolated |r_j|^2, beta_j and the
// update coefficients of every shift (cgm.nim:163-199).  One thread per
// (right-hand side, shift): every thread reads the old per-source scalars,
// then (after a barrier) shift 0's thread writes the new ones.
template <int NR>
__global__ void __launch_bounds__(1024) k_msb_fin1(RedWs red, int nb, MsbScal *T, double *hist) {
  __shared__ double lds[16 * 3 * NR];
  __shared__ double tot[3 * NR];
  double s[3 * NR];
#pragma unroll
  for (int k = 0; k < 3 * NR; k++) s[k] = 0.0;
  for (int j = threadIdx.x; j < nb; j += blockDim.x)
#pragma unroll
    for (int k = 0; k < 3 * NR; k++) s[k] += red.partials[(long)k * red.maxblocks + j];
  block_sum<3 * NR>(s, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 3 * NR; k++) tot[k] = s[k];
  __syncthreads();
  const int r = threadIdx.x / QH_MAXSHIFT, k = threadIdx.x % QH_MAXSHIFT;
  const bool mine = r < NR;
  MsbScal *t = T + (mine ? r : 0);
  const bool cont = mine && t->continuing && k < t->nshift;
  double alpha = 0.0, beta = 0.0, aim1 = 0.0, bim1 = 0.0;
  int jw = 0;  // this iteration's index in the window
  if (cont) {
    jw = t->its - t->its0;
    const double pAp = tot[3 * r], rAp = tot[3 * r + 1], ApAp = tot[3 * r + 2];
    alpha = (pAp != 0.0) ? t->r2i / pAp : 0.0;
    double r2n = t->r2i - 2.0 * alpha * rAp + alpha * alpha * ApAp;
    if (!(r2n > 0.0)) r2n = 0.0;
    beta = (t->r2i != 0.0) ? r2n / t->r2i : 0.0;
    aim1 = t->alphaim1;
    bim1 = t->betaim1;
  }
  __syncthreads();
  if (!cont) return;
  // (xc, pc1, pc2) of this iteration: x_k += xc p_k ; p_k = pc1 r_new + pc2 p_k ; a
  // frozen shift records the identity (0, 0, 1)
  double xc = 0.0, c1 = 0.0, c2 = 1.0;
  if (k == 0) {
    t->alpha = alpha;
    t->beta = beta;
    t->xc[0] = alpha;
    t->pc1[0] = 1.0;
    t->pc2[0] = beta;
    t->alphaim1 = alpha;
    t->betaim1 = beta;
    t->its += 1;
    xc = alpha;
    c1 = 1.0;
    c2 = beta;
  } else if (t->active[k]) {
    double zip1d = alpha * bim1 * (t->zim1[k] - t->zi[k]);
    zip1d += t->zim1[k] * aim1 * (1.0 + t->sg[k] * alpha);
    const double zip1 = (zip1d != 0.0) ? t->zi[k] * t->zim1[k] * aim1 / zip1d : 0.0;
    const double zr = (t->zi[k] != 0.0) ? zip1 / t->zi[k] : 0.0;
    t->xc[k] = alpha * zr;
    t->pc1[k] = zip1;
    t->pc2[k] = beta * zr * zr;
    t->zim1[k] = t->zi[k];
    t->zi[k] = zip1;
    xc = t->xc[k];
    c1 = zip1;
    c2 = t->pc2[k];
  }
  if (jw < QH_MSB_WIN) {
    double *h = hist + (((long)r * QH_MSB_WIN + jw) * t->nshift + k) * 3;
    h[0] = xc;
    h[1] = c1;
    h[2] = c2;
  } else if (k == 0) {
    t->winflow = 1;
  }
}

// v as storage precision PV holds it (FP32 storage rounds; FP64 is exact)
template <int PV> __device__ __forceinline__ double2 as_stored(double2 v) {
  if constexpr (PV == P32) return make_double2((double)(float)v.x, (double)(float)v.y);
  return v;
}

// One iteration's vector update of every right-hand side:
//   r_new = r - alpha Ap into the next ring slot, p0 = r_new + beta p0 (p0 saved
//   at the first iteration of a window), partial sums of the exact |r_new|^2.
// The shifted x_k, p_k (and x_0) wait for the end of the window (k_msb_flush).
template <int PV, int NR>
__global__ void __launch_bounds__(256) k_msb_step(Geom g, MsbPtrs P, const MsbScal *T, RedWs red) {
  __shared__ double lds[16 * NR];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double rr[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) rr[r] = 0.0;
  if (i < g.Vh) {
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const MsbScal *t = T + r;
      if (!t->continuing) continue;
      const int it = t->its - 1;  // this iteration (fin1 counted it already)
      const double alpha = t->alpha, beta = t->beta;
      const bool first = (it == t->its0);
      const double *rv = P.r[r] + (long)(it % QH_MSB_RING) * P.rstride;
      double *rw = P.r[r] + (long)((it + 1) % QH_MSB_RING) * P.rstride;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const long o = vix(g, c, i);
        const double2 a = ldc<PV, double2>(P.Ap[r], o), x = ldc<PV, double2>(rv, o), pv = ldc<PV, double2>(P.p0[r], o);
        // r_new as stored (its norm is that of the stored vector)
        const double2 rn = as_stored<PV>(make_double2(x.x - alpha * a.x, x.y - alpha * a.y));
        stc<PV, double2>(rw, o, rn);
        rr[r] += rn.x * rn.x + rn.y * rn.y;
        if (first) stc<PV, double2>(P.p0s[r], o, pv);
        stc<PV, double2>(P.p0[r], o, make_double2(rn.x + beta * pv.x, rn.y + beta * pv.y));
      }
    }
  }
  block_partials<NR>(rr, lds, red);
}

// Window coefficients, one thread per (right-hand side, shift).  Over the w
// iterations j = 0..w-1 of a window starting at x, p with residuals
// r_1..r_w after each iteration:
//   p_w = C p + sum_{i=1..w} D_i r_i ,  x_w = x + E p + sum_{i=1..w-1} F_i r_i
// with H_j = prod_{l=j..w-1} pc2_l, G_j = xc_j + pc2_j G_{j+1} (G_w = 0):
// C = H_0, E = G_0, D_i = pc1_{i-1} H_i, F_i = pc1_{i-1} G_i.  The window then
// restarts at the current iteration.
template <int NR>
__global__ void __launch_bounds__(NR * QH_MAXSHIFT) k_msb_coef(MsbPtrs P, MsbScal *T) {
  const int r = threadIdx.x / QH_MAXSHIFT, k = threadIdx.x % QH_MAXSHIFT;
  MsbScal *t = T + r;
  const int w = t->its - t->its0, its0 = t->its0, ns = t->nshift;
  __syncthreads();
  int *ci = P.cinfo + r * (2 + QH_MAXSHIFT);
  if (k == 0) {
    ci[0] = w < QH_MSB_WIN ? w : QH_MSB_WIN;
    ci[1] = its0 % QH_MSB_RING;
    t->its0 = t->its;
  }
  if (k >= ns) return;
  const int wl = w < QH_MSB_WIN ? w : QH_MSB_WIN;
  const double *h = P.hist + (long)r * QH_MSB_WIN * ns * 3;
  double *cf = P.coef + ((long)r * ns + k) * QH_MSB_NCOEF;
  double G = 0.0, H = 1.0;
  bool any = false;
  for (int j = wl; j >= 0; j--) {
    if (j < wl) {
      const double xc = h[((long)j * ns + k) * 3], c2 = h[((long)j * ns + k) * 3 + 2];
      any = any || xc != 0.0 || h[((long)j * ns + k) * 3 + 1] != 0.0 || c2 != 1.0;
      G = xc + c2 * G;
      H = c2 * H;
    }
    if (j >= 1) {
      const double c1 = h[((long)(j - 1) * ns + k) * 3 + 1];
      cf[2 + j - 1] = c1 * G;                // F_j (F_w unused)
      cf[2 + QH_MSB_WIN + j - 1] = c1 * H;   // D_j
    } else {
      cf[0] = G;  // E
      cf[1] = H;  // C
    }
  }
  ci[2 + k] = any ? 1 : 0;
}

// x_k, p_k of every shift advanced over the window (x_0 only: p_0 was kept up
// to date every iteration; its window-start copy p0s stands for p).  One color
// component at a time, so that the accumulators of up to QH_MSB_FLUSH_K shifts
// (4 doubles each) stay in registers and each residual of the window is read
// once per component for all of them.
#define QH_MSB_FLUSH_K 16
template <int PV, int NR>
__global__ void __launch_bounds__(256) k_msb_flush(Geom g, int nshift, MsbPtrs P) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
#pragma unroll 1
  for (int r = 0; r < NR; r++) {
    const int *ci = P.cinfo + r * (2 + QH_MAXSHIFT);
    const int w = ci[0], base = ci[1];
    if (w == 0) continue;
    const double *ring = P.r[r];
#pragma unroll 1
    for (int k0 = 0; k0 < nshift; k0 += QH_MSB_FLUSH_K) {
      bool act[QH_MSB_FLUSH_K];
      bool anyg = false;
#pragma unroll
      for (int q = 0; q < QH_MSB_FLUSH_K; q++) {
        act[q] = (k0 + q < nshift) && ci[2 + k0 + q];
        anyg = anyg || act[q];
      }
      if (!anyg) continue;
#pragma unroll 1
      for (int c = 0; c < 3; c++) {
        const long o = vix(g, c, i);
        double2 ax[QH_MSB_FLUSH_K], ap[QH_MSB_FLUSH_K];
#pragma unroll
        for (int q = 0; q < QH_MSB_FLUSH_K; q++) {
          if (!act[q]) continue;
          const int k = k0 + q;
          const double *cf = P.coef + ((long)r * nshift + k) * QH_MSB_NCOEF;
          const double2 xv = ldc<PV, double2>(P.x[r * nshift + k], o);
          const double2 pv = ldc<PV, double2>(k == 0 ? P.p0s[r] : P.p[r * nshift + k], o);
          ax[q] = make_double2(xv.x + cf[0] * pv.x, xv.y + cf[0] * pv.y);
          ap[q] = make_double2(cf[1] * pv.x, cf[1] * pv.y);
        }
#pragma unroll 1
        for (int s = 1; s <= w; s++) {
          const double2 rv = ldc<PV, double2>(ring + (long)((base + s) % QH_MSB_RING) * P.rstride, o);
#pragma unroll
          for (int q = 0; q < QH_MSB_FLUSH_K; q++) {
            if (!act[q]) continue;
            const double *cf = P.coef + ((long)r * nshift + k0 + q) * QH_MSB_NCOEF;
            const double F = (s < w) ? cf[2 + s - 1] : 0.0, D = cf[2 + QH_MSB_WIN + s - 1];
            ax[q].x += F * rv.x;
            ax[q].y += F * rv.y;
            ap[q].x += D * rv.x;
            ap[q].y += D * rv.y;
          }
        }
#pragma unroll
        for (int q = 0; q < QH_MSB_FLUSH_K; q++) {
          if (!act[q]) continue;
          const int k = k0 + q;
          stc<PV, double2>(P.x[r * nshift + k], o, ax[q]);
          if (k > 0) stc<PV, double2>(P.p[r * nshift + k], o, ap[q]);
        }
      }
    }
  }
}

// exact |r_j|^2: next alpha, stopping test, shift freezing (cgm.nim:176-186);
// one thread per (right-hand side, shift), reads before the barrier, writes after
template <int NR>
__global__ void __launch_bounds__(1024) k_msb_fin2(RedWs red, int nb, MsbScal *T) {
  __shared__ double lds[16 * NR];
  __shared__ double tot[NR];
  double s[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) s[k] = 0.0;
  for (int j = threadIdx.x; j < nb; j += blockDim.x)
#pragma unroll
    for (int k = 0; k < NR; k++) s[k] += red.partials[(long)k * red.maxblocks + j];
  block_sum<NR>(s, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NR; k++) tot[k] = s[k];
  __syncthreads();
  const int r = threadIdx.x / QH_MAXSHIFT, k = threadIdx.x % QH_MAXSHIFT;
  const bool mine = r < NR;
  MsbScal *t = T + (mine ? r : 0);
  const bool was = mine && t->continuing && k < t->nshift;
  bool cont = false, freeze_k = false;
  double r2 = 0.0;
  if (was) {
    r2 = tot[r];
    cont = (t->its < t->maxits) && (r2 > t->r2stop);
    freeze_k = !cont || (t->freeze && k >= 1 && t->zi[k] * t->zi[k] * r2 <= t->r2stop);
  }
  __syncthreads();
  if (!was) return;
  if (freeze_k) t->active[k] = 0;
  if (k == 0) {
    t->r2i = r2;
    t->continuing = cont;
  }
}

// Workgroups of the multi-RHS hops for chunk (see k_mrhs_hop): one per logical
// block, rounded up to 8 for the XCD order.  Fewer than 2 Vh/block + 8: a
// walk pads each t-slice to whole chunks, less than one t-slice's blocks per t.
unsigned mrhs_grid(const Geom &g, int block, long chunk) {
  const Walk wk = make_walk(g, block, chunk);
  return grid_blocks(wk.nblk, 1, 0);
}

// PV: vector storage and link precision of the hops (P64 with FP64
// arithmetic, or P32 with FP32 stencil arithmetic); A its arithmetic type
template <int PV, class A, int NL, int NR, int NB>
static void msb_hops_t(const Geom &g, const char *links, long lkblk, int par, double m2, const MsbPtrs &P,
                       const MsbScal *T, RedWs red, dim3 grid, dim3 blk, long chunk, hipStream_t st) {
  hipLaunchKernelGGL((k_mrhs_hop<PV, A, NL, NR, 0, NB>), grid, blk, 0, st, g, links, lkblk, 1 - par, 0.0, P, T, red,
                     chunk);
  hipLaunchKernelGGL((k_mrhs_hop<PV, A, NL, NR, 1, NB>), grid, blk, 0, st, g, links, lkblk, par, 4.0 * m2, P, T, red,
                     chunk);
}

template <int PV, class A, int NL, int NR>
static void msb_iter_t(const Geom &g, const char *links, long lkblk, int par, int nshift, double m2,
                       const MsbPtrs &P, MsbScal *T, RedWs red, int block, long chunk, int batches, hipStream_t st) {
  const unsigned nb = mrhs_grid(g, block, chunk);
  if (nb > (unsigned)red.maxblocks) {
    fprintf(stderr, "qexhip: multi-RHS hop grid %u above the %d partial-sum slots\n", nb, red.maxblocks);
    return;
  }
  const dim3 grid(nb), blk(block);
  if (batches >= 4 && NR >= 4)
    msb_hops_t<PV, A, NL, NR, (NR >= 4 ? NR / 4 : 1)>(g, links, lkblk, par, m2, P, T, red, grid, blk, chunk, st);
  else if (batches >= 2 && NR >= 2)
    msb_hops_t<PV, A, NL, NR, (NR >= 2 ? NR / 2 : 1)>(g, links, lkblk, par, m2, P, T, red, grid, blk, chunk, st);
  else
    msb_hops_t<PV, A, NL, NR, NR>(g, links, lkblk, par, m2, P, T, red, grid, blk, chunk, st);
  hipLaunchKernelGGL((k_msb_fin1<NR>), dim3(1), dim3(1024), 0, st, red, (int)nb, T, P.hist);
  hipLaunchKernelGGL((k_msb_step<PV, NR>), dim3((unsigned)((g.Vh + 255) / 256)), dim3(256), 0, st, g, P, T, red);
  hipLaunchKernelGGL((k_msb_fin2<NR>), dim3(1), dim3(1024), 0, st, red, (int)((g.Vh + 255) / 256), T);
}

// FP64: 1, 2, 4 or 8 right-hand sides; FP32 storage (the mixed solve's first
// phase): one
template <int NL>
static bool msb_nr(int pv, int nr, const Geom &g, const char *links, long lkblk, int par, int nshift, double m2,
                   const MsbPtrs &P, MsbScal *T, RedWs red, int block, long chunk, int batches, hipStream_t st) {
  if (pv == P32) {
    if (nr != 1) return false;
    msb_iter_t<P32, float2, NL, 1>(g, links, lkblk, par, nshift, m2, P, T, red, block, chunk, batches, st);
    return true;
  }
  if (pv != P64) return false;
  switch (nr) {
    case 1: msb_iter_t<P64, double2, NL, 1>(g, links, lkblk, par, nshift, m2, P, T, red, block, chunk, batches, st); return true;
    case 2: msb_iter_t<P64, double2, NL, 2>(g, links, lkblk, par, nshift, m2, P, T, red, block, chunk, batches, st); return true;
    case 4: msb_iter_t<P64, double2, NL, 4>(g, links, lkblk, par, nshift, m2, P, T, red, block, chunk, batches, st); return true;
    case 8: msb_iter_t<P64, double2, NL, 8>(g, links, lkblk, par, nshift, m2, P, T, red, block, chunk, batches, st); return true;
  }
  return false;
}

template <int PV, int NR>
static void msb_flush_t(const Geom &g, int nshift, const MsbPtrs &P, MsbScal *T, hipStream_t st) {
  hipLaunchKernelGGL((k_msb_coef<NR>), dim3(1), dim3(NR * QH_MAXSHIFT), 0, st, P, T);
  hipLaunchKernelGGL((k_msb_flush<PV, NR>), dim3((unsigned)((g.Vh + 255) / 256)), dim3(256), 0, st, g, nshift, P);
}

void launch_msb_flush(const Geom &g, int pv, int nr, int nshift, const MsbPtrs &P, MsbScal *T, hipStream_t st) {
  if (pv == P32 && nr == 1) {
    msb_flush_t<P32, 1>(g, nshift, P, T, st);
    return;
  }
  if (pv != P64) {
    fprintf(stderr, "qexhip: no multi-shift flush for storage %d, nr=%d\n", pv, nr);
    return;
  }
  switch (nr) {
    case 1: msb_flush_t<P64, 1>(g, nshift, P, T, st); break;
    case 2: msb_flush_t<P64, 2>(g, nshift, P, T, st); break;
    case 4: msb_flush_t<P64, 4>(g, nshift, P, T, st); break;
    case 8: msb_flush_t<P64, 8>(g, nshift, P, T, st); break;
    default: fprintf(stderr, "qexhip: no multi-shift flush for nr=%d\n", nr);
  }
}

void launch_msb_iter(const Geom &g, int pv, const char *links, long lkblk, int nl, int par, int nr, int nshift,
                     double m2, const MsbPtrs &P, MsbScal *T, RedWs red, int block, long chunk, int batches,
                     hipStream_t st) {
  const bool ok =
      (nl == 8) ? msb_nr<8>(pv, nr, g, links, lkblk, par, nshift, m2, P, T, red, block, chunk, batches, st)
                : msb_nr<4>(pv, nr, g, links, lkblk, par, nshift, m2, P, T, red, block, chunk, batches, st);
  if (!ok) fprintf(stderr, "qexhip: no multi-RHS instantiation for storage %d, nr=%d, nl=%d\n", pv, nr, nl);
}

}  // namespace qh
