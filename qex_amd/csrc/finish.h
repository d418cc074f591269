// Solver epilogues applied by the one-block finish step of every reduction
// (reduce.hip, and the batched finishes of batch.hip).
#pragma once
#include "internal.h"

namespace qh {


// FP64 true residual r = b - A x just computed (reliable update): restart the
// sloppy recursion on sigma*r with beta = |r_true|^2/|r_prev|^2 (unscaled).
// A true residual^2 more than QH_RES_GROWTH times the best one so far means
// the sloppy recursion diverged: ask the host to return x to the best iterate
// (the next call sees that residual again and accepts it), then restart the
// recursion from p = r.  After QH_MAX_RES_INCREASE such events the solve stops
// and reports the residual it has.  (CG residual norms are not monotone, so
// smaller increases are accepted.)
__device__ inline void epi_reliable(Scal *S, double s, double nreal) {
  S->r2true = s;
  if (!(s <= QH_RES_GROWTH * S->r2rel_last)) {
    S->rollback = 1;
    S->restart = 1;
    S->resinc += 1;
    S->need_rel = 0;
    S->done = S->resinc > QH_MAX_RES_INCREASE || S->its >= S->maxits;
    return;
  }
  S->rollback = 0;
  S->improved = (s <= S->r2rel_last);
  if (S->improved) S->r2rel_last = s;
  const double sig_old = S->sigma;
  const double sig_new = (s > 0.0) ? sqrt(nreal / s) : 1.0;
  const double r2old_unscaled = S->r2prev / (sig_old * sig_old);
  S->beta = (r2old_unscaled > 0.0) ? s / r2old_unscaled : 0.0;
  S->pscale = sig_new / sig_old;
  S->sigma = sig_new;
  S->r2 = s * sig_new * sig_new;
  S->r2stop_s = S->r2stop * sig_new * sig_new;
  S->maxr2 = S->r2;
  S->alpha = 0.0;
  S->since = 0;
  S->need_rel = 0;
  S->nrel += 1;
  if (S->restart) {
    S->first = 1;
    S->beta = 0.0;
    S->restart = 0;
  }
  S->done = (s <= S->r2stop) || (S->its >= S->maxits) || (S->resinc > QH_MAX_RES_INCREASE);
}

// After iteration n of the fused CG (solvers/cg.nim:174-214 recurrence):
// rr = |r_{n-1}|^2 exact, pAp = <p_n,Ap_n>, rAp = <r_{n-1},Ap_n>, ApAp = |Ap_n|^2.
//   alpha_n = rr/pAp ; |r_n|^2 = rr - 2 alpha rAp + alpha^2 ApAp ; beta_{n+1} = |r_n|^2/rr
__device__ inline void epi_cg2(Scal *S, double pAp, double rAp, double ApAp, double rr) {
  const double alpha = rr / pAp;
  double r2n = rr - 2.0 * alpha * rAp + alpha * alpha * ApAp;
  if (S->mixed && !(pAp > 0.0 && isfinite(r2n))) {
    // breakdown of the sloppy recursion (rounding made p.Ap non-positive):
    // drop this step and restart from the true residual
    S->r2prev = rr;
    S->alpha = 0.0;
    S->r2 = rr;
    S->first = 0;
    S->pscale = 1.0;
    S->its += 1;
    S->since += 1;
    S->restart = 1;
    S->need_rel = 1;
    return;
  }
  if (!(r2n > 0.0)) r2n = 0.0;
  S->r2prev = rr;
  S->pAp = pAp;
  S->alpha = alpha;
  S->flushed = 0;
  S->r2 = r2n;
  S->beta = r2n / rr;
  S->first = 0;
  S->pscale = 1.0;
  S->its += 1;
  S->since += 1;
  if (S->mixed) {
    S->need_rel = (r2n < S->delta2 * S->maxr2) || (S->since >= S->relmax) || (r2n <= S->r2stop_s) ||
                  (S->its >= S->maxits);
  } else {
    S->done = (S->its >= S->maxits) || (r2n <= S->r2stop);
  }
}

// after the pending x/r update was applied: exact |r_n|^2
__device__ inline void epi_flush(Scal *S, double rr) {
  S->r2 = rr;
  S->alpha = 0.0;
  S->beta = (S->r2prev > 0.0) ? rr / S->r2prev : 0.0;
  if (!S->mixed) S->done = (S->its >= S->maxits) || (rr <= S->r2stop);
}

// One-block finish shape: 256 threads, 8 partials per thread loaded before any
// add (2048 block partials in one trip).  Measured on the CG dot-product
// finish at 32^4: 4.7-4.8 us, against 5.5-5.6 us with 1024 threads adding one
// partial per step, 5.0 us with 128 threads and 6.5 us with one wave.  Loading
// the predicate's flags together with the partials (tested after the sum)
// gave 4.76 us, within noise; loading the whole scalar struct into thread 0's
// registers there as well, and storing it back whole, 5.68 us.
#ifndef QH_FIN_UNROLL
#define QH_FIN_UNROLL 8
#endif
#ifndef QH_FIN_THREADS
#define QH_FIN_THREADS 256
#endif

// Block total of the nb block partials of red, in a fixed order that depends
// only on nb and blockDim.x (every block running this with the same inputs
// gets the same bits); the totals land in thread 0 (and the first wave).
// lds: 16*NV doubles.
template <int NV>
__device__ __forceinline__ void sum_partials(const RedWs &red, int nb, double (&s)[NV], double *lds) {
#pragma unroll
  for (int k = 0; k < NV; k++) s[k] = 0.0;
  // QH_FIN_UNROLL strides per trip: their loads are all issued before the adds
  const int bs = blockDim.x;
  for (int j0 = threadIdx.x; j0 < nb; j0 += QH_FIN_UNROLL * bs) {
    double t[QH_FIN_UNROLL][NV];
#pragma unroll
    for (int u = 0; u < QH_FIN_UNROLL; u++) {
      const int j = j0 + u * bs;
#pragma unroll
      for (int k = 0; k < NV; k++) t[u][k] = j < nb ? red.partials[(long)k * red.maxblocks + j] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < QH_FIN_UNROLL; u++)
#pragma unroll
      for (int k = 0; k < NV; k++) s[k] += t[u][k];
  }
  block_sum<NV>(s, lds);
}

// Sum the nb block partials of F.red and apply epilogue EPI (one block).
// lds: 16*NV doubles.
template <int NV, int EPI>
__device__ __forceinline__ void finish_body(const FinArgs &F, double *lds) {
  if (F.pred == 1 && (F.S->done | F.S->need_rel)) return;
  double s[NV];
  sum_partials<NV>(F.red, F.nb, s, lds);
  if (threadIdx.x != 0) return;
  if (EPI == EPI_TMP) F.S->tmp[F.slot] = s[0];
  else if (EPI == EPI_PAP) F.S->pAp = s[0];
  else if (EPI == EPI_RELIABLE) epi_reliable(F.S, s[0], F.nreal);
  else if (EPI == EPI_CG2) epi_cg2(F.S, s[0], s[1 % NV], s[2 % NV], s[3 % NV]);
  else if (EPI == EPI_FLUSH) epi_flush(F.S, s[0]);
}

// The CG dot-product finish (EPI_CG2) on a copy of the scalars: Sout gets
// Sin with the epilogue applied, or Sin unchanged once the solve is done or
// waits for a reliable update (the double-buffered scalars of the fused CG
// form, api.hip).  One block of QH_FIN_THREADS threads.
__device__ __forceinline__ void finish_cg2_to(const RedWs &red, int nb, const Scal *Sin, Scal *Sout, double *lds) {
  double s[4];
  sum_partials<4>(red, nb, s, lds);
  if (threadIdx.x != 0) return;
  Scal sc = *Sin;
  if (!(sc.done | sc.need_rel)) epi_cg2(&sc, s[0], s[1], s[2], s[3]);
  *Sout = sc;
}

}  // namespace qh
