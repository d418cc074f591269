// Finish step of every reduction: one block adds the block partials of the
// kernel launched just before it (stream order makes them visible) in a fixed
// order, so results are bitwise reproducible, then applies the solver
// epilogue to the device-resident scalars.
#include "finish.h"

namespace qh {

template <int NV, int EPI>
__global__ void __launch_bounds__(QH_FIN_THREADS) k_finish(FinArgs F) {
  __shared__ double lds[16 * NV];
  finish_body<NV, EPI>(F, lds);
}

__global__ void __launch_bounds__(QH_FIN_THREADS) k_finish_cg2_to(RedWs red, int nb, const Scal *Sin, Scal *Sout) {
  __shared__ double lds[16 * 4];
  finish_cg2_to(red, nb, Sin, Sout, lds);
}

void launch_finish_cg2_to(const RedWs &red, int nb, const Scal *Sin, Scal *Sout, hipStream_t st) {
  hipLaunchKernelGGL(k_finish_cg2_to, dim3(1), dim3(QH_FIN_THREADS), 0, st, red, nb, Sin, Sout);
}

void launch_finish(int epi, const FinArgs &f, hipStream_t st) {
  const dim3 g(1), b(QH_FIN_THREADS);
  switch (epi) {
    case EPI_TMP: hipLaunchKernelGGL((k_finish<1, EPI_TMP>), g, b, 0, st, f); break;
    case EPI_PAP: hipLaunchKernelGGL((k_finish<1, EPI_PAP>), g, b, 0, st, f); break;
    case EPI_RELIABLE: hipLaunchKernelGGL((k_finish<1, EPI_RELIABLE>), g, b, 0, st, f); break;
    case EPI_CG2: hipLaunchKernelGGL((k_finish<4, EPI_CG2>), g, b, 0, st, f); break;
    case EPI_FLUSH: hipLaunchKernelGGL((k_finish<1, EPI_FLUSH>), g, b, 0, st, f); break;
  }
}

}  // namespace qh
