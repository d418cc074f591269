// Gauge-field kernels around the fermion force for HMC, FP64, on canonical
// link arrays U[mu][site][3][3] (QEX's single-rank site order):
//  - Wilson plaquette sum  sum_{x,mu<nu} Re tr U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
//    (the action S = -(beta/3) * that sum, QEX gauge/gaugeAction.nim:397-427 with
//    plaq coefficient beta);
//  - its force in the fermion force's convention: for U_mu(x) -> exp(w X) U_mu(x),
//    dS/dw = -1/2 Re tr(X F), so F_mu(x) = (2 beta/3) TAH(U_mu(x) A_mu(x)) with A the
//    sum of the six staples (QEX gauge/gaugeAction.nim:765-781 computes the same
//    derivative through its loop machinery);
//  - traceless anti-Hermitian Gaussian momenta, density exp(tr P^2 / 2) (the role
//    of QEX's randomTAH, gauge/gaugeUtils.nim:1474-1500);
//  - the link update U <- exp(eps P) U (QEX's axexpmuly, gauge/gaugeUtils.nim:453-475).
#include "internal.h"
#include "su3.h"

namespace qh {

__device__ __forceinline__ M3 ldm(const double2 *U, long V, int mu, long idx) {
  M3 m;
  const double2 *p = U + ((long)mu * V + idx) * 9;
#pragma unroll
  for (int e = 0; e < 9; e++) m.e[e] = p[e];
  return m;
}

__device__ __forceinline__ M3 dagger(const M3 &a) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) r.e[3 * i + j] = make_double2(a.e[3 * j + i].x, -a.e[3 * j + i].y);
  return r;
}

__device__ __forceinline__ long shifted(const Geom &g, const int c[4], int mu, int d) {
  int y[4] = {c[0], c[1], c[2], c[3]};
  y[mu] = (y[mu] + d + g.L[mu]) % g.L[mu];
  return canon_index(g, y);
}

__global__ void __launch_bounds__(256) k_plaquette(Geom g, const double2 *U, RedWs red) {
  __shared__ double lds[16];
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (idx < V) {
    int c[4];
    uint32_t lex;
    coords_of(g, idx, c, lex);
    for (int mu = 0; mu < 4; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        const M3 a = mmul(ldm(U, V, mu, idx), ldm(U, V, nu, shifted(g, c, mu, 1)));
        const M3 b = mmul(ldm(U, V, nu, idx), ldm(U, V, mu, shifted(g, c, nu, 1)));
        // Re tr(a b^+) = sum_ij Re(a_ij conj(b_ij))
        for (int e = 0; e < 9; e++) s += a.e[e].x * b.e[e].x + a.e[e].y * b.e[e].y;
      }
  }
  double v[1] = {s};
  block_partials<1>(v, lds, red);
}

__global__ void __launch_bounds__(256) k_wilson_force(Geom g, const double2 *U, double beta, double2 *F) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (idx >= V) return;
  int c[4];
  uint32_t lex;
  coords_of(g, idx, c, lex);
  const long xp = shifted(g, c, mu, 1);
  M3 A;
  for (int e = 0; e < 9; e++) A.e[e] = make_double2(0, 0);
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    const long xpn = shifted(g, c, nu, 1), xmn = shifted(g, c, nu, -1);
    int cm[4] = {c[0], c[1], c[2], c[3]};
    cm[nu] = (cm[nu] - 1 + g.L[nu]) % g.L[nu];
    const long xpmn = shifted(g, cm, mu, 1);  // x + mu - nu
    // upper staple U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
    const M3 up = mmul(mmul(ldm(U, V, nu, xp), dagger(ldm(U, V, mu, xpn))), dagger(ldm(U, V, nu, idx)));
    // lower staple U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
    const M3 lo = mmul(mmul(dagger(ldm(U, V, nu, xpmn)), dagger(ldm(U, V, mu, xmn))), ldm(U, V, nu, xmn));
    for (int e = 0; e < 9; e++) {
      A.e[e].x += up.e[e].x + lo.e[e].x;
      A.e[e].y += up.e[e].y + lo.e[e].y;
    }
  }
  const M3 T = mmul(ldm(U, V, mu, idx), A);
  const double f = 2.0 * beta / 3.0;
  double2 *out = F + ((long)mu * V + idx) * 9;
  // TAH: (T - T^+)/2 - tr/3
  double2 R[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      const double2 x = T.e[3 * i + j], y = T.e[3 * j + i];
      R[3 * i + j] = make_double2(0.5 * (x.x - y.x), 0.5 * (x.y + y.y));
    }
  const double ti = (R[0].y + R[4].y + R[8].y) / 3.0;
  R[0].y -= ti; R[4].y -= ti; R[8].y -= ti;
  R[0].x = R[4].x = R[8].x = 0.0;
  for (int e = 0; e < 9; e++) out[e] = make_double2(f * R[e].x, f * R[e].y);
}

// P = i H, H traceless Hermitian with density exp(-tr H^2 / 2): off-diagonal
// re, im ~ N(0, 1/2); diagonal = three N(0,1) minus their mean.
__global__ void k_random_momenta(Geom g, uint64_t seed, double2 *P) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (idx >= V) return;
  int c[4];
  uint32_t lex;
  coords_of(g, idx, c, lex);
  double r[10];
  for (int j = 0; j < 5; j++) gauss2(seed, lex, 32u + (uint32_t)mu, (uint32_t)j, TAG_MOM, r[2 * j], r[2 * j + 1]);
  const double s = 0.70710678118654752440084436210485, m = (r[0] + r[1] + r[2]) / 3.0;
  double2 H[9];
  H[0] = make_double2(r[0] - m, 0); H[4] = make_double2(r[1] - m, 0); H[8] = make_double2(r[2] - m, 0);
  H[1] = make_double2(s * r[3], s * r[4]); H[3] = make_double2(s * r[3], -s * r[4]);
  H[2] = make_double2(s * r[5], s * r[6]); H[6] = make_double2(s * r[5], -s * r[6]);
  H[5] = make_double2(s * r[7], s * r[8]); H[7] = make_double2(s * r[7], -s * r[8]);
  double2 *out = P + ((long)mu * V + idx) * 9;
  for (int e = 0; e < 9; e++) out[e] = make_double2(-H[e].y, H[e].x);
}

__global__ void k_link_update(Geom g, const double2 *P, double eps, double2 *U) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (idx >= V) return;
  M3 A;
  const double2 *p = P + ((long)mu * V + idx) * 9;
  for (int e = 0; e < 9; e++) A.e[e] = make_double2(eps * p[e].x, eps * p[e].y);
  const M3 u = mmul(mexp(A), ldm(U, V, mu, idx));
  double2 *out = U + ((long)mu * V + idx) * 9;
  for (int e = 0; e < 9; e++) out[e] = u.e[e];
}

static inline unsigned nb(long n) { return (unsigned)((n + 255) / 256); }

void launch_plaquette(const Geom &g, const double *U, Scal *S, int slot, RedWs red, hipStream_t st) {
  const unsigned n = nb(2 * g.Vh);
  hipLaunchKernelGGL(k_plaquette, dim3(n), dim3(256), 0, st, g, (const double2 *)U, red);
  FinArgs f{red, (int)n, S, slot, 0.0, 0};
  launch_finish(EPI_TMP, f, st);
}

void launch_wilson_force(const Geom &g, const double *U, double beta, double *F, hipStream_t st) {
  hipLaunchKernelGGL(k_wilson_force, dim3(nb(2 * g.Vh), 4), dim3(256), 0, st, g, (const double2 *)U, beta,
                     (double2 *)F);
}

void launch_random_momenta(const Geom &g, uint64_t seed, double *P, hipStream_t st) {
  hipLaunchKernelGGL(k_random_momenta, dim3(nb(2 * g.Vh), 4), dim3(256), 0, st, g, seed, (double2 *)P);
}

void launch_link_update(const Geom &g, const double *P, double eps, double *U, hipStream_t st) {
  hipLaunchKernelGGL(k_link_update, dim3(nb(2 * g.Vh), 4), dim3(256), 0, st, g, (const double2 *)P, eps,
                     (double2 *)U);
}

}  // namespace qh
