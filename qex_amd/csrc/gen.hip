// Synthetic, layout-independent inputs generated on the GPU:
//  - Philox4x32-10 counter-based generator keyed by (seed, lexicographic site,
//    stream, component block) (Salmon et al., SC'11);
//  - Box-Muller N(0,1) pairs from 53-bit uniforms;
//  - gauge links U = exp(i*0.5*H), H traceless Hermitian with unit Gaussian
//    entries; sources with E|z|^2 = 1 per color component;
//  - QEX's setBC (antiperiodic t, gauge/gaugeUtils.nim:133-140) and staggered
//    phases [8,9,11,0] (physics/stagD.nim:560-571);
//  - HISQ-form links W = 0.8U + 0.05G, L = -(1/24) U U U from the phased U,
//    interleaved [W0,L0,...] (physics/stagD.nim:603-615, gauge/fat7l.nim:145-154).
// All outputs are FP64 in QEX's canonical site order (even sites first).
#include "internal.h"
#include "su3.h"

namespace qh {

__global__ void k_gauge_random(Geom g, uint64_t seed, double2 *U) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (idx >= V) return;
  int c[4];
  uint32_t lex;
  coords_of(g, idx, c, lex);
  double r[10];
  for (int j = 0; j < 5; j++) gauss2(seed, lex, (uint32_t)mu, (uint32_t)j, TAG_GAUGE, r[2 * j], r[2 * j + 1]);
  const double tr = (r[0] + r[1] + r[2]) / 3.0;
  // A = i*0.5*H
  double2 H[9];
  H[0] = make_double2(r[0] - tr, 0); H[4] = make_double2(r[1] - tr, 0); H[8] = make_double2(r[2] - tr, 0);
  H[1] = make_double2(r[3], r[4]); H[3] = make_double2(r[3], -r[4]);
  H[2] = make_double2(r[5], r[6]); H[6] = make_double2(r[5], -r[6]);
  H[5] = make_double2(r[7], r[8]); H[7] = make_double2(r[7], -r[8]);
  M3 A;
  for (int e = 0; e < 9; e++) A.e[e] = make_double2(-0.5 * H[e].y, 0.5 * H[e].x);
  M3 E = mexp(A);
  double2 *out = U + ((long)mu * V + idx) * 9;
  for (int e = 0; e < 9; e++) out[e] = E.e[e];
}

__global__ void k_source_gauss(Geom g, uint64_t seed, int stream, int par, double2 *x) {
  const long V = 2 * g.Vh;
  const long lo = (par == 1) ? g.Vh : 0, hi = (par == 0) ? g.Vh : V;
  const long idx = lo + (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= hi) return;
  int c[4];
  uint32_t lex;
  coords_of(g, idx, c, lex);
  const double s = 0.70710678118654752440084436210485;
  for (int j = 0; j < 3; j++) {
    double a, b;
    gauss2(seed, lex, 64u + (uint32_t)stream, (uint32_t)j, TAG_SOURCE, a, b);
    x[idx * 3 + j] = make_double2(s * a, s * b);
  }
}

// setBC then stagPhase on the 4 plain link fields (both are sign flips)
__global__ void k_phase_bc(Geom g, double2 *U) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (idx >= V) return;
  int c[4];
  uint32_t lex;
  coords_of(g, idx, c, lex);
  const int masks[4] = {8, 9, 11, 0};
  int s = 0;
  for (int k = 0; k < 4; k++) s += (masks[mu] >> k) & c[k];
  int flip = s & 1;
  if (mu == 3 && c[3] == g.L[3] - 1) flip ^= 1;
  if (!flip) return;
  double2 *u = U + ((long)mu * V + idx) * 9;
  for (int e = 0; e < 9; e++) u[e] = make_double2(-u[e].x, -u[e].y);
}

__global__ void k_hisq_links(Geom g, uint64_t seed, const double2 *Uph, double2 *g8) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (idx >= V) return;
  int c[4];
  uint32_t lex;
  coords_of(g, idx, c, lex);
  const double s = 0.70710678118654752440084436210485;
  const double2 *u0 = Uph + ((long)mu * V + idx) * 9;
  double2 *w = g8 + ((long)(2 * mu) * V + idx) * 9;
  double2 *lg = g8 + ((long)(2 * mu + 1) * V + idx) * 9;
  for (int j = 0; j < 9; j++) {
    double a, b;
    gauss2(seed, lex, 16u + (uint32_t)mu, (uint32_t)j, TAG_FATG, a, b);
    w[j] = make_double2(0.8 * u0[j].x + 0.05 * s * a, 0.8 * u0[j].y + 0.05 * s * b);
  }
  int c1[4] = {c[0], c[1], c[2], c[3]}, c2[4] = {c[0], c[1], c[2], c[3]};
  c1[mu] = (c[mu] + 1) % g.L[mu];
  c2[mu] = (c[mu] + 2) % g.L[mu];
  M3 a0, a1, a2;
  const double2 *u1 = Uph + ((long)mu * V + canon_index(g, c1)) * 9;
  const double2 *u2 = Uph + ((long)mu * V + canon_index(g, c2)) * 9;
  for (int e = 0; e < 9; e++) { a0.e[e] = u0[e]; a1.e[e] = u1[e]; a2.e[e] = u2[e]; }
  M3 t = mmul(mmul(a0, a1), a2);
  const double cn = -1.0 / 24.0;
  for (int e = 0; e < 9; e++) lg[e] = make_double2(cn * t.e[e].x, cn * t.e[e].y);
}

static inline unsigned nb(long n) { return (unsigned)((n + 127) / 128); }

void launch_gauge_random(const Geom &g, uint64_t seed, double *canon, hipStream_t st) {
  hipLaunchKernelGGL(k_gauge_random, dim3(nb(2 * g.Vh), 4), dim3(128), 0, st, g, seed, (double2 *)canon);
}
void launch_source_gauss(const Geom &g, uint64_t seed, int stream, int par, double *canon, hipStream_t st) {
  const long n = (par == 2) ? 2 * g.Vh : g.Vh;
  hipLaunchKernelGGL(k_source_gauss, dim3(nb(n)), dim3(128), 0, st, g, seed, stream, par, (double2 *)canon);
}
void launch_phase_bc(const Geom &g, double *canon, int nl, hipStream_t st) {
  (void)nl;
  hipLaunchKernelGGL(k_phase_bc, dim3(nb(2 * g.Vh), 4), dim3(128), 0, st, g, (double2 *)canon);
}
void launch_hisq_links(const Geom &g, uint64_t seed, const double *Uph, double *g8, hipStream_t st) {
  hipLaunchKernelGGL(k_hisq_links, dim3(nb(2 * g.Vh), 4), dim3(128), 0, st, g, seed, (const double2 *)Uph,
                     (double2 *)g8);
}

}  // namespace qh
