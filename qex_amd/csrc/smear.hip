// HISQ link construction, FP64: the fat7 path sum with Lepage and Naik terms
// (QEX gauge/fat7l.nim:71-159, makeImpLinks), the unitary projection
// x (x^+ x)^{-1/2} by Cayley-Hamilton (maths/matrixFunctions.nim:337-352), and
// the two-level HISQ smearing built from them (physics/hisqLinks.nim:39-51).
//
// Fields are 3x3 complex per site in canonical site order.  API arrays are
// QEX's [mu][site][3][3] (site-major, 144 B per link); the intermediates are
// component-major [e][site] so that a wavefront's loads of one component are
// contiguous.  A view MF addresses both: element e of site s at p[e*es + s*ss].
//
// The path sum is evaluated in three sweeps per level, each a gather over
// neighbours with one thread per (site, mu) and the generalized staple
//   S_nu[K](x) = U_nu(x) K(x+nu) U_nu(x+mu)^+ + U_nu(x-nu)^+ K(x-nu) U_nu(x-nu+mu)
// (gauge/fat7l.nim:24-67) as the only building block:
//   k_fat3:  S_{mu,nu} = S_nu[U_mu]                       (3 fields per mu, stored)
//   k_fat5:  T_{mu,sig} = S_rho[S_{mu,nu}] + S_nu[S_{mu,rho}],  {nu,rho} = rest of {mu,sig}
//   k_fatf:  F_mu = (c1 - 6 cL) U_mu + c3 sum_nu S_{mu,nu} + c5 sum_sig T_{mu,sig}
//                   + c7 sum_sig S_sig[T_{mu,sig}] + cL sum_nu S_nu[S_{mu,nu}]
// This is the reference's nested loop (5- and 7-link staples nest the 3-link
// one; the Lepage term is the in-plane double staple whose retracing paths the
// -6 cL one-link shift cancels) with the two 5-link staples that feed the same
// 7-link direction summed before the outer staple, since S_sig is linear.
#include <algorithm>
#include <stdexcept>

#include "internal.h"
#include "su3.h"

namespace qh {

struct MF {
  double2 *p;
  long es, ss;
};

// Element e of site s of a view: the base of component e (the same for the
// whole wave) plus a 32-bit byte offset of the site, so that every load takes
// a scalar base and one offset register per view.  The views stay below
// 4 GiB: the smearing entry points refuse lattices above kMaxSmearSites
// (ws_ensure) and the projections of arbitrary link arrays run in chunks.
__device__ __forceinline__ double2 &mel(const MF &f, int e, long s) {
  char *b = (char *)(f.p + (long)e * f.es);
  return *(double2 *)(b + (uint32_t)s * (uint32_t)(f.ss * 16));
}

__device__ __forceinline__ M3 ldf(const MF &f, long s) {
  M3 m;
#pragma unroll
  for (int e = 0; e < 9; e++) m.e[e] = mel(f, e, s);
  return m;
}

__device__ __forceinline__ void stf(const MF &f, long s, const M3 &m) {
#pragma unroll
  for (int e = 0; e < 9; e++) mel(f, e, s) = m.e[e];
}

// a * b^+
__device__ __forceinline__ M3 mmul_nd(const M3 &a, const M3 &b) {
  M3 c;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double2 s = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double2 x = a.e[3 * i + k], y = b.e[3 * j + k];
        s.x += x.x * y.x + x.y * y.y;
        s.y += x.y * y.x - x.x * y.y;
      }
      c.e[3 * i + j] = s;
    }
  return c;
}

// a^+ * b
__device__ __forceinline__ M3 mmul_dn(const M3 &a, const M3 &b) {
  M3 c;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double2 s = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double2 x = a.e[3 * k + i], y = b.e[3 * k + j];
        s.x += x.x * y.x + x.y * y.y;
        s.y += x.x * y.y - x.y * y.x;
      }
      c.e[3 * i + j] = s;
    }
  return c;
}

__device__ __forceinline__ void macc(M3 &acc, double c, const M3 &a) {
#pragma unroll
  for (int e = 0; e < 9; e++) {
    acc.e[e].x += c * a.e[e].x;
    acc.e[e].y += c * a.e[e].y;
  }
}

__device__ __forceinline__ long step(const Geom &g, const int c[4], int mu, int d) {
  int y[4] = {c[0], c[1], c[2], c[3]};
  int v = y[mu] + d;
  v = v >= g.L[mu] ? v - g.L[mu] : (v < 0 ? v + g.L[mu] : v);
  y[mu] = v;
  return canon_index(g, y);
}

// acc += a L M R^+ (BACK false) or a L^+ M R (BACK true), the two halves of a
// staple.  One row of L M is formed at a time and folded into acc at once, so
// the live set is M, R, acc and a row; the chain-rule kernels, which hold
// several accumulators, use it to keep two waves per SIMD (one before).
template <bool BACK>
__device__ __forceinline__ void triple(M3 &acc, double a, const MF &L, long il, const MF &M, long im, const MF &R,
                                       long ir) {
  const M3 m = ldf(M, im);
  const M3 r = ldf(R, ir);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double2 l[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      l[k] = mel(L, BACK ? 3 * k + i : 3 * i + k, il);
      if (BACK) l[k].y = -l[k].y;
    }
    double2 t[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double2 s = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double2 y = m.e[3 * k + j];
        s.x += l[k].x * y.x - l[k].y * y.y;
        s.y += l[k].x * y.y + l[k].y * y.x;
      }
      t[j] = s;
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double2 s = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (BACK) {  // (L^+ M)_{ik} R_{kj}
          const double2 y = r.e[3 * k + j];
          s.x += t[k].x * y.x - t[k].y * y.y;
          s.y += t[k].x * y.y + t[k].y * y.x;
        } else {  // (L M)_{ik} conj(R_{jk})
          const double2 y = r.e[3 * j + k];
          s.x += t[k].x * y.x + t[k].y * y.y;
          s.y += t[k].y * y.x - t[k].x * y.y;
        }
      }
      acc.e[3 * i + j].x += a * s.x;
      acc.e[3 * i + j].y += a * s.y;
    }
  }
}

// acc += a S_nu[K](x), x with coordinates c
__device__ __forceinline__ void gstaple_acc(M3 &acc, double a, const Geom &g, const MF &Unu, const MF &K,
                                            const int c[4], long x, int mu, int nu) {
  int cm[4] = {c[0], c[1], c[2], c[3]};
  cm[nu] = c[nu] > 0 ? c[nu] - 1 : g.L[nu] - 1;
  const long xpn = step(g, c, nu, 1), xpm = step(g, c, mu, 1);
  const long xmn = canon_index(g, cm), xmnpm = step(g, cm, mu, 1);
  triple<false>(acc, a, Unu, x, K, xpn, Unu, xpm);
  triple<true>(acc, a, Unu, xmn, K, xmn, Unu, xmnpm);
}

// S_nu[K](x) as two full matrix products per half: faster in the forward
// sweeps, whose threads hold fewer other matrices than the chain-rule ones
__device__ __forceinline__ M3 gstaple(const Geom &g, const MF &Unu, const MF &K, const int c[4], long x, int mu,
                                      int nu) {
  int cm[4] = {c[0], c[1], c[2], c[3]};
  cm[nu] = c[nu] > 0 ? c[nu] - 1 : g.L[nu] - 1;
  const long xpn = step(g, c, nu, 1), xpm = step(g, c, mu, 1);
  const long xmn = canon_index(g, cm), xmnpm = step(g, cm, mu, 1);
  M3 a = mmul_nd(mmul(ldf(Unu, x), ldf(K, xpn)), ldf(Unu, xpm));
  const M3 b = mmul(mmul_dn(ldf(Unu, xmn), ldf(K, xmn)), ldf(Unu, xmnpm));
#pragma unroll
  for (int e = 0; e < 9; e++) {
    a.e[e].x += b.e[e].x;
    a.e[e].y += b.e[e].y;
  }
  return a;
}

// the k-th direction other than mu (k = 0..2), and the inverse map
__device__ __forceinline__ int other(int mu, int k) { return k < mu ? k : k + 1; }
__device__ __forceinline__ int slot(int mu, int nu) { return nu < mu ? nu : nu - 1; }

struct SmearFields {
  MF U[4];       // input links of this level
  MF S[12];      // S_{mu,nu} at [3*mu + slot(mu,nu)]
  MF T[12];      // T_{mu,sig} at [3*mu + slot(mu,sig)]
  MF out[4];     // fat links
};

struct Fat7 {
  double c1, c3, c5, c7, cL;  // c1 already shifted by -6 cL
};

// A workgroup covers 64 consecutive sites with one wavefront per direction mu
// (per sigma in k_fat5): its waves gather overlapping links of the same
// neighbourhood at the same time, so those loads hit in the CU's L1 and the
// XCD's L2 instead of returning in a later pass over the lattice.  Workgroups
// walk the sites in XCD-contiguous ranges.
//
// With bit 1 of `remap` set the groups walk t before the rest of the slice:
// each t-slice of one parity (S = hx*Ly*Lz half-sites) is cut into chunks of
// C half-sites, and consecutive groups take (parity, then t) within one chunk
// before moving to the next chunk.  The sites in flight on an XCD then span a
// few t-slices and both parities of one chunk, so their t-neighbours are read
// while still in the L2; only neighbours across chunk faces (in z) return from
// the memory side.  The host sets the bit only when S is a multiple of 64.
#ifndef QH_SMEAR_CHUNK
#define QH_SMEAR_CHUNK 1024
#endif
__device__ __forceinline__ long site_of(const Geom &g, int remap) {
  const long G = xcd_block(remap & 1);
  const int lane = threadIdx.x & 63;
  if (!(remap & 2)) return G * 64 + lane;
  if (G * 64 >= 2 * g.Vh) return 2 * g.Vh;  // padding group of the XCD-rounded grid
  const long S = (long)g.hx * g.L[1] * g.L[2];
  long C = QH_SMEAR_CHUNK;
  while (C > 64 && S % C) C >>= 1;
  const long gpc = C >> 6;
  const long gi = G % gpc;
  long r = G / gpc;
  const long p = r & 1;
  r >>= 1;
  const long t = r % g.L[3], k = r / g.L[3];
  return p * g.Vh + t * S + k * C + gi * 64 + lane;
}
__device__ __forceinline__ int wave_of() { return threadIdx.x >> 6; }

__global__ void __launch_bounds__(256) k_fat3(Geom g, SmearFields f, int remap) {
  const long V = 2 * g.Vh;
  const long x = site_of(g, remap);
  const int mu = wave_of();
  if (x >= V) return;
  int c[4];
  uint32_t lex;
  coords_of(g, x, c, lex);
#pragma unroll 1
  for (int k = 0; k < 3; k++) {
    const int nu = other(mu, k);
    stf(f.S[3 * mu + k], x, gstaple(g, f.U[nu], f.U[mu], c, x, mu, nu));
  }
}

__global__ void __launch_bounds__(192) k_fat5(Geom g, SmearFields f, int remap) {
  const long V = 2 * g.Vh;
  const long x = site_of(g, remap);
  const int mu = blockIdx.y, ks = wave_of();
  if (x >= V) return;
  const int sig = other(mu, ks);
  int nu = -1, rho = -1;
  for (int d = 0; d < 4; d++)
    if (d != mu && d != sig) {
      if (nu < 0) nu = d; else rho = d;
    }
  int c[4];
  uint32_t lex;
  coords_of(g, x, c, lex);
  M3 t = gstaple(g, f.U[rho], f.S[3 * mu + slot(mu, nu)], c, x, mu, rho);
  const M3 u = gstaple(g, f.U[nu], f.S[3 * mu + slot(mu, rho)], c, x, mu, nu);
#pragma unroll
  for (int e = 0; e < 9; e++) {
    t.e[e].x += u.e[e].x;
    t.e[e].y += u.e[e].y;
  }
  stf(f.T[3 * mu + ks], x, t);
}

// HAVE: 0 one-link only; 1 + 3-link staples; 2 + Lepage and 5/7-link (S, T valid)
template <int HAVE>
__global__ void __launch_bounds__(256) k_fatf(Geom g, SmearFields f, Fat7 c, int remap) {
  const long V = 2 * g.Vh;
  const long x = site_of(g, remap);
  const int mu = wave_of();
  if (x >= V) return;
  M3 acc = ldf(f.U[mu], x);
#pragma unroll
  for (int e = 0; e < 9; e++) {
    acc.e[e].x *= c.c1;
    acc.e[e].y *= c.c1;
  }
  if (HAVE >= 1) {
    int cc[4];
    uint32_t lex;
    coords_of(g, x, cc, lex);
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
      const int d = other(mu, k);
      macc(acc, c.c3, ldf(f.S[3 * mu + k], x));
      if (HAVE >= 2) {
        if (c.c5 != 0.0) macc(acc, c.c5, ldf(f.T[3 * mu + k], x));
        if (c.c7 != 0.0) macc(acc, c.c7, gstaple(g, f.U[d], f.T[3 * mu + k], cc, x, mu, d));
        if (c.cL != 0.0) macc(acc, c.cL, gstaple(g, f.U[d], f.S[3 * mu + k], cc, x, mu, d));
      }
    }
  }
  stf(f.out[mu], x, acc);
}

// ll_mu(x) = naik * W_mu(x) W_mu(x+mu) W_mu(x+2mu)
__global__ void __launch_bounds__(256) k_naik(Geom g, MF W0, MF W1, MF W2, MF W3, MF o0, MF o1, MF o2, MF o3,
                                              double naik, int remap) {
  const long V = 2 * g.Vh;
  const long x = site_of(g, remap);
  const int mu = wave_of();
  if (x >= V) return;
  const MF W = mu == 0 ? W0 : (mu == 1 ? W1 : (mu == 2 ? W2 : W3));
  const MF o = mu == 0 ? o0 : (mu == 1 ? o1 : (mu == 2 ? o2 : o3));
  int c[4];
  uint32_t lex;
  coords_of(g, x, c, lex);
  M3 t = mmul(mmul(ldf(W, x), ldf(W, step(g, c, mu, 1))), ldf(W, step(g, c, mu, 2)));
#pragma unroll
  for (int e = 0; e < 9; e++) {
    t.e[e].x *= naik;
    t.e[e].y *= naik;
  }
  stf(o, x, t);
}

// H^{-1/2} for H = x^+ x + 1e-20 (Hermitian positive).  The eigenvalues s_i^2
// of H come from the trigonometric solution of its characteristic cubic
// (Smith, CACM 4 (1961) 168); H^{-1/2} = f0 + f1 H + f2 H^2 is the quadratic
// through 1/s_i: with u1 = sum s_i, v = sum_{i<j} s_i s_j, w = prod s_i and
// d = w (u1 v - w): f0 = (u1 v^2 - v w - u1^2 w)/d, f1 = -(u1 tr H + w)/d,
// f2 = u1/d.
__device__ __forceinline__ M3 rsqrt_xdx(const M3 &X) {
  M3 H = mmul_dn(X, X);
#pragma unroll
  for (int k = 0; k < 3; k++) H.e[4 * k].x += 1e-20;
  const double h0 = H.e[0].x, h1 = H.e[4].x, h2 = H.e[8].x;
  const double q = (h0 + h1 + h2) * (1.0 / 3.0);
  const double b0 = h0 - q, b1 = h1 - q, b2 = h2 - q;
  const double2 z01 = H.e[1], z02 = H.e[2], z12 = H.e[5];
  const double o01 = z01.x * z01.x + z01.y * z01.y, o02 = z02.x * z02.x + z02.y * z02.y;
  const double o12 = z12.x * z12.x + z12.y * z12.y;
  const double p2 = (b0 * b0 + b1 * b1 + b2 * b2 + 2.0 * (o01 + o02 + o12)) * (1.0 / 6.0);
  double e0 = q, e1 = q, e2 = q;
  if (p2 > 0.0) {
    const double p = sqrt(p2);
    // Re(h01 h12 h20), h20 = conj(h02)
    const double2 a = make_double2(z01.x * z12.x - z01.y * z12.y, z01.x * z12.y + z01.y * z12.x);
    const double re3 = a.x * z02.x + a.y * z02.y;
    const double detb = b0 * b1 * b2 + 2.0 * re3 - b0 * o12 - b1 * o02 - b2 * o01;
    double r = detb / (2.0 * p2 * p);
    r = fmin(1.0, fmax(-1.0, r));
    const double phi = acos(r) * (1.0 / 3.0);
    e0 = q + 2.0 * p * cos(phi);
    e2 = q + 2.0 * p * cos(phi + 2.0943951023931954923);
    e1 = 3.0 * q - e0 - e2;
  }
  const double s0 = sqrt(fabs(e0)), s1 = sqrt(fabs(e1)), s2 = sqrt(fabs(e2));
  const double u1 = s0 + s1 + s2, v = s0 * s1 + s1 * s2 + s2 * s0, w = s0 * s1 * s2;
  const double di = 1.0 / (w * (u1 * v - w));
  const double f0 = (u1 * v * v - v * w - u1 * u1 * w) * di, f1 = -(u1 * 3.0 * q + w) * di, f2 = u1 * di;
  const M3 H2 = mmul(H, H);
  M3 Z;
#pragma unroll
  for (int e = 0; e < 9; e++) Z.e[e] = make_double2(f1 * H.e[e].x + f2 * H2.e[e].x, f1 * H.e[e].y + f2 * H2.e[e].y);
#pragma unroll
  for (int k = 0; k < 3; k++) Z.e[4 * k].x += f0;
  return Z;
}

// projectU: u = x (x^+ x + 1e-20)^{-1/2}
__device__ __forceinline__ M3 project_u(const M3 &X) { return mmul(X, rsqrt_xdx(X)); }

// projectU on V sites of four link fields (blockIdx.y selects the field views)
__global__ void __launch_bounds__(256) k_project(long V, MF i0, MF i1, MF i2, MF i3, MF o0, MF o1, MF o2, MF o3) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (x >= V) return;
  const MF in = mu == 0 ? i0 : (mu == 1 ? i1 : (mu == 2 ? i2 : i3));
  const MF o = mu == 0 ? o0 : (mu == 1 ? o1 : (mu == 2 ? o2 : o3));
  stf(o, x, project_u(ldf(in, x)));
}

// layout change (site-major <-> component-major)
__global__ void __launch_bounds__(256) k_relayout(long V, MF i0, MF i1, MF i2, MF i3, MF o0, MF o1, MF o2, MF o3) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (x >= V) return;
  const MF in = mu == 0 ? i0 : (mu == 1 ? i1 : (mu == 2 ? i2 : i3));
  const MF o = mu == 0 ? o0 : (mu == 1 ? o1 : (mu == 2 ? o2 : o3));
#pragma unroll
  for (int e = 0; e < 9; e++) mel(o, e, x) = mel(in, e, x);
}

// ------------------------------------------------------------- chain rule
// Derivatives with respect to the complex link entries in the pairing
// Re tr(A^+ B) (QEX gauge/fat7lderiv.nim:122-475, maths/matrixFunctions.nim:
// 361-392): given the chain C = dS/dF of F = f(U), the kernels produce dS/dU.
// The reverse of one fat7 level, per direction mu with chain C_mu:
//   CT_{mu,sig} = c5 C_mu + c7 S_sig[C_mu]
//   CS_{mu,nu}  = c3 C_mu + cL S_nu[C_mu] + sum_{sig != nu} S_rho[CT_{mu,sig}]
//   dS/dU_mu   += (c1 - 6 cL) C_mu + sum_nu S_nu[CS_{mu,nu}]
// (the staple map is its own adjoint with respect to its middle field), and
// every staple of the (mu,lam) plane adds its side-link derivative to
// dS/dU_lam (side() below).

struct DerivFields {
  MF U[4];              // level input links
  MF S[12], T[12];      // forward intermediates of this level
  MF C[4];              // chain dS/dF
  MF CT[12], CS[12];    // chains of T and S
  MF P[12];             // side-link partials at the site, [3*lam + k] from mu = other(lam, k)
  MF Q[12];             // side-link partials one site back along mu (added at y from y - mu)
  MF out[4];            // dS/dU (accumulated when acc)
};

__device__ __forceinline__ void madd(M3 &acc, const M3 &a) {
#pragma unroll
  for (int e = 0; e < 9; e++) {
    acc.e[e].x += a.e[e].x;
    acc.e[e].y += a.e[e].y;
  }
}

__device__ __forceinline__ int third_dir(int a, int b, int c) {
  int r = 0;
  for (int d = 0; d < 4; d++)
    if (d != a && d != b && d != c) r = d;
  return r;
}

// Side-link derivative of the staple S_lam[K] in the (mu,lam) plane with chain
// C, weight a, added at y:
//   C(y) U(y+mu) K(y+lam)^+ + K(y) U(y+mu) C(y+lam)^+
//   + C(y-mu)^+ U(y-mu) K(y-mu+lam) + K(y-mu)^+ U(y-mu) C(y-mu+lam),  U = U_lam.
// It is evaluated as A(y) + B(y - mu) with, at one site z,
//   A(z) = C(z) U(z+mu) K(z+lam)^+ + K(z) U(z+mu) C(z+lam)^+
//   B(z) = C(z)^+ U(z) K(z+lam) + K(z)^+ U(z) C(z+lam),
// whose four products share C(z), K(z), C(z+lam), K(z+lam).  side_half adds
// one product of A (BACK false: acc += a L(z) u R(z+lam)^+, u = U(z+mu)) or of
// B (BACK true: acc += a L(z)^+ u R(z+lam), u = U(z)), with u held by the
// caller for all terms, R loaded whole and L a row or column at a time.
template <bool BACK>
__device__ __forceinline__ void side_half(M3 &acc, double a, const M3 &u, const MF &L, long z, const MF &R,
                                          long zpl) {
  const M3 r = ldf(R, zpl);
  // one row i of x = U R^+ (A) or U R (B) at a time, folded into acc at once
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double2 x[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double2 t = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double2 uk = u.e[3 * i + k];
        if (BACK) {  // u_ik r_kj
          t.x += uk.x * r.e[3 * k + j].x - uk.y * r.e[3 * k + j].y;
          t.y += uk.x * r.e[3 * k + j].y + uk.y * r.e[3 * k + j].x;
        } else {  // u_ik conj(r_jk)
          t.x += uk.x * r.e[3 * j + k].x + uk.y * r.e[3 * j + k].y;
          t.y += uk.y * r.e[3 * j + k].x - uk.x * r.e[3 * j + k].y;
        }
      }
      x[j] = t;
    }
    double2 l[3];
    if (BACK) {  // row i of L: acc_kj += a conj(L_ik) x_j
#pragma unroll
      for (int k = 0; k < 3; k++) l[k] = mel(L, 3 * i + k, z);
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          acc.e[3 * k + j].x += a * (l[k].x * x[j].x + l[k].y * x[j].y);
          acc.e[3 * k + j].y += a * (l[k].x * x[j].y - l[k].y * x[j].x);
        }
    } else {  // column i of L: acc_kj += a L_ki x_j
#pragma unroll
      for (int k = 0; k < 3; k++) l[k] = mel(L, 3 * k + i, z);
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          acc.e[3 * k + j].x += a * (l[k].x * x[j].x - l[k].y * x[j].y);
          acc.e[3 * k + j].y += a * (l[k].x * x[j].y + l[k].y * x[j].x);
        }
    }
  }
}

// CT_{mu,sig} = c5 C_mu + c7 S_sig[C_mu]; one wave per sig, mu = blockIdx.y
__global__ void __launch_bounds__(192) k_dct(Geom g, DerivFields f, Fat7 c, int remap) {
  const long V = 2 * g.Vh;
  const long x = site_of(g, remap);
  const int mu = blockIdx.y, ks = wave_of();
  if (x >= V) return;
  const int sig = other(mu, ks);
  int cc[4];
  uint32_t lex;
  coords_of(g, x, cc, lex);
  M3 t = ldf(f.C[mu], x);
#pragma unroll
  for (int e = 0; e < 9; e++) {
    t.e[e].x *= c.c5;
    t.e[e].y *= c.c5;
  }
  if (c.c7 != 0.0) gstaple_acc(t, c.c7, g, f.U[sig], f.C[mu], cc, x, mu, sig);
  stf(f.CT[3 * mu + ks], x, t);
}

// CS_{mu,nu} = c3 C_mu + cL S_nu[C_mu] + sum_{sig != nu} S_rho[CT_{mu,sig}]
template <int HAVE5>
__global__ void __launch_bounds__(192) k_dcs(Geom g, DerivFields f, Fat7 c, int remap) {
  const long V = 2 * g.Vh;
  const long x = site_of(g, remap);
  const int mu = blockIdx.y, kn = wave_of();
  if (x >= V) return;
  const int nu = other(mu, kn);
  int cc[4];
  uint32_t lex;
  coords_of(g, x, cc, lex);
  M3 t = ldf(f.C[mu], x);
#pragma unroll
  for (int e = 0; e < 9; e++) {
    t.e[e].x *= c.c3;
    t.e[e].y *= c.c3;
  }
  if (HAVE5) {
    if (c.cL != 0.0) gstaple_acc(t, c.cL, g, f.U[nu], f.C[mu], cc, x, mu, nu);
    if (c.c5 != 0.0 || c.c7 != 0.0) {
#pragma unroll 1
      for (int k = 0; k < 3; k++) {
        const int sig = other(mu, k);
        if (sig == nu) continue;
        const int rho = third_dir(mu, nu, sig);
        gstaple_acc(t, 1.0, g, f.U[rho], f.CT[3 * mu + k], cc, x, mu, rho);
      }
    }
  }
  stf(f.CS[3 * mu + kn], x, t);
}

// side-link partials of dS/dU_lam from the staples of the (mu,lam) plane: A
// into P at the site, B into Q (k_dfin adds it one site along mu).  One
// workgroup per 64 sites and direction mu = blockIdx.y, with a pair of waves
// per lam = other(mu, k), one for A and one for B: the six waves read the same
// mu-fields (U_mu, C_mu, S_{mu,*}, T_{mu,*}, CT_{mu,*}) at the same sites, and
// each pair the same C, K at z and z + lam.
#ifndef QH_DSIDE_WAVES
#define QH_DSIDE_WAVES 1  // minimum waves per SIMD asked of the register allocator
#endif
template <int HAVE5>
__global__ void __launch_bounds__(384, QH_DSIDE_WAVES) k_dside(Geom g, DerivFields f, Fat7 c, int remap) {
  const long V = 2 * g.Vh;
  const long z = site_of(g, remap);
  const int mu = blockIdx.y, w = wave_of();
  if (z >= V) return;
  const int lam = other(mu, w >> 1);
  const bool back = w & 1;
  const int sl = 3 * mu + slot(mu, lam);
  int cc[4];
  uint32_t lex;
  coords_of(g, z, cc, lex);
  const long zpl = step(g, cc, lam, 1);
  const long zu = back ? z : step(g, cc, mu, 1);
  M3 acc;
#pragma unroll
  for (int e = 0; e < 9; e++) acc.e[e] = make_double2(0, 0);
  const M3 u = ldf(f.U[lam], zu);  // the side link, common to every term
  // terms j: 0 S_lam[U_mu] in S_{mu,lam}; 1 Lepage; 2 7-link; 3, 4 S_lam[S_{mu,rho}]
  // inside T_{mu,sig} for the two sig != lam.  Rolled loops, so that the
  // loads of one product are not hoisted into another's.
#pragma unroll 1
  for (int j = 0; j < (HAVE5 ? 5 : 1); j++) {
    double a = 1.0;
    MF Cf = f.CS[sl], Kf = f.U[mu];
    if (j == 1) {
      if (c.cL == 0.0) continue;
      a = c.cL;
      Cf = f.C[mu];
      Kf = f.S[sl];
    } else if (j == 2) {
      if (c.c7 == 0.0) continue;
      a = c.c7;
      Cf = f.C[mu];
      Kf = f.T[sl];
    } else if (j >= 3) {
      if (c.c5 == 0.0 && c.c7 == 0.0) continue;
      const int skip = slot(mu, lam);  // the k with other(mu, k) == lam
      const int ks = j == 3 ? (skip == 0 ? 1 : 0) : (skip == 2 ? 1 : 2);
      const int sig = other(mu, ks);
      Cf = f.CT[3 * mu + ks];
      Kf = f.S[3 * mu + slot(mu, third_dir(mu, lam, sig))];
    }
#pragma unroll 1
    for (int h = 0; h < 2; h++) {
      if (back) side_half<true>(acc, a, u, h ? Kf : Cf, z, h ? Cf : Kf, zpl);
      else side_half<false>(acc, a, u, h ? Kf : Cf, z, h ? Cf : Kf, zpl);
    }
  }
  const int o = 3 * lam + slot(lam, mu);
  stf(back ? f.Q[o] : f.P[o], z, acc);
}

// dS/dU_lam (+)= (c1 - 6 cL) C_lam + sum_nu S_nu[CS_{lam,nu}] + side partials
template <int HAVE3>
__global__ void __launch_bounds__(256) k_dfin(Geom g, DerivFields f, Fat7 c, int acc_out, int remap) {
  const long V = 2 * g.Vh;
  const long y = site_of(g, remap);
  const int lam = wave_of();
  if (y >= V) return;
  M3 d = ldf(f.C[lam], y);
#pragma unroll
  for (int e = 0; e < 9; e++) {
    d.e[e].x *= c.c1;
    d.e[e].y *= c.c1;
  }
  if (HAVE3) {
    int cc[4];
    uint32_t lex;
    coords_of(g, y, cc, lex);
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
      const int nu = other(lam, k);
      gstaple_acc(d, 1.0, g, f.U[nu], f.CS[3 * lam + k], cc, y, lam, nu);
      madd(d, ldf(f.P[3 * lam + k], y));
      madd(d, ldf(f.Q[3 * lam + k], step(g, cc, nu, -1)));
    }
  }
  if (acc_out) madd(d, ldf(f.out[lam], y));
  stf(f.out[lam], y, d);
}

// Naik: for ll = naik W(x) W(x+mu) W(x+2mu), dS/dW_mu(y) (+)= naik [CL(y)
// (W(y+mu) W(y+2mu))^+ + W(y-mu)^+ CL(y-mu) W(y+mu)^+ + (W(y-2mu) W(y-mu))^+ CL(y-2mu)]
__global__ void __launch_bounds__(256) k_dnaik(Geom g, MF W0, MF W1, MF W2, MF W3, MF C0, MF C1, MF C2, MF C3, MF o0,
                                               MF o1, MF o2, MF o3, double naik, int acc_out, int remap) {
  const long V = 2 * g.Vh;
  const long y = site_of(g, remap);
  const int mu = wave_of();
  if (y >= V) return;
  const MF W = mu == 0 ? W0 : (mu == 1 ? W1 : (mu == 2 ? W2 : W3));
  const MF C = mu == 0 ? C0 : (mu == 1 ? C1 : (mu == 2 ? C2 : C3));
  const MF o = mu == 0 ? o0 : (mu == 1 ? o1 : (mu == 2 ? o2 : o3));
  int c[4];
  uint32_t lex;
  coords_of(g, y, c, lex);
  const long p1 = step(g, c, mu, 1), p2 = step(g, c, mu, 2), m1 = step(g, c, mu, -1), m2 = step(g, c, mu, -2);
  const M3 wp1 = ldf(W, p1), wm1 = ldf(W, m1);
  M3 t = mmul_nd(ldf(C, y), mmul(wp1, ldf(W, p2)));
  madd(t, mmul_nd(mmul_dn(wm1, ldf(C, m1)), wp1));
  madd(t, mmul_dn(mmul(ldf(W, m2), wm1), ldf(C, m2)));
  M3 d;
#pragma unroll
  for (int e = 0; e < 9; e++) d.e[e] = make_double2(naik * t.e[e].x, naik * t.e[e].y);
  if (acc_out) madd(d, ldf(o, y));
  stf(o, y, d);
}

__device__ __forceinline__ M3 minv(const M3 &a) {
  auto mul = [](double2 x, double2 y) { return make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x); };
  auto sub = [](double2 x, double2 y) { return make_double2(x.x - y.x, x.y - y.y); };
  const double2 *A = a.e;
  M3 r;
  r.e[0] = sub(mul(A[4], A[8]), mul(A[5], A[7]));
  r.e[3] = sub(mul(A[5], A[6]), mul(A[3], A[8]));
  r.e[6] = sub(mul(A[3], A[7]), mul(A[4], A[6]));
  r.e[1] = sub(mul(A[2], A[7]), mul(A[1], A[8]));
  r.e[4] = sub(mul(A[0], A[8]), mul(A[2], A[6]));
  r.e[7] = sub(mul(A[1], A[6]), mul(A[0], A[7]));
  r.e[2] = sub(mul(A[1], A[5]), mul(A[2], A[4]));
  r.e[5] = sub(mul(A[2], A[3]), mul(A[0], A[5]));
  r.e[8] = sub(mul(A[0], A[4]), mul(A[1], A[3]));
  const double2 det = make_double2(A[0].x * r.e[0].x - A[0].y * r.e[0].y + A[1].x * r.e[3].x - A[1].y * r.e[3].y +
                                       A[2].x * r.e[6].x - A[2].y * r.e[6].y,
                                   A[0].x * r.e[0].y + A[0].y * r.e[0].x + A[1].x * r.e[3].y + A[1].y * r.e[3].x +
                                       A[2].x * r.e[6].y + A[2].y * r.e[6].x);
  const double n = 1.0 / (det.x * det.x + det.y * det.y);
  const double2 id = make_double2(det.x * n, -det.y * n);
#pragma unroll
  for (int e = 0; e < 9; e++) r.e[e] = mul(r.e[e], id);
  return r;
}

// A X + X A = C for Hermitian positive A: with p(l) = l^3 + a2 l^2 + a1 l + a0
// the characteristic polynomial of A, p(A) = 0 applied to
// A^k X = X (-A)^k + sum_j A^j C (-A)^{k-1-j} gives
// X = -[a1 C + a2 (AC - CA) + C A^2 - A C A + A^2 C] p(-A)^{-1}.
__device__ __forceinline__ M3 sylsolve(const M3 &A, const M3 &C) {
  const M3 A2 = mmul(A, A), AC = mmul(A, C), CA = mmul(C, A);
  const M3 ACA = mmul(AC, A), CA2 = mmul(C, A2), A2C = mmul(A2, C), A3 = mmul(A2, A);
  // A Hermitian: the invariants are real
  const double tr = A.e[0].x + A.e[4].x + A.e[8].x, tr2 = A2.e[0].x + A2.e[4].x + A2.e[8].x;
  const double2 z01 = A.e[1], z02 = A.e[2], z12 = A.e[5];
  const double o01 = z01.x * z01.x + z01.y * z01.y, o02 = z02.x * z02.x + z02.y * z02.y;
  const double o12 = z12.x * z12.x + z12.y * z12.y;
  const double re3 = (z01.x * z12.x - z01.y * z12.y) * z02.x + (z01.x * z12.y + z01.y * z12.x) * z02.y;
  const double det = A.e[0].x * A.e[4].x * A.e[8].x + 2.0 * re3 - A.e[0].x * o12 - A.e[4].x * o02 - A.e[8].x * o01;
  const double a2 = -tr, a1 = 0.5 * (tr * tr - tr2), a0 = -det;
  M3 N, P;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    N.e[e] = make_double2(a1 * C.e[e].x + a2 * (AC.e[e].x - CA.e[e].x) + CA2.e[e].x - ACA.e[e].x + A2C.e[e].x,
                          a1 * C.e[e].y + a2 * (AC.e[e].y - CA.e[e].y) + CA2.e[e].y - ACA.e[e].y + A2C.e[e].y);
    P.e[e] = make_double2(-A3.e[e].x + a2 * A2.e[e].x - a1 * A.e[e].x, -A3.e[e].y + a2 * A2.e[e].y - a1 * A.e[e].y);
  }
#pragma unroll
  for (int k = 0; k < 3; k++) P.e[4 * k].x += a0;
  M3 X = mmul(N, minv(P));
#pragma unroll
  for (int e = 0; e < 9; e++) X.e[e] = make_double2(-X.e[e].x, -X.e[e].y);
  return X;
}

// projectUderiv: for u = x z, z = (x^+ x)^{-1/2}, y = z^{-1}, chain c = dS/du:
// dS/dx = c z - x (t + t^+) with y t + t y = u^+ c z
__device__ __forceinline__ M3 project_u_deriv(const M3 &Uu, const M3 &X, const M3 &C) {
  const M3 Z = rsqrt_xdx(X);
  const M3 Y = minv(Z);
  M3 R = mmul(C, Z);
  const M3 t2 = sylsolve(Y, mmul_dn(Uu, R));
  M3 sh;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      sh.e[3 * i + j] = make_double2(t2.e[3 * i + j].x + t2.e[3 * j + i].x, t2.e[3 * i + j].y - t2.e[3 * j + i].y);
  const M3 xs = mmul(X, sh);
#pragma unroll
  for (int e = 0; e < 9; e++) R.e[e] = make_double2(R.e[e].x - xs.e[e].x, R.e[e].y - xs.e[e].y);
  return R;
}

__global__ void __launch_bounds__(256) k_projderiv(long V, MF u0, MF u1, MF u2, MF u3, MF x0, MF x1, MF x2, MF x3,
                                                   MF c0, MF c1, MF c2, MF c3, MF r0, MF r1, MF r2, MF r3) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (x >= V) return;
  const MF u = mu == 0 ? u0 : (mu == 1 ? u1 : (mu == 2 ? u2 : u3));
  const MF xx = mu == 0 ? x0 : (mu == 1 ? x1 : (mu == 2 ? x2 : x3));
  const MF c = mu == 0 ? c0 : (mu == 1 ? c1 : (mu == 2 ? c2 : c3));
  const MF r = mu == 0 ? r0 : (mu == 1 ? r1 : (mu == 2 ? r2 : r3));
  stf(r, x, project_u_deriv(ldf(u, x), ldf(xx, x), ldf(c, x)));
}

// F = TAH(D U^+) per link, site-major arrays; F may alias D
__global__ void __launch_bounds__(256) k_tah_mul(long n, double2 *F, const double2 *D, const double2 *U) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MF d{const_cast<double2 *>(D), 1, 9}, u{const_cast<double2 *>(U), 1, 9}, f{F, 1, 9};
  const M3 T = mmul_nd(ldf(d, i), ldf(u, i));
  M3 R;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const double2 x = T.e[3 * a + b], y = T.e[3 * b + a];
      R.e[3 * a + b] = make_double2(0.5 * (x.x - y.x), 0.5 * (x.y + y.y));
    }
  const double tri = (R.e[0].y + R.e[4].y + R.e[8].y) * (1.0 / 3.0);
  const double trr = (R.e[0].x + R.e[4].x + R.e[8].x) * (1.0 / 3.0);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    R.e[4 * k].x -= trr;
    R.e[4 * k].y -= tri;
  }
  stf(f, i, R);
}

// ------------------------------------------------------------------ host side

static MF site_major(double *base, long V, int mu) { return MF{(double2 *)base + (long)mu * V * 9, 1, 9}; }
static MF comp_major(double2 *base, long V, int k) { return MF{base + (long)k * V * 9, V, 1}; }

SmearWs::~SmearWs() {
  if (buf) (void)hipFree(buf);
}

// Workspace fields (component-major, V x 9 complex each):
//   0-3 level input U, 4-15 S, 16-27 T, 28-31 first-level fat links V1,
//   32-35 projected links W; the chain rule adds 36-47 CT, 48-59 CS,
//   60-71 side-link partials at the site, 72-75 fat chain, 76-79 long chain,
//   80-83 dS/dW, 84-87 dS/dV1, 88-99 side-link partials one site back.
//   100-107 HISQ force outer products, site-major [k][site] (smear_force_outer).
enum { WS_SMEAR = 36, WS_DERIV = 100, WS_FORCE = 108 };

// largest lattice of the smearing kernels' 32-bit view offsets (a site-major
// 3x3 field below 4 GiB); its workspace could not fit in device memory anyway
const long kMaxSmearSites = 0xffffffffL / 144;

static void ws_ensure(SmearWs &w, long V, int nfields) {
  if (V > kMaxSmearSites) throw std::runtime_error("HISQ smearing: lattice above 29826161 sites");
  if (w.buf && w.V == V && w.nfields >= nfields) return;
  if (w.buf) (void)hipFree(w.buf);
  w.buf = nullptr;
  w.V = V;
  w.nfields = 0;
  const size_t bytes = (size_t)nfields * V * 9 * sizeof(double2);
  if (hipMalloc((void **)&w.buf, bytes) != hipSuccess) {
    w.buf = nullptr;
    throw std::runtime_error("smearing workspace: out of device memory");
  }
  w.nfields = nfields;
}

static void evrec(hipEvent_t e, hipStream_t st) {
  if (hipEventRecord(e, st) != hipSuccess) throw std::runtime_error("hipEventRecord failed");
}

static unsigned nblk(long V) { return (unsigned)((V + 255) / 256); }

// workgroups of 64 sites, a multiple of 8 when they are remapped to XCDs
static unsigned site_groups(long V, int remap) {
  const unsigned n = (unsigned)((V + 63) / 64);
  return (remap & 1) ? (n + 7) & ~7u : n;
}

// remap bit 0: XCD-contiguous groups; bit 1: t-first walk (when the t-slice
// of one parity is a whole number of groups)
static int walk_mode(const Geom &g, int remap, int walk) {
  const long S = (long)g.hx * g.L[1] * g.L[2];
  return (remap ? 1 : 0) | ((walk && S % 64 == 0) ? 2 : 0);
}

// one fat7 level: input links U (4 views), output fat links out (4 views)
static void fat7_level(const Geom &g, SmearWs &w, const MF U[4], const MF out[4], const double coef[5], int remap,
                       hipStream_t st, SmearTimes *tm) {
  const long V = 2 * g.Vh;
  SmearFields f;
  for (int k = 0; k < 4; k++) {
    f.U[k] = U[k];
    f.out[k] = out[k];
  }
  for (int k = 0; k < 12; k++) {
    f.S[k] = comp_major(w.buf, V, 4 + k);
    f.T[k] = comp_major(w.buf, V, 16 + k);
  }
  const double cL = coef[4];
  const Fat7 c{coef[0] - 6.0 * cL, coef[1], coef[2], coef[3], cL};
  const bool have5 = c.c5 != 0.0 || c.c7 != 0.0 || c.cL != 0.0;
  const bool have3 = c.c3 != 0.0 || have5;
  const unsigned nb = site_groups(V, remap);
  if (tm) evrec(tm->ev[0], st);
  if (have3) hipLaunchKernelGGL(k_fat3, dim3(nb), dim3(256), 0, st, g, f, remap);
  if (tm) evrec(tm->ev[1], st);
  if (c.c5 != 0.0 || c.c7 != 0.0) hipLaunchKernelGGL(k_fat5, dim3(nb, 4), dim3(192), 0, st, g, f, remap);
  if (tm) evrec(tm->ev[2], st);
  if (have5)
    hipLaunchKernelGGL((k_fatf<2>), dim3(nb), dim3(256), 0, st, g, f, c, remap);
  else if (have3)
    hipLaunchKernelGGL((k_fatf<1>), dim3(nb), dim3(256), 0, st, g, f, c, remap);
  else
    hipLaunchKernelGGL((k_fatf<0>), dim3(nb), dim3(256), 0, st, g, f, c, remap);
  if (tm) evrec(tm->ev[3], st);
}

static void naik_links(const Geom &g, const MF W[4], const MF o[4], double naik, int remap, hipStream_t st) {
  const long V = 2 * g.Vh;
  hipLaunchKernelGGL(k_naik, dim3(site_groups(V, remap)), dim3(256), 0, st, g, W[0], W[1], W[2], W[3], o[0], o[1], o[2], o[3], naik,
                     remap);
}

static void relayout(long V, const MF in[4], const MF out[4], hipStream_t st) {
  hipLaunchKernelGGL(k_relayout, dim3(nblk(V), 4), dim3(256), 0, st, V, in[0], in[1], in[2], in[3], out[0], out[1],
                     out[2], out[3]);
}

static void project(long V, const MF in[4], const MF out[4], hipStream_t st) {
  hipLaunchKernelGGL(k_project, dim3(nblk(V), 4), dim3(256), 0, st, V, in[0], in[1], in[2], in[3], out[0], out[1],
                     out[2], out[3]);
}

void launch_make_imp_links(const Geom &g, SmearWs &w, double *fl, const double *gf, const double coef[5], double *ll,
                           const double *gl, double naik, int remap, hipStream_t st, SmearTimes *tm) {
  remap = walk_mode(g, remap & 1, remap & 2);
  const long V = 2 * g.Vh;
  ws_ensure(w, V, WS_SMEAR);
  MF in[4], U[4], out[4];
  for (int k = 0; k < 4; k++) {
    in[k] = site_major(const_cast<double *>(gf), V, k);
    U[k] = comp_major(w.buf, V, k);
    out[k] = site_major(fl, V, k);
  }
  relayout(V, in, U, st);
  fat7_level(g, w, U, out, coef, remap, st, tm);
  if (ll && naik != 0.0) {
    MF L[4], o[4];
    for (int k = 0; k < 4; k++) {
      L[k] = gl ? site_major(const_cast<double *>(gl), V, k) : U[k];
      o[k] = site_major(ll, V, k);
    }
    naik_links(g, L, o, naik, remap, st);
  }
}

// the projections of arbitrary arrays of n matrices, in views of at most
// 2^24 matrices (2.4 GB, within the views' 32-bit offsets)
const long kProjChunk = 1L << 24;

void launch_project_u(long n, double *u, const double *x, hipStream_t st) {
  for (long o = 0; o < n; o += kProjChunk) {
    const long c = std::min(kProjChunk, n - o);
    // c matrices as one view of c sites
    const MF in{(double2 *)const_cast<double *>(x) + o * 9, 1, 9}, out{(double2 *)u + o * 9, 1, 9};
    hipLaunchKernelGGL(k_project, dim3(nblk(c), 1), dim3(256), 0, st, c, in, in, in, in, out, out, out, out);
  }
}

void launch_hisq_smear(const Geom &g, SmearWs &w, const double first[5], const double second[5], double naik,
                       const double *gin, double *fl, double *ll, int remap, hipStream_t st, SmearTimes *tm) {
  remap = walk_mode(g, remap & 1, remap & 2);
  const long V = 2 * g.Vh;
  ws_ensure(w, V, WS_SMEAR);
  MF in[4], U[4], V1[4], W[4], out[4], lo[4];
  for (int k = 0; k < 4; k++) {
    in[k] = site_major(const_cast<double *>(gin), V, k);
    U[k] = comp_major(w.buf, V, k);
    V1[k] = comp_major(w.buf, V, 28 + k);
    W[k] = comp_major(w.buf, V, 32 + k);
    out[k] = site_major(fl, V, k);
    lo[k] = site_major(ll, V, k);
  }
  relayout(V, in, U, st);
  fat7_level(g, w, U, V1, first, remap, st, tm);
  project(V, V1, W, st);
  fat7_level(g, w, W, out, second, remap, st, tm ? tm + 1 : nullptr);
  if (naik != 0.0) naik_links(g, W, lo, naik, remap, st);
}

static DerivFields deriv_fields(SmearWs &w, long V, const MF U[4], const MF C[4], const MF out[4]) {
  DerivFields f;
  for (int k = 0; k < 4; k++) {
    f.U[k] = U[k];
    f.C[k] = C[k];
    f.out[k] = out[k];
  }
  for (int k = 0; k < 12; k++) {
    f.S[k] = comp_major(w.buf, V, 4 + k);
    f.T[k] = comp_major(w.buf, V, 16 + k);
    f.CT[k] = comp_major(w.buf, V, 36 + k);
    f.CS[k] = comp_major(w.buf, V, 48 + k);
    f.P[k] = comp_major(w.buf, V, 60 + k);
    f.Q[k] = comp_major(w.buf, V, 88 + k);
  }
  return f;
}

// reverse of one fat7 level: out (+)= dS/dU given C = dS/dF; rebuilds S and T
static void fat7_deriv_level(const Geom &g, SmearWs &w, const MF U[4], const MF C[4], const MF out[4],
                             const double coef[5], int acc, int remap, hipStream_t st) {
  const long V = 2 * g.Vh;
  const double cL = coef[4];
  const Fat7 c{coef[0] - 6.0 * cL, coef[1], coef[2], coef[3], cL};
  const bool have57 = c.c5 != 0.0 || c.c7 != 0.0;
  const bool have5 = have57 || c.cL != 0.0;
  const bool have3 = c.c3 != 0.0 || have5;
  const unsigned nb = site_groups(V, remap);
  SmearFields sf;
  for (int k = 0; k < 4; k++) {
    sf.U[k] = U[k];
    sf.out[k] = out[k];
  }
  for (int k = 0; k < 12; k++) {
    sf.S[k] = comp_major(w.buf, V, 4 + k);
    sf.T[k] = comp_major(w.buf, V, 16 + k);
  }
  const DerivFields f = deriv_fields(w, V, U, C, out);
  if (have3) {
    hipLaunchKernelGGL(k_fat3, dim3(nb), dim3(256), 0, st, g, sf, remap);
    if (have57) {
      hipLaunchKernelGGL(k_fat5, dim3(nb, 4), dim3(192), 0, st, g, sf, remap);
      hipLaunchKernelGGL(k_dct, dim3(nb, 4), dim3(192), 0, st, g, f, c, remap);
    }
    if (have5) {
      hipLaunchKernelGGL((k_dcs<1>), dim3(nb, 4), dim3(192), 0, st, g, f, c, remap);
      hipLaunchKernelGGL((k_dside<1>), dim3(nb, 4), dim3(384), 0, st, g, f, c, remap);
    } else {
      hipLaunchKernelGGL((k_dcs<0>), dim3(nb, 4), dim3(192), 0, st, g, f, c, remap);
      hipLaunchKernelGGL((k_dside<0>), dim3(nb, 4), dim3(384), 0, st, g, f, c, remap);
    }
    hipLaunchKernelGGL((k_dfin<1>), dim3(nb), dim3(256), 0, st, g, f, c, acc, remap);
  } else {
    hipLaunchKernelGGL((k_dfin<0>), dim3(nb), dim3(256), 0, st, g, f, c, acc, remap);
  }
}

static void naik_deriv(const Geom &g, const MF W[4], const MF C[4], const MF o[4], double naik, int acc, int remap,
                       hipStream_t st) {
  const long V = 2 * g.Vh;
  hipLaunchKernelGGL(k_dnaik, dim3(site_groups(V, remap)), dim3(256), 0, st, g, W[0], W[1], W[2], W[3], C[0], C[1],
                     C[2], C[3], o[0], o[1], o[2], o[3], naik, acc, remap);
}

void launch_fat7l_deriv(const Geom &g, SmearWs &w, double *deriv, const double *gf, const double *mid,
                        const double coef[5], double *llderiv, const double *gl, const double *llmid, double naik,
                        int remap, hipStream_t st) {
  remap = walk_mode(g, remap & 1, remap & 2);
  const long V = 2 * g.Vh;
  ws_ensure(w, V, WS_DERIV);
  MF in[4], U[4], Cin[4], C[4], out[4];
  for (int k = 0; k < 4; k++) {
    in[k] = site_major(const_cast<double *>(gf), V, k);
    U[k] = comp_major(w.buf, V, k);
    Cin[k] = site_major(const_cast<double *>(mid), V, k);
    C[k] = comp_major(w.buf, V, 72 + k);
    out[k] = site_major(deriv, V, k);
  }
  relayout(V, in, U, st);
  relayout(V, Cin, C, st);
  fat7_deriv_level(g, w, U, C, out, coef, 1, remap, st);
  if (llderiv && llmid && naik != 0.0) {
    MF L[4], Cl[4], o[4];
    for (int k = 0; k < 4; k++) {
      L[k] = gl ? site_major(const_cast<double *>(gl), V, k) : U[k];
      Cl[k] = site_major(const_cast<double *>(llmid), V, k);
      o[k] = site_major(llderiv, V, k);
    }
    naik_deriv(g, L, Cl, o, naik, 1, remap, st);
  }
}

void launch_project_u_deriv(long n, double *r, const double *u, const double *x, const double *chain,
                            hipStream_t st) {
  for (long o = 0; o < n; o += kProjChunk) {
    const long c = std::min(kProjChunk, n - o);
    const MF U{(double2 *)const_cast<double *>(u) + o * 9, 1, 9}, X{(double2 *)const_cast<double *>(x) + o * 9, 1, 9};
    const MF C{(double2 *)const_cast<double *>(chain) + o * 9, 1, 9}, R{(double2 *)r + o * 9, 1, 9};
    hipLaunchKernelGGL(k_projderiv, dim3(nblk(c), 1), dim3(256), 0, st, c, U, U, U, U, X, X, X, X, C, C, C, C, R, R,
                       R, R);
  }
}

void launch_hisq_smear_deriv(const Geom &g, SmearWs &w, const double first[5], const double second[5], double naik,
                             const double *gin, const double *dfl, const double *dll, long mstride, double *dg,
                             int acc, int remap, hipStream_t st) {
  remap = walk_mode(g, remap & 1, remap & 2);
  const long V = 2 * g.Vh;
  ws_ensure(w, V, WS_DERIV);
  MF in[4], U[4], V1[4], W[4], Cf[4], Cl[4], Cfc[4], Clc[4], t[4], t2[4], out[4];
  for (int k = 0; k < 4; k++) {
    in[k] = site_major(const_cast<double *>(gin), V, k);
    U[k] = comp_major(w.buf, V, k);
    V1[k] = comp_major(w.buf, V, 28 + k);
    W[k] = comp_major(w.buf, V, 32 + k);
    Cf[k] = MF{(double2 *)const_cast<double *>(dfl) + (long)k * mstride * 9, 1, 9};
    Cl[k] = MF{(double2 *)const_cast<double *>(dll) + (long)k * mstride * 9, 1, 9};
    Cfc[k] = comp_major(w.buf, V, 72 + k);
    Clc[k] = comp_major(w.buf, V, 76 + k);
    t[k] = comp_major(w.buf, V, 80 + k);
    t2[k] = comp_major(w.buf, V, 84 + k);
    out[k] = site_major(dg, V, k);
  }
  // forward: V1 = fat7(U; first), W = projectU(V1)
  relayout(V, in, U, st);
  fat7_level(g, w, U, V1, first, remap, st, nullptr);
  project(V, V1, W, st);
  // second level and Naik: t = dS/dW
  relayout(V, Cf, Cfc, st);
  relayout(V, Cl, Clc, st);
  fat7_deriv_level(g, w, W, Cfc, t, second, 0, remap, st);
  if (naik != 0.0) naik_deriv(g, W, Clc, t, naik, 1, remap, st);
  // projection: t2 = dS/dV1
  hipLaunchKernelGGL(k_projderiv, dim3(nblk(V), 4), dim3(256), 0, st, V, W[0], W[1], W[2], W[3], V1[0], V1[1], V1[2],
                     V1[3], t[0], t[1], t[2], t[3], t2[0], t2[1], t2[2], t2[3]);
  // first level: dg (+)= dS/dU
  fat7_deriv_level(g, w, U, t2, out, first, acc, remap, st);
}

double *smear_force_outer(SmearWs &w, long V) {
  ws_ensure(w, V, WS_FORCE);
  return (double *)(w.buf + (size_t)WS_DERIV * V * 9);
}

void launch_tah_mul(long V, double *F, const double *D, const double *U, hipStream_t st) {
  const long n = 4 * V;
  hipLaunchKernelGGL(k_tah_mul, dim3(nblk(n)), dim3(256), 0, st, n, (double2 *)F, (const double2 *)D,
                     (const double2 *)U);
}

}  // namespace qh
