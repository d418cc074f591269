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
#include <stdexcept>

#include "internal.h"
#include "su3.h"

namespace qh {

struct MF {
  double2 *p;
  long es, ss;
};

__device__ __forceinline__ M3 ldf(const MF &f, long s) {
  M3 m;
#pragma unroll
  for (int e = 0; e < 9; e++) m.e[e] = f.p[e * f.es + s * f.ss];
  return m;
}

__device__ __forceinline__ void stf(const MF &f, long s, const M3 &m) {
#pragma unroll
  for (int e = 0; e < 9; e++) f.p[e * f.es + s * f.ss] = m.e[e];
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

// S_nu[K](x), x with coordinates c
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
__device__ __forceinline__ long site_of(int remap) { return xcd_block(remap) * 64 + (threadIdx.x & 63); }
__device__ __forceinline__ int wave_of() { return threadIdx.x >> 6; }

__global__ void __launch_bounds__(256) k_fat3(Geom g, SmearFields f, int remap) {
  const long V = 2 * g.Vh;
  const long x = site_of(remap);
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
  const long x = site_of(remap);
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
  const long x = site_of(remap);
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
  const long x = site_of(remap);
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

// u = x (x^+ x + 1e-20)^{-1/2}.  The eigenvalues s_i^2 of H = x^+ x + 1e-20 come
// from the trigonometric solution of its characteristic cubic (Smith, CACM 4
// (1961) 168); H^{-1/2} = f0 + f1 H + f2 H^2 is the quadratic through 1/s_i:
// with u1 = sum s_i, v = sum_{i<j} s_i s_j, w = prod s_i, d = w (u1 v - w):
// f0 = (u1 v^2 - v w - u1^2 w)/d, f1 = -(u1 tr H + w)/d, f2 = u1/d.
__device__ __forceinline__ M3 project_u(const M3 &X) {
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
  M3 R;
#pragma unroll
  for (int e = 0; e < 9; e++) R.e[e] = make_double2(f1 * H.e[e].x + f2 * H2.e[e].x, f1 * H.e[e].y + f2 * H2.e[e].y);
#pragma unroll
  for (int k = 0; k < 3; k++) R.e[4 * k].x += f0;
  return mmul(X, R);
}

// n matrices; matrix i of field i / V at site i % V (per-field views)
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
  for (int e = 0; e < 9; e++) o.p[e * o.es + x * o.ss] = in.p[e * in.es + x * in.ss];
}

// ------------------------------------------------------------------ host side

static MF site_major(double *base, long V, int mu) { return MF{(double2 *)base + (long)mu * V * 9, 1, 9}; }
static MF comp_major(double2 *base, long V, int k) { return MF{base + (long)k * V * 9, V, 1}; }

SmearWs::~SmearWs() {
  if (buf) (void)hipFree(buf);
}

static void ws_ensure(SmearWs &w, long V) {
  if (w.buf && w.V == V) return;
  if (w.buf) (void)hipFree(w.buf);
  w.buf = nullptr;
  w.V = V;
  // U, S, T, level-1 result, projected links: 4 + 12 + 12 + 4 + 4 fields
  const size_t bytes = (size_t)36 * V * 9 * sizeof(double2);
  if (hipMalloc((void **)&w.buf, bytes) != hipSuccess) {
    w.buf = nullptr;
    throw std::runtime_error("smearing workspace: out of device memory");
  }
}

static void evrec(hipEvent_t e, hipStream_t st) {
  if (hipEventRecord(e, st) != hipSuccess) throw std::runtime_error("hipEventRecord failed");
}

static unsigned nblk(long V) { return (unsigned)((V + 255) / 256); }

// workgroups of 64 sites, a multiple of 8 when they are remapped to XCDs
static unsigned site_groups(long V, int remap) {
  const unsigned n = (unsigned)((V + 63) / 64);
  return remap ? (n + 7) & ~7u : n;
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
  const long V = 2 * g.Vh;
  ws_ensure(w, V);
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

void launch_project_u(long n, double *u, const double *x, hipStream_t st) {
  // n matrices as one view of n sites
  const MF in{(double2 *)const_cast<double *>(x), 1, 9}, out{(double2 *)u, 1, 9};
  hipLaunchKernelGGL(k_project, dim3(nblk(n), 1), dim3(256), 0, st, n, in, in, in, in, out, out, out, out);
}

void launch_hisq_smear(const Geom &g, SmearWs &w, const double first[5], const double second[5], double naik,
                       const double *gin, double *fl, double *ll, int remap, hipStream_t st, SmearTimes *tm) {
  const long V = 2 * g.Vh;
  ws_ensure(w, V);
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

}  // namespace qh
