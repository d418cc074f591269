/*
 * qexoracle.c -- CPU restatement of QEX's staggered Dslash / CG / CGM / force.
 *
 * TEST INFRASTRUCTURE ONLY (the parity reference and the CPU baseline); see
 * qexoracle.h for the conventions.  Every routine cites the QEX source it
 * restates (paths relative to QEX's repository root).
 *
 * Parity status of each routine (what pins it, tests/test_oracle_*.py):
 *  qo_philox4x32_10 : Random123 known-answer vectors.
 *  qo_stagD2/qo_stagD/qo_D/qo_Ddag/qo_stagD2xx : closed-form point-source norms
 *      (stagD.nim:735-752), free-field spectrum via FFT, dense 4^4 matrix
 *      (anti-Hermiticity, eps-symmetry, spectrum m +- i lambda), gauge covariance.
 *  qo_cg/qo_solveXX : dense LU solve on 4^4, FFT solution in the free field,
 *      true-residual recomputation.
 *  qo_cgm           : every shifted system's own residual, agreement with
 *      independent single-shift CG.
 *  qo_ferm_force    : central finite differences along random su(3)
 *      directions, gauge covariance.
 *  qo_plaquette / qo_wilson_force / qo_random_momenta / qo_link_update :
 *      unit gauge = 18 V exactly and gauge invariance; central differences
 *      through an independent numpy staple sum; anti-Hermitian traceless
 *      momenta with <|P|^2/2> = 4 per link; scipy's expm.
 *  qo_make_imp_links / qo_hisq_smear : unit-link path counts (fat7l.nim:183-217),
 *      phased unit links (9/8 U, -1/24 U U U), an explicit sum over the 90
 *      link paths, gauge covariance (tests/test_oracle_hisq.py).
 *  qo_project_u     : numpy's SVD polar factor.
 *  qo_fat7l_deriv / qo_project_u_deriv / qo_hisq_smear_deriv / qo_hisq_outer :
 *      exact (9-point) and central differences of the smeared action; the
 *      trivial smearing reproduces qo_ferm_force.
 *  qo_milc_checksum : answers worked out by hand (tests/test_oracle_milc.py).
 *  qo_crc32 / qo_scidac_checksum : the reference's CRC-32 known answer
 *      (crc32.nim:113-116), zlib's crc32, and the rank rotations recomputed
 *      site by site with zlib (tests/test_oracle_scidac.py).
 */
#include "qexoracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ------------------------------------------------------------------ layout */

long qo_volume(const int L[4]) { return (long)L[0] * L[1] * L[2] * L[3]; }

/* layout/qlayout.nim:108-136 with innerGeom = 1, one rank */
long qo_index(const int L[4], const int x[4]) {
  long V = qo_volume(L);
  long lex = x[0] + (long)L[0] * (x[1] + (long)L[1] * (x[2] + (long)L[2] * x[3]));
  int p = (x[0] + x[1] + x[2] + x[3]) & 1;
  return (lex + p * V) / 2;
}

void qo_coords(const int L[4], long idx, int x[4]) {
  long Vh = qo_volume(L) / 2;
  int p = idx >= Vh;
  long i = idx - p * Vh;
  long hx = L[0] / 2;
  long xh = i % hx, r = i / hx;
  x[1] = (int)(r % L[1]); r /= L[1];
  x[2] = (int)(r % L[2]);
  x[3] = (int)(r / L[2]);
  x[0] = (int)(2 * xh + ((p + x[1] + x[2] + x[3]) & 1));
}

static inline long nbr(const int L[4], const int c[4], int mu, int d) {
  int y[4] = {c[0], c[1], c[2], c[3]};
  y[mu] = ((y[mu] + d) % L[mu] + L[mu]) % L[mu];
  return qo_index(L, y);
}

static inline void par_range(const int L[4], int par, long *lo, long *hi) {
  long V = qo_volume(L), Vh = V / 2;
  *lo = (par == 1) ? Vh : 0;
  *hi = (par == 0) ? Vh : V;
}

/* --------------------------------------------------------------- generator */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void qo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int rnd = 0; rnd < 10; rnd++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* uniform in (0,1) from 53 bits of (a,b); Box-Muller pair */
void qo_gauss2(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
               double g[2]) {
  uint32_t ctr[4] = {c0, c1, c2, c3};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  qo_philox4x32_10(ctr, key, o);
  uint64_t a = (((uint64_t)o[0] << 32) | o[1]) >> 11;
  uint64_t b = (((uint64_t)o[2] << 32) | o[3]) >> 11;
  double u1 = ((double)a + 0.5) * 0x1.0p-53;
  double u2 = ((double)b + 0.5) * 0x1.0p-53;
  double rad = sqrt(-2.0 * log(u1));
  double th = 6.283185307179586476925286766559 * u2;
  g[0] = rad * cos(th);
  g[1] = rad * sin(th);
}

/* ------------------------------------------------------------ 3x3 algebra */
void qo_mat_mul(const double *a, const double *b, double *c) {
  const cplx *A = (const cplx *)a, *B = (const cplx *)b;
  cplx C[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cplx s = 0;
      for (int k = 0; k < 3; k++) s += A[3 * i + k] * B[3 * k + j];
      C[3 * i + j] = s;
    }
  memcpy(c, C, sizeof C);
}

/* exp(A) for a general complex 3x3 A: scaling and squaring with a degree-16
   Taylor polynomial (Horner form) */
void qo_su3_exp(const double *a, double *u) {
  const cplx *A = (const cplx *)a;
  double nrm = 0;
  for (int i = 0; i < 3; i++) {
    double s = 0;
    for (int j = 0; j < 3; j++) s += cabs(A[3 * i + j]);
    if (s > nrm) nrm = s;
  }
  int sq = 0;
  while (nrm > 0.5) { nrm *= 0.5; sq++; }
  double sc = ldexp(1.0, -sq);
  cplx B[9], T[9], W[9];
  for (int i = 0; i < 9; i++) B[i] = A[i] * sc;
  for (int i = 0; i < 9; i++) T[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int n = 16; n >= 1; n--) {
    qo_mat_mul((double *)B, (double *)T, (double *)W);
    for (int i = 0; i < 9; i++) T[i] = W[i] / (double)n + ((i % 4 == 0) ? 1.0 : 0.0);
  }
  for (int s = 0; s < sq; s++) {
    qo_mat_mul((double *)T, (double *)T, (double *)W);
    memcpy(T, W, sizeof W);
  }
  memcpy(u, T, sizeof T);
}

/* U = exp(i*s*H) */
void qo_su3_exp_i(const double *h, double s, double *u) {
  const cplx *H = (const cplx *)h;
  cplx A[9];
  for (int i = 0; i < 9; i++) A[i] = I * s * H[i];
  qo_su3_exp((double *)A, u);
}

/* maths/matrixFunctions.nim:414-419: r = (x - x^+)/2 - tr/3 */
void qo_project_tah(const double *x, double *r) {
  const cplx *X = (const cplx *)x;
  cplx R[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) R[3 * i + j] = 0.5 * (X[3 * i + j] - conj(X[3 * j + i]));
  cplx tr = (R[0] + R[4] + R[8]) / 3.0;
  R[0] -= tr; R[4] -= tr; R[8] -= tr;
  memcpy(r, R, sizeof R);
}

/* ------------------------------------------------------- synthetic inputs */
enum { TAG_GAUGE = 0, TAG_SOURCE = 1, TAG_FATG = 2, TAG_MOM = 3 };

void qo_gauge_random(const int L[4], uint64_t seed, double *U) {
  long V = qo_volume(L);
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < V; idx++) {
    int c[4];
    qo_coords(L, idx, c);
    uint32_t lex = (uint32_t)(c[0] + (long)L[0] * (c[1] + (long)L[1] * (c[2] + (long)L[2] * c[3])));
    for (int mu = 0; mu < 4; mu++) {
      double g[10];
      for (int j = 0; j < 5; j++) qo_gauss2(seed, lex, (uint32_t)mu, (uint32_t)j, TAG_GAUGE, g + 2 * j);
      cplx H[9];
      double tr = (g[0] + g[1] + g[2]) / 3.0;
      H[0] = g[0] - tr; H[4] = g[1] - tr; H[8] = g[2] - tr;
      H[1] = g[3] + I * g[4]; H[3] = conj(H[1]);
      H[2] = g[5] + I * g[6]; H[6] = conj(H[2]);
      H[5] = g[7] + I * g[8]; H[7] = conj(H[5]);
      qo_su3_exp_i((double *)H, 0.5, U + ((long)mu * V + idx) * 18);
    }
  }
}

void qo_gauge_unit(const int L[4], double *U) {
  long V = qo_volume(L);
  memset(U, 0, sizeof(double) * 4 * V * 18);
  for (long i = 0; i < 4 * V; i++) U[i * 18 + 0] = U[i * 18 + 8] = U[i * 18 + 16] = 1.0;
}

void qo_source_gauss(const int L[4], uint64_t seed, int stream, int par, double *x) {
  long lo, hi;
  par_range(L, par, &lo, &hi);
  const double s = sqrt(0.5);
#pragma omp parallel for schedule(static)
  for (long idx = lo; idx < hi; idx++) {
    int c[4];
    qo_coords(L, idx, c);
    uint32_t lex = (uint32_t)(c[0] + (long)L[0] * (c[1] + (long)L[1] * (c[2] + (long)L[2] * c[3])));
    for (int j = 0; j < 3; j++) {
      double g[2];
      qo_gauss2(seed, lex, 64u + (uint32_t)stream, (uint32_t)j, TAG_SOURCE, g);
      x[idx * 6 + 2 * j] = s * g[0];
      x[idx * 6 + 2 * j + 1] = s * g[1];
    }
  }
}

/* gauge/gaugeUtils.nim:133-140 */
void qo_set_bc(const int L[4], double *U) {
  long V = qo_volume(L);
  for (long idx = 0; idx < V; idx++) {
    int c[4];
    qo_coords(L, idx, c);
    if (c[3] == L[3] - 1)
      for (int e = 0; e < 18; e++) U[(3 * V + idx) * 18 + e] *= -1;
  }
}

/* physics/stagD.nim:560-571, phases [8,9,11,0] */
void qo_stag_phase(const int L[4], double *U) {
  static const int ph[4] = {8, 9, 11, 0};
  long V = qo_volume(L);
  for (int mu = 0; mu < 4; mu++)
    for (long idx = 0; idx < V; idx++) {
      int c[4];
      qo_coords(L, idx, c);
      int s = 0;
      for (int k = 0; k < 4; k++) s += (ph[mu] >> k) & c[k];
      if (s & 1)
        for (int e = 0; e < 18; e++) U[((long)mu * V + idx) * 18 + e] *= -1;
    }
}

/* W = 0.8 U + 0.05 G ; Lg = -(1/24) U_mu(x) U_mu(x+mu) U_mu(x+2mu), from the
   phased BC-applied U (gauge/fat7l.nim:145-154 long-link product, Naik
   coefficient -1/24 from physics/hisqLinks.nim:17-28); interleaving
   g[2mu]=fat, g[2mu+1]=long as newStag3 (physics/stagD.nim:603-615). */
void qo_hisq_links(const int L[4], uint64_t seed, const double *Uph, double *g8) {
  long V = qo_volume(L);
  const double s = sqrt(0.5);
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < V; idx++) {
    int c[4];
    qo_coords(L, idx, c);
    uint32_t lex = (uint32_t)(c[0] + (long)L[0] * (c[1] + (long)L[1] * (c[2] + (long)L[2] * c[3])));
    for (int mu = 0; mu < 4; mu++) {
      const double *u0 = Uph + ((long)mu * V + idx) * 18;
      double *w = g8 + ((long)(2 * mu) * V + idx) * 18;
      double *lg = g8 + ((long)(2 * mu + 1) * V + idx) * 18;
      for (int j = 0; j < 9; j++) {
        double g[2];
        qo_gauss2(seed, lex, 16u + (uint32_t)mu, (uint32_t)j, TAG_FATG, g);
        w[2 * j] = 0.8 * u0[2 * j] + 0.05 * s * g[0];
        w[2 * j + 1] = 0.8 * u0[2 * j + 1] + 0.05 * s * g[1];
      }
      long i1 = nbr(L, c, mu, 1), i2 = nbr(L, c, mu, 2);
      double t[18], t2[18];
      qo_mat_mul(u0, Uph + ((long)mu * V + i1) * 18, t);
      qo_mat_mul(t, Uph + ((long)mu * V + i2) * 18, t2);
      for (int e = 0; e < 18; e++) lg[e] = (-1.0 / 24.0) * t2[e];
    }
  }
}

/* ---------------------------------------------------------------- operator */
static inline void mv_add(cplx *acc, const cplx *M, const cplx *v, double sgn) {
  for (int i = 0; i < 3; i++) acc[i] += sgn * (M[3 * i] * v[0] + M[3 * i + 1] * v[1] + M[3 * i + 2] * v[2]);
}
static inline void amv_add(cplx *acc, const cplx *M, const cplx *v, double sgn) {
  for (int i = 0; i < 3; i++)
    acc[i] += sgn * (conj(M[i]) * v[0] + conj(M[3 + i]) * v[1] + conj(M[6 + i]) * v[2]);
}
static inline int hop_dir(int nl, int k) { return nl == 4 ? k : k >> 1; }
static inline int hop_len(int nl, int k) { return (nl == 8 && (k & 1)) ? 3 : 1; }

/* physics/stagD.nim:383-448 (stagD2): fwd and bwd interleaved per link */
void qo_stagD2(const int L[4], int nl, const double *g, int par,
               double *r, const double *x, double a, double b) {
  long V = qo_volume(L), lo, hi;
  par_range(L, par, &lo, &hi);
  const cplx *G = (const cplx *)g, *X = (const cplx *)x;
  cplx *R = (cplx *)r;
#pragma omp parallel for schedule(static)
  for (long idx = lo; idx < hi; idx++) {
    int c[4];
    qo_coords(L, idx, c);
    cplx acc[3];
    for (int i = 0; i < 3; i++)
      acc[i] = (a == 0) ? b * X[3 * idx + i] : a * R[3 * idx + i] + b * X[3 * idx + i];
    for (int k = 0; k < nl; k++) {
      int mu = hop_dir(nl, k), len = hop_len(nl, k);
      long jf = nbr(L, c, mu, len), jb = nbr(L, c, mu, -len);
      mv_add(acc, G + ((long)k * V + idx) * 9, X + 3 * jf, 1.0);
      amv_add(acc, G + ((long)k * V + jb) * 9, X + 3 * jb, -1.0);
    }
    for (int i = 0; i < 3; i++) R[3 * idx + i] = acc[i];
  }
}

/* physics/stagD.nim:459-462 */
void qo_stagD(const int L[4], int nl, const double *g, int par,
              double *r, const double *x, double m, double sc, double a) {
  qo_stagD2(L, nl, g, par, r, x, a / (0.5 * sc), m / (0.5 * sc));
  long lo, hi;
  par_range(L, par, &lo, &hi);
  for (long i = lo * 6; i < hi * 6; i++) r[i] *= 0.5 * sc;
}

/* physics/stagD.nim:617-622 */
void qo_D(const int L[4], int nl, const double *g, double *r, const double *x, double m) {
  qo_stagD(L, nl, g, 0, r, x, m, 1.0, 0.0);
  qo_stagD(L, nl, g, 1, r, x, m, 1.0, 0.0);
}
void qo_Ddag(const int L[4], int nl, const double *g, double *r, const double *x, double m) {
  qo_stagD(L, nl, g, 0, r, x, m, -1.0, 0.0);
  qo_stagD(L, nl, g, 1, r, x, m, -1.0, 0.0);
}

/* physics/stagD.nim:485-512 (stagD2xx): t = stagDP (rir:=0, all forward then all
   backward, :234-271); r = stagDM with rir := 4*m2*x (fwd subtract, bwd add per
   link, :312-347) */
void qo_stagD2xx(const int L[4], int nl, const double *g, int par,
                 double *r, const double *x, double m2, double *t) {
  long V = qo_volume(L), lo, hi;
  int q = 1 - par;
  const cplx *G = (const cplx *)g, *X = (const cplx *)x;
  cplx *R = (cplx *)r, *T = (cplx *)t;
  par_range(L, q, &lo, &hi);
#pragma omp parallel for schedule(static)
  for (long idx = lo; idx < hi; idx++) {
    int c[4];
    qo_coords(L, idx, c);
    cplx acc[3] = {0, 0, 0};
    for (int k = 0; k < nl; k++) {
      long jf = nbr(L, c, hop_dir(nl, k), hop_len(nl, k));
      mv_add(acc, G + ((long)k * V + idx) * 9, X + 3 * jf, 1.0);
    }
    for (int k = 0; k < nl; k++) {
      long jb = nbr(L, c, hop_dir(nl, k), -hop_len(nl, k));
      amv_add(acc, G + ((long)k * V + jb) * 9, X + 3 * jb, -1.0);
    }
    for (int i = 0; i < 3; i++) T[3 * idx + i] = acc[i];
  }
  par_range(L, par, &lo, &hi);
#pragma omp parallel for schedule(static)
  for (long idx = lo; idx < hi; idx++) {
    int c[4];
    qo_coords(L, idx, c);
    cplx acc[3];
    for (int i = 0; i < 3; i++) acc[i] = (4.0 * m2) * X[3 * idx + i];
    for (int k = 0; k < nl; k++) {
      int mu = hop_dir(nl, k), len = hop_len(nl, k);
      long jf = nbr(L, c, mu, len), jb = nbr(L, c, mu, -len);
      mv_add(acc, G + ((long)k * V + idx) * 9, T + 3 * jf, -1.0);
      amv_add(acc, G + ((long)k * V + jb) * 9, T + 3 * jb, 1.0);
    }
    for (int i = 0; i < 3; i++) R[3 * idx + i] = acc[i];
  }
}

/* ------------------------------------------------------------ reductions */
/* Deterministic: fixed chunks summed in parallel, then the chunk sums in
   order, so results do not depend on the thread count or on the order in
   which threads finish (OpenMP's reduction clause fixes neither). */
#define QO_NCHUNK 64
static double chunked_dot(long m, const double *x, const double *y) {
  double part[QO_NCHUNK];
#pragma omp parallel for schedule(static)
  for (int k = 0; k < QO_NCHUNK; k++) {
    const long lo = m * k / QO_NCHUNK, hi = m * (k + 1) / QO_NCHUNK;
    double s = 0;
    for (long i = lo; i < hi; i++) s += x[i] * y[i];
    part[k] = s;
  }
  double s = 0;
  for (int k = 0; k < QO_NCHUNK; k++) s += part[k];
  return s;
}
double qo_norm2(long n, const double *x) { return chunked_dot(6 * n, x, x); }
double qo_redot(long n, const double *x, const double *y) { return chunked_dot(6 * n, x, y); }

/* ---------------------------------------------------------------- solvers */
/* solvers/cg.nim:55-272 with precon = cpNone (z aliases r, q aliases p),
   starting from a fresh state (reset: b2=-1, r2old=rzold=1).  The operator
   is stagD2xx on parity par (physics/stagSolve.nim:64-72). */
int qo_cg(const int L[4], int nl, const double *g, int par,
          double *x, const double *b, double m, double r2req, int maxits,
          double *r2out) {
  long V = qo_volume(L), lo, hi;
  par_range(L, par, &lo, &hi);
  long n = hi - lo;
  double *r = calloc(6 * V, sizeof(double));
  double *p = calloc(6 * V, sizeof(double));
  double *Ap = calloc(6 * V, sizeof(double));
  double *t = calloc(6 * V, sizeof(double));
  const double m2 = m * m;
  double *xs = x + 6 * lo, *rs = r + 6 * lo, *ps = p + 6 * lo, *Aps = Ap + 6 * lo;
  const double *bs = b + 6 * lo;
  double b2 = qo_norm2(n, bs), r2;
  if (b2 == 0.0) {
    memset(xs, 0, sizeof(double) * 6 * n);
    r2 = 0.0;
  } else {
    qo_stagD2xx(L, nl, g, par, Ap, x, m2, t);
    for (long i = 0; i < 6 * n; i++) rs[i] = bs[i] - Aps[i];
    r2 = qo_norm2(n, rs);
  }
  double r2stop = r2req * b2;
  int itn = 0;
  double rzo = 1.0;
  while (itn < maxits && r2 > r2stop) {
    double rz = r2;
    double beta = rz / rzo;
    rzo = rz;
    if (itn == 0)
      memcpy(ps, rs, sizeof(double) * 6 * n);
    else
      for (long i = 0; i < 6 * n; i++) ps[i] = rs[i] + beta * ps[i];
    itn++;
    qo_stagD2xx(L, nl, g, par, Ap, p, m2, t);
    double pAp = qo_redot(n, ps, Aps);
    double alpha = rz / pAp;
    for (long i = 0; i < 6 * n; i++) xs[i] += alpha * ps[i];
    for (long i = 0; i < 6 * n; i++) rs[i] -= alpha * Aps[i];
    r2 = qo_norm2(n, rs);
  }
  if (r2out) *r2out = r2;
  free(r); free(p); free(Ap); free(t);
  return itn;
}

/* physics/stagSolve.nim:330-421 (solveXX) with sloppySolve = SloppyNone:
   outer defect-correction loop around solveXXInternal -> cg.  Returns the
   total iterations; *true_r2_rel = |x - A r|^2/|x|^2 on parity par. */
int qo_solveXX(const int L[4], int nl, const double *g, int par,
               double *r, const double *x, double m, double r2req, int maxits,
               double *true_r2_rel) {
  long V = qo_volume(L), lo, hi;
  par_range(L, par, &lo, &hi);
  long n = hi - lo;
  double *rho = calloc(6 * V, sizeof(double));
  double *y = calloc(6 * V, sizeof(double));
  double *t = calloc(6 * V, sizeof(double));
  memset(r, 0, sizeof(double) * 6 * V);
  memcpy(rho + 6 * lo, x + 6 * lo, sizeof(double) * 6 * n);
  double b2 = qo_norm2(n, rho + 6 * lo);
  double r2 = b2, r2stop = r2req * b2, r2prev = INFINITY;
  int its = 0;
  while (r2 > r2stop && r2 < r2prev) {
    r2prev = r2;
    int mx = maxits - its;
    if (mx <= 0) break;
    memset(y, 0, sizeof(double) * 6 * V);
    double r2i;
    its += qo_cg(L, nl, g, par, y, rho, m, r2stop / r2, mx, &r2i);
    for (long i = 6 * lo; i < 6 * hi; i++) r[i] += y[i];
    qo_stagD2xx(L, nl, g, par, rho, r, m * m, t);
    for (long i = 6 * lo; i < 6 * hi; i++) rho[i] = x[i] - rho[i];
    r2 = qo_norm2(n, rho + 6 * lo);
  }
  if (true_r2_rel) *true_r2_rel = (b2 > 0) ? r2 / b2 : 0.0;
  free(rho); free(y); free(t);
  return its;
}

/* solvers/cgm.nim:68-246 (fresh solve, b2 < 0 branch), shifted operators
   A + sigma_k with A = stagD2xx(m^2) and sigma_0 forced to 0
   (physics/stagSolve.nim:196-206). */
int qo_cgm(const int L[4], int nl, const double *g, int par,
           double *xs, const double *b, double m, const double *sigmas, int nshift,
           double r2req, int maxits, double *r2out) {
  long V = qo_volume(L), lo, hi;
  par_range(L, par, &lo, &hi);
  long n = hi - lo, o = 6 * lo;
  double *r = calloc(6 * V, sizeof(double));
  double *Ap = calloc(6 * V, sizeof(double));
  double *t = calloc(6 * V, sizeof(double));
  double *ps = calloc(6 * V * (size_t)nshift, sizeof(double));
  double *sg = calloc(nshift, sizeof(double)), *zi = calloc(nshift, sizeof(double)),
         *zim1 = calloc(nshift, sizeof(double));
  for (int k = 0; k < nshift; k++) sg[k] = (k == 0) ? 0.0 : sigmas[k];
  memcpy(r + o, b + o, sizeof(double) * 6 * n);
  for (int k = 0; k < nshift; k++) memset(xs + 6 * V * k, 0, sizeof(double) * 6 * V);
  double b2 = qo_norm2(n, b + o), r2 = b2, r2stop = r2req * b2;
  int itn = 0;
  if (r2 > r2stop) {
    double r2i = r2, r2ip1 = 0, r2r = 0, alpha = 0, beta = 0, alphaim1 = -1.0, betaim1 = 0.0;
    for (int k = 0; k < nshift; k++) {
      zim1[k] = 1.0; zi[k] = 1.0;
      memcpy(ps + 6 * V * k + o, r + o, sizeof(double) * 6 * n);
    }
    int continuing = 1;
    while (continuing) {
      for (int k = 0; k < nshift; k++) {
        double *pk = ps + 6 * V * k, *xk = xs + 6 * V * k;
        if (k == 0) {
          qo_stagD2xx(L, nl, g, par, Ap, pk, m * m, t);
          itn++;
          double pAp = qo_redot(n, pk + o, Ap + o);
          alpha = (pAp != 0.0) ? r2i / pAp : 0.0;
          for (long i = o; i < o + 6 * n; i++) r[i] -= alpha * Ap[i];
          for (long i = o; i < o + 6 * n; i++) xk[i] += alpha * pk[i];
          r2ip1 = qo_norm2(n, r + o);
          r2r = r2ip1;
          beta = (r2i != 0.0) ? r2ip1 / r2i : 0.0;
          continuing = (itn < maxits) && (r2r > r2stop);
          if (continuing)
            for (long i = o; i < o + 6 * n; i++) pk[i] = r[i] + beta * pk[i];
        } else {
          double zip1d = alpha * betaim1 * (zim1[k] - zi[k]);
          zip1d += zim1[k] * alphaim1 * (1.0 + sg[k] * alpha);
          double zip1 = (zip1d != 0.0) ? zi[k] * zim1[k] * alphaim1 / zip1d : 0.0;
          double zr = (zi[k] != 0.0) ? zip1 / zi[k] : 0.0;
          for (long i = o; i < o + 6 * n; i++) xk[i] += alpha * zr * pk[i];
          if (continuing) {
            for (long i = o; i < o + 6 * n; i++) pk[i] = zip1 * r[i] + beta * zr * zr * pk[i];
            zim1[k] = zi[k];
            zi[k] = zip1;
          }
        }
      }
      alphaim1 = alpha; betaim1 = beta; r2i = r2ip1;
      r2 = r2r;
    }
  }
  if (r2out) *r2out = r2;
  free(r); free(Ap); free(t); free(ps); free(sg); free(zi); free(zim1);
  return itn;
}

/* ------------------------------------------------------------------ force */
/* hmc/hmcAction.nim:735-771: f1[mu](n) = psi(n) psi(n+mu)^+, times -1 on odd n */
void qo_ferm_outer(const int L[4], const double *psi, double *G) {
  long V = qo_volume(L), Vh = V / 2;
  const cplx *P = (const cplx *)psi;
  cplx *F = (cplx *)G;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < V; idx++) {
    int c[4];
    qo_coords(L, idx, c);
    double eps = (idx < Vh) ? 1.0 : -1.0;
    for (int mu = 0; mu < 4; mu++) {
      long j = nbr(L, c, mu, 1);
      cplx *f = F + ((long)mu * V + idx) * 9;
      for (int a = 0; a < 3; a++)
        for (int bb = 0; bb < 3; bb++) f[3 * a + bb] = eps * P[3 * idx + a] * conj(P[3 * j + bb]);
    }
  }
}

/* hmc/hmcAction.nim:779-790: f = projectTAH(ff * U^+) with phased U */
void qo_ferm_force(const int L[4], const double *U, const double *psi, double *F) {
  long V = qo_volume(L);
  qo_ferm_outer(L, psi, F);
  const cplx *Uc = (const cplx *)U;
  cplx *Fc = (cplx *)F;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < 4 * V; i++) {
    cplx t[9];
    const cplx *u = Uc + i * 9;
    cplx *f = Fc + i * 9;
    for (int a = 0; a < 3; a++)
      for (int bb = 0; bb < 3; bb++) {
        cplx s = 0;
        for (int k = 0; k < 3; k++) s += f[3 * a + k] * conj(u[3 * bb + k]);
        t[3 * a + bb] = s;
      }
    qo_project_tah((double *)t, (double *)f);
  }
}

/* hmc/hmcAction.nim:669-721 (stagForceSolve) */
int qo_force_solve(const int L[4], int nl, const double *g, const double *phi,
                   double m, double r2req, int maxits, double *psi) {
  long V = qo_volume(L), Vh = V / 2;
  double *vphi = calloc(6 * V, sizeof(double));
  double tr2;
  int its = qo_solveXX(L, nl, g, 0, vphi, phi, m, r2req, maxits, &tr2);
  for (long i = 0; i < 6 * Vh; i++) vphi[i] *= 4.0;
  memset(psi, 0, sizeof(double) * 6 * V);
  qo_stagD2(L, nl, g, 1, psi, vphi, 0.0, 0.0);
  memcpy(psi, vphi, sizeof(double) * 6 * Vh);
  free(vphi);
  return its;
}

/* ------------------------------------------------------ gauge field (HMC) */
static inline const cplx *lnk(const double *U, long V, int mu, long idx) {
  return (const cplx *)(U + ((long)mu * V + idx) * 18);
}
static void m3mul(const cplx *a, const cplx *b, cplx *c) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cplx t = 0;
      for (int k = 0; k < 3; k++) t += a[3 * i + k] * b[3 * k + j];
      c[3 * i + j] = t;
    }
}
static void m3dag(const cplx *a, cplx *r) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) r[3 * i + j] = conj(a[3 * j + i]);
}

/* gauge/gaugeAction.nim:397-427 (plaquette term) */
double qo_plaquette(const int L[4], const double *U) {
  long V = qo_volume(L);
  double part[QO_NCHUNK];  /* deterministic chunked sum, as chunked_dot */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < QO_NCHUNK; k++) {
    double s = 0;
    for (long idx = V * k / QO_NCHUNK; idx < V * (k + 1) / QO_NCHUNK; idx++) {
      int c[4];
      qo_coords(L, idx, c);
      for (int mu = 0; mu < 4; mu++)
        for (int nu = mu + 1; nu < 4; nu++) {
          cplx a[9], b[9];
          m3mul(lnk(U, V, mu, idx), lnk(U, V, nu, nbr(L, c, mu, 1)), a);
          m3mul(lnk(U, V, nu, idx), lnk(U, V, mu, nbr(L, c, nu, 1)), b);
          for (int e = 0; e < 9; e++) s += creal(a[e] * conj(b[e]));
        }
    }
    part[k] = s;
  }
  double s = 0;
  for (int k = 0; k < QO_NCHUNK; k++) s += part[k];
  return s;
}

/* derivative of -(beta/3) sum Re tr U_P (gauge/gaugeAction.nim:765-781) */
void qo_wilson_force(const int L[4], const double *U, double beta, double *F) {
  long V = qo_volume(L);
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < V; idx++) {
    int c[4];
    qo_coords(L, idx, c);
    for (int mu = 0; mu < 4; mu++) {
      cplx A[9] = {0};
      for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        int cm[4] = {c[0], c[1], c[2], c[3]};
        cm[nu] = (cm[nu] - 1 + L[nu]) % L[nu];
        long xp = nbr(L, c, mu, 1), xpn = nbr(L, c, nu, 1), xmn = nbr(L, c, nu, -1), xpmn = nbr(L, cm, mu, 1);
        cplx t1[9], t2[9], d[9], up[9], lo[9];
        m3dag(lnk(U, V, mu, xpn), d);
        m3mul(lnk(U, V, nu, xp), d, t1);
        m3dag(lnk(U, V, nu, idx), d);
        m3mul(t1, d, up);
        m3dag(lnk(U, V, nu, xpmn), t1);
        m3dag(lnk(U, V, mu, xmn), d);
        m3mul(t1, d, t2);
        m3mul(t2, lnk(U, V, nu, xmn), lo);
        for (int e = 0; e < 9; e++) A[e] += up[e] + lo[e];
      }
      cplx T[9], R[9];
      m3mul(lnk(U, V, mu, idx), A, T);
      qo_project_tah((double *)T, (double *)R);
      cplx *f = (cplx *)(F + ((long)mu * V + idx) * 18);
      for (int e = 0; e < 9; e++) f[e] = (2.0 * beta / 3.0) * R[e];
    }
  }
}

void qo_random_momenta(const int L[4], uint64_t seed, double *P) {
  long V = qo_volume(L);
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < V; idx++) {
    int c[4];
    qo_coords(L, idx, c);
    uint32_t lex = (uint32_t)(c[0] + (long)L[0] * (c[1] + (long)L[1] * (c[2] + (long)L[2] * c[3])));
    for (int mu = 0; mu < 4; mu++) {
      double r[10];
      for (int j = 0; j < 5; j++) qo_gauss2(seed, lex, 32u + (uint32_t)mu, (uint32_t)j, TAG_MOM, r + 2 * j);
      const double s = sqrt(0.5), m = (r[0] + r[1] + r[2]) / 3.0;
      cplx H[9];
      H[0] = r[0] - m; H[4] = r[1] - m; H[8] = r[2] - m;
      H[1] = s * r[3] + I * s * r[4]; H[3] = conj(H[1]);
      H[2] = s * r[5] + I * s * r[6]; H[6] = conj(H[2]);
      H[5] = s * r[7] + I * s * r[8]; H[7] = conj(H[5]);
      cplx *p = (cplx *)(P + ((long)mu * V + idx) * 18);
      for (int e = 0; e < 9; e++) p[e] = I * H[e];
    }
  }
}

void qo_link_update(const int L[4], const double *P, double eps, double *U) {
  long V = qo_volume(L);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < 4 * V; i++) {
    cplx A[9], E[9], T[9];
    const cplx *p = (const cplx *)(P + i * 18);
    for (int e = 0; e < 9; e++) A[e] = eps * p[e];
    qo_su3_exp((double *)A, (double *)E);
    m3mul(E, (const cplx *)(U + i * 18), T);
    memcpy(U + i * 18, T, sizeof T);
  }
}

/* --------------------------------------------------- HISQ link construction */

/* Generalized staple of a matrix field K in the (mu,nu) plane,
   gauge/fat7l.nim:24-67 (computeGenStaple):
     S_nu[K](x) = U_nu(x) K(x+nu) U_nu(x+mu)^+ + U_nu(x-nu)^+ K(x-nu) U_nu(x-nu+mu) */
static void gen_staple(const int L[4], const double *U, int mu, int nu, const cplx *K, cplx *out) {
  long V = qo_volume(L);
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < V; idx++) {
    int c[4];
    qo_coords(L, idx, c);
    int cm[4] = {c[0], c[1], c[2], c[3]};
    cm[nu] = (cm[nu] - 1 + L[nu]) % L[nu];
    long xpn = nbr(L, c, nu, 1), xpm = nbr(L, c, mu, 1), xmn = nbr(L, c, nu, -1), xmnpm = nbr(L, cm, mu, 1);
    cplx t[9], d[9], a[9], b[9];
    m3mul(lnk(U, V, nu, idx), K + xpn * 9, t);
    m3dag(lnk(U, V, nu, xpm), d);
    m3mul(t, d, a);
    m3dag(lnk(U, V, nu, xmn), d);
    m3mul(d, K + xmn * 9, t);
    m3mul(t, lnk(U, V, nu, xmnpm), b);
    for (int e = 0; e < 9; e++) out[idx * 9 + e] = a[e] + b[e];
  }
}

static void axpy_field(long n, double a, const cplx *x, cplx *y) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) y[i] += a * x[i];
}

/* makeImpLinks, gauge/fat7l.nim:71-159.  coef = {oneLink, threeStaple,
   fiveStaple, sevenStaple, lepage}.  The Lepage term is the double staple in
   one plane, whose two retracing paths per direction are cancelled by moving
   6*lepage out of the one-link term (fat7l.nim:96-97); 5- and 7-link staples
   nest the 3-link staple in the remaining directions (fat7l.nim:130-141).
   Long links: ll = naik * gl_mu(x) gl_mu(x+mu) gl_mu(x+2mu) (fat7l.nim:145-154);
   skipped when ll is NULL or naik == 0; gl = gf when NULL. */
void qo_make_imp_links(const int L[4], const double *gf, const double coef[5], const double *gl, double naik,
                       double *fl, double *ll) {
  long V = qo_volume(L);
  const double c3 = coef[1], c5 = coef[2], c7 = coef[3], cL = coef[4];
  const double c1 = coef[0] - 6.0 * cL;
  const int have5 = (c5 != 0.0) || (c7 != 0.0) || (cL != 0.0);
  const int have3 = (c3 != 0.0) || have5;
  cplx *st = malloc(sizeof(cplx) * 9 * V), *t5 = malloc(sizeof(cplx) * 9 * V), *tmp = malloc(sizeof(cplx) * 9 * V);
  for (int dir = 0; dir < 4; dir++) {
    cplx *f = (cplx *)(fl + (long)dir * V * 18);
    const cplx *u = (const cplx *)(gf + (long)dir * V * 18);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < 9 * V; i++) f[i] = c1 * u[i];
    if (!have3) continue;
    for (int nu = 0; nu < 4; nu++) {
      if (nu == dir) continue;
      gen_staple(L, gf, dir, nu, u, st);
      axpy_field(9 * V, c3, st, f);
      if (cL != 0.0) {
        gen_staple(L, gf, dir, nu, st, tmp);
        axpy_field(9 * V, cL, tmp, f);
      }
      if (c5 != 0.0 || c7 != 0.0)
        for (int rho = 0; rho < 4; rho++) {
          if (rho == dir || rho == nu) continue;
          gen_staple(L, gf, dir, rho, st, t5);
          axpy_field(9 * V, c5, t5, f);
          if (c7 != 0.0)
            for (int sig = 0; sig < 4; sig++) {
              if (sig == dir || sig == nu || sig == rho) continue;
              gen_staple(L, gf, dir, sig, t5, tmp);
              axpy_field(9 * V, c7, tmp, f);
            }
        }
    }
  }
  free(st);
  free(t5);
  free(tmp);
  if (ll && naik != 0.0) {
    if (!gl) gl = gf;
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < V; idx++) {
      int c[4];
      qo_coords(L, idx, c);
      for (int mu = 0; mu < 4; mu++) {
        cplx t[9], t2[9];
        m3mul(lnk(gl, V, mu, idx), lnk(gl, V, mu, nbr(L, c, mu, 1)), t);
        m3mul(t, lnk(gl, V, mu, nbr(L, c, mu, 2)), t2);
        cplx *o = (cplx *)(ll + ((long)mu * V + idx) * 18);
        for (int e = 0; e < 9; e++) o[e] = naik * t2[e];
      }
    }
  }
}

/* Eigenvalues of a Hermitian 3x3 matrix by the trigonometric solution of its
   characteristic cubic (O. K. Smith, CACM 4 (1961) 168): with q = tr/3,
   B = H - q, p = sqrt(tr B^2 / 6), r = det(B/p)/2, the eigenvalues are
   q + 2p cos(acos(r)/3 + 2 pi k/3). */
static void herm3_eigs(const cplx *h, double ev[3]) {
  const double q = (creal(h[0]) + creal(h[4]) + creal(h[8])) / 3.0;
  const double b0 = creal(h[0]) - q, b1 = creal(h[4]) - q, b2 = creal(h[8]) - q;
  const double o01 = creal(h[1] * conj(h[1])), o02 = creal(h[2] * conj(h[2])), o12 = creal(h[5] * conj(h[5]));
  const double p2 = (b0 * b0 + b1 * b1 + b2 * b2 + 2.0 * (o01 + o02 + o12)) / 6.0;
  if (p2 <= 0.0) {
    ev[0] = ev[1] = ev[2] = q;
    return;
  }
  const double p = sqrt(p2);
  /* det B for Hermitian B: b0 b1 b2 + 2 Re(h01 h12 h20) - b0|h12|^2 - b1|h02|^2 - b2|h01|^2 */
  const double detb = b0 * b1 * b2 + 2.0 * creal(h[1] * h[5] * h[6]) - b0 * o12 - b1 * o02 - b2 * o01;
  double r = detb / (2.0 * p2 * p);
  r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
  const double phi = acos(r) / 3.0;
  ev[0] = q + 2.0 * p * cos(phi);
  ev[2] = q + 2.0 * p * cos(phi + 2.0943951023931954923);
  ev[1] = 3.0 * q - ev[0] - ev[2];
}

/* projectU by Cayley-Hamilton, maths/matrixFunctions.nim:337-352 with
   rsqrtPHM3 (:165-219): u = x (x^+ x + eps)^{-1/2}, eps = 1e-20.  The inverse
   square root of the Hermitian H = x^+ x + eps is the quadratic
   f0 + f1 H + f2 H^2 that equals 1/s_i at the eigenvalues s_i^2; with
   u1 = s0+s1+s2, v = s0 s1 + s1 s2 + s2 s0, w = s0 s1 s2 and d = w (u1 v - w):
   f0 = (u1 v^2 - v w - u1^2 w)/d, f1 = -(u1 tr H + w)/d, f2 = u1/d. */
void qo_project_u(long n, const double *x, double *u) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) {
    const cplx *X = (const cplx *)(x + i * 18);
    cplx Xd[9], H[9], H2[9], R[9];
    m3dag(X, Xd);
    m3mul(Xd, X, H);
    for (int k = 0; k < 3; k++) H[4 * k] += 1e-20;
    double ev[3];
    herm3_eigs(H, ev);
    const double s0 = sqrt(fabs(ev[0])), s1 = sqrt(fabs(ev[1])), s2 = sqrt(fabs(ev[2]));
    const double u1 = s0 + s1 + s2, v = s0 * s1 + s1 * s2 + s2 * s0, w = s0 * s1 * s2;
    const double tr = creal(H[0]) + creal(H[4]) + creal(H[8]);
    const double di = 1.0 / (w * (u1 * v - w));
    const double f0 = (u1 * v * v - v * w - u1 * u1 * w) * di, f1 = -(u1 * tr + w) * di, f2 = u1 * di;
    m3mul(H, H, H2);
    for (int e = 0; e < 9; e++) R[e] = f1 * H[e] + f2 * H2[e];
    for (int k = 0; k < 3; k++) R[4 * k] += f0;
    m3mul(X, R, (cplx *)(u + i * 18));
  }
}

/* HisqCoefs.smear, physics/hisqLinks.nim:39-51: first fat7 level (no Naik),
   unitary projection, second fat7 level with Naik long links from the
   projected links.  first/second as coef of qo_make_imp_links. */
void qo_hisq_smear(const int L[4], const double first[5], const double second[5], double naik, const double *g,
                   double *fl, double *ll) {
  long V = qo_volume(L);
  double *t1 = malloc(sizeof(double) * 72 * V), *t2 = malloc(sizeof(double) * 72 * V);
  qo_make_imp_links(L, g, first, NULL, 0.0, t1, NULL);
  qo_project_u(4 * V, t1, t2);
  qo_make_imp_links(L, t2, second, t2, naik, fl, ll);
  free(t1);
  free(t2);
}

/* ------------------------------------------- HISQ smearing chain rule (force) */
/* Derivatives are with respect to the complex link entries in QEX's pairing
   redot(A, B) = Re tr(A^+ B) (gauge/fat7lderiv.nim, maths/matrixFunctions.nim:
   361-392): given mid = dS/dF for F = f(U), deriv += dS/dU, i.e.
   sum redot(deriv, dU) = sum redot(mid, dF) to first order. */

/* Derivative of the generalized staple S_lam[K] (plane mu,lam) with respect to
   its side links U_lam, chain C, accumulated with weight a into D (direction
   lam): the four link occurrences of the upper and lower staples,
     C(y) U(y+mu) K(y+lam)^+ + K(y) U(y+mu) C(y+lam)^+
   + C(y-mu)^+ U(y-mu) K(y-mu+lam) + K(y-mu)^+ U(y-mu) C(y-mu+lam),  U = U_lam.
   The derivative with respect to K is S_lam[C] itself (the staple map is its
   own adjoint in this pairing). */
static void side_deriv(const int L[4], const double *U, int mu, int lam, const cplx *C, const cplx *K, double a,
                       cplx *D) {
  if (a == 0.0) return;
  long V = qo_volume(L);
#pragma omp parallel for schedule(static)
  for (long y = 0; y < V; y++) {
    int c[4];
    qo_coords(L, y, c);
    int cm[4] = {c[0], c[1], c[2], c[3]};
    cm[mu] = (cm[mu] - 1 + L[mu]) % L[mu];
    long ypm = nbr(L, c, mu, 1), ypl = nbr(L, c, lam, 1), ymm = qo_index(L, cm), ymmpl = nbr(L, cm, lam, 1);
    cplx t[9], d[9], r[9], acc[9] = {0};
    /* C(y) U(y+mu) K(y+lam)^+ */
    m3mul(C + y * 9, lnk(U, V, lam, ypm), t);
    m3dag(K + ypl * 9, d);
    m3mul(t, d, r);
    for (int e = 0; e < 9; e++) acc[e] += r[e];
    /* K(y) U(y+mu) C(y+lam)^+ */
    m3mul(K + y * 9, lnk(U, V, lam, ypm), t);
    m3dag(C + ypl * 9, d);
    m3mul(t, d, r);
    for (int e = 0; e < 9; e++) acc[e] += r[e];
    /* C(y-mu)^+ U(y-mu) K(y-mu+lam) */
    m3dag(C + ymm * 9, d);
    m3mul(d, lnk(U, V, lam, ymm), t);
    m3mul(t, K + ymmpl * 9, r);
    for (int e = 0; e < 9; e++) acc[e] += r[e];
    /* K(y-mu)^+ U(y-mu) C(y-mu+lam) */
    m3dag(K + ymm * 9, d);
    m3mul(d, lnk(U, V, lam, ymm), t);
    m3mul(t, C + ymmpl * 9, r);
    for (int e = 0; e < 9; e++) D[y * 9 + e] += a * (acc[e] + r[e]);
  }
}

static int third_dir(int a, int b, int c) {
  for (int d = 0; d < 4; d++)
    if (d != a && d != b && d != c) return d;
  return -1;
}

/* fat7lDeriv, gauge/fat7lderiv.nim:122-475: the chain rule of makeImpLinks.
   Per direction mu the forward pass is rebuilt (S_nu = S_nu[U_mu], T_sig =
   S_rho[S_nu] + S_nu[S_rho] for {nu,rho} = rest of {mu,sig}) and run
   backwards: chains CT_sig = c5 C + c7 S_sig[C] and CS_nu = c3 C + cL S_nu[C]
   + sum_{sig != nu} S_rho[CT_sig]; deriv_mu gets (c1 - 6 cL) C + sum_nu
   S_nu[CS_nu]; each side link U_lam gets the side derivatives of every staple
   in the (mu,lam) plane.  Naik (fat7lderiv.nim:436-475): for
   ll = naik gl(x) gl(x+mu) gl(x+2mu), llderiv_mu(y) += naik [C(y) (gl(y+mu)
   gl(y+2mu))^+ + gl(y-mu)^+ C(y-mu) gl(y+mu)^+ + (gl(y-2mu) gl(y-mu))^+ C(y-2mu)]. */
void qo_fat7l_deriv(const int L[4], const double *gf, const double *mid, const double coef[5], double *deriv,
                    const double *gl, const double *llmid, double naik, double *llderiv) {
  long V = qo_volume(L);
  const double c3 = coef[1], c5 = coef[2], c7 = coef[3], cL = coef[4];
  const double c1 = coef[0] - 6.0 * cL;
  cplx *S[4], *T[4], *CS[4], *CT[4], *tmp = malloc(sizeof(cplx) * 9 * V);
  for (int k = 0; k < 4; k++) {
    S[k] = calloc(9 * V, sizeof(cplx));
    T[k] = calloc(9 * V, sizeof(cplx));
    CS[k] = calloc(9 * V, sizeof(cplx));
    CT[k] = calloc(9 * V, sizeof(cplx));
  }
  for (int mu = 0; mu < 4; mu++) {
    const cplx *u = (const cplx *)(gf + (long)mu * V * 18);
    const cplx *C = (const cplx *)(mid + (long)mu * V * 18);
    cplx *Dm = (cplx *)(deriv + (long)mu * V * 18);
    axpy_field(9 * V, c1, C, Dm);
    for (int nu = 0; nu < 4; nu++)
      if (nu != mu) gen_staple(L, gf, mu, nu, u, S[nu]);
    for (int sig = 0; sig < 4; sig++) {
      if (sig == mu) continue;
      int nu = -1, rho = -1;
      for (int d = 0; d < 4; d++)
        if (d != mu && d != sig) {
          if (nu < 0) nu = d; else rho = d;
        }
      gen_staple(L, gf, mu, rho, S[nu], T[sig]);
      gen_staple(L, gf, mu, nu, S[rho], tmp);
      axpy_field(9 * V, 1.0, tmp, T[sig]);
      /* CT_sig = c5 C + c7 S_sig[C] */
      gen_staple(L, gf, mu, sig, C, CT[sig]);
      for (long i = 0; i < 9 * V; i++) CT[sig][i] = c7 * CT[sig][i] + c5 * C[i];
    }
    for (int nu = 0; nu < 4; nu++) {
      if (nu == mu) continue;
      gen_staple(L, gf, mu, nu, C, CS[nu]);
      for (long i = 0; i < 9 * V; i++) CS[nu][i] = cL * CS[nu][i] + c3 * C[i];
      for (int sig = 0; sig < 4; sig++) {
        if (sig == mu || sig == nu) continue;
        gen_staple(L, gf, mu, third_dir(mu, nu, sig), CT[sig], tmp);
        axpy_field(9 * V, 1.0, tmp, CS[nu]);
      }
    }
    for (int nu = 0; nu < 4; nu++) {
      if (nu == mu) continue;
      gen_staple(L, gf, mu, nu, CS[nu], tmp);
      axpy_field(9 * V, 1.0, tmp, Dm);
    }
    for (int lam = 0; lam < 4; lam++) {
      if (lam == mu) continue;
      cplx *Dl = (cplx *)(deriv + (long)lam * V * 18);
      side_deriv(L, gf, mu, lam, CS[lam], u, 1.0, Dl);     /* S_lam[U_mu] */
      side_deriv(L, gf, mu, lam, C, S[lam], cL, Dl);       /* S_lam[S_lam] (Lepage) */
      side_deriv(L, gf, mu, lam, C, T[lam], c7, Dl);       /* S_lam[T_lam] (7-link) */
      for (int sig = 0; sig < 4; sig++) {                   /* S_lam[S_other] inside T_sig */
        if (sig == mu || sig == lam) continue;
        side_deriv(L, gf, mu, lam, CT[sig], S[third_dir(mu, lam, sig)], 1.0, Dl);
      }
    }
  }
  for (int k = 0; k < 4; k++) {
    free(S[k]); free(T[k]); free(CS[k]); free(CT[k]);
  }
  free(tmp);
  if (llderiv && llmid && naik != 0.0) {
    if (!gl) gl = gf;
#pragma omp parallel for schedule(static)
    for (long y = 0; y < V; y++) {
      int c[4];
      qo_coords(L, y, c);
      for (int mu = 0; mu < 4; mu++) {
        const cplx *C = (const cplx *)(llmid + (long)mu * V * 18);
        long p1 = nbr(L, c, mu, 1), p2 = nbr(L, c, mu, 2), m1 = nbr(L, c, mu, -1), m2 = nbr(L, c, mu, -2);
        cplx a[9], b[9], t[9], r[9], acc[9];
        /* C(y) (gl(y+mu) gl(y+2mu))^+ */
        m3mul(lnk(gl, V, mu, p1), lnk(gl, V, mu, p2), a);
        m3dag(a, b);
        m3mul(C + y * 9, b, acc);
        /* gl(y-mu)^+ C(y-mu) gl(y+mu)^+ */
        m3dag(lnk(gl, V, mu, m1), a);
        m3mul(a, C + m1 * 9, t);
        m3dag(lnk(gl, V, mu, p1), b);
        m3mul(t, b, r);
        for (int e = 0; e < 9; e++) acc[e] += r[e];
        /* (gl(y-2mu) gl(y-mu))^+ C(y-2mu) */
        m3mul(lnk(gl, V, mu, m2), lnk(gl, V, mu, m1), a);
        m3dag(a, b);
        m3mul(b, C + m2 * 9, r);
        cplx *o = (cplx *)(llderiv + ((long)mu * V + y) * 18);
        for (int e = 0; e < 9; e++) o[e] += naik * (acc[e] + r[e]);
      }
    }
  }
}

static void m3inv(const cplx *a, cplx *r) {
  cplx c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
  cplx det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  cplx id = 1.0 / det;
  r[0] = c00 * id; r[3] = c01 * id; r[6] = c02 * id;
  r[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  r[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  r[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  r[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  r[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  r[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

/* Solves A X + X A = C for Hermitian positive A (maths/projUderiv.nim:96,
   sylsolve) in closed form: with p(l) = l^3 + a2 l^2 + a1 l + a0 the
   characteristic polynomial of A, Cayley-Hamilton (p(A) = 0) applied to
   A^k X = X (-A)^k + sum_j A^j C (-A)^{k-1-j} gives
   X = -[a1 C + a2 (AC - CA) + C A^2 - A C A + A^2 C] p(-A)^{-1}. */
static void sylsolve3(const cplx *A, const cplx *C, cplx *X) {
  cplx A2[9], AC[9], CA[9], CA2[9], ACA[9], A2C[9], A3[9], N[9], P[9], Pi[9];
  m3mul(A, A, A2);
  m3mul(A, C, AC);
  m3mul(C, A, CA);
  m3mul(C, A2, CA2);
  m3mul(AC, A, ACA);
  m3mul(A2, C, A2C);
  m3mul(A2, A, A3);
  const cplx tr = A[0] + A[4] + A[8], tr2 = A2[0] + A2[4] + A2[8];
  const cplx det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                   A[2] * (A[3] * A[7] - A[4] * A[6]);
  const cplx a2 = -tr, a1 = 0.5 * (tr * tr - tr2), a0 = -det;
  for (int e = 0; e < 9; e++) {
    N[e] = a1 * C[e] + a2 * (AC[e] - CA[e]) + CA2[e] - ACA[e] + A2C[e];
    P[e] = -A3[e] + a2 * A2[e] - a1 * A[e];
  }
  for (int k = 0; k < 3; k++) P[4 * k] += a0;
  m3inv(P, Pi);
  m3mul(N, Pi, X);
  for (int e = 0; e < 9; e++) X[e] = -X[e];
}

/* projectUderiv, maths/matrixFunctions.nim:361-392: for u = x z,
   z = (x^+ x)^{-1/2}, y = z^{-1}, chain c = dS/du:
   dS/dx = c z - x (t + t^+) with y t + t y = u^+ c z. */
void qo_project_u_deriv(long n, const double *u, const double *x, const double *chain, double *r) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) {
    const cplx *X = (const cplx *)(x + i * 18), *Uu = (const cplx *)(u + i * 18), *C = (const cplx *)(chain + i * 18);
    cplx Xd[9], H[9], H2[9], Z[9], Y[9], R[9], Ud[9], t1[9], t2[9], s[9], xs[9];
    m3dag(X, Xd);
    m3mul(Xd, X, H);
    for (int k = 0; k < 3; k++) H[4 * k] += 1e-20;
    double ev[3];
    herm3_eigs(H, ev);
    const double s0 = sqrt(fabs(ev[0])), s1 = sqrt(fabs(ev[1])), s2 = sqrt(fabs(ev[2]));
    const double u1 = s0 + s1 + s2, v = s0 * s1 + s1 * s2 + s2 * s0, w = s0 * s1 * s2;
    const double tr = creal(H[0]) + creal(H[4]) + creal(H[8]);
    const double di = 1.0 / (w * (u1 * v - w));
    const double f0 = (u1 * v * v - v * w - u1 * u1 * w) * di, f1 = -(u1 * tr + w) * di, f2 = u1 * di;
    m3mul(H, H, H2);
    for (int e = 0; e < 9; e++) Z[e] = f1 * H[e] + f2 * H2[e];
    for (int k = 0; k < 3; k++) Z[4 * k] += f0;
    m3inv(Z, Y);
    m3mul(C, Z, R);
    m3dag(Uu, Ud);
    m3mul(Ud, R, t1);
    sylsolve3(Y, t1, t2);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) s[3 * a + b] = t2[3 * a + b] + conj(t2[3 * b + a]);
    m3mul(X, s, xs);
    cplx *o = (cplx *)(r + i * 18);
    for (int e = 0; e < 9; e++) o[e] = R[e] - xs[e];
  }
}

/* HisqCoefs.smearGetForce, gauge/hisqsmear.nim:54-91: dS/dg for fl, ll =
   smear(g), given dfl = dS/dfl and dll = dS/dll: second-level fat7 and Naik
   chain to the projected links w, projectUderiv to the first-level links v,
   first-level chain to g.  dg is accumulated. */
void qo_hisq_smear_deriv(const int L[4], const double first[5], const double second[5], double naik, const double *g,
                         const double *dfl, const double *dll, double *dg) {
  long V = qo_volume(L);
  double *v = malloc(sizeof(double) * 72 * V), *w = malloc(sizeof(double) * 72 * V);
  double *t = calloc(72 * V, sizeof(double)), *t2 = malloc(sizeof(double) * 72 * V);
  qo_make_imp_links(L, g, first, NULL, 0.0, v, NULL);
  qo_project_u(4 * V, v, w);
  qo_fat7l_deriv(L, w, dfl, second, t, w, dll, naik, t);
  qo_project_u_deriv(4 * V, w, v, t, t2);
  qo_fat7l_deriv(L, g, t2, first, dg, NULL, NULL, 0.0, NULL);
  free(v); free(w); free(t); free(t2);
}

/* Outer products of the HISQ fermion force (hmc/hmcAction.nim:735-771 with
   the 3-hop term of physics/stagD.nim:640-715): G[2mu](x) = eps(x) psi(x)
   psi(x+mu)^+, G[2mu+1](x) = eps(x) psi(x) psi(x+3mu)^+, eps = -1 on odd x. */
void qo_hisq_outer(const int L[4], const double *psi, double *G) {
  long V = qo_volume(L), Vh = V / 2;
  const cplx *P = (const cplx *)psi;
  cplx *F = (cplx *)G;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < V; idx++) {
    int c[4];
    qo_coords(L, idx, c);
    double eps = (idx < Vh) ? 1.0 : -1.0;
    for (int k = 0; k < 8; k++) {
      long j = nbr(L, c, k >> 1, (k & 1) ? 3 : 1);
      cplx *f = F + ((long)k * V + idx) * 9;
      for (int a = 0; a < 3; a++)
        for (int bb = 0; bb < 3; bb++) f[3 * a + bb] = eps * P[3 * idx + a] * conj(P[3 * j + bb]);
    }
  }
}

/* ----------------------------------------------------- MILC gauge-file checksum */
/* src/new/milcIO.nim:43-51: sum29 ^= rotl(w_k, k mod 29), sum31 ^= rotl(w_k, k mod 31)
   over the 32-bit data words w_k, k counted from offset.  The reference writes
   the rotation as (t shl k) + (t shr (32-k)); at k = 0 that shifts by 32, whose
   result depends on the platform, so this follows MILC's file format, which
   defines the rotation (a plain t at k = 0). */
void qo_milc_checksum(const uint32_t *w, long n, long offset, uint32_t out[2]) {
  uint32_t a = 0, b = 0;
  for (long i = 0; i < n; i++) {
    const long k = offset + i;
    const int k29 = (int)(k % 29), k31 = (int)(k % 31);
    const uint32_t t = w[i];
    a ^= k29 ? (t << k29) | (t >> (32 - k29)) : t;
    b ^= k31 ? (t << k31) | (t >> (32 - k31)) : t;
  }
  out[0] = a;
  out[1] = b;
}

/* ------------------------------------------------------ SciDAC record checksum */
/* CRC-32 as src/io/crc32.nim:1-40 defines it: table of the reflected polynomial
   0xedb88320, register started at ~0, one table step per byte, complemented at
   the end.  The reference's own check (crc32.nim:113-116) is the known answer
   0x414FA339 for "The quick brown fox jumps over the lazy dog". */
static uint32_t crc_tbl[256];
static int crc_tbl_ready;

uint32_t qo_crc32(const unsigned char *p, long n) {
  uint32_t *tbl = crc_tbl;
  if (!crc_tbl_ready) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t r = i;
      for (int j = 0; j < 8; j++) r = (r & 1) ? (r >> 1) ^ 0xedb88320u : r >> 1;
      tbl[i] = r;
    }
    crc_tbl_ready = 1;
  }
  uint32_t c = 0xffffffffu;
  for (long i = 0; i < n; i++) c = (c >> 8) ^ tbl[(c ^ p[i]) & 0xff];
  return ~c;
}

/* The SciDAC checksum of a binary record, as the qiolite package QEX reads and
   writes its files with (qex.nimble:35 pins github.com/usqcd-software/qiolite
   at #head; it follows USQCD QIO's DML checksum): every site's bytes, as they
   stand in the file (big-endian), get their own CRC-32 c; with r the site's
   global lexicographic rank (x fastest), suma ^= rotl(c, r mod 29) and
   sumb ^= rotl(c, r mod 31). */
void qo_scidac_checksum(const unsigned char *data, long nsites, long site_bytes, long rank0, uint32_t out[2]) {
  uint32_t a = 0, b = 0;
  for (long k = 0; k < nsites; k++) {
    const uint32_t c = qo_crc32(data + k * site_bytes, site_bytes);
    const long r = rank0 + k;
    const int r29 = (int)(r % 29), r31 = (int)(r % 31);
    a ^= r29 ? (c << r29) | (c >> (32 - r29)) : c;
    b ^= r31 ? (c << r31) | (c >> (32 - r31)) : c;
  }
  out[0] = a;
  out[1] = b;
}
