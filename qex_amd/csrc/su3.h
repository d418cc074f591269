// Shared by the input generators and the gauge-field kernels: the Philox4x32-10
// counter-based generator with Box-Muller normals, small FP64 3x3 complex
// matrix algebra, and canonical-order site helpers.
#pragma once
#include "common.h"

namespace qh {

// Philox stream tags (the fourth counter word)
enum { TAG_GAUGE = 0, TAG_SOURCE = 1, TAG_FATG = 2, TAG_MOM = 3 };

__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int rnd = 0; rnd < 10; rnd++) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
}

__device__ __forceinline__ void gauss2(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       double &g0, double &g1) {
  uint32_t c[4] = {c0, c1, c2, c3};
  philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t a = ((((uint64_t)c[0]) << 32) | c[1]) >> 11;
  const uint64_t b = ((((uint64_t)c[2]) << 32) | c[3]) >> 11;
  const double u1 = ((double)a + 0.5) * 0x1.0p-53;
  const double u2 = ((double)b + 0.5) * 0x1.0p-53;
  const double rad = sqrt(-2.0 * log(u1));
  const double th = 6.283185307179586476925286766559 * u2;
  g0 = rad * cos(th);
  g1 = rad * sin(th);
}

struct M3 { double2 e[9]; };

__device__ __forceinline__ M3 mmul(const M3 &a, const M3 &b) {
  M3 c;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double2 s = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double2 x = a.e[3 * i + k], y = b.e[3 * k + j];
        s.x += x.x * y.x - x.y * y.y;
        s.y += x.x * y.y + x.y * y.x;
      }
      c.e[3 * i + j] = s;
    }
  return c;
}

// exp(A), general complex 3x3: scaling and squaring, degree-16 Taylor (Horner)
__device__ M3 mexp(const M3 &A) {
  double nrm = 0;
  for (int i = 0; i < 3; i++) {
    double s = 0;
    for (int j = 0; j < 3; j++) s += hypot(A.e[3 * i + j].x, A.e[3 * i + j].y);
    nrm = fmax(nrm, s);
  }
  int sq = 0;
  while (nrm > 0.5) { nrm *= 0.5; sq++; }
  const double sc = ldexp(1.0, -sq);
  M3 B, T;
  for (int i = 0; i < 9; i++) {
    B.e[i] = make_double2(A.e[i].x * sc, A.e[i].y * sc);
    T.e[i] = make_double2((i % 4 == 0) ? 1.0 : 0.0, 0.0);
  }
  for (int n = 16; n >= 1; n--) {
    M3 W = mmul(B, T);
    const double inv = 1.0 / (double)n;
    for (int i = 0; i < 9; i++)
      T.e[i] = make_double2(W.e[i].x * inv + ((i % 4 == 0) ? 1.0 : 0.0), W.e[i].y * inv);
  }
  for (int s = 0; s < sq; s++) T = mmul(T, T);
  return T;
}

__device__ __forceinline__ void coords_of(const Geom &g, long idx, int c[4], uint32_t &lex) {
  const int par = idx >= g.Vh;
  const Site s = decode(g, idx - par * g.Vh, par);
  c[0] = s.x; c[1] = s.y; c[2] = s.z; c[3] = s.t;
  lex = (uint32_t)(s.x + (long)g.L[0] * s.row);
}

__device__ __forceinline__ long canon_index(const Geom &g, const int c[4]) {
  const long V = 2 * g.Vh;
  const long lex = c[0] + (long)g.L[0] * (c[1] + (long)g.L[1] * (c[2] + (long)g.L[2] * c[3]));
  const int p = (c[0] + c[1] + c[2] + c[3]) & 1;
  return (lex + p * V) / 2;
}

}  // namespace qh
