// Shared device-side definitions for the qexhip library (gfx950 / CDNA4).
//
// Internal field layout in HBM (one parity = one half-lattice "checkerboard"):
//   color vector field, parity p:   [3 colors][stride] complex   (SoA)
//   link field k, parity p:         [ncomp][stride] complex      (SoA)
// where the site index inside a parity is QEX's canonical checkerboard index
// i = floor(lex/2), lex = x + Lx*(y + Ly*(z + Lz*t))  (QEX layout/qlayout.nim:108-136),
// so a wavefront of 64 consecutive sites reads 64 consecutive complex words per
// component.  stride = Vh rounded up to a multiple of 256 sites.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace qh {

// ------------------------------------------------------------ precision tags
enum Prec : int { P64 = 0, P32 = 1, P16 = 2, PBF = 3 };

struct c64 { double re, im; };
struct c32 { float re, im; };

template <class T> struct acc2;
template <> struct acc2<double> { using type = double2; };
template <> struct acc2<float> { using type = float2; };

// storage word for one complex number per precision
struct half2s { __half2 v; };
struct bf2s { uint32_t v; };

template <int P> struct Store;
template <> struct Store<P64> { using word = double2; static constexpr int bytes = 16; };
template <> struct Store<P32> { using word = float2; static constexpr int bytes = 8; };
template <> struct Store<P16> { using word = __half2; static constexpr int bytes = 4; };
template <> struct Store<PBF> { using word = uint32_t; static constexpr int bytes = 4; };

__device__ __forceinline__ float bf16_to_f(uint32_t b) { return __uint_as_float(b << 16); }
// round-to-nearest-even float -> bf16 (finite inputs)
__device__ __forceinline__ uint32_t f_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

// load one complex (as accumulation type A) from storage word
template <int P, class A> __device__ __forceinline__ A ldc(const void *base, long off);
template <> __device__ __forceinline__ double2 ldc<P64, double2>(const void *b, long o) {
  return ((const double2 *)b)[o];
}
template <> __device__ __forceinline__ float2 ldc<P32, float2>(const void *b, long o) {
  return ((const float2 *)b)[o];
}
template <> __device__ __forceinline__ double2 ldc<P32, double2>(const void *b, long o) {
  float2 f = ((const float2 *)b)[o];
  return make_double2(f.x, f.y);
}
template <> __device__ __forceinline__ float2 ldc<P64, float2>(const void *b, long o) {
  double2 d = ((const double2 *)b)[o];
  return make_float2((float)d.x, (float)d.y);
}
template <> __device__ __forceinline__ float2 ldc<P16, float2>(const void *b, long o) {
  return __half22float2(((const __half2 *)b)[o]);
}
template <> __device__ __forceinline__ float2 ldc<PBF, float2>(const void *b, long o) {
  uint32_t w = ((const uint32_t *)b)[o];
  return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}

template <int P, class A> __device__ __forceinline__ void stc(void *base, long off, A v);
template <> __device__ __forceinline__ void stc<P64, double2>(void *b, long o, double2 v) {
  ((double2 *)b)[o] = v;
}
template <> __device__ __forceinline__ void stc<P32, float2>(void *b, long o, float2 v) {
  ((float2 *)b)[o] = v;
}
template <> __device__ __forceinline__ void stc<P32, double2>(void *b, long o, double2 v) {
  ((float2 *)b)[o] = make_float2((float)v.x, (float)v.y);
}
template <> __device__ __forceinline__ void stc<P64, float2>(void *b, long o, float2 v) {
  ((double2 *)b)[o] = make_double2(v.x, v.y);
}
template <> __device__ __forceinline__ void stc<P16, float2>(void *b, long o, float2 v) {
  ((__half2 *)b)[o] = __float22half2_rn(v);
}
template <> __device__ __forceinline__ void stc<PBF, float2>(void *b, long o, float2 v) {
  ((uint32_t *)b)[o] = f_to_bf16(v.x) | (f_to_bf16(v.y) << 16);
}

// ----------------------------------------------------------- complex algebra
template <class A> struct cx;
template <> struct cx<double2> {
  using R = double;
};
template <> struct cx<float2> {
  using R = float;
};

template <class A> __device__ __forceinline__ A mk(typename cx<A>::R x, typename cx<A>::R y) {
  A a; a.x = x; a.y = y; return a;
}
// acc += a*b
template <class A> __device__ __forceinline__ void cmac(A &acc, A a, A b) {
  acc.x += a.x * b.x - a.y * b.y;
  acc.y += a.x * b.y + a.y * b.x;
}
// acc += conj(a)*b
template <class A> __device__ __forceinline__ void cjmac(A &acc, A a, A b) {
  acc.x += a.x * b.x + a.y * b.y;
  acc.y += a.x * b.y - a.y * b.x;
}

// ------------------------------------------------------------------ geometry
struct Geom {
  int L[4];      // Lx, Ly, Lz, Lt
  int hx;        // Lx/2
  long Vh;       // sites per parity
  long stride;   // padded sites per parity in the SoA layout
};

struct Site {
  int x, y, z, t;
  long row;      // y + Ly*(z + Lz*t)
};

__device__ __forceinline__ Site decode(const Geom &g, long i, int par) {
  Site s;
  int xh = (int)(i % g.hx);
  long r = i / g.hx;
  s.row = r;
  s.y = (int)(r % g.L[1]);
  long r2 = r / g.L[1];
  s.z = (int)(r2 % g.L[2]);
  s.t = (int)(r2 / g.L[2]);
  s.x = 2 * xh + ((par + s.y + s.z + s.t) & 1);
  return s;
}

__device__ __forceinline__ int wrap(int v, int L) {
  v = (v >= L) ? v - L : v;
  return (v < 0) ? v + L : v;
}

// checkerboard index (in the opposite parity) of the site s + d*mu_hat
__device__ __forceinline__ long nbr(const Geom &g, const Site &s, int mu, int d) {
  if (mu == 0) {
    int x = wrap(s.x + d, g.L[0]);
    return (x >> 1) + (long)g.hx * s.row;
  }
  int y = s.y, z = s.z, t = s.t;
  if (mu == 1) y = wrap(y + d, g.L[1]);
  else if (mu == 2) z = wrap(z + d, g.L[2]);
  else t = wrap(t + d, g.L[3]);
  return (s.x >> 1) + (long)g.hx * (y + (long)g.L[1] * (z + (long)g.L[2] * t));
}

// ---------------------------------------------------------------- reductions
// wave64 sum (DPP/permute via shfl_xor), then across the waves of the block in
// LDS.  Result valid in thread 0.  blockDim.x must be a multiple of 64, <=1024.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *lds /* 16*NV */) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) lds[w * NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = 0;
      for (int j = 0; j < nw; j++) s += lds[j * NV + k];
      v[k] = s;
    }
  }
  __syncthreads();
}

}  // namespace qh
