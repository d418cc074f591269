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
#include <type_traits>
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

// Stencil loop form (stencil.h): >= 0 software-pipeline depth in links;
// -1 / -2 rolled loop over hops / links with separate forward and backward
// accumulators; -3 rolled loop over links with one accumulator (fewest
// registers: 100 VGPRs for the FP64 hop; fastest measured).
#ifndef QH_PIPE
#define QH_PIPE -3
#endif
// the same for the stencils of 16-bit link formats.  Measured at 32^4 (Q16
// 12-real hop / mixed CG iteration): -3 22.8 / 70.4 us, -1 24.5 / 73.5 us,
// -4 25.3 / 78.3 us; the directions fully unrolled (40% fewer vector
// instructions, all loads in flight, 4 waves per SIMD instead of 8) were 3%
// slower in the CG.  Waves per SIMD, not instruction issue or loads in flight
// per wave, set these kernels' speed.
#ifndef QH_PIPE16
#define QH_PIPE16 -5
#endif

namespace qh {

// ------------------------------------------------------------ precision tags
// P64/P32/P16/PBF: IEEE double, single, half, bfloat16.  PQ16 (links only):
// 16-bit fixed point, value = int16 * scale/32767 with one scale per link
// field (the largest |re|,|im| of that field), i.e. absolute precision
// scale/32767 instead of FP16's 11-bit relative precision.
enum Prec : int { P64 = 0, P32 = 1, P16 = 2, PBF = 3, PQ16 = 4 };

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
// element offsets are 32-bit (fields are < 4 GiB): scalar base + vector offset
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
template <int P, class A> __device__ __forceinline__ A ldc(const void *base, long off);
template <> __device__ __forceinline__ double2 ldc<P64, double2>(const void *b, long o) {
  const d2v v = *(const d2v *)((const char *)b + (uint32_t)o * 16u);
  return make_double2(v.x, v.y);
}
template <> __device__ __forceinline__ float2 ldc<P32, float2>(const void *b, long o) {
  const f2v v = *(const f2v *)((const char *)b + (uint32_t)o * 8u);
  return make_float2(v.x, v.y);
}
template <> __device__ __forceinline__ double2 ldc<P32, double2>(const void *b, long o) {
  const f2v v = *(const f2v *)((const char *)b + (uint32_t)o * 8u);
  return make_double2(v.x, v.y);
}
template <> __device__ __forceinline__ float2 ldc<P64, float2>(const void *b, long o) {
  const d2v v = *(const d2v *)((const char *)b + (uint32_t)o * 16u);
  return make_float2((float)v.x, (float)v.y);
}
template <> __device__ __forceinline__ float2 ldc<P16, float2>(const void *b, long o) {
  return __half22float2(*(const __half2 *)((const char *)b + (uint32_t)o * 4u));
}
template <> __device__ __forceinline__ float2 ldc<PBF, float2>(const void *b, long o) {
  const uint32_t w = *(const uint32_t *)((const char *)b + (uint32_t)o * 4u);
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

// Output store of the stencil kernels: the FP64 result is written with a
// non-temporal hint (written once, not re-read by the same kernel; +2% on the
// 32^4 hop).  QH_NO_NT_STORE restores plain stores.
template <int P, class A> __device__ __forceinline__ void stc_out(void *b, long o, A v) {
#ifndef QH_NO_NT_STORE
  if constexpr (P == P64 && std::is_same<A, double2>::value) {
    typedef double d2v __attribute__((ext_vector_type(2)));
    d2v t = {v.x, v.y};
    __builtin_nontemporal_store(t, (d2v *)b + o);
    return;
  }
#endif
  stc<P, A>(b, o, v);
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
// Complex multiply-accumulates as chains of fused multiply-adds on the
// accumulator (four per product; written as acc += (p - q) they would cost six).
template <class A> __device__ __forceinline__ void cmac(A &acc, A a, A b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
// acc += conj(a)*b
template <class A> __device__ __forceinline__ void cjmac(A &acc, A a, A b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
}
// acc -= conj(a)*b
template <class A> __device__ __forceinline__ void cjmsub(A &acc, A a, A b) {
  acc.x = fma(-a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(-a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

// ------------------------------------------------------------------ geometry
struct Geom {
  int L[4];      // Lx, Ly, Lz, Lt
  int hx;        // Lx/2
  long Vh;       // sites per parity
  long stride;   // padded sites per parity in the SoA layout
  long lkcache;  // link bytes the stencils read with cache-allocating loads (link_cached_blocks)
};

// Element index of color component c of site i in a vector field.  Blocked
// layout (default): the three components of 64 consecutive sites together
// (192 elements per block), one contiguous region per wave; measured at 32^4
// against component planes a stride apart: FP64 hop 103.1 -> 101.3 us, FP16
// and Q16 12-real hops -4 to -5%, FP64 CG -1.8%, FP32-inner CG -3%; the 8-link
// HISQ FP64 hop at 32^3 x 64 397 -> 412 us with its links in planes, 386 us
// with them blocked as well.  QH_VEC_BLOCKED=0 gives the plane layout.
#ifndef QH_VEC_BLOCKED
#define QH_VEC_BLOCKED 1
#endif
__host__ __device__ __forceinline__ long vix(const Geom &g, int c, long i) {
  if (QH_VEC_BLOCKED) return (i >> 6) * 192 + c * 64 + (i & 63);
  return c * g.stride + i;
}

struct Site {
  int x, y, z, t;
  long row;      // y + Ly*(z + Lz*t)
};

__host__ __device__ __forceinline__ Site decode(const Geom &g, long i, int par) {
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

__host__ __device__ __forceinline__ int wrap(int v, int L) {
  v = (v >= L) ? v - L : v;
  return (v < 0) ? v + L : v;
}

// checkerboard index (in the opposite parity) of the site s + d*mu_hat
__host__ __device__ __forceinline__ long nbr(const Geom &g, const Site &s, int mu, int d) {
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

// Sum over the block of each v[k]; the total lands in thread 0 (blockDim.x a
// multiple of 64, at most 1024).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *lds /* 16*NV */) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) lds[w * NV + k] = v[k];
  __syncthreads();
  // the first wave adds the (at most 16) wave sums: lanes 0-15 each take one,
  // then a 4-step butterfly leaves the total in every one of them (one thread
  // adding them in turn made the 1024-thread finish of the CG dot products
  // 6.2 us instead of 5.4 us)
  if (w == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = lane < nw ? lds[lane * NV + k] : 0.0;
#pragma unroll
      for (int o = 8; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
      v[k] = s;
    }
  }
  __syncthreads();
}

// Reduction workspace shared by consecutive kernels on one stream.
struct RedWs {
  double *partials;   // [4][maxblocks]
  int maxblocks;
};

// Block partial sums of a reduction: the block's total of each of the NV
// values goes to partials[k*maxblocks + block]; a one-block finish kernel
// (reduce.hip) then adds the partials in a fixed order (deterministic, and no
// grid-wide fence in the producing kernel).
template <int NV>
__device__ __forceinline__ void block_partials_at(double (&v)[NV], double *lds, const RedWs &red, long bid) {
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) red.partials[(long)k * red.maxblocks + bid] = v[k];
}

template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double *lds, const RedWs &red) {
  block_partials_at<NV>(v, lds, red, blockIdx.y * gridDim.x + blockIdx.x);
}

// XCD-aware block order: the hardware deals workgroups round-robin over the 8
// XCDs (block b runs on XCD b%8); this remap gives each XCD one contiguous run
// of sites, so stencil neighbours (y, z, t rows) are read through the same L2.
__device__ __forceinline__ long xcd_block(int remap) {
  const long b = blockIdx.x;
  if (!remap) return b;
  const long nb = gridDim.x, q = nb >> 3, r = nb & 7;
  const long xcd = b & 7, idx = b >> 3;
  return xcd * q + (xcd < r ? xcd : r) + idx;
}

// Logical blocks served by this workgroup.  The launch has gridDim.x blocks,
// a multiple of 8 when remap is set, and may be smaller than the nblk logical
// blocks (persistent form): with remap, XCD x (= blockIdx.x % 8) owns the
// contiguous logical range [x*Q, (x+1)*Q), Q = ceil(nblk/8), which its
// gridDim.x/8 workgroups walk with stride gridDim.x/8; without remap the
// blocks stride over all logical blocks.  With gridDim.x = nblk rounded up to
// 8 every workgroup serves at most one logical block (the xcd_block order).
struct BlockIter {
  long cur, end, step;
};
__device__ __forceinline__ BlockIter block_iter(long nblk, int remap) {
  const long G = gridDim.x;
  if (!remap) return BlockIter{(long)blockIdx.x, nblk, G};
  const long per = G >> 3, x = blockIdx.x & 7, lane = blockIdx.x >> 3;
  const long Q = (nblk + 7) >> 3;
  const long end = (x + 1) * Q < nblk ? (x + 1) * Q : nblk;
  return BlockIter{x * Q + lane, end, per};
}

// Site walk of a one-parity stencil launch.  chunk = 0: logical block lb
// covers sites [lb*spb, (lb+1)*spb) in canonical order.  chunk = C > 0 (a
// multiple of spb): the S = Vh/Lt sites of one t-slice are cut into chunks of
// C consecutive sites, and the logical blocks walk t fastest, then the chunk:
// block lb = (k*Lt + t)*B + w covers sites t*S + k*C + w*spb + [0, spb), B =
// C/spb.  The t -/+ 1 neighbours of a chunk are then read again while the
// previous and next t-slices of the same chunk are in flight on the same XCD
// (with the XCD-contiguous block ranges of block_iter), instead of 2 S sites
// of traffic later.  The last chunk of a slice may be partial.
struct Walk {
  long nblk;  // logical blocks
  long S;     // sites per t-slice of one parity
  long C;     // chunk (0: canonical order)
  long B;     // logical blocks per chunk and t
  int Lt;
};
__host__ __device__ __forceinline__ Walk make_walk(const Geom &g, long spb, long chunk) {
  Walk w;
  w.Lt = g.L[3];
  w.S = g.Vh / g.L[3];
  w.C = (chunk > 0 && chunk % spb == 0 && chunk < w.S) ? chunk : 0;
  if (w.C) {
    w.B = w.C / spb;
    w.nblk = (w.S + w.C - 1) / w.C * w.Lt * w.B;
  } else {
    w.B = 0;
    w.nblk = (g.Vh + spb - 1) / spb;
  }
  return w;
}
// site of local index l (0 <= l < spb) in logical block lb, or -1 past the lattice
__host__ __device__ __forceinline__ long walk_site(const Walk &w, long lb, long spb, long l) {
  if (!w.C) return lb * spb + l;  // the caller bounds it by Vh
  const long wb = lb % w.B, r = lb / w.B;
  const long t = r % w.Lt, k = r / w.Lt;
  const long off = k * w.C + wb * spb + l;
  return off < w.S ? t * w.S + off : -1;
}

// grid for nblk logical blocks: capped at cap workgroups when cap > 0,
// rounded up to a multiple of 8 for the XCD-aware order
inline unsigned grid_blocks(long nblk, int remap, int cap) {
  long G = (cap > 0 && cap < nblk) ? cap : nblk;
  if (remap) G = (G + 7) & ~7L;
  return (unsigned)(G > 0 ? G : 1);
}

// ------------------------------------------------------------- link storage
// One link = RC reals (18, or 12 = first two rows) in precision PL, stored as
// 16-byte planes plus one narrower remainder plane:
//   block (k, parity): plane p (0 <= p < nfull) at byte (p*stride + i)*16,
//                      remainder plane at byte nfull*16*stride + i*rem.
// Rows 0,1 of a 12-real link hold the phased link; row 2 is rebuilt as
// sign * conj(row0 x row1), the sign being the staggered phase times the
// antiperiodic-t boundary sign of that link (phased SU(3) = sign * SU(3)).
template <int PL, int RC> struct LFmt {
  static constexpr int bpr = (PL == P64) ? 8 : (PL == P32) ? 4 : 2;
  static constexpr int bytes = RC * bpr;
  static constexpr int nfull = bytes / 16;
  static constexpr int rem = bytes % 16;
  static constexpr int words = bytes / 4;
};

inline long link_block_bytes(int prec, int rc, long stride) {
  const int bpr = (prec == P64) ? 8 : (prec == P32) ? 4 : 2;
  return (long)rc * bpr * stride;
}

typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef unsigned int u2v __attribute__((ext_vector_type(2)));

// Blocked link storage (BLK): the planes of each run of 64 sites lie
// together, plane p of site i at (i/64)*64*bytes + p*1024 + (i%64)*16 and the
// remainder plane at (i/64)*64*bytes + nfull*1024 + (i%64)*rem, so that one
// wave's link loads come from one contiguous region instead of nfull+1
// regions a plane apart (stride is a multiple of 64).  Used for FP64 and FP32
// fields, plain and HISQ (one-parity hop at 32^4: FP64 110 -> 103.5 us,
// 12-real 79 -> 69 us, FP32 58.6 -> 54 us; the 8-link HISQ FP64 hop at
// 32^3 x 64: 410 -> 386 us with blocked color vectors, though 401 -> 421 us
// with planar ones).  The 16-bit formats moved within +-2% and keep the plane
// layout.  QH_LBLK_HISQ=0 keeps HISQ link sets in planes.
#ifndef QH_LBLK_HISQ
#define QH_LBLK_HISQ 1
#endif
template <int PL, int NL> constexpr bool link_blocked() {
  return (NL == 4 || QH_LBLK_HISQ) && (PL == P64 || PL == P32);
}
#ifndef QH_LBLK_LOG2
#define QH_LBLK_LOG2 6  // sites per block: 64
#endif
constexpr int LBLK = 1 << QH_LBLK_LOG2;
template <int PL, int RC, bool BLK> __host__ __device__ __forceinline__ long lplane_off(long stride, long i, int p) {
  using F = LFmt<PL, RC>;
  if (BLK) return (i >> QH_LBLK_LOG2) * ((long)LBLK * F::bytes) + (long)p * LBLK * 16 + (i & (LBLK - 1)) * 16;
  return ((long)p * stride + i) * 16;
}
template <int PL, int RC, bool BLK> __host__ __device__ __forceinline__ long lrem_off(long stride, long i) {
  using F = LFmt<PL, RC>;
  if (BLK)
    return (i >> QH_LBLK_LOG2) * ((long)LBLK * F::bytes) + (long)F::nfull * LBLK * 16 + (i & (LBLK - 1)) * F::rem;
  return (long)F::nfull * 16 * stride + i * F::rem;
}

template <int PL, int RC, bool BLK>
__device__ __forceinline__ void st_link_words(char *blk, long stride, long i, const uint32_t (&w)[LFmt<PL, RC>::words]) {
  using F = LFmt<PL, RC>;
#pragma unroll
  for (int p = 0; p < F::nfull; p++) {
    u4v v;
    v.x = w[4 * p]; v.y = w[4 * p + 1]; v.z = w[4 * p + 2]; v.w = w[4 * p + 3];
    *(u4v *)(blk + lplane_off<PL, RC, BLK>(stride, i, p)) = v;
  }
  if (F::rem == 8) {
    u2v v;
    v.x = w[4 * F::nfull]; v.y = w[4 * F::nfull + 1];
    *(u2v *)(blk + lrem_off<PL, RC, BLK>(stride, i)) = v;
  } else if (F::rem == 4) {
    *(unsigned *)(blk + lrem_off<PL, RC, BLK>(stride, i)) = w[4 * F::nfull];
  }
}

typedef float f4v __attribute__((ext_vector_type(4)));

template <class T, bool NT> __device__ __forceinline__ T ldv(const char *p) {
  return NT ? __builtin_nontemporal_load((const T *)p) : *(const T *)p;
}

// Where ld_link reads plane p (p == nfull: the remainder plane) of the link of
// site i: a base pointer and a 32-bit offset from it.  While one link field's
// block (stride x bytes per link) stays below 4 GiB the base is the block
// pointer and the offset holds everything (scalar base + 32-bit vector offset
// addressing).  A larger blocked field (FP64 18-real links from 29.8M sites
// per parity on) is "wide": the base moves to the site's 64-site block in
// 64-bit arithmetic and the offset stays below one block (9 KiB).  The plane
// layout (16-bit formats) is never wide: qexhip_lattice_create bounds it
// (36 B x stride).  The host test of the address arithmetic calls these too.
template <int PL, int RC, bool BLK>
__host__ __device__ __forceinline__ bool link_wide(long stride) {
  return BLK && (unsigned long)stride * (unsigned long)LFmt<PL, RC>::bytes > 0xffffffffUL;
}
template <int PL, int RC, bool BLK, bool WIDE>
__host__ __device__ __forceinline__ const char *link_base(const char *blk, long i) {
  using F = LFmt<PL, RC>;
  return (BLK && WIDE) ? blk + (size_t)(i >> QH_LBLK_LOG2) * (size_t)(LBLK * F::bytes) : blk;
}
template <int PL, int RC, bool BLK, bool WIDE>
__host__ __device__ __forceinline__ uint32_t link_off(long stride, long i, int p) {
  using F = LFmt<PL, RC>;
  const uint32_t base = (BLK && !WIDE) ? (uint32_t)(i >> QH_LBLK_LOG2) * (uint32_t)(LBLK * F::bytes) : 0u;
  const uint32_t lane = (uint32_t)(BLK ? (i & (LBLK - 1)) : i);
  const uint32_t pstep = BLK ? (uint32_t)LBLK * 16u : (uint32_t)stride * 16u;
  return base + (p < F::nfull ? (uint32_t)p * pstep + lane * 16u : (uint32_t)F::nfull * pstep + lane * (uint32_t)F::rem);
}

// The stored reals of one link (LFmt layout) into U: rows 0-1 for 12-real
// storage, all of it for 18.  QS = false leaves PQ16 values unscaled (the
// integers as floats; the caller applies the field's scale).
template <int PL, int RC, bool NT, class A, bool BLK, bool WIDE, bool QS = true>
__device__ __forceinline__ void ld_link_planes(const char *blk0, long stride, long i, A (&U)[9], float qscale) {
  using F = LFmt<PL, RC>;
  const char *blk = link_base<PL, RC, BLK, WIDE>(blk0, i);
  const uint32_t rbase = link_off<PL, RC, BLK, WIDE>(stride, i, F::nfull);
  auto poff = [&](int p) { return link_off<PL, RC, BLK, WIDE>(stride, i, p); };
  if constexpr (PL == P64) {
#pragma unroll
    for (int c = 0; c < RC / 2; c++) {
      const d2v v = ldv<d2v, NT>(blk + poff(c));
      U[c] = mk<A>(v.x, v.y);
    }
  } else if constexpr (PL == P32) {
#pragma unroll
    for (int p = 0; p < F::nfull; p++) {
      const f4v v = ldv<f4v, NT>(blk + poff(p));
      U[2 * p] = mk<A>(v.x, v.y);
      U[2 * p + 1] = mk<A>(v.z, v.w);
    }
    if constexpr (F::rem == 8) {
      const f2v v = ldv<f2v, NT>(blk + rbase);
      U[8] = mk<A>(v.x, v.y);
    }
  } else {
    uint32_t w[F::words];
#pragma unroll
    for (int p = 0; p < F::nfull; p++) {
      const u4v v = ldv<u4v, NT>(blk + poff(p));
      w[4 * p] = v.x; w[4 * p + 1] = v.y; w[4 * p + 2] = v.z; w[4 * p + 3] = v.w;
    }
    if constexpr (F::rem == 8) {
      const u2v v = ldv<u2v, NT>(blk + rbase);
      w[4 * F::nfull] = v.x; w[4 * F::nfull + 1] = v.y;
    } else if constexpr (F::rem == 4) {
      w[4 * F::nfull] = ldv<unsigned, NT>(blk + rbase);
    }
#pragma unroll
    for (int c = 0; c < RC / 2; c++) {
      if constexpr (PL == P16) {
        __half2 h;
        __builtin_memcpy(&h, &w[c], 4);
        const float2 f = __half22float2(h);
        U[c] = mk<A>(f.x, f.y);
      } else if constexpr (PL == PQ16) {
        if constexpr (QS)
          U[c] = mk<A>((float)(int16_t)(w[c] & 0xffffu) * qscale, (float)(int16_t)(w[c] >> 16) * qscale);
        else
          U[c] = mk<A>((float)(int16_t)(w[c] & 0xffffu), (float)(int16_t)(w[c] >> 16));
      } else {
        U[c] = mk<A>(__uint_as_float(w[c] << 16), __uint_as_float(w[c] & 0xffff0000u));
      }
    }
  }
}

// Load one link (LFmt layout, see above) straight into 3x3 complex form,
// rebuilding row 2 for 12-real storage: row2_k = sgn*conj(row0 x row1).  The
// wide/narrow choice depends only on stride, a kernel argument, so the branch
// is uniform over the wave.
template <int PL, int RC, bool NT, class A, bool BLK>
__device__ __forceinline__ void ld_link(const char *blk0, long stride, long i, A (&U)[9], typename cx<A>::R sgn,
                                        float qscale = 0.0f) {
  if (link_wide<PL, RC, BLK>(stride))
    ld_link_planes<PL, RC, NT, A, BLK, true>(blk0, stride, i, U, qscale);
  else
    ld_link_planes<PL, RC, NT, A, BLK, false>(blk0, stride, i, U, qscale);
  if constexpr (RC == 12) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const A a1 = U[(k + 1) % 3], a2 = U[(k + 2) % 3], b1 = U[3 + (k + 1) % 3], b2 = U[3 + (k + 2) % 3];
      // (a1 b2 - a2 b1), as one fused chain per component
      typename cx<A>::R re = a1.x * b2.x, im = a1.x * b2.y;
      re = fma(-a1.y, b2.y, re);
      re = fma(-a2.x, b1.x, re);
      re = fma(a2.y, b1.y, re);
      im = fma(a1.y, b2.x, im);
      im = fma(-a2.x, b1.y, im);
      im = fma(-a2.y, b1.x, im);
      U[6 + k] = mk<A>(sgn * re, -sgn * im);
    }
  }
}

// staggered phase (QEX masks [8,9,11,0], physics/stagD.nim:560-571) of a link
// in direction mu at site s, times the antiperiodic-t sign of the forward
// (fwd=1, link leaving s) or backward (fwd=0, link arriving at s) link.
__device__ __forceinline__ float link_sign(const Geom &g, const Site &s, int mu, int fwd) {
  int e = 0;
  if (mu == 0) e = s.t;
  else if (mu == 1) e = s.x + s.t;
  else if (mu == 2) e = s.x + s.y + s.t;
  else e = fwd ? (s.t == g.L[3] - 1) : (s.t == 0);
  return (e & 1) ? -1.0f : 1.0f;
}

}  // namespace qh
