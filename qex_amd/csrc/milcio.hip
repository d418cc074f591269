// MILC binary gauge configurations (QEX src/new/milcIO.nim, milcFileField.nim):
// a 96-byte header {int32 magic 20103, int32 dims[4], char time_stamp[64],
// int32 order, uint32 sum29, uint32 sum31} followed, for order 0, by the
// sites in lexicographic order (x fastest), each holding the four links
// U_0..U_3 as 3x3 complex single-precision numbers (288 bytes per site).
// The checksum XORs every 32-bit data word w_k rotated left by k mod 29 and
// k mod 31 (milcIO.nim:43-51).  Files written on a machine of the other byte
// order carry the byte-swapped magic number; their words are swapped back on
// reading.
//
// The host reads or writes the bytes and the checksum; one kernel converts
// between the file's site order and the library's canonical FP64 link arrays
// [mu][site][3][3] on the device.
#include <stdio.h>
#include <string.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "internal.h"
#include "su3.h"

namespace qh {

namespace {

constexpr int32_t MILC_MAGIC = 20103;

struct MilcHeader {
  int32_t magic;
  int32_t dims[4];
  char stamp[64];
  int32_t order;
  uint32_t sum29, sum31;
};
static_assert(sizeof(MilcHeader) == 96, "MILC header is 96 bytes");

uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

uint32_t rotl(uint32_t v, int k) { return k ? (v << k) | (v >> (32 - k)) : v; }

struct File {
  FILE *f;
  explicit File(FILE *p) : f(p) {}
  ~File() {
    if (f) fclose(f);
  }
};

}  // namespace

void milc_checksum(const uint32_t *w, size_t n, uint32_t *s29, uint32_t *s31) {
  uint32_t a = 0, b = 0;
  int k29 = 0, k31 = 0;
  for (size_t k = 0; k < n; k++) {
    a ^= rotl(w[k], k29);
    b ^= rotl(w[k], k31);
    if (++k29 == 29) k29 = 0;
    if (++k31 == 31) k31 = 0;
  }
  *s29 = a;
  *s31 = b;
}

// file (lexicographic site, mu, 18 floats) -> canonical [mu][site][9] double2
__global__ void __launch_bounds__(256) k_milc_to_canon(Geom g, const float *raw, double2 *U) {
  const long V = 2 * g.Vh;
  const long lex = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (lex >= V) return;
  int c[4];
  long r = lex;
  for (int d = 0; d < 4; d++) {
    c[d] = (int)(r % g.L[d]);
    r /= g.L[d];
  }
  const long idx = canon_index(g, c);
  const float *src = raw + (lex * 4 + mu) * 18;
  double2 *dst = U + ((long)mu * V + idx) * 9;
#pragma unroll
  for (int e = 0; e < 9; e++) dst[e] = make_double2((double)src[2 * e], (double)src[2 * e + 1]);
}

__global__ void __launch_bounds__(256) k_canon_to_milc(Geom g, const double2 *U, float *raw) {
  const long V = 2 * g.Vh;
  const long lex = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (lex >= V) return;
  int c[4];
  long r = lex;
  for (int d = 0; d < 4; d++) {
    c[d] = (int)(r % g.L[d]);
    r /= g.L[d];
  }
  const long idx = canon_index(g, c);
  const double2 *src = U + ((long)mu * V + idx) * 9;
  float *dst = raw + (lex * 4 + mu) * 18;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    dst[2 * e] = (float)src[e].x;
    dst[2 * e + 1] = (float)src[e].y;
  }
}

void milc_read_header(const char *path, MilcInfo *info) {
  File f(fopen(path, "rb"));
  if (!f.f) throw std::runtime_error(std::string("cannot open ") + path);
  MilcHeader h;
  if (fread(&h, sizeof h, 1, f.f) != 1) throw std::runtime_error("MILC file: short header");
  bool swap = false;
  if (h.magic != MILC_MAGIC) {
    if ((int32_t)bswap32((uint32_t)h.magic) != MILC_MAGIC) throw std::runtime_error("not a MILC gauge file");
    swap = true;
    h.magic = MILC_MAGIC;
    for (int d = 0; d < 4; d++) h.dims[d] = (int32_t)bswap32((uint32_t)h.dims[d]);
    h.order = (int32_t)bswap32((uint32_t)h.order);
    h.sum29 = bswap32(h.sum29);
    h.sum31 = bswap32(h.sum31);
  }
  for (int d = 0; d < 4; d++) info->dims[d] = h.dims[d];
  memcpy(info->stamp, h.stamp, 64);
  info->stamp[64] = 0;
  info->order = h.order;
  info->byteswapped = swap ? 1 : 0;
  info->sum29 = h.sum29;
  info->sum31 = h.sum31;
}

void milc_read(const Geom &g, const char *path, double *U, MilcInfo *info, hipStream_t st) {
  milc_read_header(path, info);
  for (int d = 0; d < 4; d++)
    if (info->dims[d] != g.L[d]) throw std::runtime_error("MILC file dimensions differ from the lattice");
  if (info->order != 0) throw std::runtime_error("MILC file: only site-ordered (order 0) files are supported");
  const long V = 2 * g.Vh;
  const size_t nwords = (size_t)V * 4 * 18;
  std::vector<uint32_t> w(nwords);
  {
    File f(fopen(path, "rb"));
    if (!f.f) throw std::runtime_error(std::string("cannot open ") + path);
    if (fseek(f.f, (long)sizeof(MilcHeader), SEEK_SET) != 0) throw std::runtime_error("MILC file: seek failed");
    if (fread(w.data(), sizeof(uint32_t), nwords, f.f) != nwords)
      throw std::runtime_error("MILC file: shorter than its dimensions require (single-precision data expected)");
  }
  if (info->byteswapped)
    for (auto &x : w) x = bswap32(x);
  uint32_t s29, s31;
  milc_checksum(w.data(), nwords, &s29, &s31);
  info->checksum_ok = (s29 == info->sum29 && s31 == info->sum31) ? 1 : 0;
  if (!info->checksum_ok) throw std::runtime_error("MILC file: checksum mismatch");
  float *raw = nullptr;
  if (hipMalloc((void **)&raw, nwords * sizeof(float)) != hipSuccess)
    throw std::runtime_error("MILC read: out of device memory");
  struct Free {
    float *p;
    ~Free() { (void)hipFree(p); }
  } guard{raw};
  if (hipMemcpyAsync(raw, w.data(), nwords * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess)
    throw std::runtime_error("MILC read: copy to device failed");
  hipLaunchKernelGGL(k_milc_to_canon, dim3((unsigned)((V + 255) / 256), 4), dim3(256), 0, st, g, raw, (double2 *)U);
  if (hipStreamSynchronize(st) != hipSuccess) throw std::runtime_error("MILC read: conversion failed");
}

void milc_write(const Geom &g, const char *path, const double *U, const char *stamp, hipStream_t st) {
  const long V = 2 * g.Vh;
  const size_t nwords = (size_t)V * 4 * 18;
  std::vector<uint32_t> w(nwords);
  float *raw = nullptr;
  if (hipMalloc((void **)&raw, nwords * sizeof(float)) != hipSuccess)
    throw std::runtime_error("MILC write: out of device memory");
  struct Free {
    float *p;
    ~Free() { (void)hipFree(p); }
  } guard{raw};
  hipLaunchKernelGGL(k_canon_to_milc, dim3((unsigned)((V + 255) / 256), 4), dim3(256), 0, st, g, (const double2 *)U,
                     raw);
  if (hipMemcpyAsync(w.data(), raw, nwords * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    throw std::runtime_error("MILC write: conversion failed");
  MilcHeader h;
  memset(&h, 0, sizeof h);
  h.magic = MILC_MAGIC;
  for (int d = 0; d < 4; d++) h.dims[d] = g.L[d];
  if (stamp) strncpy(h.stamp, stamp, sizeof h.stamp);
  h.order = 0;
  milc_checksum(w.data(), nwords, &h.sum29, &h.sum31);
  File f(fopen(path, "wb"));
  if (!f.f) throw std::runtime_error(std::string("cannot create ") + path);
  if (fwrite(&h, sizeof h, 1, f.f) != 1 || fwrite(w.data(), sizeof(uint32_t), nwords, f.f) != nwords)
    throw std::runtime_error("MILC write: write failed");
}

}  // namespace qh
