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
// The host reads and writes the header; the site data stream through the
// chunk pipeline of iopipe.h, and one kernel per chunk converts between the
// file's words and the library's canonical FP64 link arrays [mu][site][3][3]
// while computing the checksum partials of its 64 sites.
#include <fcntl.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "internal.h"
#include "iopipe.h"
#include "su3.h"

namespace qh {

namespace {

using namespace iop;

constexpr int32_t MILC_MAGIC = 20103;
constexpr size_t MILC_SITE_BYTES = 4 * 18 * sizeof(float);

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

struct Fd {
  int fd;
  explicit Fd(int d) : fd(d) {}
  ~Fd() {
    if (fd >= 0) close(fd);
  }
  Fd(const Fd &) = delete;
  Fd &operator=(const Fd &) = delete;
};

__device__ __forceinline__ long canon_of_lex(const Geom &g, long lex) {
  int c[4];
  long r = lex;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    c[d] = (int)(r % g.L[d]);
    r /= g.L[d];
  }
  return canon_index(g, c);
}

// XOR of the block's 256 lanes into part[2*blockIdx.x ..]
__device__ __forceinline__ void block_xor(uint32_t a, uint32_t b, uint32_t *part) {
  __shared__ uint32_t red[8];
  a = wave_xor(a);
  b = wave_xor(b);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[2 * w] = a;
    red[2 * w + 1] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0] ^ red[2] ^ red[4] ^ red[6];
    part[2 * blockIdx.x + 1] = red[1] ^ red[3] ^ red[5] ^ red[7];
  }
}

// One lane per link: site lex0 + blockIdx.x*64 + threadIdx.x/4, mu = threadIdx.x%4.
// Its 18 words have data word indices k0 .. k0+17, k0 = ((lex*4 + mu)*18).
// file words of a chunk -> canonical links, with the (sum29, sum31) partials
__global__ void __launch_bounds__(256)
    k_milc_unpack(Geom g, const uint32_t *raw, long lex0, long n, int swap, double2 *U, uint32_t *part) {
  const long V = 2 * g.Vh;
  const long s = (long)blockIdx.x * SITES_PER_BLOCK + (threadIdx.x >> 2);
  const int mu = threadIdx.x & 3;
  uint32_t a = 0, b = 0;
  if (s < n) {
    const long lex = lex0 + s;
    const uint32_t *src = raw + (s * 4 + mu) * 18;
    double2 *dst = U + ((long)mu * V + canon_of_lex(g, lex)) * 9;
    const long k0 = (lex * 4 + mu) * 18;
    int r29 = (int)(k0 % 29), r31 = (int)(k0 % 31);
#pragma unroll
    for (int e = 0; e < 18; e++) {
      uint32_t w = src[e];
      if (swap) w = __builtin_bswap32(w);
      a ^= rotl32(w, r29);
      b ^= rotl32(w, r31);
      r29 = r29 == 28 ? 0 : r29 + 1;
      r31 = r31 == 30 ? 0 : r31 + 1;
      if (e & 1)
        dst[e >> 1].y = (double)__uint_as_float(w);
      else
        dst[e >> 1].x = (double)__uint_as_float(w);
    }
  }
  block_xor(a, b, part);
}

// canonical links -> file words of a chunk (this machine's byte order)
__global__ void __launch_bounds__(256)
    k_milc_pack(Geom g, const double2 *U, long lex0, long n, uint32_t *raw, uint32_t *part) {
  const long V = 2 * g.Vh;
  const long s = (long)blockIdx.x * SITES_PER_BLOCK + (threadIdx.x >> 2);
  const int mu = threadIdx.x & 3;
  uint32_t a = 0, b = 0;
  if (s < n) {
    const long lex = lex0 + s;
    uint32_t *dst = raw + (s * 4 + mu) * 18;
    const double2 *src = U + ((long)mu * V + canon_of_lex(g, lex)) * 9;
    const long k0 = (lex * 4 + mu) * 18;
    int r29 = (int)(k0 % 29), r31 = (int)(k0 % 31);
#pragma unroll
    for (int e = 0; e < 18; e++) {
      const double v = (e & 1) ? src[e >> 1].y : src[e >> 1].x;
      const uint32_t w = __float_as_uint((float)v);
      dst[e] = w;
      a ^= rotl32(w, r29);
      b ^= rotl32(w, r31);
      r29 = r29 == 28 ? 0 : r29 + 1;
      r31 = r31 == 30 ? 0 : r31 + 1;
    }
  }
  block_xor(a, b, part);
}

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
  Fd fd(open(path, O_RDONLY));
  if (fd.fd < 0) throw std::runtime_error(std::string("cannot open ") + path);
  struct stat sb;
  if (fstat(fd.fd, &sb) != 0 || (long long)sb.st_size < (long long)sizeof(MilcHeader) + (long long)V * 288)
    throw std::runtime_error("MILC file: shorter than its dimensions require (single-precision data expected)");
  const Chunking ch(g, MILC_SITE_BYTES);
  Pipe p(st, (size_t)ch.per * ch.site_bytes, ch.nblocks, 0, "MILC read");
  long blk = 0;
  for (long c = 0; c < ch.nchunks; c++) {
    const int b = (int)(c & 1);
    const size_t bytes = ch.bytes(g, c);
    if (c >= 2 && hipEventSynchronize(p.ev[b]) != hipSuccess) throw std::runtime_error("MILC read: device error");
    pio(false, fd.fd, (char *)p.host[b], bytes, (off_t)sizeof(MilcHeader) + (off_t)(c * ch.per) * (off_t)ch.site_bytes,
        "MILC read");
    if (hipMemcpyAsync(p.dev[b], p.host[b], bytes, hipMemcpyHostToDevice, st) != hipSuccess)
      throw std::runtime_error("MILC read: copy to device failed");
    const long nb = ch.blocks(g, c);
    hipLaunchKernelGGL(k_milc_unpack, dim3((unsigned)nb), dim3(256), 0, st, g, p.dev[b], c * ch.per,
                       ch.sites(g, c), info->byteswapped, (double2 *)U, p.part + 2 * blk);
    if (hipGetLastError() != hipSuccess || hipEventRecord(p.ev[b], st) != hipSuccess)
      throw std::runtime_error("MILC read: conversion failed");
    blk += nb;
  }
  if (hipStreamSynchronize(st) != hipSuccess) throw std::runtime_error("MILC read: conversion failed");
  uint32_t s29, s31;
  p.checksum(ch.nblocks, &s29, &s31, "MILC read");
  info->checksum_ok = (s29 == info->sum29 && s31 == info->sum31) ? 1 : 0;
  if (!info->checksum_ok) throw std::runtime_error("MILC file: checksum mismatch");
}

void milc_write(const Geom &g, const char *path, const double *U, const char *stamp, hipStream_t st) {
  const long V = 2 * g.Vh;
  Fd fd(open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644));
  if (fd.fd < 0) throw std::runtime_error(std::string("cannot create ") + path);
  (void)posix_fallocate(fd.fd, 0, (off_t)sizeof(MilcHeader) + (off_t)V * 288);  // a refusal is harmless
  const Chunking ch(g, MILC_SITE_BYTES);
  Pipe p(st, (size_t)ch.per * ch.site_bytes, ch.nblocks, 0, "MILC write");
  long blk = 0;
  auto drain = [&](long c) {  // write chunk c once its copy to the host is done
    const int b = (int)(c & 1);
    if (hipEventSynchronize(p.ev[b]) != hipSuccess) throw std::runtime_error("MILC write: device error");
    pio(true, fd.fd, (char *)p.host[b], ch.bytes(g, c),
        (off_t)sizeof(MilcHeader) + (off_t)(c * ch.per) * (off_t)ch.site_bytes, "MILC write");
  };
  for (long c = 0; c < ch.nchunks; c++) {
    const int b = (int)(c & 1);
    if (c >= 2) drain(c - 2);
    const long nb = ch.blocks(g, c);
    hipLaunchKernelGGL(k_milc_pack, dim3((unsigned)nb), dim3(256), 0, st, g, (const double2 *)U, c * ch.per,
                       ch.sites(g, c), p.dev[b], p.part + 2 * blk);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(p.host[b], p.dev[b], ch.bytes(g, c), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipEventRecord(p.ev[b], st) != hipSuccess)
      throw std::runtime_error("MILC write: conversion failed");
    blk += nb;
  }
  for (long c = ch.nchunks >= 2 ? ch.nchunks - 2 : 0; c < ch.nchunks; c++) drain(c);
  MilcHeader h;
  memset(&h, 0, sizeof h);
  h.magic = MILC_MAGIC;
  for (int d = 0; d < 4; d++) h.dims[d] = g.L[d];
  if (stamp) strncpy(h.stamp, stamp, sizeof h.stamp);
  h.order = 0;
  p.checksum(ch.nblocks, &h.sum29, &h.sum31, "MILC write");
  pio(true, fd.fd, (char *)&h, sizeof h, 0, "MILC write");
}

}  // namespace qh
