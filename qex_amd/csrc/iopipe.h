// Host side of the lattice-file readers and writers (milcio.hip, scidac.hip):
// the file's site data move in whole t-slices of at most 64 MiB through two
// pinned host buffers and two device buffers; the host's read or write of one
// chunk overlaps the device conversion and PCIe copy of the next, and each
// chunk's file transfer is split over up to 8 threads.  The conversion
// kernels leave one pair of 32-bit checksum partials per 64-site block, which
// the host XORs together at the end.
#pragma once
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "common.h"

namespace qh {
namespace iop {

constexpr size_t CHUNK_BYTES = size_t(64) << 20;
constexpr int SITES_PER_BLOCK = 64;

// n bytes between buf and the file at offset off, split over up to 8 host
// threads of at least 8 MiB each (one thread copying into the page cache
// moves about 6 GB/s; parallel file systems also take parallel streams)
inline void pio(bool write, int fd, char *buf, size_t n, off_t off, const char *what) {
  constexpr size_t MIN_PART = size_t(8) << 20;
  const size_t nparts = n / MIN_PART < 1 ? 1 : (n / MIN_PART > 8 ? 8 : n / MIN_PART);
  const size_t part = (n + nparts - 1) / nparts;
  std::vector<int> err(nparts, 0);
  auto run = [&](size_t k) {
    size_t done = k * part;
    const size_t end = done + part < n ? done + part : n;
    while (done < end) {
      const ssize_t r = write ? pwrite(fd, buf + done, end - done, off + (off_t)done)
                              : pread(fd, buf + done, end - done, off + (off_t)done);
      if (r <= 0) {
        err[k] = 1;
        return;
      }
      done += (size_t)r;
    }
  };
  std::vector<std::thread> th;
  for (size_t k = 1; k < nparts; k++) th.emplace_back(run, k);
  run(0);
  for (auto &t : th) t.join();
  for (int e : err)
    if (e) throw std::runtime_error(std::string(what) + (write ? ": write failed" : ": short read"));
}

// device and pinned buffers of the two-stage pipeline, the checksum partials
// and a small auxiliary device table (naux words)
struct Pipe {
  hipStream_t st;
  uint32_t *dev[2] = {nullptr, nullptr};
  uint32_t *host[2] = {nullptr, nullptr};
  uint32_t *part = nullptr, *aux = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  Pipe(hipStream_t s, size_t chunk_bytes, long nblocks, size_t naux, const char *what) : st(s) {
    for (int b = 0; b < 2; b++) {
      if (hipMalloc((void **)&dev[b], chunk_bytes) != hipSuccess ||
          hipHostMalloc((void **)&host[b], chunk_bytes, hipHostMallocDefault) != hipSuccess ||
          hipEventCreateWithFlags(&ev[b], hipEventDisableTiming) != hipSuccess)
        throw std::runtime_error(std::string(what) + ": out of memory for the transfer buffers");
    }
    if (hipMalloc((void **)&part, (size_t)nblocks * 2 * sizeof(uint32_t)) != hipSuccess ||
        (naux && hipMalloc((void **)&aux, naux * sizeof(uint32_t)) != hipSuccess))
      throw std::runtime_error(std::string(what) + ": out of device memory");
  }
  ~Pipe() {
    (void)hipStreamSynchronize(st);  // nothing may still be reading or writing the buffers
    for (int b = 0; b < 2; b++) {
      if (dev[b]) (void)hipFree(dev[b]);
      if (host[b]) (void)hipHostFree(host[b]);
      if (ev[b]) (void)hipEventDestroy(ev[b]);
    }
    if (part) (void)hipFree(part);
    if (aux) (void)hipFree(aux);
  }
  Pipe(const Pipe &) = delete;
  Pipe &operator=(const Pipe &) = delete;

  // XOR of the nblocks partial pairs
  void checksum(long nblocks, uint32_t *a, uint32_t *b, const char *what) const {
    std::vector<uint32_t> h((size_t)nblocks * 2);
    if (hipMemcpy(h.data(), part, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error(std::string(what) + ": checksum copy failed");
    uint32_t x = 0, y = 0;
    for (long k = 0; k < nblocks; k++) {
      x ^= h[2 * k];
      y ^= h[2 * k + 1];
    }
    *a = x;
    *b = y;
  }
};

// whole t-slices per chunk
struct Chunking {
  long slab, per, nchunks, nblocks;  // sites per t-slice and per chunk, chunks, 64-site blocks in all
  size_t site_bytes;
  Chunking(const Geom &g, size_t sb) : site_bytes(sb) {
    slab = (long)g.L[0] * g.L[1] * g.L[2];
    long nt = (long)(CHUNK_BYTES / ((size_t)slab * sb));
    if (nt < 1) nt = 1;
    if (nt > g.L[3]) nt = g.L[3];
    per = nt * slab;
    nchunks = (g.L[3] + nt - 1) / nt;
    nblocks = 0;
    for (long c = 0; c < nchunks; c++) nblocks += blocks(g, c);
  }
  long sites(const Geom &g, long c) const {
    const long V = 2 * g.Vh, s = V - c * per;
    return s < per ? s : per;
  }
  long blocks(const Geom &g, long c) const { return (sites(g, c) + SITES_PER_BLOCK - 1) / SITES_PER_BLOCK; }
  size_t bytes(const Geom &g, long c) const { return (size_t)sites(g, c) * site_bytes; }
};

__device__ __forceinline__ uint32_t rotl32(uint32_t v, int k) { return k ? (v << k) | (v >> (32 - k)) : v; }

__device__ __forceinline__ uint32_t wave_xor(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v ^= (uint32_t)__shfl_xor((int)v, o);
  return v;
}

}  // namespace iop
}  // namespace qh
