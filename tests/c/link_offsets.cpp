// Host check of the link-load address arithmetic (qex_amd/csrc/common.h):
// for every link storage format, the address ld_link reads plane p of site i
// from (link_base + link_off, in the narrow or wide form link_wide picks) must
// be the byte the packer wrote it to (lplane_off / lrem_off, 64-bit), on both
// sides of the 4 GiB block size, at lattices whose FP64 link blocks pass it
// (96^4: 42 467 328 sites per parity) and at the largest lattice
// qexhip_lattice_create accepts.  Built for the host only
// (no GPU code, no device call); exits non-zero on the first mismatch.
#include <cstdint>
#include <cstdio>
#include "common.h"

using namespace qh;

static int fails = 0;

template <int PL, int RC, bool BLK>
static void check(long Vh, const char *name) {
  using F = LFmt<PL, RC>;
  const long stride = (Vh + 255) / 256 * 256;
  const char *blk = (const char *)(uintptr_t)0x10000;
  const long sites[] = {0, 1, 63, 64, 65, Vh / 2, Vh / 2 + 37, Vh - 65, Vh - 64, Vh - 1};
  for (long i : sites) {
    for (int p = 0; p <= F::nfull; p++) {
      if (p == F::nfull && F::rem == 0) continue;
      // the form ld_link takes at this size, and the narrow form wherever its
      // 32-bit offsets suffice
      const bool wide = link_wide<PL, RC, BLK>(stride);
      const uintptr_t got = wide ? (uintptr_t)(link_base<PL, RC, BLK, true>(blk, i) + link_off<PL, RC, BLK, true>(stride, i, p))
                                 : (uintptr_t)(link_base<PL, RC, BLK, false>(blk, i) + link_off<PL, RC, BLK, false>(stride, i, p));
      const long want_off = p < F::nfull ? lplane_off<PL, RC, BLK>(stride, i, p) : lrem_off<PL, RC, BLK>(stride, i);
      const uintptr_t want = (uintptr_t)blk + (uintptr_t)want_off;
      if (wide != ((unsigned long)stride * F::bytes > 0xffffffffUL && BLK)) {
        std::printf("WRONG FORM %s Vh=%ld: wide=%d\n", name, Vh, (int)wide);
        fails++;
      }
      if (got != want) {
        std::printf("MISMATCH %s Vh=%ld site %ld plane %d: read at +%lu, stored at +%ld\n", name, Vh, i, p,
                    (unsigned long)(got - (uintptr_t)blk), want_off);
        fails++;
      }
    }
  }
}

int main() {
  const long vh96 = 96L * 96 * 96 * 96 / 2;
  const long vhmax = 89478400;  // largest parity qexhip_lattice_create accepts
  // narrow (32^4, 48^4, and the last parity below 4 GiB of FP64 18-real links)
  // and wide sizes
  for (long Vh : {524288L, 2654208L, 29826048L, 29826304L, vh96, 33554432L /* 64^3 x 256 */, vhmax}) {
    check<P64, 18, true>(Vh, "fp64-18-blocked");
    check<P64, 12, true>(Vh, "fp64-12-blocked");
    check<P32, 18, true>(Vh, "fp32-18-blocked");
    check<P32, 12, true>(Vh, "fp32-12-blocked");
    check<P16, 18, false>(Vh, "fp16-18-planes");
    check<P16, 12, false>(Vh, "fp16-12-planes");
    check<PQ16, 18, false>(Vh, "q16-18-planes");
    check<PBF, 12, false>(Vh, "bf16-12-planes");
  }
  // the FP64 vector offset (ldc: 32-bit element offset x 16 B) of the last
  // component of the last site at the largest accepted parity
  const Geom g{{2, 2, 2, 2}, 1, vhmax, vhmax};
  const uint64_t last = (uint64_t)vix(g, 2, vhmax - 1) * 16u;
  if (last >= (1ull << 32) || (uint64_t)((uint32_t)vix(g, 2, vhmax - 1) * 16u) != last) {
    std::printf("MISMATCH fp64 vector offset %llu\n", (unsigned long long)last);
    fails++;
  }
  if (!fails) std::printf("link offsets ok\n");
  return fails ? 1 : 0;
}
