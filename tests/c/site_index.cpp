// Host check of the site and neighbour index arithmetic of the stencils
// (decode and nbr, qex_amd/csrc/common.h) against coordinates computed the
// plain way: for every site of both parities of several lattices with
// unequal extents, decode must return the site's coordinates and nbr, for
// hops of 1 and 3 sites in each direction, the checkerboard index of the
// shifted site (QEX's canonical index floor(lex/2), layout/qlayout.nim:108-136).  Built for the host only
// (no GPU code, no device call); exits non-zero on the first mismatch.
#include <cstdint>
#include <cstdio>
#include "common.h"

using namespace qh;

static int fails = 0;

static void check(int Lx, int Ly, int Lz, int Lt) {
  Geom g;
  g.L[0] = Lx; g.L[1] = Ly; g.L[2] = Lz; g.L[3] = Lt;
  g.hx = Lx / 2;
  g.Vh = (long)Lx * Ly * Lz * Lt / 2;
  g.stride = (g.Vh + 255) / 256 * 256;
  const int L[4] = {Lx, Ly, Lz, Lt};
  int minL = Lx;
  for (int d = 1; d < 4; d++) minL = L[d] < minL ? L[d] : minL;
  for (int t = 0; t < Lt; t++)
    for (int z = 0; z < Lz; z++)
      for (int y = 0; y < Ly; y++)
        for (int x = 0; x < Lx; x++) {
          const int par = (x + y + z + t) & 1;
          const long lex = x + (long)Lx * (y + (long)Ly * (z + (long)Lz * t));
          const long i = lex / 2;
          const Site s = decode(g, i, par);
          if (s.x != x || s.y != y || s.z != z || s.t != t || s.row != lex / Lx) {
            std::printf("DECODE %dx%dx%dx%d site %ld: got (%d,%d,%d,%d)\n", Lx, Ly, Lz, Lt, i, s.x, s.y, s.z, s.t);
            if (++fails > 10) return;
          }
          for (int mu = 0; mu < 4; mu++)
            for (int d : {1, -1, 3, -3}) {
              if (d == 3 || d == -3) {
                if (minL < 4) continue;  // 3-hop links need extents >= 4
              }
              int c[4] = {x, y, z, t};
              c[mu] = ((c[mu] + d) % L[mu] + L[mu]) % L[mu];
              const long nlex = c[0] + (long)Lx * (c[1] + (long)Ly * (c[2] + (long)Lz * c[3]));
              const long got = nbr(g, s, mu, d);
              if (got != nlex / 2) {
                std::printf("NBR %dx%dx%dx%d site %ld mu %d d %d: got %ld want %ld\n", Lx, Ly, Lz, Lt, i, mu, d, got,
                            nlex / 2);
                if (++fails > 10) return;
              }
            }
        }
}

int main() {
  check(4, 4, 4, 4);
  check(8, 8, 4, 8);
  check(8, 4, 6, 4);
  check(2, 2, 2, 2);
  check(2, 4, 2, 6);
  check(6, 4, 8, 4);
  check(12, 4, 4, 6);
  check(16, 8, 8, 4);
  if (fails) {
    std::printf("%d mismatches\n", fails);
    return 1;
  }
  std::printf("site indices ok\n");
  return 0;
}
