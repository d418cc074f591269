// QEX's own SIMD field layout on one rank <-> the library's canonical order,
// on the device, so that a QEX shim can upload its fields' raw storage and
// let the GPU reorder it (the role of the host repack loops in QEX's QUDA
// wrapper, quda/qudaWrapperImpl.nim:204-246).
//
// Layout (layout/qlayout.nim:10-150): outerGeom = L / innerGeom; a site x goes
// to index = (oi div 2) * nSitesInner + ii with lane
// ii = sum_d ((x_d div og_d) mod ig_d) * prod_{e<d} ig_e, outer position
// oi = sum_d (x_d mod og_d) * prod_{e<d} og_e (+ nSitesOuter for odd sites),
// and, when sum_d ((x_d div og_d) mod ig_d) * og_d is odd, the coordinate in
// direction innerCbDir shifted by innerCb (mod its outer extent) so that every
// SIMD vector holds sites of one parity.  Storage (field/fieldET.nim:18-22,
// maths/complexType.nim:11-13, maths/matrixConcept.nim:124-125): per outer
// site, for each complex component c, the real parts of all lanes, then the
// imaginary parts: double ((o * n + c) * 2 + r) * nSitesInner + l.
#include <stdexcept>

#include "internal.h"

namespace qh {

namespace {

struct QexLayout {
  int og[4], ig[4];
  int innerCb, innerCbDir, nin;
  long nout;
};

__device__ __forceinline__ long qex_index(const QexLayout &q, const int x[4]) {
  long ii = 0, oi = 0, mi = 1, mo = 1, mcb = 0;
  int ib = 0, p = 0, ocb = 0;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const int o = x[d] % q.og[d];
    const int k = (x[d] / q.og[d]) % q.ig[d];
    ii += k * mi;
    mi *= q.ig[d];
    if (d == q.innerCbDir) {
      ocb = o;
      mcb = mo;
    }
    oi += o * mo;
    mo *= q.og[d];
    ib += k * q.og[d];
    p += x[d];
  }
  if (ib & 1) oi += (long)((ocb + q.innerCb) % q.og[q.innerCbDir] - ocb) * mcb;
  if (p & 1) oi += q.nout;
  return (oi / 2) * q.nin + ii;
}

// canonical site idx -> coordinates
__device__ __forceinline__ void canon_coords(const Geom &g, long idx, int x[4]) {
  const int par = idx >= g.Vh ? 1 : 0;
  long lex = 2 * (idx - par * g.Vh);  // the even member of the site's lex pair {2j, 2j+1}
  long r = lex;
  int s = 0;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    x[d] = (int)(r % g.L[d]);
    r /= g.L[d];
    s += x[d];
  }
  if ((s & 1) != par) x[0] += 1;  // the other member: x0 + 1 (Lx even, same row)
}

// one thread per (canonical site, field): FROM = true reads QEX storage
template <bool FROM>
__global__ void __launch_bounds__(256)
    k_qex_layout(Geom g, QexLayout q, double *qex, int n, double2 *canon) {
  const long V = 2 * g.Vh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (idx >= V) return;
  int x[4];
  canon_coords(g, idx, x);
  const long j = qex_index(q, x);
  const long o = j / q.nin, l = j - o * q.nin;
  double *base = qex + (long)f * q.nout * n * 2 * q.nin + o * n * 2 * q.nin + l;
  double2 *c = canon + ((long)f * V + idx) * n;
  for (int e = 0; e < n; e++) {
    if (FROM) {
      c[e] = make_double2(base[(2 * e) * q.nin], base[(2 * e + 1) * q.nin]);
    } else {
      const double2 z = c[e];
      base[(2 * e) * q.nin] = z.x;
      base[(2 * e + 1) * q.nin] = z.y;
    }
  }
}

// layoutSetupQ (qlayout.nim:10-67) for one rank
QexLayout setup(const Geom &g, const int ig[4]) {
  QexLayout q;
  int icb = 0, icbd = -1;
  long nout = 1;
  int nin = 1;
  for (int d = 0; d < 4; d++) {
    if (ig[d] < 1 || g.L[d] % ig[d]) throw std::runtime_error("QEX layout: innerGeom must divide the lattice");
    // newLayout splits no direction into more than 2 SIMD parts (layout/layoutX.nim:124-131)
    if (ig[d] > 2) throw std::runtime_error("QEX layout: innerGeom entries above 2 are not QEX layouts");
    q.ig[d] = ig[d];
    q.og[d] = g.L[d] / ig[d];
    nout *= q.og[d];
    nin *= ig[d];
    if (ig[d] > 1 && (q.og[d] & 1)) icb++;
    if (ig[d] == 1 && !(q.og[d] & 1)) icbd = d;
  }
  if (icb == 0) {
    icbd = 0;
  } else {
    if (icbd < 0) throw std::runtime_error("QEX layout: not enough 2's in localGeom");
    icb = q.og[icbd] / 2;
    if (!(icb & 1)) throw std::runtime_error("QEX layout: error in cb choice");
  }
  q.innerCb = icb;
  q.innerCbDir = icbd;
  q.nin = nin;
  q.nout = nout;
  return q;
}

}  // namespace

void qex_layout_convert(const Geom &g, const int inner_geom[4], int from_qex, double *qex, int nfields, int ncomp,
                        double *canon, hipStream_t st) {
  if (nfields < 1 || nfields > 65535) throw std::runtime_error("QEX layout: nfields out of range");
  if (ncomp < 1 || ncomp > 64) throw std::runtime_error("QEX layout: ncomp out of range");
  const QexLayout q = setup(g, inner_geom);
  const long V = 2 * g.Vh;
  const dim3 grid((unsigned)((V + 255) / 256), (unsigned)nfields);
  if (from_qex)
    hipLaunchKernelGGL(k_qex_layout<true>, grid, dim3(256), 0, st, g, q, qex, ncomp, (double2 *)canon);
  else
    hipLaunchKernelGGL(k_qex_layout<false>, grid, dim3(256), 0, st, g, q, qex, ncomp, (double2 *)canon);
  if (hipGetLastError() != hipSuccess) throw std::runtime_error("QEX layout: kernel launch failed");
}

}  // namespace qh
