/* Plain-C caller of the qexhip ABI: the call sequence a QEX shim makes
 * (init, lattice, links, host-array parity solve, finalize), checked against
 * the free-field closed form.  With unit links (staggered phases, antiperiodic
 * t) an even-projected plane wave b is an eigenvector of 4(m^2 - D_eo D_oe)
 * with eigenvalue 4(m^2 + sum_mu sin^2 p_mu), so x = b / (4(m^2 + lambda)).
 * Prints "OK" and returns 0 on success. */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "qexhip.h"

static long cidx(const int L[4], const int c[4]) {
  long V = (long)L[0] * L[1] * L[2] * L[3];
  long lex = c[0] + (long)L[0] * (c[1] + (long)L[1] * (c[2] + (long)L[2] * c[3]));
  return (lex + ((c[0] + c[1] + c[2] + c[3]) & 1) * V) / 2;
}

#define CHECK(x)                                                         \
  do {                                                                   \
    if ((x) != 0) {                                                      \
      fprintf(stderr, "%s failed: %s\n", #x, qexhip_last_error());      \
      return 1;                                                          \
    }                                                                    \
  } while (0)

static int solve_check(qexhip_links *links, const int L[4], int sloppy, int slinks, double tol) {
  const long V = (long)L[0] * L[1] * L[2] * L[3];
  const double PI = 3.14159265358979323846;
  double p[4] = {2 * PI * 1 / L[0], 2 * PI * 2 / L[1], 2 * PI * 3 / L[2], PI * (2 * 1 + 1) / L[3]};
  double lam = 0, m = 0.2;
  for (int mu = 0; mu < 4; mu++) lam += sin(p[mu]) * sin(p[mu]);
  double complex *b = calloc(V * 3, sizeof(double complex)), *x = calloc(V * 3, sizeof(double complex));
  int c[4];
  for (c[3] = 0; c[3] < L[3]; c[3]++)
    for (c[2] = 0; c[2] < L[2]; c[2]++)
      for (c[1] = 0; c[1] < L[1]; c[1]++)
        for (c[0] = 0; c[0] < L[0]; c[0]++) {
          if ((c[0] + c[1] + c[2] + c[3]) & 1) continue;
          double ph = p[0] * c[0] + p[1] * c[1] + p[2] * c[2] + p[3] * c[3];
          b[3 * cidx(L, c) + 1] = cexp(I * ph);
        }
  qexhip_solver_params sp;
  memset(&sp, 0, sizeof sp);
  sp.r2req = 1e-24;
  sp.maxits = 2000;
  sp.sloppy = sloppy;
  sp.sloppy_links = slinks;
  sp.recon = 18;
  sp.reliable_delta = 0.1;
  sp.reliable_max_its = 100;
  CHECK(qexhip_solve_ee_host(links, (double *)x, (const double *)b, m, &sp));
  double err = 0, nrm = 0;
  for (long i = 0; i < 3 * V / 2; i++) {
    double complex want = b[i] / (4 * (m * m + lam));
    err += cabs(x[i] - want) * cabs(x[i] - want);
    nrm += cabs(want) * cabs(want);
  }
  printf("sloppy=%d links=%d: iterations %d, true r2 %.3e, rel err %.3e\n", sloppy, slinks, sp.iterations, sp.r2,
         sqrt(err / nrm));
  free(b);
  free(x);
  return (sqrt(err / nrm) < tol && sp.r2 <= 1e-24) ? 0 : 1;
}

int main(void) {
  const int L[4] = {8, 8, 8, 8};
  const long V = 4096;
  CHECK(qexhip_init(0));
  qexhip_lattice *lat;
  CHECK(qexhip_lattice_create(L, &lat));
  /* unit links with QEX's phases [8,9,11,0] and the antiperiodic t boundary */
  double *g = calloc(4 * V * 18, sizeof(double));
  int c[4];
  for (c[3] = 0; c[3] < L[3]; c[3]++)
    for (c[2] = 0; c[2] < L[2]; c[2]++)
      for (c[1] = 0; c[1] < L[1]; c[1]++)
        for (c[0] = 0; c[0] < L[0]; c[0]++) {
          const int e[4] = {c[3], c[0] + c[3], c[0] + c[1] + c[3], c[3] == L[3] - 1};
          for (int mu = 0; mu < 4; mu++) {
            double s = (e[mu] & 1) ? -1.0 : 1.0;
            double *u = g + ((long)mu * V + cidx(L, c)) * 18;
            u[0] = u[8] = u[16] = s;
          }
        }
  qexhip_links *links;
  CHECK(qexhip_links_create(lat, 4, g, 0, &links));
  int ok = 1;
  int rok = 0;
  CHECK(qexhip_links_recon_ok(links, &rok));
  ok &= (rok == 1);
  ok &= solve_check(links, L, QEXHIP_FP64, -1, 1e-10) == 0;
  ok &= solve_check(links, L, QEXHIP_FP16, QEXHIP_Q16, 1e-9) == 0;
  /* the error contract: odd extents are refused with a message */
  const int bad[4] = {3, 4, 4, 4};
  qexhip_lattice *lat2 = NULL;
  ok &= (qexhip_lattice_create(bad, &lat2) != 0) && strlen(qexhip_last_error()) > 0;
  CHECK(qexhip_links_destroy(links));
  CHECK(qexhip_lattice_destroy(lat));
  CHECK(qexhip_finalize());
  free(g);
  printf(ok ? "OK\n" : "FAILED\n");
  return ok ? 0 : 1;
}
