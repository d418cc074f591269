/*
 * qexoracle.h -- CPU restatement of QEX's even-odd staggered Dslash, CG,
 * multi-shift CG and staggered fermion force, in plain C (C99 + OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity reference for the HIP
 * implementation in qex_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * Conventions (all restated from the reference, file:line into QEX's src/):
 *  - Lattice L[4] = (Lx,Ly,Lz,Lt), all even.  lex = x + Lx*(y + Ly*(z + Lz*t)).
 *    Canonical site index idx = (lex + p*V)/2 with p = (x+y+z+t)&1: even sites
 *    first, then odd (layout/qlayout.nim:108-136 with innerGeom=1).
 *  - Color vector: 3 complex = 6 doubles (re,im) per site.
 *  - Link field k: V x 3x3 complex row-major = 18 doubles per site.
 *    nlinks = 4 (plain, g[mu]) or 8 (HISQ: g[2mu]=fat hop 1, g[2mu+1]=long
 *    hop 3; physics/stagD.nim:40-51,603-615).
 *  - Links are passed PHASED and BC-applied, as QEX's Staggered object holds
 *    them (physics/stagD.nim:560-571, gauge/gaugeUtils.nim:133-140).
 *  - Parity argument: 0 even, 1 odd, 2 all.
 */
#ifndef QEXORACLE_H
#define QEXORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- layout ---- */
long qo_volume(const int L[4]);
long qo_index(const int L[4], const int x[4]);
void qo_coords(const int L[4], long idx, int x[4]);

/* ---- counter-based generator (synthetic inputs) ---- */
void qo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* two N(0,1) deviates from counter (c0,c1,c2,c3) and 64-bit seed */
void qo_gauss2(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
               double g[2]);
/* U_mu(x) = exp(i*0.5*H), H traceless Hermitian, unit Gaussian entries */
void qo_gauge_random(const int L[4], uint64_t seed, double *U /*4*V*18*/);
/* identity links */
void qo_gauge_unit(const int L[4], double *U);
/* complex Gaussian color vectors, E|z|^2=1 per component, on parity par */
void qo_source_gauss(const int L[4], uint64_t seed, int stream, int par, double *x /*V*6*/);
/* antiperiodic t boundary: flip g[3] on t=Lt-1 (gauge/gaugeUtils.nim:133-140) */
void qo_set_bc(const int L[4], double *U);
/* staggered phases, masks [8,9,11,0] (physics/stagD.nim:560-571) */
void qo_stag_phase(const int L[4], double *U);
/* HISQ-form links from phased U: W = 0.8U + 0.05G, Lg = -(1/24) U U U,
   interleaved g8 = [W0,L0,W1,L1,W2,L2,W3,L3] (stagD.nim:603-615, fat7l.nim:145-154) */
void qo_hisq_links(const int L[4], uint64_t seed, const double *Uph, double *g8);

/* ---- operator ---- */
/* r = a*r + b*x + sum_k [g_k(x) x(x+d_k) - g_k(x-d_k)^+ x(x-d_k)] on parity par
   (physics/stagD.nim:383-448) */
void qo_stagD2(const int L[4], int nl, const double *g, int par,
               double *r, const double *x, double a, double b);
/* r = a*r + m*x + sc*D*x with D = (1/2)(2D)  (stagD.nim:459-462) */
void qo_stagD(const int L[4], int nl, const double *g, int par,
              double *r, const double *x, double m, double sc, double a);
/* r_p = 4*m2*x_p - (2D)_{p,q} (2D)_{q,p} x_p  (stagD.nim:485-520) */
void qo_stagD2xx(const int L[4], int nl, const double *g, int par,
                 double *r, const double *x, double m2, double *tscratch);
/* D(m) = m + D and D^+(m) = m - D on all sites (stagD.nim:617-622) */
void qo_D(const int L[4], int nl, const double *g, double *r, const double *x, double m);
void qo_Ddag(const int L[4], int nl, const double *g, double *r, const double *x, double m);

/* ---- reductions (FP64, fieldET.nim:648-778) ---- */
double qo_norm2(long nsites, const double *x);
double qo_redot(long nsites, const double *x, const double *y);

/* ---- solvers ---- */
/* CG (cpNone) on A = 4(m^2 - D_pq D_qp) (solvers/cg.nim:55-272), x starts at
   the given value; returns iterations; *r2out = recursive r2. */
int qo_cg(const int L[4], int nl, const double *g, int par,
          double *x, const double *b, double m, double r2req, int maxits,
          double *r2out);
/* solveXX outer defect-correction loop (physics/stagSolve.nim:330-421) */
int qo_solveXX(const int L[4], int nl, const double *g, int par,
               double *r, const double *x, double m, double r2req, int maxits,
               double *true_r2_rel);
/* multi-shift CG (solvers/cgm.nim:68-246): solves (A + sigma_k) x_k = b,
   sigma_0 = 0 for the base system; xs = nshift fields of V*6 doubles */
int qo_cgm(const int L[4], int nl, const double *g, int par,
           double *xs, const double *b, double m, const double *sigmas, int nshift,
           double r2req, int maxits, double *r2out);

/* ---- force ---- */
/* F_mu(x) = TAH(eps(x) psi(x) psi(x+mu)^+ U_mu(x)^+) with phased U
   (hmc/hmcAction.nim:723-792, unsmeared); F: 4*V*18 doubles */
void qo_ferm_force(const int L[4], const double *U, const double *psi, double *F);
/* outer product only: G_mu(x) = eps(x) psi(x) psi(x+mu)^+  */
void qo_ferm_outer(const int L[4], const double *psi, double *G);
/* psi = [Y_e ; (2D)_oe Y_e] with Y_e = (m^2 - D_eo D_oe)^{-1} phi_e
   (hmc/hmcAction.nim:669-721) */
int qo_force_solve(const int L[4], int nl, const double *g, const double *phi,
                   double m, double r2req, int maxits, double *psi);

/* ---- gauge field pieces of the staggered HMC ---- */
/* sum_{x,mu<nu} Re tr U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ ; QEX's Wilson
   action is -(beta/3) times this (gauge/gaugeAction.nim:397-427) */
double qo_plaquette(const int L[4], const double *U);
/* F = (2 beta/3) TAH(U * staples): dS/dw = -1/2 Re tr(X F) for U -> exp(wX)U */
void qo_wilson_force(const int L[4], const double *U, double beta, double *F);
/* P = i H, H traceless Hermitian, density exp(-tr H^2/2) (gaugeUtils.nim:1474-1500) */
void qo_random_momenta(const int L[4], uint64_t seed, double *P);
/* U <- exp(eps P) U (gaugeUtils.nim:453-475) */
void qo_link_update(const int L[4], const double *P, double eps, double *U);

/* ---- HISQ link construction (gauge/fat7l.nim, physics/hisqLinks.nim) ---- */
/* makeImpLinks: coef = {oneLink, threeStaple, fiveStaple, sevenStaple, lepage};
   fl, gf: 4*V*18; ll = naik * gl gl gl (gl = gf when NULL), skipped when ll is
   NULL or naik == 0 */
void qo_make_imp_links(const int L[4], const double *gf, const double coef[5], const double *gl, double naik,
                       double *fl, double *ll);
/* u = x (x^+ x + 1e-20)^{-1/2} for n 3x3 matrices (projectU, Cayley-Hamilton) */
void qo_project_u(long n, const double *x, double *u);
/* HisqCoefs.smear: fat7(first) -> projectU -> fat7(second) + Naik */
void qo_hisq_smear(const int L[4], const double first[5], const double second[5], double naik, const double *g,
                   double *fl, double *ll);

/* ---- HISQ smearing chain rule (derivatives in the redot pairing Re tr(A^+ B)) ---- */
/* fatl7Deriv: deriv += dS/dgf given mid = dS/dfl; llderiv += dS/dgl given llmid = dS/dll
   (Naik part skipped when llderiv or llmid is NULL or naik == 0; gl = gf when NULL) */
void qo_fat7l_deriv(const int L[4], const double *gf, const double *mid, const double coef[5], double *deriv,
                    const double *gl, const double *llmid, double naik, double *llderiv);
/* projectUderiv on n matrices: r = dS/dx given u = projectU(x) and chain = dS/du */
void qo_project_u_deriv(long n, const double *u, const double *x, const double *chain, double *r);
/* smearGetForce: dg += dS/dg given dfl = dS/dfl, dll = dS/dll for (fl, ll) = smear(g) */
void qo_hisq_smear_deriv(const int L[4], const double first[5], const double second[5], double naik, const double *g,
                         const double *dfl, const double *dll, double *dg);
/* HISQ force outer products: G[2mu] = eps psi psi(x+mu)^+, G[2mu+1] = eps psi psi(x+3mu)^+ */
void qo_hisq_outer(const int L[4], const double *psi, double *G);

/* ---- MILC gauge files (src/new/milcIO.nim) ---- */
void qo_milc_checksum(const uint32_t *w, long n, long offset, uint32_t out[2]);

/* ---- SciDAC/ILDG gauge and field files (src/io/readerQiolite.nim, writerQiolite.nim) ---- */
/* CRC-32 of n bytes, reflected polynomial 0xedb88320, initial ~0, final
   complement (src/io/crc32.nim:1-40) */
uint32_t qo_crc32(const unsigned char *p, long n);
/* SciDAC record checksum over nsites sites of site_bytes file bytes each, site
   k having global lexicographic rank rank0 + k: out[0] = suma, out[1] = sumb */
void qo_scidac_checksum(const unsigned char *data, long nsites, long site_bytes, long rank0, uint32_t out[2]);

/* ---- 3x3 helpers used by tests ---- */
void qo_su3_exp_i(const double *H /*18, Hermitian*/, double s, double *U /*18*/);
void qo_su3_exp(const double *A /*18, general*/, double *U /*18*/);
void qo_mat_mul(const double *a, const double *b, double *c);
void qo_project_tah(const double *x, double *r);

#ifdef __cplusplus
}
#endif
#endif
