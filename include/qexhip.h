/*
 * qexhip.h -- C ABI of the MI355X (gfx950) staggered Dslash / CG / multi-shift
 * CG / fermion-force library.  This is the drop-in seam for QEX's solver
 * backend switch (QEX src/solvers/solverBase.nim:6-7, dispatched in
 * src/physics/stagSolve.nim:98-157), in the same place QEX plugs in QUDA
 * (src/quda/qudaWrapperImpl.nim:171-273) and Grid (src/grid/gridStag.nim:5-83).
 *
 * Conventions (identical to QEX's):
 *  - Site order: QEX's canonical single-rank Layout[1] order, even sites first:
 *    idx = (lex + p*V)/2, lex = x + Lx*(y + Ly*(z + Lz*t)), p = (x+y+z+t)&1
 *    (src/layout/qlayout.nim:108-136).  All lattice extents must be even.
 *  - Host "canonical" arrays: color-vector fields are V x 3 complex doubles
 *    (re,im interleaved) = 6 doubles/site; link fields are nl x V x 3x3
 *    complex doubles, row-major = 18 doubles/site/link.
 *  - Links are PHASED and BC-applied, exactly as QEX's Staggered.g holds them
 *    (src/physics/stagD.nim:560-571, src/gauge/gaugeUtils.nim:133-140).
 *    nl = 4: plain, g[mu];  nl = 8: HISQ form, g[2mu] fat (1 hop),
 *    g[2mu+1] long (3 hops) (src/physics/stagD.nim:40-51,603-615).
 *  - The operator solved by the parity solvers is QEX's
 *    A = 4(m^2 - D_pq D_qp) (stagD2ee/stagD2oo, src/physics/stagD.nim:485-520).
 *  - Parity: 0 even, 1 odd, 2 all.
 *
 * Error contract: every function returns 0 on success and a non-zero code on
 * failure; qexhip_last_error() returns the message (thread-local).  QEX's shim
 * maps a non-zero return to qexError (src/base/qexInternal.nim:39-48).
 * Calls are synchronous with respect to the host (results are complete on
 * return) unless stated otherwise, and not re-entrant, like QEX's backend
 * calls (made on the master thread outside threads:, stagSolve.nim:396).
 * Device pointers passed in must hold completed data.
 */
#ifndef QEXHIP_H
#define QEXHIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct qexhip_lattice qexhip_lattice;
typedef struct qexhip_links qexhip_links;
typedef struct qexhip_field qexhip_field;

/* storage precisions.  QEXHIP_Q16 (links only): 16-bit fixed point with one
   scale per link field, value = int16 * max|entry|/32767; used with FP16
   color vectors. */
enum { QEXHIP_FP64 = 0, QEXHIP_FP32 = 1, QEXHIP_FP16 = 2, QEXHIP_BF16 = 3, QEXHIP_Q16 = 4 };
/* solver_params.sloppy only: FP16 storage for every sloppy vector, the residual
   and the solution increment included (QEXHIP_FP16 keeps those two in FP32) */
enum { QEXHIP_FP16_ALL = 5 };
enum { QEXHIP_EVEN = 0, QEXHIP_ODD = 1, QEXHIP_ALL = 2 };

/* Solver parameters: inputs and outputs mirror QEX's SolverParams
   (src/solvers/solverBase.nim:10-28). */
typedef struct {
  /* inputs */
  double r2req;           /* relative squared residual |b-Ax|^2/|b|^2 (QEX r2req) */
  int maxits;             /* iteration cap (QEX maxits) */
  int sloppy;             /* QEXHIP_FP64: plain FP64 CG (SloppyNone); FP32/FP16/BF16:
                             FP32-arithmetic inner CG with this spinor/link storage and
                             FP64 reliable updates (QEX sloppySolve, solverBase.nim:8-9) */
  int sloppy_links;       /* link storage of the inner operator; -1 = same as sloppy;
                             QEXHIP_Q16 needs sloppy = FP16 or FP32 */
  int recon;              /* inner-operator links: 18 reals, or 12 (rows 0-1 stored, row 2
                             rebuilt; plain phased-SU(3) links only); 0 = 18 */
  double reliable_delta;  /* reliable update when |r| < delta*|r_last_update| (0.1) */
  int reliable_max_its;   /* ... or after this many iterations (100) */
  int use_prev_soln;      /* 1: the solution field holds the initial guess (QEX usePrevSoln) */
  int verbosity;          /* QEX verbosity */
  /* outputs */
  int iterations;         /* QEX iterations */
  int reliable_updates;   /* FP64 reliable updates performed (mixed mode) */
  double r2;              /* final FP64 true |b-Ax|^2/|b|^2 (multi-shift: of the base system
                             sigma_0 = 0, recomputed from the returned solution) */
  double seconds;         /* wall time of the call */
  double flops;           /* QEX flop count (nl*4*72+60)*nEven*its (stagSolve.nim:121) */
  /* multi-shift solves only: NULL, or an array of nshift doubles that receives
     each shifted system's FP64 true |b-(A+sigma_k)x_k|^2/|b|^2 (the residuals
     cgm.nim:208-239 prints at verbosity 3); the caller owns it */
  double *shift_r2;
} qexhip_solver_params;

/* ---- library lifetime (QEX qexGlobalInitializers/Finalizers pattern,
        src/quda/qudaWrapperImpl.nim:71-92) ---- */
/* replaces qudaInit (qudaWrapperImpl.nim:71-84): selects the GPU, creates the stream */
int qexhip_init(int device);
/* replaces qudaFini (qudaWrapperImpl.nim:86-89) */
int qexhip_finalize(void);
const char *qexhip_last_error(void);
/* wait for all work queued by the library */
int qexhip_sync(void);
/* the library's HIP stream (hipStream_t), for callers that order work against it */
void *qexhip_stream(void);

/* ---- lattice (replaces qudaSetup's layout, qudaWrapperImpl.nim:94-127) ---- */
int qexhip_lattice_create(const int dims[4], qexhip_lattice **out);
int qexhip_lattice_destroy(qexhip_lattice *lat);

/* ---- links: device-resident copy of Staggered.g (newStag/newStag3,
        src/physics/stagD.nim:573-615; repack as qudaWrapperImpl.nim:222-249) ---- */
int qexhip_links_create(qexhip_lattice *lat, int nl, const double *canon, int canon_on_device,
                        qexhip_links **out);
/* new gauge generation in place (same nl) */
int qexhip_links_update(qexhip_links *l, const double *canon, int canon_on_device);
int qexhip_links_destroy(qexhip_links *l);

/* ---- color-vector fields (QEX Field[V, ColorVector], src/field/fieldET.nim:17-22) ---- */
int qexhip_field_create(qexhip_lattice *lat, int prec, qexhip_field **out);
int qexhip_field_destroy(qexhip_field *f);
/* canonical V x 6 doubles; only the sites of parity par are read / written */
int qexhip_field_import(qexhip_field *f, int par, const double *canon, int canon_on_device);
int qexhip_field_export(const qexhip_field *f, int par, double *canon, int canon_on_device);
/* FP64-accumulated norm2 / redot on a parity (fieldET.nim:648-778) */
int qexhip_field_norm2(const qexhip_field *f, int par, double *out);
int qexhip_field_redot(const qexhip_field *x, const qexhip_field *y, int par, double *out);
/* y = a*x + b*y on a parity (QEX field expressions such as r.odd += b/m,
   stagD.nim:633-637); FP64 fields */
int qexhip_field_axpby(const qexhip_field *x, qexhip_field *y, int par, double a, double b);

/* ---- operator (src/physics/stagD.nim) ----
   All fields of one call share one precision; arithmetic is FP64 for FP64
   fields and FP32 otherwise. */
/* stagD2 (stagD.nim:383-448): r = a*r + b*x + (2D)x on parity par */
int qexhip_stagD2(qexhip_links *l, int par, qexhip_field *r, const qexhip_field *x, double a, double b);
/* stagD (stagD.nim:459-462): r = a*r + m*x + sc*D*x, D = (2D)/2 */
int qexhip_stagD(qexhip_links *l, int par, qexhip_field *r, const qexhip_field *x, double m, double sc,
                 double a);
/* stagD2ee / stagD2oo (stagD.nim:485-520): r_p = 4*m2*x_p - (2D)_pq (2D)_qp x_p */
int qexhip_stagD2xx(qexhip_links *l, int par, qexhip_field *r, const qexhip_field *x, double m2);
/* No counterpart in QEX: qexhip_stagD2xx for callers that cannot allocate
   device memory (code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc).  Same operator, arguments and result as
   qexhip_stagD2xx; it calls no hipMalloc/hipFree and no allocating routine:
   the odd-parity temporary lives in a static 128 MiB device buffer that each
   GPU holds for the life of the process (lattices up to 2 796 032 sites per
   parity in FP64, 5 592 064 in FP32; larger ones are refused with an error),
   the launch shape is the tuner's earlier choice or the default (no tuning
   run), and the links must already hold a copy in the fields' precision
   (FP64 always; FP32 after any FP32 operator call or solve on them, else an
   error).  FP64 or FP32 fields.  All calls on one GPU share the buffer:
   calls that overlap on one GPU (other host threads, other streams) race. */
int qexhip_stagD2xx_noalloc(qexhip_links *l, int par, qexhip_field *r, const qexhip_field *x, double m2);
/* multi-RHS stagD2 (stagDPN/stagDN, stagD.nim:132-199,451-456): r_j = a*r_j + b*x_j + (2D)x_j
   for j < n on parity par, each link loaded once for up to 8 fields; FP64 or FP32 fields.
   QEX's stagDN(sd, r, g, x, m) is a = 0, b = m. */
int qexhip_stagD2_multi(qexhip_links *l, int par, int n, qexhip_field *const *r, const qexhip_field *const *x,
                        double a, double b);

/* ---- solvers ---- */
/* solveEE / solveOO (src/physics/stagSolve.nim:330-427): solve A r = x on
   parity par with A = 4(m^2 - D_pq D_qp); FP64 fields.  Outer defect-
   correction loop as in solveXX; inner solve per sp->sloppy. */
int qexhip_solve_xx(qexhip_links *l, int par, qexhip_field *r, const qexhip_field *x, double m,
                    qexhip_solver_params *sp);
/* QUDA-adapter-shaped host entry (qudaSolveEE/qudaSolveOO,
   qudaWrapperImpl.nim:171-273): canonical host arrays of V x 6 doubles. */
int qexhip_solve_ee_host(qexhip_links *l, double *r, const double *t, double m, qexhip_solver_params *sp);
int qexhip_solve_oo_host(qexhip_links *l, double *r, const double *t, double m, qexhip_solver_params *sp);
/* multi-shift CG on a parity (solveXXInternal multi-shift, stagSolve.nim:164-229,
   and CgmState.solve, src/solvers/cgm.nim:68-246): solves
   (A + sigma_k) x_k = b, k = 0..nshift-1, with sigma_0 treated as 0.
   sp->sloppy == QEXHIP_FP64: QEX's FP64 recursion, every shift updated until
   the base system converges; the shifted x_k, p_k are advanced once per 16
   iterations from that window's residuals (the same sums, reassociated; the
   batched solve's machinery with one right-hand side).  Otherwise (no QEX
   counterpart) the recursion runs on FP32 storage until |r|^2 <=
   max(r2req, 1e-12) |b|^2 (QEXHIP_MS_SLOPPY_R2 sets the 1e-12), and each shift is then
   refined to r2req by defect correction around the mixed CG of
   qexhip_solve_xx (sp's sloppy storage); iterations count both phases.
   sp->r2 and sp->shift_r2 are FP64 true residuals either way. */
int qexhip_solve_multishift(qexhip_links *l, int par, qexhip_field *const *xs, const qexhip_field *b, double m,
                            const double *sigmas, int nshift, qexhip_solver_params *sp);

/* ---- batched solves (no single QEX counterpart: each system is one QEX solve,
        run together on one GPU so that one launch per solver stage serves all
        of them).  flags: QEXHIP_BATCH_GRAPH replays the iterations from a HIP
        graph; QEXHIP_BATCH_TUNE picks the launch shape by timing candidates
        (cached per shape); QEXHIP_BATCH_NOFREEZE keeps updating shifted
        systems that already converged (QEX's cgm.nim behaviour). ---- */
#define QEXHIP_BATCH_GRAPH 1
#define QEXHIP_BATCH_TUNE 2
#define QEXHIP_BATCH_NOFREEZE 4
/* n independent FP64 CG solves A_k x_k = b_k, A_k = 4(m_k^2 - D_pq D_qp) with
   links[k] (lattices may differ; all plain or all HISQ form), each as
   qexhip_solve_xx's inner CG from x = 0 (solvers/cg.nim:55-272); sp[k] gives
   r2req and maxits and receives iterations and the FP64 true residual. */
int qexhip_solve_batch(int n, qexhip_links *const *links, int par, qexhip_field *const *x,
                       qexhip_field *const *b, const double *masses, qexhip_solver_params *sp, int flags);
/* multi-shift CG (src/solvers/cgm.nim:68-246) for nrhs right-hand sides on one
   set of links: (A + sigma_k) x[j*nshift+k] = b[j], sigma_0 treated as 0; each
   link load serves up to 8 right-hand sides.  sp[j] per right-hand side. */
int qexhip_solve_multishift_batch(qexhip_links *l, int par, int nrhs, qexhip_field *const *xs,
                                  qexhip_field *const *b, double m, const double *sigmas, int nshift,
                                  qexhip_solver_params *sp, int flags);
/* launch-shape autotuner (no QEX counterpart): on/off (default on unless
   QEXHIP_TUNE=0 or QEXHIP_BLOCK is set), and a text table of the shapes chosen */
int qexhip_set_autotune(int on);
int qexhip_autotune_report(char *buf, int len);
/* CG iteration form: 2 = two stencil kernels forming p on the fly, 3 = a
   streaming update kernel plus two stencil kernels, 4 = form 3 with the dot-
   product finish folded into the next iteration's streaming kernel (single
   systems; batches take 3); 0 = chosen by the tuner */
int qexhip_set_cg_form(int form);
/* link bytes that the stencils of lattice lat read with cache-allocating
   loads, no QEX counterpart: the first link blocks (one direction and parity
   each) that fit in this many bytes are read so that they can stay in the
   last-level cache from one sweep to the next; the rest, with the
   non-temporal hint.  0 = all non-temporal; -1 = chosen by the tuner per
   kernel family and link format (the default).  With
   lat = NULL, sets the default for lattices created afterwards (-1 restores
   the built-in one) */
int qexhip_set_link_cache(qexhip_lattice *lat, long bytes);
/* site walk of the one-parity hop of every operator and solver call, no QEX
   counterpart: 0 = canonical order, a multiple of 256 = t-slices cut into
   chunks of that many sites and walked t first (their t +- 1 neighbours are
   read again while still in L2), -1 = chosen by the tuner */
int qexhip_set_hop_walk(long chunk);
/* t-march form of the one-parity hop of plain links (operator calls and the
   solvers' FP64 residuals), no QEX counterpart: each thread walks this many
   t-slices of one half-site index and keeps the forward t-neighbour it loaded
   for the next slice but one, where it is the backward one; 0 = off (the site
   walk of qexhip_set_hop_walk applies), -1 = chosen by the tuner */
int qexhip_set_hop_tmarch(int slices);
/* stencil loop of the one-parity hop, no QEX counterpart: 1 = both links of a
   hop per trip (more loads in flight per wave, for lattices too small to fill
   the device), 0 = one link per trip, -1 = chosen by the tuner */
int qexhip_set_hop_two(int two);
/* threads per output site of the one-parity hop, no QEX counterpart: 2 or 4
   split each site's links between that many threads, whose sums are added by
   lane shuffles (more waves for lattices too small to fill the device); 0 or
   1 = one thread per site; -1 = chosen by the tuner */
int qexhip_set_hop_split(int split);
/* site walk of the two multi-RHS hops of the batched multi-shift solve
   (qexhip_solve_multishift_batch), no QEX counterpart: same values as
   qexhip_set_hop_walk; -1 = chosen by the tuner (with QEXHIP_BATCH_TUNE, else
   the canonical order) */
int qexhip_set_msb_walk(long chunk);
/* multi-RHS hop (qexhip_stagD2_multi) launch shape, no QEX counterpart: the
   batches in which a thread loads its right-hand sides' vectors per link (1
   or 2; 0 = chosen by the tuner) and the site walk (0: canonical order; a
   multiple of 256: t-slices cut into chunks of that many sites, walked t
   first; -1 = chosen by the tuner) */
int qexhip_set_hop_multi_shape(int batches, long chunk);
/* block size of the other-parity stencil kernel of CG forms 3 and 4 (64,
   128 or 256; 0 = chosen by the tuner), independent of the other kernels' */
int qexhip_set_cg_block1(int block1);
/* workgroups of the persistent kernels of CG forms 3 and 4, no QEX
   counterpart: form 3 runs both stencil kernels on at most this many
   workgroups (0 = one workgroup per block), form 4 its streaming kernel (0 =
   512); -1 = chosen by the tuner */
int qexhip_set_cg_cap(int cap);
/* CG launches: 1 (default) replays HIP graphs, in the mixed-precision CG one per
   reliable-update interval (its iterations, the FP64 reliable update and the
   best-iterate copy); 2 replays graphs of 16 iterations with the reliable
   updates launched from the host; 0 launches kernel by kernel
   (QEXHIP_CG_GRAPH=0 and QEXHIP_CG_INTERVALS=0 at init select 0 and 2) */
int qexhip_set_cg_graph(int on);

/* ---- fermion force (src/hmc/hmcAction.nim:669-792) ---- */
/* stagForceSolve (hmcAction.nim:669-721): psi_e = (m^2 - D_eo D_oe)^-1 phi_e,
   psi_o = (2D)_oe psi_e */
int qexhip_force_solve(qexhip_links *l, qexhip_field *psi, const qexhip_field *phi, double m,
                       qexhip_solver_params *sp);
/* the outer-product accumulation of stagD2deriv / stagDeriv (stagD.nim:640-715,
   matrixConcept.nim:718-725): for every site y and link k, F_k(y) += c a(y)
   b(y + d_k)^+, c = ce on even and co on odd sites, d_k one site in
   direction mu (three for k = 2mu+1 when nl = 8); F is canonical
   [nl][site][3][3] FP64 on the device and is accumulated into */
int qexhip_outer_acc(qexhip_links *l, const qexhip_field *a, const qexhip_field *b, double ce, double co, double *F);
/* fermForce, unsmeared (hmcAction.nim:723-792): F_mu(x) = TAH(eps(x) psi(x) psi(x+mu)^+ U_mu(x)^+);
   outer_only=1 returns eps(x) psi(x) psi(x+mu)^+ instead.  F: canonical 4 x V x 18 doubles. */
int qexhip_force(qexhip_links *l, const qexhip_field *psi, double *F, int F_on_device, int outer_only);

/* ---- gauge field for HMC (FP64 canonical arrays U[4][V][3][3] in device
        memory; the pieces a staggered HMC needs around the fermion force) ---- */
/* sum_{x,mu<nu} Re tr U_P; QEX's Wilson action is -(beta/3) times this
   (gauge/gaugeAction.nim:397-427) */
int qexhip_plaquette(qexhip_lattice *lat, const double *U_dev, double *sum);
/* F_mu(x) = (2 beta/3) TAH(U_mu(x) * staples), so that dS_W/dw = -1/2 Re tr(X F) for
   U -> exp(wX)U, the convention of qexhip_force (gauge/gaugeAction.nim:765-781) */
int qexhip_wilson_force(qexhip_lattice *lat, const double *U_dev, double beta, double *F_dev);
/* traceless anti-Hermitian Gaussian momenta, density exp(tr P^2/2)
   (gauge/gaugeUtils.nim:1474-1500) */
int qexhip_random_momenta(qexhip_lattice *lat, uint64_t seed, double *P_dev);
/* U <- exp(eps P) U (gauge/gaugeUtils.nim:453-475) */
int qexhip_link_update(qexhip_lattice *lat, const double *P_dev, double eps, double *U_dev);

/* ---- synthetic inputs (no QEX counterpart; QEX's random fields use its own
        RNG, src/rng).  Outputs are canonical FP64 arrays in device memory. ---- */
/* ---- HISQ link construction (device arrays, canonical [mu][site][3][3] complex) ---- */

/* Fat7lCoefs, gauge/fat7l.nim:5-11 */
typedef struct {
  double one_link, three_staple, five_staple, seven_staple, lepage;
} qexhip_fat7_coefs;
/* HisqCoefs without the projection policy (Cayley-Hamilton), physics/hisqLinks.nim:4-8 */
typedef struct {
  qexhip_fat7_coefs fat7first, fat7second;
  double naik;
} qexhip_hisq_coefs;
/* HisqCoefs.init defaults, physics/hisqLinks.nim:17-28 (setHisqFat7 :10-15) */
void qexhip_hisq_coefs_init(qexhip_hisq_coefs *c);
/* makeImpLinks, gauge/fat7l.nim:71-159: fl = fat7 path sum of gf; when ll is not
   NULL and naik != 0, ll_mu(x) = naik gl_mu(x) gl_mu(x+mu) gl_mu(x+2mu) with
   gl = gf_long, or gf when gf_long is NULL.  Outputs must not alias gf. */
int qexhip_make_imp_links(qexhip_lattice *lat, double *fl_dev, const double *gf_dev, const qexhip_fat7_coefs *coef,
                          double *ll_dev, const double *gf_long_dev, double naik);
/* projectU, maths/matrixFunctions.nim:337-352: u = x (x^+ x + 1e-20)^{-1/2} on n 3x3 matrices */
int qexhip_project_u(long n, double *u_dev, const double *x_dev);
/* HisqCoefs.smear, physics/hisqLinks.nim:39-51: fat7(first) -> projectU ->
   fat7(second) and Naik long links; c = NULL takes the defaults.  g is the
   phased, BC-applied gauge field (tests/examples/testStagProp.nim:24-33). */
int qexhip_hisq_smear(qexhip_lattice *lat, const qexhip_hisq_coefs *c, const double *g_dev, double *fl_dev,
                      double *ll_dev);
/* ---- HISQ smearing chain rule: derivatives with respect to the complex link
   entries in QEX's pairing redot(A, B) = Re tr(A^+ B); outputs accumulate ---- */
/* fat7lDeriv, gauge/fat7lderiv.nim:122-475: deriv += dS/dgauge given mid = dS/dfl
   for fl = makeImpLinks(gauge, coef); when llderiv and llmid are not NULL and
   naik != 0, llderiv += dS/dllgauge given llmid = dS/dll (llgauge = gauge when
   NULL; llderiv may be deriv) */
int qexhip_fat7l_deriv(qexhip_lattice *lat, double *deriv_dev, const double *gauge_dev, const double *mid_dev,
                       const qexhip_fat7_coefs *coef, double *llderiv_dev, const double *llgauge_dev,
                       const double *llmid_dev, double naik);
/* projectUderiv, maths/matrixFunctions.nim:361-392: r = dS/dx for u = projectU(x),
   chain = dS/du, on n 3x3 matrices */
int qexhip_project_u_deriv(long n, double *r_dev, const double *u_dev, const double *x_dev, const double *chain_dev);
/* smearGetForce's returned proc, gauge/hisqsmear.nim:54-91: dg += dS/dg given
   dfl = dS/dfl and dll = dS/dll for (fl, ll) = smear(g); c = NULL: defaults */
int qexhip_hisq_smear_deriv(qexhip_lattice *lat, const qexhip_hisq_coefs *c, const double *g_dev,
                            const double *dfl_dev, const double *dll_dev, double *dg_dev);
/* HISQ fermion force, hmc/hmcAction.nim:851-876 with fermForce's smearing chain
   (gauge/hisqsmear.nim:54-91): F_mu(x) = TAH(2 dS/dg_mu(x) g_mu(x)^+) for
   S = phi^+ (m^2 - D_eo D_oe)^-1 phi, psi from qexhip_force_solve on the HISQ
   links l = newStag3(smear(g)), g the phased BC-applied gauge field.  Same
   convention as qexhip_force: for g -> exp(wX)g, dS/dw = -1/2 Re tr(X F). */
int qexhip_hisq_force(qexhip_links *l, const qexhip_hisq_coefs *c, const double *g_dev, const qexhip_field *psi,
                      double *F_dev);
/* timing: ms[0] per smear with default coefficients; ms[1..3] k_fat3, k_fat5,
   k_fatf of the first level, ms[4..6] of the second */
int qexhip_time_hisq_smear(qexhip_lattice *lat, const double *g_dev, double *fl_dev, double *ll_dev, int nrep,
                           double *ms);

/* ---- MILC binary gauge files (src/new/milcIO.nim, src/new/milcFileField.nim):
   96-byte header (magic 20103, dims, 64-char time stamp, order, sum29/sum31
   checksums), then per site in lexicographic order the four links as 3x3
   complex float32.  Links on the device are canonical FP64 [mu][site][3][3]. */
typedef struct {
  int dims[4];
  char timestamp[65];
  int order;
  int byteswapped;   /* file written with the other byte order */
  uint32_t sum29, sum31;
  int checksum_ok;   /* set by qexhip_milc_read */
} qexhip_milc_info;
/* header only (MilcHeader, milcIO.nim:9-25) */
int qexhip_milc_read_info(const char *path, qexhip_milc_info *info);
/* openMilcFileGaugeField (milcFileField.nim:46-57) + conversion to the device
   layout; fails on a dimension or checksum mismatch (info filled either way) */
int qexhip_milc_read(qexhip_lattice *lat, const char *path, double *U_dev, qexhip_milc_info *info);
/* newMilcFileGaugeField + close (milcFileField.nim:30-37,58-69): writes single
   precision with the checksum; timestamp may be NULL */
int qexhip_milc_write(qexhip_lattice *lat, const char *path, const double *U_dev, const char *timestamp);

/* ---- SciDAC / ILDG lattice files (LIME containers), QEX's default gauge and
   field format (src/gauge/gaugeUtils.nim:96-131; src/io/readerQiolite.nim,
   src/io/writerQiolite.nim over the qiolite package).  Field data on the
   device are canonical complex128 arrays [nobj][site][nelem]: nelem 9 for
   3x3 matrices (a gauge field is nobj = 4), 3 for color vectors, 1 for
   complex numbers; nobj 1..4.  The per-site CRC-32 SciDAC checksum is
   computed on the device while converting and verified on reading. */
typedef struct {
  int dims[4];
  int nrecords;          /* field records in the file */
  char datatype[64];     /* of the selected record, e.g. QDP_D3_ColorMatrix */
  char date[64];
  char precision;        /* 'F' or 'D' */
  int colors, spins, typesize, datacount;
  int ildg;              /* binary record is ildg-binary-data */
  int has_checksum;
  uint32_t suma, sumb;            /* stored in the file */
  uint32_t calc_suma, calc_sumb;  /* computed from the data read or written */
  int checksum_ok;
  long long filemd_len, recmd_len;  /* metadata lengths without the NUL */
} qexhip_scidac_info;
/* getFileLattice + newReader + recordMetadata (readerQiolite.nim:11-17,38-51,
   60-61): file and record description of field record `record` (0-based);
   filemd/recmd (may be NULL) receive up to cap-1 bytes of the file and record
   metadata, NUL-terminated */
int qexhip_scidac_read_info(const char *path, int record, qexhip_scidac_info *info, char *filemd,
                            long long filemd_cap, char *recmd, long long recmd_cap);
/* Reader.read (readerQiolite.nim:95-161,174-183): fails when the lattice,
   datacount or element count differ from the record's, or on a checksum
   mismatch; single-precision records are widened to FP64 */
int qexhip_scidac_read(qexhip_lattice *lat, const char *path, int record, double *data_dev, int nobj, int nelem,
                       qexhip_scidac_info *info);
/* newWriter (writerQiolite.nim:29-42): creates the file and writes its
   file message with metadata filemd */
int qexhip_scidac_create(const char *path, const int dims[4], const char *filemd);
/* Writer.write (writerQiolite.nim:83-160): appends one field record in
   precision 'F' or 'D' with metadata recmd; ildg != 0 adds the ildg-format
   record and names the data ildg-binary-data (gauge fields only); date NULL
   means the current UTC time.  info (may be NULL) receives the checksums. */
int qexhip_scidac_append(qexhip_lattice *lat, const char *path, const char *recmd, const double *data_dev, int nobj,
                         int nelem, char precision, int ildg, const char *date, qexhip_scidac_info *info);
/* ---- QEX's own field storage on one rank <-> canonical order, on the device
   (the host repack of quda/qudaWrapperImpl.nim:204-246).  A QEX field of
   Layout[V] (layout/qlayout.nim:10-150, field/fieldET.nim:18-22) with
   inner_geom = Layout.innerGeom holds, per outer site o and complex component
   c, the real parts of its V = prod(inner_geom) SIMD lanes, then the
   imaginary parts: double ((o*ncomp + c)*2 + r)*V + lane.  qex_dev holds
   nfields such fields back to back (ncomp = 9 for a ColorMatrix, 3 for a
   ColorVector); canon_dev is canonical complex128 [nfields][site][ncomp].
   Layouts QEX refuses (innerGeom entries not dividing the lattice or above
   2, no valid checkerboard fix-up) are refused. */
int qexhip_from_qex_layout(qexhip_lattice *lat, const int inner_geom[4], const double *qex_dev, int nfields,
                           int ncomp, double *canon_dev);
int qexhip_to_qex_layout(qexhip_lattice *lat, const int inner_geom[4], const double *canon_dev, int nfields,
                         int ncomp, double *qex_dev);

/* timing: ms[0], ms[1] = device time per whole-lattice launch of the
   conversion-plus-checksum kernels of Writer.write and Reader.read (no file) */
int qexhip_time_scidac(qexhip_lattice *lat, const double *data_dev, int nobj, int nelem, char precision, int nrep,
                       double *ms);

int qexhip_gen_gauge(qexhip_lattice *lat, uint64_t seed, double *U_dev);
int qexhip_gen_source(qexhip_lattice *lat, uint64_t seed, int stream, int par, double *x_dev);
int qexhip_gen_phase_bc(qexhip_lattice *lat, double *U_dev);
int qexhip_gen_hisq(qexhip_lattice *lat, uint64_t seed, const double *Uph_dev, double *g8_dev);

/* ---- measurement (the in-module bench of stagD.nim:767-810) ----
   average duration (ms, HIP events on the library stream) of one one-parity
   hop r_e = (2D)_eo x_o with links and fields stored in precision prec
   (QEXHIP_Q16: fixed-point links, FP16 fields) and links of recon (18 or 12)
   reals. */
int qexhip_time_hop(qexhip_links *l, int prec, int recon, int nrep, double *ms_per_hop);
/* the same for the multi-RHS hop over nrhs fields (FP64 or FP32, 18-real links):
   ms for one application to all nrhs fields */
int qexhip_time_hop_multi(qexhip_links *l, int prec, int nrhs, int nrep, double *ms);
/* device read bandwidth (GB/s) of a grid-stride non-temporal 16-byte-load kernel over
   mbytes MiB: the practical ceiling a streaming kernel reaches (no QEX counterpart) */
int qexhip_stream_read(long mbytes, int nrep, double *gbs);
/* whether the links are phased SU(3) matrices, so 12-real storage is exact
   (1) or not (0) */
int qexhip_links_recon_ok(qexhip_links *l, int *ok);

#ifdef __cplusplus
}
#endif
#endif
