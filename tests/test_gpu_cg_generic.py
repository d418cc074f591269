"""QEX's generic CG over a user-supplied operator (qex_amd.cg: CgState,
cgSolve; solvers/cg.nim:13-283), driven with the even-even staggered normal
operator: the solution equals the oracle's solveXX and the fused solveEE,
with the oracle's iteration count; a Hermitian preconditioner gives the same
solution; a zero source returns x = 0; a solve stopped by maxits continues
where it stopped.  Also the short form solve(s, x, b, m, res)
(stagSolve.nim:683-693)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 8181


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


class EvenOp:
    """A = 4 m^2 - D_eo D_oe on the even sites (stagD2ee)."""

    def __init__(self, Q, s, m, precon=False):
        self.Q, self.s, self.m2 = Q, s, m * m
        if precon:
            from qex_amd.cg import CgPrecon
            self.precon = CgPrecon.cpHerm

    def apply(self, Ap, p):
        self.Q.stagD2ee(self.s.se, self.s.so, Ap, self.s.g, p, self.m2)

    def applyPrecon(self, z, r):  # a constant positive scaling: Hermitian, changes no solution
        z.copy_from(r, "even")
        z.axpby(0.0, z, 0.25, "even")


def _setup(oracle, gpu, L):
    lo = gpu.Layout(L)
    U = oracle.phased(L, oracle.gauge_random(L, SEED))
    return lo, U, gpu.Staggered(lo, U)


def test_cg_solve_matches_oracle_and_fused(oracle, gpu):
    import qex_amd as Q
    L, m = (8, 8, 8, 8), 0.1
    lo, U, s = _setup(oracle, gpu, L)
    b = oracle.source_gauss(L, 1, 0, 0)
    sp = Q.SolverParams()
    sp.r2req, sp.maxits, sp.verbosity, sp.subsetName = 1e-20, 1000, 0, "even"
    x = lo.ColorVector()
    Q.cgSolve(x, lo.ColorVector().set(b), EvenOp(Q, s, m), sp)
    want, its, _ = oracle.solveXX(L, U, 0, b, m, 1e-20)
    assert rel(x.get(), want) < 1e-9
    assert abs(sp.iterations - its) <= 1
    xf = lo.ColorVector()
    spf = Q.SolverParams()
    spf.r2req, spf.verbosity = 1e-20, 0
    Q.solveEE(s, xf, lo.ColorVector().set(b), m, spf)
    assert rel(x.get(), xf.get()) < 1e-9
    # Hermitian preconditioning: same solution
    sp2 = Q.SolverParams()
    sp2.r2req, sp2.maxits, sp2.verbosity, sp2.subsetName = 1e-20, 1000, 0, "even"
    xp = lo.ColorVector()
    Q.cgSolve(xp, lo.ColorVector().set(b), EvenOp(Q, s, m, precon=True), sp2)
    assert rel(xp.get(), want) < 1e-9


def test_cg_state_zero_source_and_continuation(oracle, gpu):
    import qex_amd as Q
    L, m = (8, 8, 4, 8), 0.2
    lo, U, s = _setup(oracle, gpu, L)
    op = EvenOp(Q, s, m)
    sp = Q.SolverParams()
    sp.r2req, sp.maxits, sp.verbosity, sp.subsetName = 1e-20, 1000, 0, "even"
    x = lo.ColorVector().set(oracle.source_gauss(L, 5, 0, 0))
    Q.cgSolve(x, lo.ColorVector(), op, sp)
    assert sp.iterations == 0 and np.abs(x.get()).max() == 0.0
    b = lo.ColorVector().set(oracle.source_gauss(L, 6, 0, 0))
    x1 = lo.ColorVector()
    Q.cgSolve(x1, b, op, sp)
    full = sp.iterations
    x2 = lo.ColorVector()
    st = Q.newCgState(x2, b)
    sp.maxits = 10
    st.solve(op, sp)
    assert sp.iterations == 10
    sp.maxits = 1000
    st.solve(op, sp)                       # continues from iteration 10
    assert sp.iterations == full
    assert rel(x2.get(), x1.get()) < 1e-12


def test_solve_short_form(oracle, gpu):
    import qex_amd as Q
    L, m = (8, 4, 4, 8), 0.3
    lo, U, s = _setup(oracle, gpu, L)
    b = oracle.source_gauss(L, 7, 0, 2)
    x = lo.ColorVector()
    sp = Q.solve(s, x, lo.ColorVector().set(b), m, 1e-20)
    r = lo.ColorVector()
    Q.D(s, r, x, m)
    assert rel(r.get(), b) < 1e-9 and sp.r2 <= 1e-19
    with pytest.raises(Q.QexHipError, match="sbQex"):
        Q.solve(s, x, lo.ColorVector().set(b), m, 1e-20, cpuonly=True)


def test_cgm_state_shifts(oracle, gpu):
    """Generic multi-shift CG with A = 4 m^2 - D_eo D_oe: every shifted solution
    equals an independent single-shift solve at mass sqrt(m^2 + sigma/4), and
    the unshifted one the fused solveMulti's."""
    import qex_amd as Q
    L, m = (8, 8, 4, 8), 0.1
    lo, U, s = _setup(oracle, gpu, L)
    op = EvenOp(Q, s, m)
    b = oracle.source_gauss(L, 8, 0, 0)
    sig = [0.0, 0.01, 0.1, 1.0]
    xs = [lo.ColorVector() for _ in sig]
    sp = Q.SolverParams()
    sp.r2req, sp.maxits, sp.verbosity, sp.subsetName = 1e-20, 2000, 0, "even"
    st = Q.newCgmState(xs, lo.ColorVector().set(b), sig)
    st.solve(lambda a, p, shift=0.0: op.apply(a, p), sp)
    for k, sgm in enumerate(sig):
        want, _, _ = oracle.solveXX(L, U, 0, b, float(np.sqrt(m * m + sgm / 4.0)), 1e-22)
        assert rel(xs[k].get(), want) < 1e-8, sgm
    xm = [lo.ColorVector() for _ in sig]
    spm = Q.SolverParams()
    spm.r2req, spm.verbosity = 1e-20, 0
    Q.solveMulti(s, xm, lo.ColorVector().set(b), m, sig, spm)
    assert abs(spm.iterations - sp.iterations) <= 2
