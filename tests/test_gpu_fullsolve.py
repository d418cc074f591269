"""Full-lattice drivers of the host mirror over the C ABI: D(m) x = b with
even-odd preconditioning (stagSolve.nim:469-557) and the multi-mass solve
(stagSolve.nim:560-668), checked with the oracle's D."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 4711


def resid2(O, L, g, x, b, m):
    r = b - O.D(L, g, x, m)
    return O.norm2(r) / O.norm2(b)


@pytest.mark.parametrize("hisq", [False, True])
@pytest.mark.parametrize("sloppy", [0, 2])
def test_full_solve(oracle, gpu, hisq, sloppy):
    O = oracle
    L = (8, 8, 8, 8)
    U = O.phased(L, O.gauge_random(L, SEED))
    g = O.hisq_links(L, SEED + 1, U) if hisq else U
    b = O.source_gauss(L, 3, 0, 2)
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, g)
    fb = lo.ColorVector().set(b)
    fx = lo.ColorVector()
    sp = gpu.SolverParams()
    sp.r2req, sp.verbosity, sp.sloppySolve = 1e-20, 0, sloppy
    if sloppy:
        sp.sloppyLinks = gpu.SloppyType.SloppyHalfFixed
    gpu.solve(s, fx, fb, 0.1, sp)
    assert sp.r2 <= 1e-20
    assert resid2(O, L, g, fx.get(), b, 0.1) <= 1.01e-20


def test_eo_reconstruct(oracle, gpu):
    O = oracle
    L = (4, 4, 4, 8)
    U = O.phased(L, O.gauge_random(L, SEED))
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, U)
    x = O.source_gauss(L, 5, 0, 2)
    b = O.source_gauss(L, 6, 0, 2)
    fx, fb = lo.ColorVector().set(x), lo.ColorVector().set(b)
    gpu.eoReconstruct(s, fx, fb, 0.3)
    V = lo.nSites
    want = x.copy()
    Dx = O.D(L, U, np.concatenate([x[: V // 2], np.zeros_like(x[V // 2:])]), 0.0)
    want[V // 2:] = (b[V // 2:] - Dx[V // 2:]) / 0.3
    got = fx.get()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-13


def test_multi_mass_solve(oracle, gpu):
    O = oracle
    L = (8, 8, 8, 8)
    U = O.phased(L, O.gauge_random(L, SEED))
    b = O.source_gauss(L, 9, 0, 2)
    ms = [0.1, 0.2, 0.5, 1.0]
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, U)
    fb = lo.ColorVector().set(b)
    xs = [lo.ColorVector() for _ in ms]
    sp = gpu.SolverParams()
    sp.r2req, sp.verbosity = 1e-20, 0
    gpu.solveMultiMass(s, xs, fb, ms, sp)
    for f, m in zip(xs, ms):
        assert resid2(O, L, U, f.get(), b, m) <= 1e-18
