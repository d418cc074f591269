"""Edge cases of the C ABI: zero sources (QEX's cg.nim:138-143 returns x = 0
without iterating), the smallest lattices, oversized and mismatched inputs,
and out-of-range batch and shift counts.  Every solve is bounded by maxits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 1717


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _setup(oracle, gpu, L, seed=SEED):
    lo = gpu.Layout(L)
    U = oracle.phased(L, oracle.gauge_random(L, seed))
    return lo, U, gpu.Staggered(lo, U)


def _sp(gpu, r2req=1e-20, maxits=2000):
    sp = gpu.SolverParams()
    sp.r2req, sp.maxits, sp.verbosity = r2req, maxits, 0
    return sp


def test_zero_source_single_solves(oracle, gpu):
    L = (8, 8, 8, 8)
    lo, U, s = _setup(oracle, gpu, L)
    S = gpu.SloppyType
    for sloppy, links, rc in [(S.SloppyNone, None, 18), (S.SloppySingle, None, 18),
                              (S.SloppyHalf, S.SloppyHalfFixed, 12)]:
        b = lo.ColorVector()
        x = lo.ColorVector().set(oracle.source_gauss(L, 3, 0, 2))   # stale content must be cleared
        sp = _sp(gpu)
        sp.sloppySolve, sp.sloppyLinks, sp.recon = sloppy, links, rc
        gpu.solveEE(s, x, b, 0.1, sp)
        got = x.get(sub=0)
        assert np.all(np.isfinite(got)) and np.abs(got).max() == 0.0, sloppy
        assert sp.iterations == 0 and sp.r2 == 0.0


def test_zero_source_multishift(oracle, gpu):
    L = (8, 8, 8, 8)
    lo, U, s = _setup(oracle, gpu, L)
    sig = [0.0, 0.01, 0.1]
    xs = [lo.ColorVector().set(oracle.source_gauss(L, 4 + k, 0, 2)) for k in range(3)]
    sp = _sp(gpu)
    gpu.solveMulti(s, xs, lo.ColorVector(), 0.1, sig, sp)
    for x in xs:
        got = x.get(sub=0)
        assert np.all(np.isfinite(got)) and np.abs(got).max() == 0.0
    assert sp.iterations == 0


def test_zero_source_inside_batches(oracle, gpu):
    L = (8, 8, 8, 8)
    lo, U, s = _setup(oracle, gpu, L)
    b1 = lo.ColorVector().set(oracle.source_gauss(L, 8, 0, 0))
    zero = lo.ColorVector()
    # uneven-batch CG: the zero system stays zero, the other one converges
    xs = [lo.ColorVector(), lo.ColorVector()]
    sps = [_sp(gpu), _sp(gpu)]
    gpu.solveBatch([s, s], xs, [zero, b1], [0.1, 0.1], sps)
    assert np.abs(xs[0].get(sub=0)).max() == 0.0 and sps[0].iterations == 0
    assert sps[1].r2 <= 1e-20 and np.all(np.isfinite(xs[1].get(sub=0)))
    # multi-RHS multi-shift: a zero right-hand side among nonzero ones
    sig = [0.0, 0.05]
    xm = [[lo.ColorVector() for _ in sig] for _ in range(2)]
    spm = [_sp(gpu), _sp(gpu)]
    gpu.solveMultiBatch(s, xm, [zero, b1], 0.1, sig, spm)
    for x in xm[0]:
        assert np.abs(x.get(sub=0)).max() == 0.0
    for x in xm[1]:
        assert np.all(np.isfinite(x.get(sub=0)))
    assert spm[1].r2 <= 1e-20


@pytest.mark.parametrize("L", [(2, 2, 2, 2), (2, 4, 2, 6)])
def test_smallest_lattices(oracle, gpu, L):
    """Extent 2: the forward and backward neighbours of a direction coincide."""
    lo, U, s = _setup(oracle, gpu, L)
    x = oracle.source_gauss(L, 21, 0, 0)
    fr = lo.ColorVector()
    gpu.stagD2ee(s.se, s.so, fr, s.g, lo.ColorVector().set(x), 0.3)
    assert rel(fr.get(sub=0), oracle.stagD2xx(L, U, 0, x, 0.3)) < 1e-12
    b = lo.ColorVector().set(x)
    xs = lo.ColorVector()
    sp = _sp(gpu)
    gpu.solveEE(s, xs, b, 0.3, sp)
    want, its, _ = oracle.solveXX(L, U, 0, x, 0.3, 1e-20)
    assert rel(xs.get(sub=0), want) < 1e-8


def test_invalid_inputs_are_refused(oracle, gpu):
    from qex_amd import QexHipError
    with pytest.raises(QexHipError, match="too large"):
        gpu.Layout((65536, 65536, 2, 2))
    L = (4, 4, 4, 4)
    lo, U, s = _setup(oracle, gpu, L)
    other = gpu.Layout((4, 4, 4, 6))
    with pytest.raises(QexHipError, match="different lattices"):
        gpu.stagD2(s.se, other.ColorVector(), s.g, lo.ColorVector(), 0.0, 0.0)
    b = lo.ColorVector().set(oracle.source_gauss(L, 2, 0, 0))
    with pytest.raises(QexHipError, match="nshift"):
        gpu.solveMulti(s, [lo.ColorVector() for _ in range(65)], b, 0.1, [0.01 * k for k in range(65)], _sp(gpu))
    with pytest.raises(QexHipError, match="n out of range"):
        gpu.solveBatch([], [], [], [], [])
    with pytest.raises(QexHipError, match="HISQ"):
        from qex_amd import hisq
        import torch
        Ug = torch.from_numpy(U).cuda()
        hisq.hisqForce(s, None, Ug, lo.ColorVector())
