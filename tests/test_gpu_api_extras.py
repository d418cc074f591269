"""The remaining operator entry points of QEX's staggered API in the mirror:
stagD1/stagD1x/stagDb (stagD.nim:467-478), peqDdag (:623-625), rephase
(:110-113) and the derivative primitives stagD2deriv/stagDeriv (:640-715),
against the oracle's operator and, for the derivatives, against the exact
linear identity they stand for: sum_k Re tr(delta_k^+ G_k) = Re<c, (2D)_delta x>
with (2D)_delta the hop built from arbitrary matrices delta."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 7373


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_stagD1_stagD1x_stagDb_peqDdag(oracle, gpu):
    import qex_amd as Q
    L = (8, 4, 6, 4)
    lo = gpu.Layout(L)
    U = oracle.phased(L, oracle.gauge_random(L, SEED))
    s = gpu.Staggered(lo, U)
    x = oracle.source_gauss(L, 1, 0, 2)
    r0 = oracle.source_gauss(L, 2, 0, 2)
    fx = lo.ColorVector().set(x)
    for par, sd in ((0, s.se), (1, s.so)):
        want = r0.copy()
        oracle.stagD2(L, U, par, want, x, 0.0, 0.0)
        fr = lo.ColorVector().set(r0)
        Q.stagD1(sd, fr, s.g, fx)
        assert rel(fr.get(), want) < 1e-13            # the other parity keeps r0
        neg = r0.copy()
        oracle.stagD(L, U, par, neg, x, 0.0, -2.0)      # -(2D)x on the subset
        fr = lo.ColorVector().set(r0)
        Q.stagD1x(sd, fr, s.g, fx)
        assert rel(fr.get(), neg) < 1e-13
        want = r0.copy()
        oracle.stagD2(L, U, par, want, x, 0.0, 0.3 / (0.5 * 2.0))
        fr = lo.ColorVector().set(r0)
        Q.stagDb(sd, fr, s.g, fx, 0.3, 2.0)
        assert rel(fr.get(), want) < 1e-13
    # peqDdag: r += m x - D x on all sites
    want = r0.copy()
    oracle.stagD(L, U, 2, want, x, 0.2, -1.0, 1.0)
    fr = lo.ColorVector().set(r0)
    Q.peqDdag(s, fr, fx, 0.2)
    assert rel(fr.get(), want) < 1e-13


@pytest.mark.parametrize("nl", [4, 8])
def test_stagD2deriv_identity(oracle, gpu, nl):
    """G = stagD2deriv accumulates d/dU Re<c, (2D)x>: for arbitrary matrices
    delta, sum Re tr(delta^+ G) equals Re<c, (2D)_delta x> on all sites."""
    import torch
    import qex_amd as Q
    L = (8, 8, 4, 8)
    lo = gpu.Layout(L)
    V = lo.nSites
    rng = np.random.default_rng(SEED + nl)
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 1))
    g = U if nl == 4 else oracle.hisq_links(L, SEED + 2, U)
    s = gpu.Staggered(lo, g)
    c = oracle.source_gauss(L, 3, 0, 2)
    x = oracle.source_gauss(L, 4, 0, 2)
    G0 = rng.normal(size=(nl, V, 3, 3)) + 1j * rng.normal(size=(nl, V, 3, 3))
    G = _cuda(G0)
    Q.stagD2deriv(s, G, lo.ColorVector().set(c), lo.ColorVector().set(x))
    dG = G.cpu().numpy() - G0                               # the accumulated derivative
    delta = rng.normal(size=(nl, V, 3, 3)) + 1j * rng.normal(size=(nl, V, 3, 3))
    lhs = float(np.sum(np.real(np.conj(delta) * dG)))      # sum Re tr(delta^+ G)
    hop = np.zeros((V, 3), dtype=np.complex128)
    oracle.stagD2(L, np.ascontiguousarray(delta), 2, hop, x, 0.0, 0.0)
    rhs = float(np.real(np.vdot(c, hop)))
    assert abs(lhs - rhs) < 1e-11 * max(abs(rhs), 1.0)
    # the same through the library's own hop with links delta
    sd = gpu.Staggered(lo, delta)
    fh = lo.ColorVector()
    Q.stagD1(sd.sa, fh, sd.g, lo.ColorVector().set(x))
    assert abs(float(np.real(np.vdot(c, fh.get()))) - rhs) < 1e-11 * max(abs(rhs), 1.0)


def test_stagDeriv_and_rephase(oracle, gpu):
    import torch
    import qex_amd as Q
    L = (8, 4, 4, 8)
    lo = gpu.Layout(L)
    V = lo.nSites
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 5))
    s = gpu.Staggered(lo, U)
    x = oracle.source_gauss(L, 6, 0, 2)
    F0 = oracle.gauge_random(L, SEED + 6)                     # prior content, rephased with the rest
    f = _cuda(F0)
    Q.stagDeriv(s, f, lo.ColorVector().set(x))
    outer = oracle.hisq_outer(L, x)[0::2]                     # eps x x(+mu)^+
    want = oracle.phased(L, F0 + outer)
    assert rel(f.cpu().numpy(), want) < 1e-13
    # rephase alone is the oracle's setBC + stagPhase
    g = _cuda(F0)
    Q.rephase(s, g)
    assert np.array_equal(g.cpu().numpy(), oracle.phased(L, F0))
    with pytest.raises(Q.QexHipError, match="plain"):
        Q.stagDeriv(gpu.Staggered(lo, oracle.hisq_links(L, 1, U)), torch.zeros((8, V, 3, 3), dtype=torch.complex128,
                                                                                 device="cuda"), lo.ColorVector())


def test_toSingle(oracle, gpu):
    """toSingle: the operator applied to single-precision fields runs the FP32
    links and arithmetic, within single-precision rounding of FP64."""
    import qex_amd as Q
    L = (8, 8, 4, 8)
    lo = gpu.Layout(L)
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 9))
    s = gpu.Staggered(lo, U)
    assert Q.toSingle(s) is s
    x = oracle.source_gauss(L, 10, 0, 2)
    fx = lo.ColorVector().set(x)
    fx1 = Q.toSingle(fx)
    assert fx1.prec == Q.FP32 and rel(fx1.get(), x) < 1e-7
    r1 = lo.ColorVector(prec=Q.FP32)
    Q.stagD2(s.se, r1, s.g, fx1, 0.0, 0.3)
    want = np.zeros_like(x)
    oracle.stagD2(L, U, 0, want, x, 0.0, 0.3)
    assert rel(r1.get(sub="even"), np.where(np.arange(lo.nSites)[:, None] < lo.nSites // 2, want, 0)) < 1e-6
