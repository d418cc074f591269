"""GPU parity of the staggered hop (stagD2/stagD/stagD2xx/D/Ddag) against the
CPU oracle, through the C ABI.  FP64: 1e-12 relative (the FP64 parity bar);
FP32 arithmetic: 1e-5; FP16/BF16 storage: 4e-3 / 3e-2."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20240601


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make_links(O, L, hisq):
    U = O.phased(L, O.gauge_random(L, SEED))
    return O.hisq_links(L, SEED + 1, U) if hisq else U


def sd_of(s, par):
    return {0: s.se, 1: s.so, 2: s.sa}[par]


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (6, 4, 8, 4), (8, 8, 8, 8)])
@pytest.mark.parametrize("hisq", [False, True])
def test_generators_match_oracle(oracle, gpu, L, hisq):
    from qex_amd import gauge
    O = oracle
    lo = gpu.Layout(L)
    Ug = gauge.random_gauge(lo, SEED)
    Uo = O.gauge_random(L, SEED)
    assert np.abs(Ug.cpu().numpy() - Uo).max() < 1e-13
    gauge.phase_bc(lo, Ug)
    Uph = O.phased(L, Uo)
    assert np.abs(Ug.cpu().numpy() - Uph).max() < 1e-13
    if hisq:
        g8g = gauge.hisq_links(lo, SEED + 1, Ug).cpu().numpy()
        assert np.abs(g8g - O.hisq_links(L, SEED + 1, Uph)).max() < 1e-13
    for par in (0, 1, 2):
        xg = gauge.source(lo, 77, stream=3, par=par).cpu().numpy()
        assert np.abs(xg - O.source_gauss(L, 77, 3, par)).max() < 1e-14


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (6, 4, 8, 4), (8, 8, 8, 8)])
@pytest.mark.parametrize("hisq", [False, True])
@pytest.mark.parametrize("par", [0, 1, 2])
@pytest.mark.parametrize("ab", [(0.0, 0.0), (0.0, 0.37), (1.5, -0.25)])
def test_stagD2_fp64(oracle, gpu, L, hisq, par, ab):
    O = oracle
    g = make_links(O, L, hisq)
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, g)
    x = O.source_gauss(L, 11, 0, 2)
    r0 = O.source_gauss(L, 12, 1, 2)
    fx = lo.ColorVector().set(x)
    fr = lo.ColorVector().set(r0)
    gpu.stagD2(sd_of(s, par), fr, s.g, fx, *ab)
    got = fr.get()
    want = r0.copy()
    O.stagD2(L, g, par, want, x, *ab)
    assert rel(got, want) < 1e-12


@pytest.mark.parametrize("hisq", [False, True])
@pytest.mark.parametrize("par", [0, 1])
def test_stagD2xx_fp64(oracle, gpu, hisq, par):
    O = oracle
    L = (8, 8, 8, 8)
    g = make_links(O, L, hisq)
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, g)
    x = O.source_gauss(L, 21, 0, par)
    fx = lo.ColorVector().set(x)
    fr = lo.ColorVector()
    (gpu.stagD2ee if par == 0 else gpu.stagD2oo)(s.se, s.so, fr, s.g, fx, 0.01)
    got = fr.get(sub=par)
    want = O.stagD2xx(L, g, par, x, 0.01)
    assert rel(got, want) < 1e-12


@pytest.mark.parametrize("hisq", [False, True])
def test_D_Ddag_stagD(oracle, gpu, hisq):
    O = oracle
    L = (4, 6, 4, 8)
    g = make_links(O, L, hisq)
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, g)
    x = O.source_gauss(L, 31, 0, 2)
    fx = lo.ColorVector().set(x)
    fr = lo.ColorVector()
    gpu.D(s, fr, fx, 0.3)
    assert rel(fr.get(), O.D(L, g, x, 0.3)) < 1e-12
    gpu.Ddag(s, fr, fx, 0.3)
    assert rel(fr.get(), O.Ddag(L, g, x, 0.3)) < 1e-12
    r0 = O.source_gauss(L, 32, 0, 2)
    fr.set(r0)
    gpu.stagD(s.so, fr, s.g, fx, 0.2, -1.0, 0.5)
    want = r0.copy()
    O.stagD(L, g, 1, want, x, 0.2, -1.0, 0.5)
    assert rel(fr.get(), want) < 1e-12


@pytest.mark.parametrize("prec,tol", [(1, 1e-5), (2, 4e-3), (3, 3e-2)])
@pytest.mark.parametrize("hisq", [False, True])
def test_stagD2_reduced_precision(oracle, gpu, prec, tol, hisq):
    O = oracle
    L = (8, 8, 8, 8)
    g = make_links(O, L, hisq)
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, g)
    x = O.source_gauss(L, 41, 0, 2)
    fx = lo.ColorVector(prec).set(x)
    fr = lo.ColorVector(prec)
    gpu.stagD2(s.se, fr, s.g, fx, 0.0, 0.0)
    want = np.zeros_like(x)
    O.stagD2(L, g, 0, want, x, 0.0, 0.0)
    got = fr.get(sub=0)
    assert rel(got[: lo.nEven], want[: lo.nEven]) < tol


@pytest.mark.parametrize("L", [(32, 32, 32, 32), (16, 16, 16, 16)])
def test_large_volume_properties(gpu, L):
    """At full size: anti-Hermiticity Re<y,Hx> = -Re<Hy,x> and linearity, FP64."""
    from qex_amd import gauge
    lo = gpu.Layout(L)
    U = gauge.random_gauge(lo, SEED)
    gauge.phase_bc(lo, U)
    s = gpu.Staggered(lo, U)
    x = lo.ColorVector().set(gauge.source(lo, 5, 0, 2))
    y = lo.ColorVector().set(gauge.source(lo, 6, 0, 2))
    Hx, Hy = lo.ColorVector(), lo.ColorVector()
    gpu.stagD2(s.sa, Hx, s.g, x, 0.0, 0.0)
    gpu.stagD2(s.sa, Hy, s.g, y, 0.0, 0.0)
    a = y.redot(Hx)
    b = Hy.redot(x)
    scale = np.sqrt(y.norm2() * Hx.norm2())
    assert abs(a + b) / scale < 1e-13
    # linearity: H(x + 2y) = Hx + 2Hy
    xa = x.get(device="cuda")
    ya = y.get(device="cuda")
    z = lo.ColorVector().set((xa + 2 * ya).contiguous())
    Hz = lo.ColorVector()
    gpu.stagD2(s.sa, Hz, s.g, z, 0.0, 0.0)
    lhs = Hz.get(device="cuda")
    rhs = Hx.get(device="cuda") + 2 * Hy.get(device="cuda")
    import torch
    assert float(torch.linalg.norm(lhs - rhs) / torch.linalg.norm(rhs)) < 1e-14


@pytest.mark.parametrize("shape", [(0, -1), (1, 0), (2, 0)])
@pytest.mark.parametrize("n", [1, 3, 8, 11])
@pytest.mark.parametrize("hisq", [False, True])
@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-6)])
def test_stagD2_multi_rhs(oracle, gpu, n, hisq, prec, tol, shape):
    """Multi-RHS stagD2 (each link loaded once per group of up to 8 fields)
    equals the oracle's stagD2 applied field by field, with the vectors loaded
    in 1 or 2 batches per link and with the tuner's choice."""
    gpu.set_hop_multi_shape(*shape)
    try:
        _multi_rhs_case(oracle, gpu, n, hisq, prec, tol)
    finally:
        gpu.set_hop_multi_shape(0, -1)


@pytest.mark.parametrize("L,chunk", [((12, 12, 12, 8), 256), ((12, 12, 12, 8), 512), ((16, 8, 8, 6), 256)])
@pytest.mark.parametrize("hisq", [False, True])
@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-6)])
def test_stagD2_multi_rhs_t_walk(oracle, gpu, L, chunk, hisq, prec, tol):
    """The multi-RHS hop over the t-first site walk: t-slices of 864 sites per
    parity (12^3) cut into chunks of 256 or 512 sites leave a partial last
    chunk; 16x8x8 slices (512 sites) take two exact chunks of 256.  Every site
    of both parities still equals the oracle's stagD2."""
    gpu.set_hop_multi_shape(1, chunk)
    try:
        _multi_rhs_case(oracle, gpu, 8, hisq, prec, tol, L=L)
    finally:
        gpu.set_hop_multi_shape(0, -1)


def _multi_rhs_case(oracle, gpu, n, hisq, prec, tol, L=(8, 4, 6, 8)):
    O = oracle
    g = make_links(O, L, hisq)
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, g)
    a, b = 0.5, -0.3
    xs = [O.source_gauss(L, 100 + j, 0, 2) for j in range(n)]
    r0 = [O.source_gauss(L, 200 + j, 1, 2) for j in range(n)]
    fx = [lo.ColorVector(prec).set(x) for x in xs]
    fr = [lo.ColorVector(prec).set(r) for r in r0]
    for par in (0, 1):
        gpu.stagD2multi(sd_of(s, par), fr, s.g, fx, a, b)
    for j in range(n):
        want = r0[j].copy()
        for par in (0, 1):
            O.stagD2(L, g, par, want, xs[j], a, b)
        assert rel(fr[j].get(), want) < tol, j


def test_stagDN_matches_stagD2(oracle, gpu):
    """QEX's stagDN: r_i = m x_i + (2D) x_i (the sc argument is not applied)."""
    O = oracle
    L = (6, 6, 6, 6)
    g = make_links(O, L, False)
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, g)
    xs = [O.source_gauss(L, 300 + j, 0, 2) for j in range(5)]
    fx = [lo.ColorVector().set(x) for x in xs]
    fr = [lo.ColorVector() for _ in xs]
    gpu.stagDN([s.se] * 5, fr, s.g, fx, 0.2, sc=3.0)
    for j, x in enumerate(xs):
        want = np.zeros_like(x)
        O.stagD2(L, g, 0, want, x, 0.0, 0.2)
        assert rel(fr[j].get(sub=0)[: lo.nEven], want[: lo.nEven]) < 1e-12


@pytest.mark.parametrize("nrhs", [1, 8])
def test_time_hop_multi_runs(gpu, nrhs):
    from qex_amd import gauge
    lo = gpu.Layout((16, 16, 16, 16))
    U = gauge.random_gauge(lo, 1)
    gauge.phase_bc(lo, U)
    s = gpu.Staggered(lo, U)
    assert gpu.time_hop_multi(s, 0, nrhs, 5) > 0.0


def test_stream_read_probe(gpu):
    gbs = gpu.stream_read_gbs(256, 5)
    assert 500.0 < gbs < 8000.0


@pytest.mark.parametrize("L", [(8, 4, 6, 4), (8, 8, 4, 8), (16, 16, 16, 16)])
@pytest.mark.parametrize("hisq", [False, True])
def test_hop_two_links_per_trip(oracle, gpu, L, hisq):
    """The hop's stencil loop taking both links of a hop per trip (the form the
    tuner picks for small lattices) equals the oracle: stagD2 on each parity
    and both in FP64 and FP32, and FP64 and mixed CG solves (whose true
    residuals run the reducing hops) to the oracle's residual."""
    O = oracle
    g = make_links(O, L, hisq)
    gpu.set_hop_two(1)
    try:
        lo = gpu.Layout(L)
        s = gpu.Staggered(lo, g)
        x = O.source_gauss(L, 61, 0, 2)
        r0 = O.source_gauss(L, 62, 1, 2)
        for prec, tol in ((gpu.FP64, 1e-12), (gpu.FP32, 1e-5)):
            for par in (0, 1, 2):
                fx = lo.ColorVector(prec).set(x)
                fr = lo.ColorVector(prec).set(r0)
                gpu.stagD2(sd_of(s, par), fr, s.g, fx, 0.5, -0.25)
                want = r0.copy()
                O.stagD2(L, g, par, want, x, 0.5, -0.25)
                assert rel(fr.get(), want) < tol, (prec, par)
        b = O.source_gauss(L, 63, 0, 0)
        fb = lo.ColorVector().set(b)
        for sloppy in (0, 1):
            fs = lo.ColorVector()
            sp = gpu.SolverParams()
            sp.r2req, sp.verbosity, sp.sloppySolve = 1e-20, 0, sloppy
            gpu.solveEE(s, fs, fb, 0.1, sp)
            r = b - O.stagD2xx(L, g, 0, fs.get(), 0.01)
            assert O.norm2(r[: lo.nEven]) / O.norm2(b) <= 1.0001e-20, sloppy
    finally:
        gpu.set_hop_two(-1)


@pytest.mark.parametrize("L", [(8, 4, 6, 4), (8, 8, 4, 8), (16, 16, 16, 16)])
@pytest.mark.parametrize("hisq", [False, True])
@pytest.mark.parametrize("split", [2, 4])
def test_hop_split_threads_per_site(oracle, gpu, L, hisq, split):
    """The hop's split form (2 or 4 threads per output site, each taking a
    share of its links, the sums added by lane shuffles; the tuner's pick for
    small lattices) equals the oracle: stagD2 on each parity in FP64, FP32 and
    FP16 storage, and FP64 and mixed CG solves, whose true residuals run the
    reducing hops."""
    O = oracle
    g = make_links(O, L, hisq)
    gpu.set_hop_split(split)
    try:
        lo = gpu.Layout(L)
        s = gpu.Staggered(lo, g)
        x = O.source_gauss(L, 81, 0, 2)
        r0 = O.source_gauss(L, 82, 1, 2)
        for prec, tol in ((gpu.FP64, 1e-12), (gpu.FP32, 1e-5), (gpu.FP16, 4e-3)):
            for par in (0, 1, 2):
                fx = lo.ColorVector(prec).set(x)
                fr = lo.ColorVector(prec).set(r0)
                gpu.stagD2(sd_of(s, par), fr, s.g, fx, 0.5, -0.25)
                want = r0.copy()
                O.stagD2(L, g, par, want, x, 0.5, -0.25)
                assert rel(fr.get(), want) < tol, (prec, par)
        b = O.source_gauss(L, 83, 0, 0)
        fb = lo.ColorVector().set(b)
        for sloppy in (0, 1):
            fs = lo.ColorVector()
            sp = gpu.SolverParams()
            sp.r2req, sp.verbosity, sp.sloppySolve = 1e-20, 0, sloppy
            gpu.solveEE(s, fs, fb, 0.1, sp)
            r = b - O.stagD2xx(L, g, 0, fs.get(), 0.01)
            assert O.norm2(r[: lo.nEven]) / O.norm2(b) <= 1.0001e-20, sloppy
    finally:
        gpu.set_hop_split(-1)


@pytest.mark.parametrize("L,hisq", [((8, 8, 4, 8), False), ((8, 4, 6, 4), True)])
def test_link_cache_budgets_bitwise(oracle, gpu, L, hisq):
    """The link-cache budget (the first link blocks read with
    cache-allocating loads, the rest non-temporally) changes no arithmetic.
    Budgets that cover none, some (odd and even counts) or all of the blocks
    of the 16-bit 12-real, 16-bit 18-real, FP32 and FP64 link fields give
    stagD2 (FP64, FP32, each parity), the 8-field multi-RHS hop, and the
    FP64, FP32-, FP16- and Q16-link
    (18- and, for plain links, 12-real) mixed CG solves bitwise the results
    of budget 0, which equal the oracle's."""
    O = oracle
    g = make_links(O, L, hisq)
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, g)
    stride = -(-(int(np.prod(L)) // 2) // 256) * 256
    # one link block (one direction, one parity) is 24, 36, 72 or 144 B x stride
    # for 16-bit 12-real, 16-bit 18-real, FP32 and FP64 links: odd and even
    # block counts in every format, so that a direction with one cached block
    # occurs for both output parities
    budgets = [0] + [n * 24 * stride for n in (1, 3, 5, 7)] + [n * 144 * stride for n in (1, 3, 5, 7)] + [1 << 30]
    x = O.source_gauss(L, 71, 0, 2)
    r0 = O.source_gauss(L, 72, 1, 2)
    b = O.source_gauss(L, 73, 0, 0)
    S = gpu.SloppyType
    solves = [(S.SloppyNone, None, 18), (S.SloppySingle, None, 18), (S.SloppyHalf, None, 18),
              (S.SloppyHalf, S.SloppyHalfFixed, 18)]
    if not hisq:
        solves += [(S.SloppyHalf, S.SloppyHalfFixed, 12), (S.SloppyHalf, None, 12)]
    ref = None
    try:
        for bud in budgets:
            gpu.set_link_cache(lo, bud)
            got = []
            for prec in (gpu.FP64, gpu.FP32):
                for par in (0, 1):
                    fx = lo.ColorVector(prec).set(x)
                    fr = lo.ColorVector(prec).set(r0)
                    gpu.stagD2(sd_of(s, par), fr, s.g, fx, 0.5, -0.25)
                    got.append(fr.get())
                # the multi-RHS hop: 8 fields per link load
                fxs = [lo.ColorVector(prec).set(np.roll(x, j)) for j in range(8)]
                frs = [lo.ColorVector(prec).set(r0) for _ in range(8)]
                gpu.stagD2multi(s.sa, frs, s.g, fxs, 0.5, -0.25)
                got += [f.get() for f in frs]
            for sl, sll, rc in solves:
                fs = lo.ColorVector()
                sp = gpu.SolverParams()
                sp.r2req, sp.verbosity, sp.sloppySolve, sp.recon = 1e-16, 0, sl, rc
                if sll is not None:
                    sp.sloppyLinks = sll
                gpu.solveEE(s, fs, lo.ColorVector().set(b), 0.1, sp)
                got.append(fs.get())
                got.append(np.array([sp.iterations]))
            if ref is None:
                ref = got
                k = 0
                for prec, tol in ((gpu.FP64, 1e-12), (gpu.FP32, 1e-5)):
                    for par in (0, 1):
                        want = r0.copy()
                        O.stagD2(L, g, par, want, x, 0.5, -0.25)
                        assert rel(ref[k], want) < tol, (prec, par)
                        k += 1
                    for j in range(8):
                        want = r0.copy()
                        for par in (0, 1):
                            O.stagD2(L, g, par, want, np.roll(x, j), 0.5, -0.25)
                        assert rel(ref[k], want) < tol, (prec, "multi", j)
                        k += 1
                for _ in solves:
                    r = b - O.stagD2xx(L, g, 0, ref[k], 0.01)
                    assert O.norm2(r[: lo.nEven]) / O.norm2(b) <= 1.0001e-16
                    k += 2
            else:
                for k, (a, c) in enumerate(zip(got, ref)):
                    assert np.array_equal(a, c), (bud, k)
    finally:
        gpu.set_link_cache(lo, -1)


@pytest.mark.parametrize("L", [(8, 4, 6, 4), (6, 4, 8, 10), (8, 8, 4, 8), (4, 4, 4, 2)])
@pytest.mark.parametrize("tmarch", [1, 3, 4, 16])
def test_hop_tmarch(oracle, gpu, L, tmarch):
    """The t-march form of the plain-link hop (each thread walks tmarch
    t-slices and keeps its forward t-neighbour for the slice after next; 3 and
    4 leave a partial last segment at Lt = 10, 16 exceeds every Lt, Lt = 2
    makes both t-neighbours one site) equals the oracle: stagD2 on each parity
    and both, FP64 and FP32, and FP64 and mixed CG solves, whose reliable
    updates and true residuals run the reducing hops, to the oracle's
    residual."""
    O = oracle
    U = O.phased(L, O.gauge_random(L, 4321))
    gpu.set_hop_tmarch(tmarch)
    try:
        lo = gpu.Layout(L)
        s = gpu.Staggered(lo, U)
        x = O.source_gauss(L, 51, 0, 2)
        r0 = O.source_gauss(L, 52, 1, 2)
        for prec, tol in ((gpu.FP64, 1e-12), (gpu.FP32, 1e-5)):
            for par in (0, 1, 2):
                fx = lo.ColorVector(prec).set(x)
                fr = lo.ColorVector(prec).set(r0)
                gpu.stagD2(sd_of(s, par), fr, s.g, fx, 0.5, -0.25)
                want = r0.copy()
                O.stagD2(L, U, par, want, x, 0.5, -0.25)
                assert rel(fr.get(), want) < tol, (prec, par)
        b = O.source_gauss(L, 53, 0, 0)
        fb = lo.ColorVector().set(b)
        for sloppy in (0, 1):
            fs = lo.ColorVector()
            sp = gpu.SolverParams()
            sp.r2req, sp.verbosity, sp.sloppySolve = 1e-20, 0, sloppy
            gpu.solveEE(s, fs, fb, 0.1, sp)
            r = b - O.stagD2xx(L, U, 0, fs.get(), 0.01)
            assert O.norm2(r[: lo.nEven]) / O.norm2(b) <= 1.0001e-20, sloppy
    finally:
        gpu.set_hop_tmarch(-1)


@pytest.mark.parametrize("L,chunk", [((12, 12, 12, 8), 512), ((16, 16, 16, 8), 1024), ((8, 8, 8, 8), 0)])
@pytest.mark.parametrize("hisq", [False, True])
def test_hop_t_walk(oracle, gpu, L, chunk, hisq):
    """The one-parity hop over the t-first site walk (12^3 t-slices: 864 sites
    per parity, chunks of 512 leave a partial one; 16^3: two chunks of 1024;
    chunk 0: canonical order) equals the oracle: stagD2 on each parity and
    both, FP64 and FP32, and an FP64 and a mixed CG solve (whose reliable
    updates and true residuals run reducing hops) to the oracle's residual."""
    O = oracle
    g = make_links(O, L, hisq)
    gpu.set_hop_walk(chunk)
    try:
        lo = gpu.Layout(L)
        s = gpu.Staggered(lo, g)
        x = O.source_gauss(L, 41, 0, 2)
        r0 = O.source_gauss(L, 42, 1, 2)
        for prec, tol in ((gpu.FP64, 1e-12), (gpu.FP32, 1e-5)):
            for par in (0, 1, 2):
                fx = lo.ColorVector(prec).set(x)
                fr = lo.ColorVector(prec).set(r0)
                gpu.stagD2(sd_of(s, par), fr, s.g, fx, 0.5, -0.25)
                want = r0.copy()
                O.stagD2(L, g, par, want, x, 0.5, -0.25)
                assert rel(fr.get(), want) < tol, (prec, par)
        b = O.source_gauss(L, 43, 0, 0)
        fb = lo.ColorVector().set(b)
        for sloppy in (0, 1):
            fs = lo.ColorVector()
            sp = gpu.SolverParams()
            sp.r2req, sp.verbosity, sp.sloppySolve = 1e-20, 0, sloppy
            gpu.solveEE(s, fs, fb, 0.1, sp)
            r = b - O.stagD2xx(L, g, 0, fs.get(), 0.01)
            assert O.norm2(r[: lo.nEven]) / O.norm2(b) <= 1.0001e-20, sloppy
    finally:
        gpu.set_hop_walk(-1)
