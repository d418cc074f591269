"""QEX's own SIMD field storage <-> canonical order on the device
(qexhip_from_qex_layout / qexhip_to_qex_layout) against the oracle's
restatement of layoutIndexQ and QEX's field storage: bit-exact both ways on
layouts with and without the checkerboard fix-up, exact round trips at 32^4
for 4- and 16-lane layouts, and refusal of the layouts QEX refuses."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 6060


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("L,ig,nf,n", [
    ((4, 4, 4, 4), (1, 1, 1, 1), 1, 3),
    ((4, 4, 4, 8), (1, 1, 2, 2), 4, 9),
    ((4, 4, 6, 6), (1, 1, 1, 2), 2, 9),     # odd outer extent: the innerCb shift is active
    ((8, 4, 4, 4), (2, 2, 2, 2), 8, 9),     # 16 lanes, a HISQ link set
    ((6, 4, 2, 2), (2, 1, 1, 1), 1, 3),
])
def test_matches_oracle(oracle, gpu, L, ig, nf, n):
    from qex_amd import gauge
    print(f"qex layout L={L} innerGeom={ig}", flush=True)
    lo = gpu.Layout(L)
    V = oracle.volume(L)
    rng = np.random.default_rng(SEED)
    f = rng.normal(size=(nf, V, n)) + 1j * rng.normal(size=(nf, V, n))
    raw = oracle.to_qex_layout(L, ig, f)
    got = gauge.toQexLayout(lo, ig, _cuda(f))
    assert np.array_equal(got.cpu().numpy(), raw)
    back = gauge.fromQexLayout(lo, ig, _cuda(raw), n)
    assert np.array_equal(back.cpu().numpy(), f)


@pytest.mark.parametrize("ig", [(1, 1, 2, 2), (2, 2, 2, 2)])
def test_full_size_roundtrip(oracle, gpu, ig):
    import time
    import torch
    from qex_amd import gauge
    L = (32, 32, 32, 32)
    lo = gpu.Layout(L)
    U = gauge.random_gauge(lo, SEED).reshape(4, lo.nSites, 9)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    raw = gauge.toQexLayout(lo, ig, U)
    t1 = time.perf_counter()
    back = gauge.fromQexLayout(lo, ig, raw, 9)
    t2 = time.perf_counter()
    gb = U.numel() * 16 * 2 / 1e9
    print(f"32^4 x 4 links innerGeom={ig}: to {gb / (t1 - t0):.0f} GB/s, from {gb / (t2 - t1):.0f} GB/s "
          f"(host clock, read + write)", flush=True)
    assert torch.equal(back, U)
    # spot checks of single sites against the oracle's index map
    q = oracle.qex_layout_setup(L, ig)
    C = oracle.all_coords(L)
    ni = q["nSitesInner"]
    r = raw.reshape(4, q["nSitesOuter"], 9, 2, ni)
    rng = np.random.default_rng(SEED + 1)
    for s in rng.integers(0, lo.nSites, 50):
        o, l = divmod(oracle.qex_layout_index(q, C[s]), ni)
        assert r[2, o, 4, 0, l].item() == U[2, s, 4].real.item()
        assert r[2, o, 4, 1, l].item() == U[2, s, 4].imag.item()


def test_refused(gpu):
    import torch
    from qex_amd import QexHipError, gauge
    lo = gpu.Layout((4, 4, 4, 6))
    f = torch.zeros((1, lo.nSites, 3), dtype=torch.complex128, device="cuda")
    with pytest.raises(QexHipError, match="above 2"):
        gauge.toQexLayout(lo, (1, 1, 4, 1), f)
    with pytest.raises(QexHipError, match="cb choice"):
        gauge.toQexLayout(lo, (1, 1, 1, 2), f)


def test_solve_from_qex_storage(oracle, gpu):
    """The path of a QEX shim: HISQ-form links and a source in QEX's own
    4-lane storage are uploaded as they are, converted on the device, solved,
    and the solution converted back; it equals the oracle's solution laid out
    by the oracle's restatement of QEX's storage."""
    import torch
    from qex_amd import gauge
    L, ig = (8, 8, 4, 8), (1, 1, 2, 2)
    lo = gpu.Layout(L)
    V = lo.nSites
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 2))
    g8 = oracle.hisq_links(L, SEED + 3, U)                         # (8, V, 3, 3) fat/long interleaved
    b = oracle.source_gauss(L, 9, 0, 0)                             # even sites
    raw_g = oracle.to_qex_layout(L, ig, g8.reshape(8, V, 9))
    raw_b = oracle.to_qex_layout(L, ig, b.reshape(1, V, 3))
    g_dev = gauge.fromQexLayout(lo, ig, _cuda(raw_g), 9).reshape(8, V, 3, 3)
    b_dev = gauge.fromQexLayout(lo, ig, _cuda(raw_b), 3).reshape(V, 3)
    s = gpu.Staggered(lo, g_dev)
    x = lo.ColorVector()
    sp = gpu.SolverParams()
    sp.r2req, sp.maxits, sp.verbosity = 1e-20, 2000, 0
    gpu.solveEE(s, x, lo.ColorVector().set(b_dev), 0.1, sp)
    raw_x = gauge.toQexLayout(lo, ig, x.get(device=torch.device("cuda")).reshape(1, V, 3).contiguous())
    want, its, _ = oracle.solveXX(L, g8, 0, b, 0.1, 1e-20)
    want_raw = oracle.to_qex_layout(L, ig, want.reshape(1, V, 3))
    got = raw_x.cpu().numpy()
    assert np.linalg.norm(got - want_raw) / np.linalg.norm(want_raw) < 1e-9
    assert abs(sp.iterations - its) <= 1
