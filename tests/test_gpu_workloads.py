"""GPU parity at the full benchmark workloads, against the C oracle.

The inputs are the oracle's own seeded arrays, uploaded, so the GPU and the
oracle solve bit-identical systems; where the oracle would take minutes on the
host its results come from tests/golden/workloads.json (written by
tests/golden/make_workload_fixtures.py).

- Plain hop, FP64 and FP32, at 16^4, 32^4 and 48^4 against O.stagD2
  (FP64 1e-12 relative; FP32 arithmetic 1e-5).
- CG at 32^4, m = 0.01, to r2req = 1e-20: FP64 iterations equal to the
  oracle's (+-1); the FP64 true residual of every variant recomputed with the
  oracle's operator; the solution against the oracle's on 64 sampled sites.
- HISQ form at 32^3 x 64, m = 0.005: FP64, FP16, BF16 and 16-bit fixed-point
  link storage, each true residual recomputed by the oracle, each error
  against the FP64 solve.
- Batched multi-shift, HISQ form 24^3 x 48, 8 sources x 12 shifts: every
  shifted system's residual recomputed by the oracle and equal to the one the
  solver reports; source 0's iteration count equal to the oracle's CGM.
- 16 independent plain solves on 8^4, 12^4, 16^4, 16^3 x 32: iteration counts
  against the oracle's CG, true residuals by the oracle.
- 24^4 HMC (beta = 5.6, m = 0.05, tau = 1, 20 leapfrog steps): reversibility.

Error bars on solutions: two solutions whose true residuals are below
sqrt(r2req)|b| differ by at most 2 kappa sqrt(r2req) |x| (kappa the condition
number of A = 4(m^2 - D_eo D_oe), at most (4m^2 + 64)/(4m^2) for unitary
links); the tests assert that bound and, below it, a measured-scale bar.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SEED = 20240601
R2REQ = 1e-20


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def fixtures():
    path = os.path.join(HERE, "golden", "workloads.json")
    with open(path) as f:
        return json.load(f)


def need(fixtures, key):
    if key not in fixtures:
        pytest.skip(f"tests/golden/workloads.json has no {key} (run tests/golden/make_workload_fixtures.py {key})")
    return fixtures[key]


def true_r2(O, L, g, b, x, m, sigma=0.0):
    ne = O.volume(L) // 2
    r = b - O.stagD2xx(L, g, 0, x, m * m) - sigma * x
    return O.norm2(r[:ne]) / O.norm2(b[:ne])


def sample(x, idx):
    return np.asarray(x)[idx]


def fixture_sample(fx):
    a = np.asarray(fx, dtype=np.float64)
    return a[..., 0] + 1j * a[..., 1]


# ---------------------------------------------------------------- hops
@pytest.mark.parametrize("L", [(16, 16, 16, 16), (32, 32, 32, 32), (48, 48, 48, 48)])
def test_hop_fp64_fp32_full_size(oracle, gpu, L):
    O = oracle
    print(f"hop parity {L}", flush=True)
    U = O.phased(L, O.gauge_random(L, SEED))
    x = O.source_gauss(L, 11, 0, 1)
    want = np.zeros_like(x)
    O.stagD2(L, U, 0, want, x, 0.0, 0.0)
    ne = O.volume(L) // 2
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, U)
    del U
    for prec, bar in ((gpu.FP64, 1e-12), (gpu.FP32, 1e-5)):
        fx = lo.ColorVector(prec).set(x)
        fr = lo.ColorVector(prec)
        gpu.stagD2(s.se, fr, s.g, fx, 0.0, 0.0)
        got = fr.get(sub="even")[:ne]
        err = rel(got, want[:ne])
        print(f"  prec {prec}: rel err {err:.3e}", flush=True)
        assert err < bar, (L, prec, err)


# ---------------------------------------------------------------- config 2 CG
@pytest.fixture(scope="module")
def plain32(oracle, gpu, fixtures):
    O = oracle
    fx = need(fixtures, "plain_32x32x32x32")
    L = tuple(fx["lattice"])
    U = O.phased(L, O.gauge_random(L, fx["gauge_seed"]))
    b = O.source_gauss(L, fx["source_seed"], 0, 0)
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, U)
    fb = lo.ColorVector().set(b)
    x64 = lo.ColorVector()
    sp = gpu.SolverParams()
    sp.r2req, sp.verbosity = fx["r2req"], 0
    gpu.solveEE(s, x64, fb, fx["mass"], sp)
    return dict(L=L, U=U, b=b, lo=lo, s=s, fb=fb, x64=x64, sp=sp, fx=fx)


def test_cg_32x32x32x32_fp64_matches_oracle(oracle, plain32):
    O, c = oracle, plain32
    fx, sp, L = c["fx"], c["sp"], c["L"]
    print(f"CG 32^4 FP64: {sp.iterations} its (oracle {fx['iterations']}), reported r2 {sp.r2:.3e}", flush=True)
    assert abs(sp.iterations - fx["iterations"]) <= 1, (sp.iterations, fx["iterations"])
    assert sp.r2 <= R2REQ
    x = c["x64"].get()
    r2 = true_r2(O, L, c["U"], c["b"], x, fx["mass"])
    print(f"  oracle-recomputed true r2 {r2:.3e}", flush=True)
    assert r2 <= 1.0001 * R2REQ
    # 64 sampled sites against the oracle's own FP64 solution
    idx = fx["sample_sites"]
    want = fixture_sample(fx["x_sample"])
    got = sample(x, idx)
    xn = np.sqrt(fx["x_norm2"])
    m2 = fx["mass"] ** 2
    kappa = (4 * m2 + 64.0) / (4 * m2)
    d = float(np.linalg.norm(got - want))
    print(f"  sample |dx| / |x| = {d / xn:.3e} (bound {2 * kappa * np.sqrt(R2REQ):.1e}), "
          f"sample rel {d / np.linalg.norm(want):.3e}", flush=True)
    assert d <= 2 * kappa * np.sqrt(R2REQ) * xn
    # the two solves follow the same recurrence on the same inputs: they agree
    # far inside that bound (measured 4e-14)
    assert d / np.linalg.norm(want) < 1e-10
    n2 = c["x64"].norm2("even")
    assert abs(n2 - fx["x_norm2"]) <= 1e-6 * fx["x_norm2"]


@pytest.mark.parametrize("variant", ["fp32_inner", "q16_12_fp16", "fp16_12"])
def test_cg_32x32x32x32_mixed_true_residual(oracle, gpu, plain32, variant):
    O, c = oracle, plain32
    S = gpu.SloppyType
    sloppy, links, recon = {"fp32_inner": (S.SloppySingle, None, 18),
                            "q16_12_fp16": (S.SloppyHalf, S.SloppyHalfFixed, 12),
                            "fp16_12": (S.SloppyHalf, None, 12)}[variant]
    lo, fx, L = c["lo"], c["fx"], c["L"]
    x = lo.ColorVector()
    sp = gpu.SolverParams()
    sp.r2req, sp.verbosity = R2REQ, 0
    sp.sloppySolve, sp.sloppyLinks, sp.recon = sloppy, links, recon
    gpu.solveEE(c["s"], x, c["fb"], fx["mass"], sp)
    xv = x.get()
    r2 = true_r2(O, L, c["U"], c["b"], xv, fx["mass"])
    ne = lo.nEven
    x64 = c["x64"].get()[:ne]
    err = rel(xv[:ne], x64)
    print(f"CG 32^4 {variant}: {sp.iterations} its, {sp.reliableUpdates} reliable updates, "
          f"reported r2 {sp.r2:.3e}, oracle r2 {r2:.3e}, err vs FP64 {err:.3e}", flush=True)
    assert sp.r2 <= R2REQ and r2 <= 1.0001 * R2REQ
    assert sp.reliableUpdates >= 1
    # FP32 arithmetic keeps the FP64 iteration count; 16-bit storage may add some
    cap = {"fp32_inner": 1.02, "q16_12_fp16": 1.05, "fp16_12": 1.6}[variant]
    assert sp.iterations <= cap * fx["iterations"], (sp.iterations, fx["iterations"])
    m2 = fx["mass"] ** 2
    assert err <= 2 * (4 * m2 + 64.0) / (4 * m2) * np.sqrt(R2REQ)
    # measured 3e-13 (FP32 inner), 5e-12 (Q16 12-real), 3e-11 (FP16 12-real)
    assert err < 1e-9


# ---------------------------------------------------------------- config 3
@pytest.fixture(scope="module")
def hisq32(oracle, gpu, fixtures):
    O = oracle
    fx = need(fixtures, "hisq_32x32x32x64")
    L = tuple(fx["lattice"])
    U = O.phased(L, O.gauge_random(L, fx["gauge_seed"]))
    g = O.hisq_links(L, fx["hisq_seed"], U)
    del U
    b = O.source_gauss(L, fx["source_seed"], 0, 0)
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, g)
    fb = lo.ColorVector().set(b)
    x64 = lo.ColorVector()
    sp = gpu.SolverParams()
    sp.r2req, sp.verbosity = fx["r2req"], 0
    gpu.solveEE(s, x64, fb, fx["mass"], sp)
    return dict(L=L, g=g, b=b, lo=lo, s=s, fb=fb, x64=x64, sp=sp, fx=fx)


def test_hisq_32x32x32x64_fp64_matches_oracle(oracle, hisq32):
    O, c = oracle, hisq32
    fx, sp, L = c["fx"], c["sp"], c["L"]
    print(f"HISQ CG FP64: {sp.iterations} its (oracle {fx['iterations']}), r2 {sp.r2:.3e}", flush=True)
    assert abs(sp.iterations - fx["iterations"]) <= 1, (sp.iterations, fx["iterations"])
    x = c["x64"].get()
    r2 = true_r2(O, L, c["g"], c["b"], x, fx["mass"])
    print(f"  oracle-recomputed true r2 {r2:.3e}", flush=True)
    assert sp.r2 <= R2REQ and r2 <= 1.0001 * R2REQ
    want = fixture_sample(fx["x_sample"])
    got = sample(x, fx["sample_sites"])
    e = rel(got, want)
    print(f"  sample rel err vs oracle {e:.3e}", flush=True)
    assert e < 1e-6


@pytest.mark.parametrize("storage", ["fp16", "bf16", "q16links_fp16vec"])
def test_hisq_32x32x32x64_storage_true_residual(oracle, gpu, hisq32, storage):
    O, c = oracle, hisq32
    S = gpu.SloppyType
    sloppy, links = {"fp16": (S.SloppyHalf, None), "bf16": (S.SloppyBF16, None),
                     "q16links_fp16vec": (S.SloppyHalf, S.SloppyHalfFixed)}[storage]
    lo, fx, L = c["lo"], c["fx"], c["L"]
    x = lo.ColorVector()
    sp = gpu.SolverParams()
    sp.r2req, sp.verbosity, sp.maxits = R2REQ, 0, 200000
    sp.sloppySolve, sp.sloppyLinks = sloppy, links
    gpu.solveEE(c["s"], x, c["fb"], fx["mass"], sp)
    xv = x.get()
    r2 = true_r2(O, L, c["g"], c["b"], xv, fx["mass"])
    ne = lo.nEven
    err = rel(xv[:ne], c["x64"].get()[:ne])
    print(f"HISQ CG {storage}: {sp.iterations} its, reported r2 {sp.r2:.3e}, oracle r2 {r2:.3e}, "
          f"err vs FP64 {err:.3e}", flush=True)
    assert sp.r2 <= R2REQ and r2 <= 1.0001 * R2REQ
    assert err < 1e-8


# ---------------------------------------------------------------- config 4(a)
def test_multishift_batch_24x24x24x48_eight_sources(oracle, gpu, fixtures):
    O = oracle
    fx = need(fixtures, "cgm_24x24x24x48")
    L = tuple(fx["lattice"])
    U = O.phased(L, O.gauge_random(L, fx["gauge_seed"]))
    g = O.hisq_links(L, fx["hisq_seed"], U)
    del U
    m, shifts = fx["mass"], fx["shifts"]
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, g)
    nr = 8
    bs = [O.source_gauss(L, fx["source_seed"] + j, 0, 0) for j in range(nr)]
    fb = [lo.ColorVector().set(b) for b in bs]
    xs = [[lo.ColorVector() for _ in shifts] for _ in range(nr)]
    sps = [gpu.SolverParams() for _ in range(nr)]
    for sp in sps:
        sp.r2req, sp.verbosity, sp.maxits = R2REQ, 0, 100000
    gpu.solveMultiBatch(s, xs, fb, m, shifts, sps)
    print(f"multi-shift batch 24^3x48: iterations {[sp.iterations for sp in sps]} "
          f"(oracle source 0: {fx['iterations']})", flush=True)
    assert abs(sps[0].iterations - fx["iterations"]) <= 1
    ne = lo.nEven
    worst = 0.0
    for j in range(nr):
        for k, sg in enumerate(shifts):
            x = xs[j][k].get()
            r2 = true_r2(O, L, g, bs[j], x, m, sg)
            worst = max(worst, r2)
            # the solver's own FP64 recomputation agrees with the oracle's
            assert abs(sps[j].shiftR2[k] - r2) <= 1e-2 * r2 + 1e-15 * np.sqrt(r2), (j, k, sps[j].shiftR2[k], r2)
            assert r2 <= 1e-19, (j, k, r2)
            if j == 0:
                want = fixture_sample(fx["x_sample_per_shift"][k])
                e = rel(sample(x, fx["sample_sites"]), want)
                assert e < 1e-6, (k, e)
        assert sps[j].r2 == sps[j].shiftR2[0]
    print(f"  worst oracle-recomputed shifted r2 {worst:.3e}", flush=True)
    assert max(fx["true_r2_per_shift"]) <= 1e-19


@pytest.mark.parametrize("sloppy,slinks", [(1, -1), (2, 4)])
def test_mixed_multishift_24x24x24x48(oracle, gpu, fixtures, sloppy, slinks):
    """Source 0 of the same workload through the mixed-precision multi-shift
    solve (FP32 recursion, every shift refined by the mixed CG): each shifted
    system's oracle-recomputed true residual at r2req and the fixture's
    solution sample."""
    import time
    import torch
    O = oracle
    fx = need(fixtures, "cgm_24x24x24x48")
    L = tuple(fx["lattice"])
    U = O.phased(L, O.gauge_random(L, fx["gauge_seed"]))
    g = O.hisq_links(L, fx["hisq_seed"], U)
    del U
    m, shifts = fx["mass"], fx["shifts"]
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, g)
    b = O.source_gauss(L, fx["source_seed"], 0, 0)
    fb = lo.ColorVector().set(b)
    xs = [lo.ColorVector() for _ in shifts]
    sp = gpu.SolverParams()
    sp.r2req, sp.verbosity, sp.maxits = R2REQ, 0, 100000
    sp.sloppySolve = gpu.SloppyType(sloppy)
    sp.sloppyLinks = None if slinks < 0 else gpu.SloppyType(slinks)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gpu.solveMulti(s, xs, fb, m, shifts, sp)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    print(f"mixed multi-shift 24^3x48 sloppy={sloppy} links={slinks}: {sp.iterations} iterations, "
          f"{sp.reliableUpdates} reliable updates, {sec:.3f} s (FP64 CGM: {fx['iterations']} iterations)", flush=True)
    worst = 0.0
    for k, sg in enumerate(shifts):
        x = xs[k].get()
        r2 = true_r2(O, L, g, b, x, m, sg)
        worst = max(worst, r2)
        assert abs(sp.shiftR2[k] - r2) <= 1e-2 * r2 + 1e-15 * np.sqrt(r2), (k, sp.shiftR2[k], r2)
        assert r2 <= 1.0001 * R2REQ, (k, r2)
        e = rel(sample(x, fx["sample_sites"]), fixture_sample(fx["x_sample_per_shift"][k]))
        assert e < 1e-6, (k, e)
    print(f"  worst oracle-recomputed shifted r2 {worst:.3e}", flush=True)


# ---------------------------------------------------------------- config 4(b)
def test_uneven_batch_sixteen_lattices(oracle, gpu, fixtures):
    O = oracle
    fx = need(fixtures, "batch16")
    ss, lats, fbs, bs, xs, gs = [], [tuple(v) for v in fx["lattices"]], [], [], [], []
    for k, L in enumerate(lats):
        U = O.phased(L, O.gauge_random(L, fx["gauge_seeds"][k]))
        lo = gpu.Layout(L)
        ss.append(gpu.Staggered(lo, U))
        gs.append(U)
        b = O.source_gauss(L, fx["source_seeds"][k], 0, 0)
        bs.append(b)
        fbs.append(lo.ColorVector().set(b))
        xs.append(lo.ColorVector())
    sps = [gpu.SolverParams() for _ in lats]
    for sp in sps:
        sp.r2req, sp.verbosity = R2REQ, 0
    gpu.solveBatch(ss, xs, fbs, fx["masses"], sps)
    its = [sp.iterations for sp in sps]
    print(f"uneven batch: iterations {its}\n  oracle {fx['iterations']}", flush=True)
    for k, L in enumerate(lats):
        ne = O.volume(L) // 2
        bar = 1 if fx["iterations"][k] <= 0.25 * 3 * ne else 2
        assert abs(its[k] - fx["iterations"][k]) <= bar, (k, its[k], fx["iterations"][k])
        r2 = true_r2(O, L, gs[k], bs[k], xs[k].get(), fx["masses"][k])
        assert sps[k].r2 <= R2REQ and r2 <= 1.0001 * R2REQ, (k, sps[k].r2, r2)


# ---------------------------------------------------------------- config 5
def test_hmc_24x24x24x24_reversible(gpu):
    from qex_amd import gauge, hmc
    L = (24, 24, 24, 24)
    lo = gpu.Layout(L)
    U = gauge.unit_gauge(lo, device="cuda")
    h = hmc.StaggeredHMC(lo, U, 5.6, 0.05, r2req_force=1e-26, r2req_action=1e-26)
    dU, dP = h.reversibility(41, tau=1.0, nsteps=20)
    print(f"24^4 leapfrog tau=1, 20 steps: reversibility |dU| {dU:.2e} |dP| {dP:.2e}", flush=True)
    assert dU < 1e-10 and dP < 1e-8, (dU, dP)
