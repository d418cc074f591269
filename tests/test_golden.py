"""Golden fixtures (tests/golden/golden_4x4x4x4.npz, written by
tests/golden/make_golden.py): the CPU oracle reproduces them, and the GPU path
reproduces them through the C ABI using its own input generator."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_4x4x4x4.npz")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD, allow_pickle=False))


def test_oracle_reproduces_golden(oracle, gold):
    L = tuple(int(v) for v in gold["L"])
    seed = int(gold["seed"])
    U = oracle.phased(L, oracle.gauge_random(L, seed))
    g8 = oracle.hisq_links(L, seed + 1, U)
    x = gold["x"]
    for name, g in (("plain", U), ("hisq", g8)):
        r = np.zeros_like(x)
        oracle.stagD2(L, g, 0, r, x, 0.0, 0.0)
        assert rel(r, gold[f"{name}_hop_even"]) < 1e-14
        assert rel(oracle.D(L, g, x, 0.1), gold[f"{name}_D_m0.1"]) < 1e-14
        assert rel(oracle.stagD2xx(L, g, 0, x, 0.01), gold[f"{name}_D2ee_m2_0.01"]) < 1e-14
        sol, its, _ = oracle.solveXX(L, g, 0, gold[f"{name}_cg_b"], 0.2, 1e-22)
        assert rel(sol, gold[f"{name}_cg_x"]) < 1e-12 and its == int(gold[f"{name}_cg_its"])
    assert rel(oracle.ferm_force(L, U, gold["force_psi"]), gold["force_F"]) < 1e-14
    xs, its, _ = oracle.cgm(L, U, 0, gold["cgm_b"], 0.1, gold["cgm_sigmas"], 1e-22)
    assert rel(xs, gold["cgm_x"]) < 1e-12 and its == int(gold["cgm_its"])
    Uu = oracle.gauge_random(L, seed)
    assert abs(oracle.plaquette(L, Uu) - float(gold["plaquette"])) < 1e-12 * abs(float(gold["plaquette"]))
    assert rel(oracle.wilson_force(L, Uu, 5.6), gold["wilson_force_b5.6"]) < 1e-14
    P = oracle.random_momenta(L, seed + 2)
    assert np.array_equal(P, gold["momenta"])
    oracle.link_update(L, P, 0.1, Uu)
    assert rel(Uu, gold["link_update_0.1"]) < 1e-14


@pytest.mark.gpu
def test_gpu_reproduces_golden(gpu, gold):
    import torch
    from qex_amd import gauge
    L = tuple(int(v) for v in gold["L"])
    seed = int(gold["seed"])
    lo = gpu.Layout(L)
    U = gauge.random_gauge(lo, seed)
    gauge.phase_bc(lo, U)
    g8 = gauge.hisq_links(lo, seed + 1, U)
    x = lo.ColorVector().set(gold["x"])
    ne = lo.nEven
    for name, g in (("plain", U), ("hisq", g8)):
        s = gpu.Staggered(lo, g)
        r = lo.ColorVector()
        gpu.stagD2(s.se, r, s.g, x, 0.0, 0.0)
        assert rel(r.get()[:ne], gold[f"{name}_hop_even"][:ne]) < 1e-12
        gpu.D(s, r, x, 0.1)
        assert rel(r.get(), gold[f"{name}_D_m0.1"]) < 1e-12
        gpu.stagD2ee(s.se, s.so, r, s.g, x, 0.01)
        assert rel(r.get(sub="even")[:ne], gold[f"{name}_D2ee_m2_0.01"][:ne]) < 1e-12
        b = lo.ColorVector().set(gold[f"{name}_cg_b"])
        sol = lo.ColorVector()
        sp = gpu.SolverParams()
        sp.r2req, sp.verbosity = 1e-22, 0
        gpu.solveEE(s, sol, b, 0.2, sp)
        assert rel(sol.get()[:ne], gold[f"{name}_cg_x"][:ne]) < 1e-10
        assert abs(sp.iterations - int(gold[f"{name}_cg_its"])) <= 1
    s = gpu.Staggered(lo, U)
    psi = lo.ColorVector().set(gold["force_psi"])
    F = gpu.fermForce(s, psi, device="cuda")
    torch.cuda.synchronize()
    assert rel(F.cpu().numpy(), gold["force_F"]) < 1e-12
    # multi-shift, single and batched
    sig = [float(v) for v in gold["cgm_sigmas"]]
    bm = lo.ColorVector().set(gold["cgm_b"])
    ys = [lo.ColorVector() for _ in sig]
    sp = gpu.SolverParams()
    sp.r2req, sp.verbosity = 1e-22, 0
    gpu.solveMulti(s, ys, bm, 0.1, sig, sp)
    assert abs(sp.iterations - int(gold["cgm_its"])) <= 1
    for k in range(len(sig)):
        assert rel(ys[k].get()[:ne], gold["cgm_x"][k][:ne]) < 1e-10
    zs = [[lo.ColorVector() for _ in sig]]
    sps = [gpu.SolverParams()]
    sps[0].r2req, sps[0].verbosity = 1e-22, 0
    gpu.solveMultiBatch(s, zs, [bm], 0.1, sig, sps, freeze=False)
    for k in range(len(sig)):
        assert rel(zs[0][k].get()[:ne], gold["cgm_x"][k][:ne]) < 1e-10
    # gauge-field pieces
    from qex_amd import hmc
    Uu = gauge.random_gauge(lo, seed)
    assert abs(hmc.plaquette(lo, Uu) - float(gold["plaquette"])) < 1e-12 * abs(float(gold["plaquette"]))
    assert rel(hmc.wilson_force(lo, Uu, 5.6).cpu().numpy(), gold["wilson_force_b5.6"]) < 1e-12
    P = hmc.random_momenta(lo, seed + 2)
    assert np.abs(P.cpu().numpy() - gold["momenta"]).max() < 1e-15
    hmc.link_update(lo, P, 0.1, Uu)
    assert rel(Uu.cpu().numpy(), gold["link_update_0.1"]) < 1e-13
