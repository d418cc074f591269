"""HISQ smearing chain rule and HISQ fermion force on the GPU
(qexhip_fat7l_deriv, qexhip_project_u_deriv, qexhip_hisq_smear_deriv,
qexhip_hisq_force) against the CPU oracle, plus a finite difference of the
GPU action:

 - fat7lDeriv of every term, with Naik long links from a separate field, and
   its accumulation into the output, equal the oracle's;
 - projectUderiv equals the oracle's;
 - the two-level smearing derivative equals the oracle's;
 - the HISQ fermion force TAH(2 dS/dg g^+) equals the oracle's chain on the
   same psi, and its contraction with a random su(3) field equals a 9-point
   central difference of S = phi^+ (m^2 - D_eo D_oe)^-1 phi along g -> exp(wX) g.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 515
TERMS = [
    (1.0, 0, 0, 0, 0),
    (0, 1.0, 0, 0, 0),
    (0, 0, 1.0, 0, 0),
    (0, 0, 0, 1.0, 0),
    (0, 0, 0, 0, 1.0),
    (0.3, -0.7, 0.45, -0.2, 0.15),
]
W9 = np.array([1 / 280, -4 / 105, 1 / 5, -4 / 5, 0, 4 / 5, -1 / 5, 4 / 105, -1 / 280])


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _crand(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


@pytest.mark.parametrize("L", [(8, 8, 8, 8), (4, 6, 4, 8)])
def test_fat7_deriv_matches_oracle(oracle, gpu, L):
    import torch
    from qex_amd import hisq
    print(f"fat7 deriv L={L}", flush=True)
    lo = gpu.Layout(L)
    V = lo.nSites
    rng = np.random.default_rng(SEED)
    U = oracle.gauge_random(L, SEED)
    C = _crand(rng, 4, V, 3, 3)
    for c in TERMS:
        d = torch.zeros((4, V, 3, 3), dtype=torch.complex128, device="cuda")
        hisq.fat7lDeriv(lo, d, _cuda(U), _cuda(C), hisq.Fat7lCoefs(*c))
        want, _ = oracle.fat7l_deriv(L, U, C, oracle.fat7_coefs(*c))
        assert rel(d.cpu().numpy(), want) < 1e-12, c


def test_fat7_deriv_naik_and_accumulation(oracle, gpu):
    import torch
    from qex_amd import hisq
    L = (8, 8, 8, 8)
    lo = gpu.Layout(L)
    V = lo.nSites
    rng = np.random.default_rng(SEED + 1)
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 1))
    W = oracle.gauge_random(L, SEED + 2)
    C, Cl, D0 = _crand(rng, 4, V, 3, 3), _crand(rng, 4, V, 3, 3), _crand(rng, 4, V, 3, 3)
    second = oracle.hisq_coefs()[1]
    d, ld = _cuda(D0), _cuda(D0)
    hisq.fat7lDeriv(lo, d, _cuda(U), _cuda(C), hisq.Fat7lCoefs(*second), ld, _cuda(W), _cuda(Cl), -1.0 / 24)
    wd, wld = oracle.fat7l_deriv(L, U, C, second, naik=-1.0 / 24, gl=W, llmid=Cl)
    assert rel(d.cpu().numpy() - D0, wd) < 1e-12
    assert rel(ld.cpu().numpy() - D0, wld) < 1e-12
    # llderiv may be deriv itself (the reference's fat7lDeriv(deriv, ..., deriv, ...))
    d = _cuda(D0)
    hisq.fat7lDeriv(lo, d, _cuda(U), _cuda(C), hisq.Fat7lCoefs(*second), d, None, _cuda(Cl), -1.0 / 24)
    wd2, wld2 = oracle.fat7l_deriv(L, U, C, second, naik=-1.0 / 24, llmid=Cl)
    assert rel(d.cpu().numpy() - D0, wd2 + wld2) < 1e-12


def test_project_u_deriv_matches_oracle(oracle, gpu):
    from qex_amd import hisq
    rng = np.random.default_rng(SEED + 2)
    X = np.eye(3) + 0.3 * _crand(rng, 4, 500, 3, 3)
    C = _crand(rng, 4, 500, 3, 3)
    u = oracle.project_u(X)
    got = hisq.projectUderiv(_cuda(u), _cuda(X), _cuda(C)).cpu().numpy()
    assert rel(got, oracle.project_u_deriv(u, X, C)) < 1e-12


@pytest.mark.parametrize("L", [(8, 8, 8, 8), (4, 4, 6, 8)])
def test_hisq_smear_deriv_matches_oracle(oracle, gpu, L):
    import torch
    from qex_amd import hisq
    print(f"smear deriv L={L}", flush=True)
    lo = gpu.Layout(L)
    V = lo.nSites
    rng = np.random.default_rng(SEED + 3)
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 3))
    Cf, Cl = _crand(rng, 4, V, 3, 3), _crand(rng, 4, V, 3, 3)
    fl, ll, f = hisq.smearGetForce(lo, hisq.newHISQ(), _cuda(U))
    dg = torch.zeros_like(fl)
    f(dg, _cuda(Cf), _cuda(Cl))
    want = oracle.hisq_smear_deriv(L, U, Cf, Cl)
    # projectU's derivative is evaluated at first-level links whose condition
    # numbers reach a few hundred on these rough fields
    assert rel(dg.cpu().numpy(), want) < 1e-11


def _g8(oracle, L, U):
    fl, ll = oracle.hisq_smear(L, U)
    return np.ascontiguousarray(np.stack([t for mu in range(4) for t in (fl[mu], ll[mu])]))


def test_hisq_force_matches_oracle_chain(oracle, gpu):
    from qex_amd import hisq
    L, m = (8, 8, 8, 8), 0.1
    lo = gpu.Layout(L)
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 4))
    Ug = _cuda(U)
    s, fl, ll = hisq.hisqStag(lo, Ug)
    phi = lo.ColorVector().set(oracle.source_gauss(L, 9, 0, 0))
    psi = lo.ColorVector()
    sp = gpu.SolverParams()
    sp.r2req, sp.verbosity = 1e-26, 0
    gpu.stagForceSolve(s, psi, phi, m, sp)
    F = hisq.hisqForce(s, None, Ug, psi).cpu().numpy()
    # the oracle's chain on the same psi
    G8 = oracle.hisq_outer(L, psi.get())
    GU = oracle.hisq_smear_deriv(L, U, 0.5 * G8[0::2], 0.5 * G8[1::2])
    T = 2 * GU @ np.conj(np.swapaxes(U, -1, -2))
    want = 0.5 * (T - np.conj(np.swapaxes(T, -1, -2)))
    want -= np.trace(want, axis1=-2, axis2=-1)[..., None, None] / 3 * np.eye(3)
    assert rel(F, want) < 1e-11
    # GPU outer products of HISQ links match the oracle's
    G8g = gpu.fermForce(s, psi, outer_only=True, device="cuda").cpu().numpy()
    assert rel(G8g, G8) < 1e-14


def test_hisq_force_against_action_differences(oracle, gpu):
    from qex_amd import hisq, hmc
    L, m = (8, 8, 8, 8), 0.2
    lo = gpu.Layout(L)
    V = lo.nSites
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 5))
    Ug = _cuda(U)
    phi = lo.ColorVector().set(oracle.source_gauss(L, 11, 0, 0))

    def action(Ux):
        s, _, _ = hisq.hisqStag(lo, Ux)
        x = lo.ColorVector()
        sp = gpu.SolverParams()
        sp.r2req, sp.verbosity = 1e-28, 0
        gpu.solveEE(s, x, phi, m, sp)
        # solveEE solves 4(m^2 - D_eo D_oe) x = phi
        return 4.0 * float(np.vdot(phi.get()[: V // 2], x.get()[: V // 2]).real)

    s, _, _ = hisq.hisqStag(lo, Ug)
    psi = lo.ColorVector()
    sp = gpu.SolverParams()
    sp.r2req, sp.verbosity = 1e-28, 0
    gpu.stagForceSolve(s, psi, phi, m, sp)
    F = hisq.hisqForce(s, None, Ug, psi)
    rng = np.random.default_rng(SEED + 6)
    A = _crand(rng, 4, V, 3, 3)
    X = 0.5 * (A - np.conj(np.swapaxes(A, -1, -2)))
    X -= np.trace(X, axis1=-2, axis2=-1)[..., None, None] / 3 * np.eye(3)
    analytic = -0.5 * float(np.real(np.sum(np.trace(X @ F.cpu().numpy(), axis1=-2, axis2=-1))))
    Xg = _cuda(X)
    h = 5e-5
    vals = []
    for k in range(-4, 5):
        Uk = Ug.clone()
        if k:
            hmc.link_update(lo, Xg, k * h, Uk)
        vals.append(action(Uk))
    fd = float(np.dot(W9, vals) / h)
    print(f"HISQ force: analytic {analytic:.12e} fd {fd:.12e}", flush=True)
    # measured agreement 2.4e-10; the north-star bar for the force is 1e-8
    assert abs(fd - analytic) < 1e-8 * abs(analytic)
