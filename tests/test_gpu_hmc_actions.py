"""The Pauli-Villars and ratio (Hasenbusch) actions of the HMC, and its nested
gauge level, on the GPU through the C ABI (QEX hmc/hmcAction.nim:947-1152,
examples/generichmc.nim:93-111):

 - right after its heatbath each action is Gamma distributed with mean equal
   to the number of even-site colour components (the heatbaths are exact);
 - each action equals its closed form evaluated by the CPU oracle (the ratio
   action's full-lattice solves at negated masses included);
 - the weighted fermion force is the directional derivative of the summed
   action along U -> exp(w X) U for a random su(3) field X (plain and HISQ);
 - Hasenbusch + Pauli-Villars HMC with the gauge force on an inner level:
   dH scales as dt^2 and the integrator is reversible; one inner step is the
   one-level leapfrog.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = (8, 8, 8, 8)


def _hmc(gpu, actions, hisq=False, **kw):
    from qex_amd import gauge, hmc
    lo = gpu.Layout(L)
    h = hmc.StaggeredHMC(lo, gauge.unit_gauge(lo, device="cuda"), 5.6, 0.0, r2req_force=1e-26,
                         r2req_action=1e-28, actions=actions, hisq=hisq, **kw)
    # a rough configuration away from the unit gauge
    hmc.link_update(lo, hmc.random_momenta(lo, 3), 0.5, h.U)
    h._links_changed()
    return h


def _kinds():
    from qex_amd import hmc
    return {"plain": lambda: [hmc.StaggeredFermionAction(0.2)],
            "pv": lambda: [hmc.StaggeredPauliVillarsAction(0.5)],
            "ratio": lambda: [hmc.StaggeredRatioAction(0.1, 0.4)]}


@pytest.mark.parametrize("kind", ["plain", "pv", "ratio"])
def test_heatbath_is_exact(gpu, kind):
    h = _hmc(gpu, _kinds()[kind]())
    ne = 3 * h.lo.nEven
    S = []
    for seed in range(4):
        h.heatbath(100 + seed)
        S.append(h.fermion_action())
    print(f"{kind}: S/N_e after heatbath {[round(s / ne, 5) for s in S]}", flush=True)
    # S ~ Gamma(ne, 1): std sqrt(ne) = 78 at 8^4; the mean of 4 within 5 sigma
    assert abs(np.mean(S) - ne) < 5 * np.sqrt(ne / 4), S
    assert all(abs(s - ne) < 6 * np.sqrt(ne) for s in S), S


def test_actions_match_oracle_closed_forms(oracle, gpu):
    from qex_amd import hmc
    pv = hmc.StaggeredPauliVillarsAction(0.5)
    ra = hmc.StaggeredRatioAction(0.1, 0.4)
    pl = hmc.StaggeredFermionAction(0.2)
    h = _hmc(gpu, [pl, pv, ra])
    h.heatbath(7)
    got = [a.action() for a in (pl, pv, ra)]
    g = oracle.phased(L, h.U.cpu().numpy())
    ne = h.lo.nEven
    want = []
    phi = pl.phi.get()
    y, _, _ = oracle.solveXX(L, g, 0, phi, 0.2, 1e-28)
    want.append(4.0 * np.vdot(phi[:ne], y[:ne]).real)
    phi = pv.phi.get()
    want.append(0.25 * np.vdot(phi[:ne], oracle.stagD2xx(L, g, 0, phi, 0.25)[:ne]).real)
    phi = ra.phi.get()
    y, _, _ = oracle.solveXX(L, g, 0, phi, 0.1, 1e-28)
    want.append(np.vdot(phi[:ne], phi[:ne]).real + (0.16 - 0.01) * 4.0 * np.vdot(phi[:ne], y[:ne]).real)
    print(f"actions {got} oracle {want}", flush=True)
    for a, b in zip(got, want):
        assert abs(a - b) < 1e-10 * abs(b), (got, want)
    assert all(abs(p.phi.get()[ne:]).max() == 0 for p in (pl, pv, ra))


def _directional(h, seed=77, step=2e-4):
    """(analytic, Richardson central difference) of sum S_f along U -> exp(w X) U."""
    import torch
    from qex_amd import hmc
    X = hmc.random_momenta(h.lo, seed)
    F = h.fermion_force().clone()
    an = -0.5 * float((X * F.transpose(-1, -2)).sum().real)
    U0 = h.U.clone()
    S = {}
    for w in (step, -step, 2 * step, -2 * step):
        h.U.copy_(U0)
        hmc.link_update(h.lo, X, w, h.U)
        h._links_changed()
        S[w] = h.fermion_action()
    h.U.copy_(U0)
    h._links_changed()
    d1 = (S[step] - S[-step]) / (2 * step)
    d2 = (S[2 * step] - S[-2 * step]) / (4 * step)
    torch.cuda.synchronize()
    return an, (4.0 * d1 - d2) / 3.0


@pytest.mark.parametrize("kind", ["plain", "pv", "ratio"])
def test_force_is_directional_derivative(gpu, kind):
    h = _hmc(gpu, _kinds()[kind]())
    h.heatbath(9)
    an, fd = _directional(h)
    print(f"{kind}: dS/dw analytic {an:.12e} fd {fd:.12e}", flush=True)
    assert abs(an) > 1.0
    assert abs(fd - an) < 1e-6 * abs(an), (an, fd)


def test_hisq_weighted_forces_are_directional_derivative(gpu):
    from qex_amd import hmc
    h = _hmc(gpu, [hmc.StaggeredRatioAction(0.1, 0.4), hmc.StaggeredPauliVillarsAction(0.5)], hisq=True)
    h.heatbath(13)
    an, fd = _directional(h)
    print(f"HISQ ratio + PV: dS/dw analytic {an:.12e} fd {fd:.12e}", flush=True)
    assert abs(fd - an) < 1e-6 * abs(an), (an, fd)


def _hasenbusch(gpu, **kw):
    from qex_amd import hmc
    return _hmc(gpu, [hmc.StaggeredFermionAction(0.4), hmc.StaggeredRatioAction(0.1, 0.4),
                      hmc.StaggeredPauliVillarsAction(0.8)], **kw)


def _dH(h, n, tau=0.5):
    import torch
    from qex_amd import hmc
    U0 = h.U.clone()
    P = hmc.random_momenta(h.lo, 5)
    h.heatbath(6)
    H0 = h.hamiltonian(P)
    P = h.integrate(P, tau, n)
    dH = h.hamiltonian(P) - H0
    h.U.copy_(U0)
    h._links_changed()
    torch.cuda.synchronize()
    return dH


@pytest.mark.parametrize("integrator", ["leapfrog", "2MN"])
def test_hasenbusch_pv_nested_dH_scales_as_dt2(gpu, integrator):
    h = _hasenbusch(gpu, integrator=integrator, gauge_steps=3)
    dH = [_dH(h, n) for n in (6, 12)]
    print(f"{integrator} Hasenbusch + PV, gauge level x3: dH {dH}", flush=True)
    assert abs(dH[0]) > 1e-5, dH
    assert 3.2 < dH[0] / dH[1] < 4.8, dH


def test_hasenbusch_pv_nested_reversible(gpu):
    h = _hasenbusch(gpu, integrator="2MN", gauge_steps=2)
    dU, dP = h.reversibility(11, tau=0.5, nsteps=6)
    assert dU < 1e-10 and dP < 1e-9, (dU, dP)


def test_one_inner_gauge_step_is_one_level_leapfrog(gpu):
    import torch
    from qex_amd import hmc
    out = []
    for gs in (0, 1):
        h = _hasenbusch(gpu, integrator="leapfrog", gauge_steps=gs)
        P = hmc.random_momenta(h.lo, 21)
        h.heatbath(22)
        P = h.integrate(P, 0.3, 3)
        torch.cuda.synchronize()
        out.append((h.U.clone(), P))
    assert float(torch.max(torch.abs(out[0][0] - out[1][0]))) < 1e-13
    assert float(torch.max(torch.abs(out[0][1] - out[1][1]))) < 1e-11
