"""Staggered HMC (plain or HISQ fermions) with a Wilson gauge action, on the GPU
through the C ABI.

Mirrors QEX's staggered HMC (src/examples/staghmc.nim:81-353,
src/examples/generichmc.nim:71-111, and the fermion actions of
src/hmc/hmcAction.nim:794-1213):

- pseudofermion fields on even sites, one per fermion action: the plain
  StaggeredFermionAction (phi = (D^+(m) eta)_e, S_f = phi^+ (m^2 - D_eo D_oe)^{-1}
  phi), the Pauli-Villars action and the Hasenbusch ratio action, eta complex
  Gaussian;
- S_g = -(beta/3) sum_P Re tr U_P (gauge/gaugeAction.nim:397-427);
- momenta P traceless anti-Hermitian, H = 1/2 sum |P|^2 + S_g + sum S_f;
- leapfrog or Omelyan 2MN (QEX's mdevolve integrators,
  src/algorithms/integrator.nim:32-34, with the update order of staghmc.nim
  mdt/mdv/mdvf), on one level or with the gauge force on a nested inner level.

Force convention (shared by qexhip_force and qexhip_wilson_force): for
U -> exp(w X) U, dS/dw = -1/2 Re tr(X F), so Pdot = -F/2 conserves H.

Fields are canonical torch complex128 tensors on the GPU; every kernel call
goes through libqexhip.so.
"""
import ctypes
import time

import numpy as np
import torch

from . import gauge as G
from . import hisq as H
from . import stag as Q
from ._lib import call


# mdevolve's default 2MN parameter (listed in algorithms/integrator.nim:60-61)
OMELYAN_LAMBDA = 0.1931833275037836


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def plaquette(lo, U):
    torch.cuda.synchronize()
    v = ctypes.c_double()
    call("qexhip_plaquette", lo.h, _vp(U), ctypes.byref(v))
    return v.value


def wilson_force(lo, U, beta, out=None):
    F = torch.empty_like(U) if out is None else out
    torch.cuda.synchronize()
    call("qexhip_wilson_force", lo.h, _vp(U), float(beta), _vp(F))
    return F


def random_momenta(lo, seed, out=None):
    P = torch.empty((4, lo.nSites, 3, 3), dtype=torch.complex128, device="cuda") if out is None else out
    torch.cuda.synchronize()
    call("qexhip_random_momenta", lo.h, ctypes.c_uint64(seed), _vp(P))
    return P


def link_update(lo, P, eps, U):
    """U <- exp(eps P) U in place."""
    torch.cuda.synchronize()
    call("qexhip_link_update", lo.h, _vp(P), float(eps), _vp(U))
    return U


def kinetic(P):
    return 0.5 * float(torch.sum(P.real ** 2 + P.imag ** 2))


def _gaussian(lo, seed, k):
    """Complex Gaussian eta on all sites, E|eta|^2 = 1 per component (density
    exp(-eta^+ eta)); action k of a trajectory draws from its own stream."""
    return lo.ColorVector().set(G.source(lo, seed, stream=7 if k == 0 else 100 + k, par=2))


class StaggeredFermionAction:
    """QEX's StaggeredFermionAction (hmcAction.nim:794-885): one pseudofermion on
    even sites,
        heatbath  phi_e = (D^+(m) eta)_e,
        action    S = phi^+ M(m)^-1 phi,  M(m) = m^2 - D_eo D_oe,
        force     fermForce of stagForceSolve(phi, m), weight +1.
    QEX writes the heatbath as D(-m) eta and the action as 1/2 |D(-m)^-1 phi|^2
    for unit-variance real Gaussians; with E|eta|^2 = 1 per complex component
    the same distribution reads as above."""

    name = "StaggeredFermionAction"

    def __init__(self, mass):
        self.mass = float(mass)

    def describe(self):
        return f"{self.name}(mass: {self.mass})"

    def bind(self, hmc):
        self.h = hmc
        self.phi = hmc.lo.ColorVector()

    def heatbath(self, eta):
        Q.Ddag(self.h.s, self.phi, eta, self.mass)
        self.phi.zero("odd")

    def action(self):
        h = self.h
        sp = h._sp(h.r2a)
        Q.solveEE(h.s, h.y, self.phi, self.mass, sp)
        h.solver_its += sp.iterations
        return 4.0 * self.phi.redot(h.y, "even")

    def force(self, out):
        """Writes F of psi = stagForceSolve(phi, m) into out; returns its weight."""
        h = self.h
        sp = h._sp(h.r2f)
        Q.stagForceSolve(h.s, h.psi, self.phi, self.mass, sp)
        h.solver_its += sp.iterations
        h._force_of(h.psi, out)
        return 1.0


class StaggeredPauliVillarsAction(StaggeredFermionAction):
    """QEX's StaggeredPauliVillarsAction (hmcAction.nim:947-1033), a bosonic
    field that cancels one quartet's determinant at the heavy mass m:
        heatbath  phi_e = (D(m)^-1 eta)_e  (a full-lattice solve, :962-964),
        action    S = |D(m) phi_e|^2 = phi^+ M(m) phi  (:983-1002),
        force     fermForce of psi = [phi_e ; (2D)_oe phi_e], weight -1
                  (QEX's sc = -0.25 dtau, :1004-1033)."""

    name = "StaggeredPauliVillarsAction"

    def heatbath(self, eta):
        h = self.h
        sp = h._sp(h.r2a)
        Q.solve(h.s, self.phi, eta, self.mass, sp)
        h.solver_its += sp.iterations
        self.phi.zero("odd")

    def action(self):
        h = self.h
        Q.D(h.s, h.y, self.phi, self.mass)
        return h.y.norm2()

    def force(self, out):
        h = self.h
        Q.stagD2(h.s.so, h.psi, h.s.g, self.phi, 0.0, 0.0)
        h.psi.copy_from(self.phi, "even")
        h._force_of(h.psi, out)
        return -1.0


class StaggeredRatioAction:
    """QEX's StaggeredRatioAction (hmcAction.nim:1095-1152), the Hasenbusch
    ratio det M(massNum) / det M(massDen):
        heatbath  phi_e = (D(-massDen)^-1 D(-massNum) eta)_e  (:1104-1108),
        action    S = |D(-massNum)^-1 D(-massDen) phi_e|^2  (:1113-1131)
                    = |phi|^2 + (massDen^2 - massNum^2) phi^+ M(massNum)^-1 phi,
        force     fermForce of stagForceSolve(phi, massNum),
                  weight massDen^2 - massNum^2  (:1133-1152).
    Both solves are full-lattice solves at negated masses, as in QEX."""

    name = "StaggeredRatioAction"

    def __init__(self, massNum, massDen):
        self.massNum, self.massDen = float(massNum), float(massDen)

    def describe(self):
        return f"{self.name}(massNum: {self.massNum}, massDen: {self.massDen})"

    def bind(self, hmc):
        self.h = hmc
        self.phi = hmc.lo.ColorVector()

    def heatbath(self, eta):
        h = self.h
        Q.D(h.s, h.y, eta, -self.massNum)
        sp = h._sp(h.r2a)
        Q.solve(h.s, self.phi, h.y, -self.massDen, sp)
        h.solver_its += sp.iterations
        self.phi.zero("odd")

    def action(self):
        h = self.h
        Q.D(h.s, h.y, self.phi, -self.massDen)
        sp = h._sp(h.r2a)
        Q.solve(h.s, h.psi, h.y, -self.massNum, sp)
        h.solver_its += sp.iterations
        return h.psi.norm2()

    def force(self, out):
        h = self.h
        sp = h._sp(h.r2f)
        Q.stagForceSolve(h.s, h.psi, self.phi, self.massNum, sp)
        h.solver_its += sp.iterations
        h._force_of(h.psi, out)
        return self.massDen * self.massDen - self.massNum * self.massNum


class StaggeredHMC:
    """Staggered HMC; hisq=True smears the phased links into HISQ fat and long
    links (physics/hisqLinks.nim:39-51) for the fermion operator and chains the
    fermion force through the smearing (gauge/hisqsmear.nim:54-91).

    actions lists the fermion actions (default: one StaggeredFermionAction at
    mass), all on the one operator of these links.  gauge_steps > 0 puts the
    gauge force on an inner level of its own, gauge_steps integrator steps per
    outer position step, as QEX's nested action levels do
    (hmc/hmcAction.nim ActionLevel, examples/generichmc.nim:93-111)."""

    def __init__(self, lo, U, beta, mass, r2req_force=1e-20, r2req_action=1e-24, sloppy=Q.SloppyType.SloppyNone,
                 integrator="leapfrog", lam=OMELYAN_LAMBDA, sloppy_links=None, recon=18, hisq=False,
                 actions=None, gauge_steps=0, gauge_integrator=None, gauge_lam=OMELYAN_LAMBDA):
        for name in (integrator, gauge_integrator or integrator):
            if name not in ("leapfrog", "2MN"):
                raise ValueError("integrator must be 'leapfrog' or '2MN'")
        self.integrator, self.lam = integrator, float(lam)
        self.gauge_steps = int(gauge_steps)
        self.gauge_integrator, self.gauge_lam = gauge_integrator or integrator, float(gauge_lam)
        self.lo = lo
        self.U = U.contiguous()
        self.beta = float(beta)
        self.mass = float(mass)
        self.Uph = self.U.clone()
        G.phase_bc(lo, self.Uph)
        self.hisq = H.newHISQ() if hisq else None
        self.s = Q.Staggered(lo, self._fermion_links())
        self._stale = False
        self.psi = lo.ColorVector()
        self.y = lo.ColorVector()
        self.r2f, self.r2a = r2req_force, r2req_action
        self.sloppy, self.sloppy_links, self.recon = sloppy, sloppy_links, recon
        self.solver_its = 0
        self.Ff = torch.empty_like(self.U)
        self.Fg = torch.empty_like(self.U)
        self._Ft = None
        self.actions = list(actions) if actions is not None else [StaggeredFermionAction(mass)]
        if not self.actions:
            raise ValueError("at least one fermion action is needed")
        for a in self.actions:
            a.bind(self)

    def describe(self):
        return [a.describe() for a in self.actions]

    def _sp(self, r2req):
        sp = Q.SolverParams()
        sp.r2req, sp.verbosity, sp.maxits = r2req, 0, 100000
        sp.sloppySolve, sp.sloppyLinks, sp.recon = self.sloppy, self.sloppy_links, self.recon
        return sp

    def _fermion_links(self):
        if self.hisq is None:
            return self.Uph
        fl, ll = self.hisq.smear(self.lo, self.Uph)
        return torch.stack([t for mu in range(4) for t in (fl[mu], ll[mu])]).contiguous()

    def _links_changed(self):
        """U has changed: the fermion operator (and its smearing) is rebuilt
        before the next fermion action, force or heatbath needs it."""
        self._stale = True

    def _fermion_ready(self):
        if self._stale:
            self.Uph.copy_(self.U)
            G.phase_bc(self.lo, self.Uph)
            self.s.update(self._fermion_links())
            self._stale = False

    def _force_of(self, psi, out):
        if self.hisq is None:
            Q.fermForce(self.s, psi, out=out)
        else:
            # force with respect to the phased links; the staggered signs drop out of TAH(2 dS/dU U^+)
            H.hisqForce(self.s, self.hisq, self.Uph, psi, F=out)

    def heatbath(self, seed):
        """A fresh pseudofermion for every action (hmcAction.nim:794-818, :947-981, :1095-1111)."""
        self._fermion_ready()
        for k, a in enumerate(self.actions):
            a.heatbath(_gaussian(self.lo, seed, k))

    def fermion_action(self):
        self._fermion_ready()
        return sum(a.action() for a in self.actions)

    def gauge_action(self):
        return -(self.beta / 3.0) * plaquette(self.lo, self.U)

    def fermion_force(self):
        """F of all fermion actions, each scaled by its weight (QEX's addForce
        with sc = weight * dtau / 4 per action)."""
        self._fermion_ready()
        for k, a in enumerate(self.actions):
            if k == 0:
                w = a.force(self.Ff)
                if w != 1.0:
                    self.Ff.mul_(w)
            else:
                if self._Ft is None:
                    self._Ft = torch.empty_like(self.U)
                w = a.force(self._Ft)
                self.Ff.add_(self._Ft, alpha=w)
        return self.Ff

    def gauge_force(self):
        return wilson_force(self.lo, self.U, self.beta, out=self.Fg)

    def pdot(self):
        """-1/2 (F_gauge + F_fermion)."""
        Ff = self.fermion_force()
        return -0.5 * (self.gauge_force() + Ff)

    def hamiltonian(self, P):
        return kinetic(P) + self.gauge_action() + self.fermion_action()

    def _levels(self, nsteps):
        if self.gauge_steps <= 0:
            return [(self.pdot, nsteps, self.integrator, self.lam)]
        return [(lambda: -0.5 * self.fermion_force(), nsteps, self.integrator, self.lam),
                (lambda: -0.5 * self.gauge_force(), self.gauge_steps, self.gauge_integrator, self.gauge_lam)]

    def _evolve(self, P, tau, levels):
        """Integrate the first level's force over tau; its position steps are
        the evolution of the remaining levels, the innermost one the link
        update U <- exp(dt P) U."""
        if not levels:
            link_update(self.lo, P, tau, self.U)
            self._links_changed()
            return P
        pdot, nsteps, name, lam = levels[0]
        inner = levels[1:]
        dt = tau / nsteps
        if name == "leapfrog":
            P = P + (0.5 * dt) * pdot()
            for i in range(nsteps):
                P = self._evolve(P, dt, inner)
                P = P + (dt if i < nsteps - 1 else 0.5 * dt) * pdot()
            return P
        # Omelyan's second-order minimum-norm integrator, steps of
        # V(lam dt) T(dt/2) V((1-2 lam) dt) T(dt/2) V(lam dt) with adjacent V
        # updates merged (mdevolve's 2MN, selected by the name "2MN[,lam]" in
        # algorithms/integrator.nim:32-34); lam = 0 is the position leapfrog
        if lam != 0.0:
            P = P + (lam * dt) * pdot()
        for i in range(nsteps):
            P = self._evolve(P, 0.5 * dt, inner)
            P = P + ((1.0 - 2.0 * lam) * dt) * pdot()
            P = self._evolve(P, 0.5 * dt, inner)
            if lam != 0.0:
                P = P + ((2.0 * lam if i < nsteps - 1 else lam) * dt) * pdot()
        return P

    def leapfrog(self, P, tau, nsteps):
        return self._evolve(P, tau, [(self.pdot, nsteps, "leapfrog", 0.0)])

    def omelyan2mn(self, P, tau, nsteps, lam=OMELYAN_LAMBDA):
        return self._evolve(P, tau, [(self.pdot, nsteps, "2MN", lam)])

    def integrate(self, P, tau, nsteps):
        return self._evolve(P, tau, self._levels(nsteps))

    def trajectory(self, seed, tau=1.0, nsteps=20, rng=None, metropolis=True):
        """One HMC trajectory, with a Metropolis test unless metropolis is False
        (thermalization); returns a dict of diagnostics."""
        t0 = time.perf_counter()
        U0 = self.U.clone()
        P = random_momenta(self.lo, seed)
        self.heatbath(seed + 1)
        H0 = self.hamiltonian(P)
        P = self.integrate(P, tau, nsteps)
        H1 = self.hamiltonian(P)
        dH = H1 - H0
        u = (rng or np.random.default_rng(seed)).random()
        accept = bool(u < np.exp(-dH)) if metropolis else True
        if not accept:
            self.U.copy_(U0)
            self._links_changed()
        torch.cuda.synchronize()
        return {"dH": dH, "accept": accept, "seconds": time.perf_counter() - t0,
                "plaq": plaquette(self.lo, self.U) / (18.0 * self.lo.nSites)}

    def reversibility(self, seed, tau=1.0, nsteps=20):
        """Integrate forward, flip the momenta, integrate back: max |U - U0|, max |P + P0|."""
        U0 = self.U.clone()
        P0 = random_momenta(self.lo, seed)
        self.heatbath(seed + 1)
        P1 = self.integrate(P0.clone(), tau, nsteps)
        P2 = self.integrate(-P1, tau, nsteps)
        dU = float(torch.max(torch.abs(self.U - U0)))
        dP = float(torch.max(torch.abs(P2 + P0)))
        self.U.copy_(U0)
        self._links_changed()
        return dU, dP


# ------------------------------------------------------------------ checks
def _coords(L, idx):
    """Coordinates of canonical site index idx (even sites first)."""
    V = int(np.prod(L))
    p = 1 if idx >= V // 2 else 0
    for lex in (2 * idx - p * V, 2 * idx - p * V + 1):
        c = [lex % L[0], (lex // L[0]) % L[1], (lex // (L[0] * L[1])) % L[2], lex // (L[0] * L[1] * L[2])]
        if (sum(c) & 1) == p:
            return c
    raise ValueError(idx)


def _index(L, c):
    V = int(np.prod(L))
    c = [int(c[k]) % L[k] for k in range(4)]
    lex = c[0] + L[0] * (c[1] + L[1] * (c[2] + L[2] * c[3]))
    return (lex + (sum(c) & 1) * V) // 2


def _staples(L, U, mu, site):
    """Sum A of the six staples of U_mu(site): sum over plaquettes through the link
    of Re tr U_P = Re tr(U_mu(site) A)."""
    c = _coords(L, site)

    def at(nu, d):
        y = list(c)
        for k, v in d:
            y[k] += v
        return _index(L, y)

    A = np.zeros((3, 3), complex)
    for nu in range(4):
        if nu == mu:
            continue
        m = lambda k, s: U[k, s].cpu().numpy()
        A += m(nu, at(nu, [(mu, 1)])) @ m(mu, at(nu, [(nu, 1)])).conj().T @ m(nu, site).conj().T
        A += (m(nu, at(nu, [(mu, 1), (nu, -1)])).conj().T @ m(mu, at(nu, [(nu, -1)])).conj().T
              @ m(nu, at(nu, [(nu, -1)])))
    return A


def pseudofermion(lo, s, mass, seed):
    """phi = (D^+(m) eta)_e with eta complex Gaussian on all sites (hmcAction.nim:794-818);
    canonical (V, 3) torch tensor, zero on odd sites."""
    eta = lo.ColorVector().set(G.source(lo, seed, stream=7, par=2))
    phi = lo.ColorVector()
    Q.Ddag(s, phi, eta, mass)
    phi.zero("odd")
    return phi.get(device="cuda")


def _su3_alg(rng):
    a = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    x = 0.5 * (a - a.conj().T)
    return x - np.trace(x) / 3 * np.eye(3)


def _expm1_su3(A):
    """exp(A) - 1 for a small 3x3 matrix, by its power series (no cancellation)."""
    out = np.zeros((3, 3), complex)
    term = np.eye(3, dtype=complex)
    for k in range(1, 12):
        term = term @ A / k
        out += term
    return out


def force_fd_check(lo, Uph, phi, mass, nlinks=64, seed=1, h=1e-5, r2req=1e-28, beta=None, U=None, details=None,
                   richardson=False):
    """Finite-difference check of the fermion force (and, with beta and plain links
    U, of the Wilson gauge force) on nlinks seeded links.

    For U_mu(x) -> exp(w X) U_mu(x), X a random su(3) element, compares
    -1/2 Re tr(X F_mu(x)) with the central difference (S(+h) - S(-h))/2h.  The
    fermion action difference uses the exact identity
        S(U') - S(U) = -Y'^+ (M' - M) Y,   M = m^2 - D_eo D_oe,  Y = M^-1 phi.
    (M' - M) Y is applied directly through the link difference dU (the
    operator is linear in the links: (2D')(2D') - (2D)(2D) = (2D)(2D_dU) +
    (2D_dU)(2D')), so no O(1) quantities are subtracted; Y' = Y + d with
    M' d = phi - M' Y solved as a correction, so the error of Y cancels from
    Y'.  The gauge action is linear in each link: its difference is
    -(beta/3) Re tr(dU A) with A the staple sum.

    richardson=True also takes the central difference at 2h and reports
    (4 D(h) - D(2h))/3, which removes the O(h^2) truncation term.

    Returns (worst relative error of the fermion force, worst relative error
    of the gauge force or None, seconds); relative to max(|analytic|, 1e-3).
    details, if a list, receives (mu, site, finite difference, analytic,
    1/2 |X| |F_mu(x)|) per link."""
    t0 = time.perf_counter()
    L = lo.physGeom
    s = Q.Staggered(lo, Uph)
    sp = Q.SolverParams()
    sp.r2req, sp.verbosity, sp.maxits = r2req, 0, 100000
    psi = lo.ColorVector()
    fphi = lo.ColorVector().set(phi)
    Q.stagForceSolve(s, psi, fphi, mass, sp)
    F = Q.fermForce(s, psi, device=Uph.device)
    Fg = wilson_force(lo, U, beta) if beta is not None else None
    Up = Uph.clone()
    s_p = Q.Staggered(lo, Up)
    Z = torch.zeros_like(Uph)
    s_d = Q.Staggered(lo, Z)
    a, t1, t2, u1, u2, rhs, dd = (lo.ColorVector() for _ in range(7))
    m2 = mass * mass
    Q.stagD2ee(s.se, s.so, a, s.g, psi, m2)  # 4 M Y
    spd = Q.SolverParams()
    spd.r2req, spd.verbosity, spd.maxits = 1e-18, 0, 100000
    rng = np.random.default_rng(seed)
    worst_f, worst_g = 0.0, 0.0
    for _ in range(nlinks):
        mu, site = int(rng.integers(4)), int(rng.integers(lo.nSites))
        X = _su3_alg(rng)
        u0 = Uph[mu, site].cpu().numpy()
        dS = []
        steps = (h, -h, 2 * h, -2 * h) if richardson else (h, -h)
        for w in steps:
            dU = _expm1_su3(w * X) @ u0
            Up[mu, site] = Uph[mu, site] + torch.from_numpy(dU).to(Uph.device)
            s_p.update(Up)
            Z[mu, site] = torch.from_numpy(dU).to(Uph.device)
            s_d.update(Z)
            Q.stagD2(s_p.so, t1, s_p.g, psi, 0.0, 0.0)  # (2D')_oe Y
            Q.stagD2(s_d.se, u1, s_d.g, t1, 0.0, 0.0)  # (2D_dU)_eo (2D')_oe Y
            Q.stagD2(s_d.so, t2, s_d.g, psi, 0.0, 0.0)  # (2D_dU)_oe Y
            Q.stagD2(s.se, u2, s.g, t2, 0.0, 0.0)  # (2D)_eo (2D_dU)_oe Y
            u1.axpby(1.0, u2, 1.0, "even")  # -4 (M' - M) Y
            rhs.copy_from(a, "even")
            rhs.axpby(4.0, fphi, -1.0, "even")
            rhs.axpby(1.0, u1, 1.0, "even")  # 4 (phi - M' Y)
            Q.solveEE(s_p, dd, rhs, mass, spd)  # d = M'^-1 (phi - M' Y)
            dd.axpby(1.0, psi, 1.0, "even")  # Y' = Y + d
            dS.append(0.25 * dd.redot(u1, "even"))
            Z[mu, site] = 0.0
        Up[mu, site] = Uph[mu, site]
        fd = (dS[0] - dS[1]) / (2 * h)
        if richardson:
            fd = (4.0 * fd - (dS[2] - dS[3]) / (4 * h)) / 3.0
        Fx = F[mu, site].cpu().numpy()
        an = -0.5 * np.trace(X @ Fx).real
        worst_f = max(worst_f, abs(fd - an) / max(abs(an), 1e-3))
        if details is not None:
            details.append((mu, site, fd, an, 0.5 * np.linalg.norm(X) * np.linalg.norm(Fx)))
        if Fg is not None:
            A = _staples(L, U, mu, site)
            u = U[mu, site].cpu().numpy()
            def gd(hh):
                dUg = (_expm1_su3(hh * X) - _expm1_su3(-hh * X)) @ u
                return -(beta / 3.0) * np.trace(dUg @ A).real / (2 * hh)
            fdg = (4.0 * gd(h) - gd(2 * h)) / 3.0 if richardson else gd(h)
            ang = -0.5 * np.trace(X @ Fg[mu, site].cpu().numpy()).real
            worst_g = max(worst_g, abs(fdg - ang) / max(abs(ang), 1e-3))
    torch.cuda.synchronize()
    return worst_f, (worst_g if Fg is not None else None), time.perf_counter() - t0


def main(argv=None):
    """Staggered HMC run in the manner of staghmc.nim and generichmc.nim (cold
    start or a loaded configuration, Metropolis), one JSON line per trajectory;
    --fd-links N first runs force_fd_check."""
    import argparse
    import json

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--lat", type=int, nargs=4, default=[8, 8, 8, 8])
    ap.add_argument("--beta", type=float, default=5.6)
    ap.add_argument("--mass", type=float, default=0.05)
    ap.add_argument("--tau", type=float, default=1.0)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--trajs", type=int, default=10)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--frsq", type=float, default=1e-20, help="force solve r2req")
    ap.add_argument("--arsq", type=float, default=1e-24, help="action solve r2req")
    ap.add_argument("--sloppy", type=int, default=0,
                    help="SloppyType of the solves (vectors): 0 FP64, 1 FP32, 2 FP16, 3 BF16")
    ap.add_argument("--sloppy-links", type=int, default=-1,
                    help="SloppyType of the sloppy links (-1: as the vectors; 4: 16-bit fixed point)")
    ap.add_argument("--recon", type=int, default=18, help="reals per sloppy link (18 or 12)")
    ap.add_argument("--fd-links", type=int, default=0)
    ap.add_argument("--reversibility", action="store_true")
    ap.add_argument("--integrator", default="leapfrog", help="leapfrog or 2MN[,lambda]")
    ap.add_argument("--warmup", type=int, default=0, help="trajectories without Metropolis first")
    ap.add_argument("--warmup-integrator", default="",
                    help="integrator of the warmup trajectories (default: --integrator)")
    ap.add_argument("--warmup-steps", type=int, default=0, help="MD steps per warmup trajectory (default: --steps)")
    ap.add_argument("--warmup-sloppy", type=int, default=-1,
                    help="SloppyType of the warmup solves (default: --sloppy); links per --sloppy-links/--recon")
    ap.add_argument("--save", default="", help="write the configuration after the warmup (SciDAC, FP64)")
    ap.add_argument("--load", default="", help="start from this gauge file (saveGauge / SciDAC / ILDG) instead "
                                               "of the cold start")
    ap.add_argument("--hisq", action="store_true", help="HISQ fermions: smeared fat + Naik links")
    ap.add_argument("--hasenbusch-mass", type=float, default=0.0,
                    help="split the quartet as StaggeredFermionAction(this mass) x "
                         "StaggeredRatioAction(--mass / this mass), as generichmc.nim does")
    ap.add_argument("--pv-mass", type=float, default=0.0, help="add --num-pv StaggeredPauliVillarsAction(this mass)")
    ap.add_argument("--num-pv", type=int, default=1)
    ap.add_argument("--gauge-steps", type=int, default=0,
                    help="gauge force on an inner level with this many steps per outer step (0: one level)")
    ap.add_argument("--gauge-integrator", default="", help="integrator of the inner gauge level (default: --integrator)")
    args = ap.parse_args(argv)

    from ._lib import init
    init(0)
    lo = Q.Layout(tuple(args.lat))
    if args.fd_links:
        U = G.random_gauge(lo, args.seed)
        Uph = G.phase_bc(lo, U.clone())
        phi = pseudofermion(lo, Q.Staggered(lo, Uph), args.mass, args.seed + 1)
        wf, wg, sec = force_fd_check(lo, Uph, phi, args.mass, nlinks=args.fd_links, seed=args.seed,
                                     beta=args.beta, U=U)
        print(json.dumps({"check": "force_fd", "lat": args.lat, "mass": args.mass, "links": args.fd_links,
                          "fermion_worst_rel": wf, "gauge_worst_rel": wg, "seconds": round(sec, 2)}), flush=True)
    def parse_integrator(spec):
        name, *rest = spec.split(",")
        return name, float(rest[0]) if rest else OMELYAN_LAMBDA

    from . import fileio
    U0 = G.unit_gauge(lo)
    if args.load:
        fileio.loadGauge(lo, U0, args.load)
    name, lam = parse_integrator(args.integrator)
    if args.hasenbusch_mass > 0:
        actions = [StaggeredFermionAction(args.hasenbusch_mass),
                   StaggeredRatioAction(args.mass, args.hasenbusch_mass)]
    else:
        actions = [StaggeredFermionAction(args.mass)]
    if args.pv_mass > 0:
        actions += [StaggeredPauliVillarsAction(args.pv_mass) for _ in range(args.num_pv)]
    gname, glam = parse_integrator(args.gauge_integrator or args.integrator)
    h = StaggeredHMC(lo, U0, args.beta, args.mass, r2req_force=args.frsq,
                     r2req_action=args.arsq, sloppy=Q.SloppyType(args.sloppy), integrator=name,
                     lam=lam, sloppy_links=None if args.sloppy_links < 0 else Q.SloppyType(args.sloppy_links),
                     recon=args.recon, hisq=args.hisq, actions=actions, gauge_steps=args.gauge_steps,
                     gauge_integrator=gname, gauge_lam=glam)
    print(json.dumps({"actions": h.describe(), "gauge_steps": args.gauge_steps}), flush=True)
    rng = np.random.default_rng(args.seed)
    # warmup: its own integrator, step count and solver storage if given
    prod = (h.integrator, h.lam, h.sloppy)
    if args.warmup_integrator:
        h.integrator, h.lam = parse_integrator(args.warmup_integrator)
    if args.warmup_sloppy >= 0:
        h.sloppy = Q.SloppyType(args.warmup_sloppy)
    wsteps = args.warmup_steps or args.steps
    for n in range(args.warmup):
        its0 = h.solver_its
        r = h.trajectory(1000 * (n + 1) + args.seed, args.tau, wsteps, rng=rng, metropolis=False)
        r.update({"traj": n + 1 - args.warmup, "warmup": True, "solver_its": h.solver_its - its0})
        print(json.dumps(r), flush=True)
    h.integrator, h.lam, h.sloppy = prod
    if args.save:
        fileio.saveGauge(lo, h.U, args.save, "D")
        back = torch.empty_like(h.U)
        fileio.loadGauge(lo, back, args.save)
        print(json.dumps({"saved": args.save, "roundtrip_exact": bool(torch.equal(back, h.U)),
                          "plaq": plaquette(lo, h.U) / (18.0 * lo.nSites)}), flush=True)
        del back
    rows = []
    for n in range(args.warmup, args.warmup + args.trajs):
        its0 = h.solver_its
        r = h.trajectory(1000 * (n + 1) + args.seed, args.tau, args.steps, rng=rng, metropolis=True)
        r.update({"traj": n + 1 - args.warmup, "warmup": False, "solver_its": h.solver_its - its0})
        if args.reversibility:
            dU, dP = h.reversibility(777 + n, args.tau, args.steps)
            r.update({"rev_dU": dU, "rev_dP": dP})
        rows.append(r)
        print(json.dumps(r), flush=True)
    if rows:
        dh = np.array([r["dH"] for r in rows])
        ex = np.exp(-dh)
        summary = {"summary": True, "lat": args.lat, "beta": args.beta, "mass": args.mass, "tau": args.tau,
                   "steps": args.steps, "integrator": args.integrator, "actions": h.describe(),
                   "gauge_steps": args.gauge_steps, "trajectories": len(rows),
                   "acceptance": float(np.mean([r["accept"] for r in rows])),
                   "dH_mean": float(dh.mean()), "dH_abs_max": float(np.abs(dh).max()),
                   "exp_minus_dH_mean": float(ex.mean()),
                   "exp_minus_dH_stderr": float(ex.std(ddof=1) / np.sqrt(len(ex))) if len(ex) > 1 else None,
                   "seconds_per_trajectory": float(np.mean([r["seconds"] for r in rows])),
                   "solver_its_per_trajectory": float(np.mean([r["solver_its"] for r in rows])),
                   "plaq_last": rows[-1]["plaq"]}
        if args.reversibility:
            summary["rev_dU_max"] = max(r["rev_dU"] for r in rows)
            summary["rev_dP_max"] = max(r["rev_dP"] for r in rows)
        print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
