"""Plain-staggered HMC with a Wilson gauge action, on the GPU through the C ABI.

Mirrors QEX's staggered HMC (src/examples/staghmc.nim:81-353, and the action
pieces of src/hmc/hmcAction.nim:794-885):

- one pseudofermion field on even sites, phi = (D^+(m) eta)_e with eta complex
  Gaussian (heatbath, hmcAction.nim:794-818);
- S_f = phi^+ (m^2 - D_eo D_oe)^{-1} phi;
- S_g = -(beta/3) sum_P Re tr U_P (gauge/gaugeAction.nim:397-427);
- momenta P traceless anti-Hermitian, H = 1/2 sum |P|^2 + S_g + S_f;
- leapfrog (QEX's mdevolve leapfrog, src/algorithms/integrator.nim:32-34, with
  the update order of staghmc.nim mdt/mdv/mdvf).

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
from . import stag as Q
from ._lib import call


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


class StaggeredHMC:
    def __init__(self, lo, U, beta, mass, r2req_force=1e-20, r2req_action=1e-24, sloppy=Q.SloppyType.SloppyNone):
        self.lo = lo
        self.U = U.contiguous()
        self.beta = float(beta)
        self.mass = float(mass)
        self.Uph = self.U.clone()
        G.phase_bc(lo, self.Uph)
        self.s = Q.Staggered(lo, self.Uph)
        self.phi = lo.ColorVector()
        self.psi = lo.ColorVector()
        self.y = lo.ColorVector()
        self.r2f, self.r2a = r2req_force, r2req_action
        self.sloppy = sloppy
        self.solver_its = 0
        self.Ff = torch.empty_like(self.U)
        self.Fg = torch.empty_like(self.U)

    def _sp(self, r2req):
        sp = Q.SolverParams()
        sp.r2req, sp.verbosity, sp.maxits = r2req, 0, 100000
        sp.sloppySolve = self.sloppy
        return sp

    def _links_changed(self):
        self.Uph.copy_(self.U)
        G.phase_bc(self.lo, self.Uph)
        self.s.update(self.Uph)

    def heatbath(self, seed):
        """phi_e = (D^+(m) eta)_e (hmcAction.nim:794-818)."""
        eta = self.lo.ColorVector().set(G.source(self.lo, seed, stream=7, par=2))
        Q.Ddag(self.s, self.phi, eta, self.mass)
        self.phi.zero("odd")

    def fermion_action(self):
        sp = self._sp(self.r2a)
        Q.solveEE(self.s, self.y, self.phi, self.mass, sp)
        self.solver_its += sp.iterations
        return 4.0 * self.phi.redot(self.y, "even")

    def gauge_action(self):
        return -(self.beta / 3.0) * plaquette(self.lo, self.U)

    def pdot(self):
        """-1/2 (F_gauge + F_fermion)."""
        sp = self._sp(self.r2f)
        Q.stagForceSolve(self.s, self.psi, self.phi, self.mass, sp)
        self.solver_its += sp.iterations
        Q.fermForce(self.s, self.psi, out=self.Ff)
        wilson_force(self.lo, self.U, self.beta, out=self.Fg)
        return -0.5 * (self.Fg + self.Ff)

    def hamiltonian(self, P):
        return kinetic(P) + self.gauge_action() + self.fermion_action()

    def leapfrog(self, P, tau, nsteps):
        dt = tau / nsteps
        P = P + (0.5 * dt) * self.pdot()
        for i in range(nsteps):
            link_update(self.lo, P, dt, self.U)
            self._links_changed()
            P = P + (dt if i < nsteps - 1 else 0.5 * dt) * self.pdot()
        return P

    def trajectory(self, seed, tau=1.0, nsteps=20, rng=None):
        """One HMC trajectory with Metropolis; returns a dict of diagnostics."""
        t0 = time.perf_counter()
        U0 = self.U.clone()
        P = random_momenta(self.lo, seed)
        self.heatbath(seed + 1)
        H0 = self.hamiltonian(P)
        P = self.leapfrog(P, tau, nsteps)
        H1 = self.hamiltonian(P)
        dH = H1 - H0
        u = (rng or np.random.default_rng(seed)).random()
        accept = bool(u < np.exp(-dH))
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
        P1 = self.leapfrog(P0.clone(), tau, nsteps)
        P2 = self.leapfrog(-P1, tau, nsteps)
        dU = float(torch.max(torch.abs(self.U - U0)))
        dP = float(torch.max(torch.abs(P2 + P0)))
        self.U.copy_(U0)
        self._links_changed()
        return dU, dP
