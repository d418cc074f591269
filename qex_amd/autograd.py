"""torch.autograd binding of the staggered fermion action and its force.

S(U) = phi_e^+ (m^2 - D_eo D_oe)^{-1} phi_e for phased links U (canonical
(4, V, 3, 3) complex128 tensor on the GPU).  forward() solves with the HIP CG;
backward() uses the hand-written outer-product kernel (qexhip_force with
outer_only=1, the derivative of hmc/hmcAction.nim:735-771 before the TAH
projection):

  dS = Re sum conj(dU) . G,   G_mu(x) = 1/2 eps(x) psi(x) psi(x+mu)^+,

with psi = [Y_e ; (2D)_oe Y_e], Y = (m^2 - D_eo D_oe)^{-1} phi.  That is torch's
gradient of a real function of complex inputs (d/dRe + i d/dIm), so
U.grad == grad_output * G.
"""
import torch

from . import stag as Q


class StaggeredFermionAction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, phi, mass, lo, r2req):
        s = Q.Staggered(lo, U.detach().contiguous())
        fphi = lo.ColorVector().set(phi.detach().contiguous())
        psi = lo.ColorVector()
        sp = Q.SolverParams()
        sp.r2req, sp.verbosity = r2req, 0
        Q.stagForceSolve(s, psi, fphi, mass, sp)
        # S = phi_e^+ Y_e with Y_e = psi_e
        S = fphi.redot(psi, "even")
        Gf = Q.fermForce(s, psi, outer_only=True, device=U.device)
        ctx.save_for_backward(Gf)
        return torch.tensor(S, dtype=torch.float64, device=U.device)

    @staticmethod
    def backward(ctx, grad_out):
        (Gf,) = ctx.saved_tensors
        return grad_out * 0.5 * Gf, None, None, None, None


def staggered_action(U, phi, mass, lo, r2req=1e-26):
    return StaggeredFermionAction.apply(U, phi, mass, lo, r2req)
