"""QEX's generic conjugate gradient over a user-supplied operator
(solvers/cg.nim:13-283): CgState, newCgState, CgState.solve, cgSolve.

The operator is any object with apply(Ap, p) (and, for preconditioning,
precon = CgPrecon.cpHerm with applyPrecon(z, r)); vectors are this library's
FP64 color-vector Fields, so every vector operation is a HIP kernel.  The host
drives the iteration as QEX does (one reduction read back per step).  For the
staggered operator itself, solveEE / solveOO run the fused device-resident CG
instead.  The state persists between calls: a second solve continues from
the first one's iteration, as the reference's does (cg.nim:139-270)."""
import enum

from ._lib import QexHipError


class CgPrecon(enum.IntEnum):
    cpNone = 0
    cpHerm = 1


class CgState:
    """cg.nim:13-19, 21-27 (reset), 42-50 (newCgState)."""

    def __init__(self, x, b):
        lo = b.l
        self.x, self.b = x, b
        self.r = lo.ColorVector()
        self.Ap = lo.ColorVector()
        self.p = lo.ColorVector()
        self.z = self.r
        self.precon = CgPrecon.cpNone
        self.reset()

    def reset(self):
        self.b2 = -1.0
        self.iterations = 0
        self.r2 = 1.0
        self.r2old = 1.0
        self.rzold = 1.0
        self.r2stop = 0.0
        self.rz = 0.0
        self.alpha = 0.0

    def _init_precon(self):
        self.z = self.r if self.precon == CgPrecon.cpNone else self.r.l.ColorVector()

    def solve(self, op, sp, x=None, b=None):
        """Solve A x = b on sp.subsetName (cg.nim:55-270; with x, b given, the
        state is reset first, cg.nim:272-277)."""
        if x is not None or b is not None:
            self.x = x if x is not None else self.x
            self.b = b if b is not None else self.b
            self.reset()
        sub = getattr(sp, "subsetName", "all")
        precon = CgPrecon(getattr(op, "precon", CgPrecon.cpNone))
        if precon != self.precon:
            self.precon = precon
            self._init_precon()
            self.reset()
        if precon == CgPrecon.cpHerm and not hasattr(op, "applyPrecon"):
            raise QexHipError("cg.solve: precon == cpHerm but op.applyPrecon not found")
        r, p, Ap, x, b, z = self.r, self.p, self.Ap, self.x, self.b, self.z
        if self.b2 < 0:  # first call
            self.b2 = b.norm2(sub)
            if self.b2 == 0.0:
                x.zero(sub)
                r.zero(sub)
                self.r2 = self.rz = 0.0
            else:
                op.apply(Ap, x)
                r.copy_from(b, sub)
                r.axpby(-1.0, Ap, 1.0, sub)                    # r = b - A x
                self.r2 = r.norm2(sub)
                p.zero(sub)
        self.r2stop = sp.r2req * self.b2
        itn, r2, rzo, alpha = self.iterations, self.r2, self.rzold, self.alpha
        while itn < sp.maxits and r2 > self.r2stop:
            if precon == CgPrecon.cpHerm:
                op.applyPrecon(z, r)
                rz = r.redot(z, sub)
            else:
                rz = r2
            beta = rz / rzo
            rzo = rz
            if itn == 0:
                p.copy_from(z, sub)
            else:
                p.axpby(1.0, z, beta, sub)                     # p = z + beta p
            itn += 1
            op.apply(Ap, p)
            pAp = p.redot(Ap, sub)
            alpha = rz / pAp
            x.axpby(alpha, p, 1.0, sub)
            r.axpby(-alpha, Ap, 1.0, sub)
            r2 = r.norm2(sub)
            if sp.verbosity >= 2:
                print(f"CG iteration: {itn}   r2/b2: {r2 / self.b2}")
        self.iterations, self.r2, self.rzold, self.rz, self.alpha = itn, r2, rzo, rzo, alpha
        sp.iterations = itn
        sp.r2 = r2 / self.b2 if self.b2 > 0 else 0.0
        if sp.verbosity >= 1:
            print(f"CG final iterations: {itn}  r2/b2: {sp.r2}")
        return sp


def newCgState(x, b):
    return CgState(x, b)


def cgSolve(x, b, op, sp):
    """cgSolve (cg.nim:279-281)."""
    cg = newCgState(x, b)
    return cg.solve(op, sp, x, b)
