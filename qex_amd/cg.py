"""QEX's generic conjugate gradient over a user-supplied operator
(solvers/cg.nim:13-283): CgState, newCgState, CgState.solve, cgSolve; and its
generic multi-shift CG (solvers/cgm.nim:13-246): CgmState, newCgmState,
CgmState.solve.

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


class CgmState:
    """Multi-shift CG state (cgm.nim:13-66): solutions xs[k] of (A + sigma_k) x = b,
    sigma_0 taken as 0 (cgm.nim:112-115); recycle restarts from xs[0]."""

    def __init__(self, xs, b, sigmas, recycle=False):
        if len(xs) != len(sigmas) or not xs:
            raise QexHipError("CgmState: one solution field per shift")
        lo = b.l
        self.xs, self.b, self.sigmas = list(xs), b, [float(v) for v in sigmas]
        self.ps = [lo.ColorVector() for _ in xs]
        self.r = lo.ColorVector()
        self.Ap = lo.ColorVector()
        self.reset(recycle)

    def reset(self, recycle=False):
        self.iterations = 0
        self.b2 = 1.0 if recycle else -1.0
        self.r2 = 1.0
        self.r2stop = 0.0

    def solve(self, op, sp):
        """Jegerlehner's multi-shift CG (hep-lat/9612014; cgm.nim:68-246).  op is
        called as op(Ap, p) for the unshifted operator."""
        sub = getattr(sp, "subsetName", "all")
        xs, ps, r, Ap, b = self.xs, self.ps, self.r, self.Ap, self.b
        sg = [0.0] + self.sigmas[1:]
        if self.b2 < 0.0:  # new solution
            r.copy_from(b, sub)
            for x in xs:
                x.zero(sub)
            b2 = b.norm2(sub)
            r2 = b2
        else:  # restart from xs[0]
            op(Ap, xs[0])
            r.copy_from(b, sub)
            r.axpby(-1.0, Ap, 1.0, sub)
            b2 = self.b2
            r2 = r.norm2(sub)
        r2stop = sp.r2req * b2
        self.r2stop = r2stop
        itn = 0
        if r2 > r2stop:
            n = len(xs)
            ri = r2
            alpha_prev, beta_prev = -1.0, 0.0
            z_prev, z = [1.0] * n, [1.0] * n
            for k in range(n):
                ps[k].copy_from(r, sub)
            going = True
            while going:
                # unshifted system
                op(Ap, ps[0])
                itn += 1
                pAp = ps[0].redot(Ap, sub)
                alpha = ri / pAp if pAp != 0.0 else 0.0
                r.axpby(-alpha, Ap, 1.0, sub)
                xs[0].axpby(alpha, ps[0], 1.0, sub)
                rn = r.norm2(sub)
                beta = rn / ri if ri != 0.0 else 0.0
                going = itn < sp.maxits and rn > r2stop
                if going:
                    ps[0].axpby(1.0, r, beta, sub)         # p = r + beta p
                # shifted systems: zeta recurrences
                for k in range(1, n):
                    den = alpha * beta_prev * (z_prev[k] - z[k]) + z_prev[k] * alpha_prev * (1.0 + sg[k] * alpha)
                    zn = z[k] * z_prev[k] * alpha_prev / den if den != 0.0 else 0.0
                    zr = zn / z[k] if z[k] != 0.0 else 0.0
                    xs[k].axpby(alpha * zr, ps[k], 1.0, sub)
                    if going:
                        ps[k].axpby(zn, r, beta * zr * zr, sub)
                        z_prev[k], z[k] = z[k], zn
                alpha_prev, beta_prev, ri = alpha, beta, rn
                r2 = rn
                if sp.verbosity >= 2:
                    print(f"CG iteration: {itn}  r2/b2: {r2 / b2}")
        self.r2, self.b2, self.iterations = r2, b2, itn
        sp.iterations = itn
        sp.r2 = r2 / b2 if b2 > 0 else 0.0
        return sp


def newCgmState(xs, b, sigmas, recycle=False):
    return CgmState(xs, b, sigmas, recycle)
