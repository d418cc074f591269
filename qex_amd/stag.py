"""Host-side mirror of QEX's staggered-fermion API over the qexhip C ABI.

Names, argument meaning and error behaviour follow QEX's Nim procedures so
that callers and tests read like QEX code:

  Layout            ~ newLayout (src/layout/layoutX.nim:71-200), one rank
  Field             ~ Field[V, ColorVector] (src/field/fieldET.nim:17-22)
  Staggered         ~ Staggered[G,T] with se/so (src/physics/stagD.nim:15-35,68-75)
  newStag/newStag3  ~ src/physics/stagD.nim:573-615
  stagD2/stagD/stagD2ee/stagD2oo/D/Ddag ~ src/physics/stagD.nim:383-637
  SolverParams      ~ src/solvers/solverBase.nim:10-63
  solveEE/solveOO   ~ src/physics/stagSolve.nim:330-427
  solveMulti        ~ multi-shift solveXXInternal, src/physics/stagSolve.nim:164-229
  stagForceSolve/fermForce ~ src/hmc/hmcAction.nim:669-792

Data live on the GPU in the library's layout; `Field.set`/`Field.get` move
canonical arrays (V x 3 complex128, QEX's single-rank site order, even sites
first) in from numpy (host) or torch (device) arrays.
"""
import ctypes
import enum

import numpy as np

from . import _lib
from ._lib import ALL, BF16, EVEN, FP16, FP32, FP64, ODD, Q16, QexHipError, SolverParamsC, call

try:  # torch is plumbing for device arrays; numpy arrays work without it
    import torch
except Exception:  # pragma: no cover
    torch = None

_PAR = {"even": EVEN, "odd": ODD, "all": ALL, EVEN: EVEN, ODD: ODD, ALL: ALL}


def _par(p):
    try:
        return _PAR[p]
    except KeyError:
        raise QexHipError(f"unknown subset {p!r}")


def _ptr(a):
    """(pointer, on_device) of a contiguous complex128/float64 numpy or torch array.

    The library works on its own HIP stream and requires device data to be
    complete when a call starts, so torch's pending work is waited for here."""
    if torch is not None and isinstance(a, torch.Tensor):
        if not a.is_contiguous():
            raise QexHipError("tensor must be contiguous")
        if a.dtype not in (torch.complex128, torch.float64):
            raise QexHipError("tensor must be complex128 or float64")
        if a.is_cuda:
            torch.cuda.current_stream(a.device).synchronize()
        return ctypes.c_void_p(a.data_ptr()), int(a.is_cuda)
    a = np.asarray(a)
    if not a.flags["C_CONTIGUOUS"] or a.dtype not in (np.complex128, np.float64):
        raise QexHipError("array must be C-contiguous complex128 or float64")
    return ctypes.c_void_p(a.ctypes.data), 0


class Layout:
    """One-rank lattice (QEX newLayout).  All extents must be even."""

    def __init__(self, lat):
        _lib.init()
        self.physGeom = tuple(int(v) for v in lat)
        if len(self.physGeom) != 4:
            raise QexHipError("4-d lattices only")
        self.nSites = int(np.prod(self.physGeom))
        self.nEven = self.nSites // 2
        h = ctypes.c_void_p()
        call("qexhip_lattice_create", (ctypes.c_int * 4)(*self.physGeom), ctypes.byref(h))
        self.h = h

    def ColorVector(self, prec=FP64):
        return Field(self, prec)

    def __del__(self):
        if getattr(self, "h", None) and _lib._lib is not None:
            _lib._lib.qexhip_lattice_destroy(self.h)
            self.h = None


def newLayout(lat):
    return Layout(lat)


class Field:
    """Color-vector field on the GPU (both parities)."""

    def __init__(self, lo, prec=FP64):
        self.l = lo
        self.prec = prec
        h = ctypes.c_void_p()
        call("qexhip_field_create", lo.h, prec, ctypes.byref(h))
        self.h = h

    def set(self, arr, sub="all"):
        p, dev = _ptr(arr)
        call("qexhip_field_import", self.h, _par(sub), p, dev)
        return self

    def get(self, sub="all", out=None, device=None):
        """Canonical (V, 3) complex128 copy; numpy unless device is a torch device."""
        V = self.l.nSites
        if out is None:
            if device is not None:
                out = torch.zeros((V, 3), dtype=torch.complex128, device=device)
            else:
                out = np.zeros((V, 3), dtype=np.complex128)
        p, dev = _ptr(out)
        call("qexhip_field_export", self.h, _par(sub), p, dev)
        return out

    def norm2(self, sub="all"):
        v = ctypes.c_double()
        call("qexhip_field_norm2", self.h, _par(sub), ctypes.byref(v))
        return v.value

    def axpby(self, a, x, b=1.0, sub="all"):
        """self = a*x + b*self on a subset (FP64 fields)."""
        call("qexhip_field_axpby", x.h, self.h, _par(sub), float(a), float(b))
        return self

    def copy_from(self, x, sub="all"):
        return self.axpby(1.0, x, 0.0, sub)

    def zero(self, sub="all"):
        return self.axpby(0.0, self, 0.0, sub)

    def redot(self, other, sub="all"):
        v = ctypes.c_double()
        call("qexhip_field_redot", self.h, other.h, _par(sub), ctypes.byref(v))
        return v.value

    def __del__(self):
        if getattr(self, "h", None) and _lib._lib is not None:
            _lib._lib.qexhip_field_destroy(self.h)
            self.h = None


class StaggeredD:
    """One checkerboard of a Staggered operator (QEX StaggeredD: sub/subset)."""

    def __init__(self, owner, sub):
        self.owner = owner
        self.sub = sub
        self.par = _par(sub)


class Staggered:
    """QEX Staggered: links g (phased, BC-applied) plus even/odd descriptors.

    g: canonical link array (nl, V, 3, 3) complex128, numpy or torch (device);
    nl = 4 plain, 8 = [fat0, long0, fat1, long1, ...] (newStag3 interleave).
    """

    def __init__(self, lo, g):
        self.l = lo
        nl = int(g.shape[0])
        if nl not in (4, 8):
            raise QexHipError("Staggered: need 4 (plain) or 8 (fat+long) link fields")
        if tuple(g.shape[1:]) != (lo.nSites, 3, 3):
            raise QexHipError("Staggered: link array shape must be (nl, V, 3, 3)")
        p, dev = _ptr(g)
        h = ctypes.c_void_p()
        call("qexhip_links_create", lo.h, nl, p, dev, ctypes.byref(h))
        self.h = h
        self.nl = nl
        self.g = self  # QEX passes s.g alongside s.se/s.so; the handle carries the links
        self.se = StaggeredD(self, "even")
        self.so = StaggeredD(self, "odd")
        self.sa = StaggeredD(self, "all")

    def update(self, g):
        """New gauge generation (same shape)."""
        p, dev = _ptr(g)
        call("qexhip_links_update", self.h, p, dev)

    def __del__(self):
        if getattr(self, "h", None) and _lib._lib is not None:
            _lib._lib.qexhip_links_destroy(self.h)
            self.h = None


def newStag(lo, g):
    """newStag (stagD.nim:573-592): plain links g[0..3]."""
    return Staggered(lo, g)


def newStag3(lo, g, g3):
    """newStag3 (stagD.nim:603-615): interleave fat g[mu] and long g3[mu]."""
    if torch is not None and isinstance(g, torch.Tensor):
        gg = torch.stack([t for mu in range(4) for t in (g[mu], g3[mu])]).contiguous()
    else:
        gg = np.ascontiguousarray(np.stack([t for mu in range(4) for t in (g[mu], g3[mu])]))
    return Staggered(lo, gg)


def _sd(sd):
    if isinstance(sd, StaggeredD):
        return sd.owner, sd.par
    raise QexHipError("expected a StaggeredD (s.se, s.so or s.sa)")


def stagD2(sd, r, g, x, a, b):
    """r = a*r + b*x + (2D)x on sd's subset (stagD.nim:383-448)."""
    s, par = _sd(sd)
    call("qexhip_stagD2", s.h, par, r.h, x.h, float(a), float(b))


def stagD(sd, r, g, x, m, sc=1.0, a=0.0):
    """r = a*r + m*x + sc*D*x (stagD.nim:459-462)."""
    s, par = _sd(sd)
    call("qexhip_stagD", s.h, par, r.h, x.h, float(m), float(sc), float(a))


def stagD1(sd, r, g, x):
    """r = (2D)x on sd's subset (stagD.nim:467-469, stagDP with r initialised to 0)."""
    stagD2(sd, r, g, x, 0.0, 0.0)


def stagD1x(sd, r, g, x):
    """r = -(2D)x on sd's subset (stagD.nim:471-473, stagDM)."""
    stagD(sd, r, g, x, 0.0, -2.0)


def stagDb(sd, r, g, x, m, sc=1.0):
    """r = (m/(0.5 sc)) x + (2D)x: stagD without the final 0.5 sc scaling
    (stagD.nim:476-478)."""
    stagD2(sd, r, g, x, 0.0, m / (0.5 * sc))


def stagD2multi(sd, rs, g, xs, a, b):
    """r_j = a*r_j + b*x_j + (2D)x_j for every field pair on sd's subset, each link
    loaded once for up to 8 fields (the multi-RHS form of stagD2)."""
    s, par = _sd(sd)
    if len(rs) != len(xs):
        raise QexHipError("stagD2multi: one output per input")
    call("qexhip_stagD2_multi", s.h, par, len(rs), _vec(rs), _vec(xs), float(a), float(b))


def stagDN(sds, rs, g, xs, m, sc=1.0):
    """QEX's stagDN (stagD.nim:451-456): r_i = m*x_i + (2D)x_i for several fields.
    As in QEX, sc is accepted and not applied (the 0.5*sc scaling there is
    commented out)."""
    sd = sds[0] if isinstance(sds, (list, tuple)) else sds
    stagD2multi(sd, rs, g, xs, 0.0, m)


def stagD2ee(se, so, r, g, x, m2):
    """r_e = 4*m2*x_e - (2D)_eo (2D)_oe x_e (stagD.nim:514-516)."""
    call("qexhip_stagD2xx", se.owner.h, EVEN, r.h, x.h, float(m2))


def stagD2oo(se, so, r, g, x, m2):
    """r_o = 4*m2*x_o - (2D)_oe (2D)_eo x_o (stagD.nim:518-520)."""
    call("qexhip_stagD2xx", se.owner.h, ODD, r.h, x.h, float(m2))


def D(s, r, x, m):
    """D(m) = m + D on all sites (stagD.nim:617-619)."""
    stagD(s.se, r, s.g, x, m)
    stagD(s.so, r, s.g, x, m)


def Ddag(s, r, x, m):
    """D^+(m) = m - D on all sites (stagD.nim:620-622)."""
    stagD(s.se, r, s.g, x, m, -1.0)
    stagD(s.so, r, s.g, x, m, -1.0)


def peqDdag(s, r, x, m):
    """r += D^+(m) x on all sites (stagD.nim:623-625)."""
    stagD(s.se, r, s.g, x, m, -1.0, 1.0)
    stagD(s.so, r, s.g, x, m, -1.0, 1.0)


def eoReduce(s, r, b, m):
    """r.even = (D^+ b).even (stagD.nim:626-632)."""
    stagD(s.se, r, s.g, b, m, -1.0)


def eoReconstruct(s, r, b, m):
    """r.odd = (b.odd - D_oe r.even)/m (stagD.nim:633-637)."""
    stagD(s.so, r, s.g, r, 0.0, -1.0 / m)
    r.axpby(1.0 / m, b, 1.0, "odd")


class SolverBackend(enum.IntEnum):
    """solverBase.nim:6-7 plus the HIP backend this library provides."""
    sbQex = 0
    sbQuda = 1
    sbGrid = 2
    sbHip = 3


class SloppyType(enum.IntEnum):
    """solverBase.nim:8-9, plus BF16 storage and 16-bit fixed-point links
    (SloppyHalfFixed: valid for sloppyLinks only, with FP16 or FP32 vectors)."""
    SloppyNone = 0
    SloppySingle = 1
    SloppyHalf = 2
    SloppyBF16 = 3
    SloppyHalfFixed = 4
    SloppyHalfAll = 5  # FP16 for every sloppy vector, residual and solution increment included


_SLOPPY_PREC = {SloppyType.SloppyNone: FP64, SloppyType.SloppySingle: FP32,
                SloppyType.SloppyHalf: FP16, SloppyType.SloppyBF16: BF16,
                SloppyType.SloppyHalfFixed: Q16, SloppyType.SloppyHalfAll: 5}


class SolverParams:
    """QEX SolverParams (solverBase.nim:10-63) with the HIP backend's extras."""

    def __init__(self):
        # inputs (defaults of solverBase.nim:49-59)
        self.r2req = 1e-6
        self.maxits = 50000
        self.backend = SolverBackend.sbHip
        self.sloppySolve = SloppyType.SloppyNone
        self.sloppyLinks = None  # None: same storage as the spinors
        self.recon = 18          # 12: 12-real links in the inner operator (plain SU(3) links)
        self.usePrevSoln = False
        self.verbosity = 1
        self.subsetName = "all"
        self.reliableDelta = 0.1
        self.reliableMaxIts = 100
        self.resetStats()

    def resetStats(self):
        self.calls = 0
        self.iterations = 0
        self.iterationsMax = 0
        self.seconds = 0.0
        self.flops = 0.0
        self.r2 = 0.0
        self.reliableUpdates = 0
        # multi-shift solves: FP64 true |b-(A+sigma_k)x_k|^2/|b|^2 per shift
        self.shiftR2 = []

    @property
    def finalIterations(self):
        return self.iterations

    def _c(self):
        if self.backend != SolverBackend.sbHip:
            raise QexHipError(f"backend {self.backend.name} is not provided by this library")
        c = SolverParamsC()
        c.r2req = self.r2req
        c.maxits = self.maxits
        c.sloppy = _SLOPPY_PREC[SloppyType(self.sloppySolve)]
        c.sloppy_links = -1 if self.sloppyLinks is None else _SLOPPY_PREC[SloppyType(self.sloppyLinks)]
        c.recon = int(self.recon)
        c.reliable_delta = self.reliableDelta
        c.reliable_max_its = self.reliableMaxIts
        c.use_prev_soln = int(self.usePrevSoln)
        c.verbosity = self.verbosity
        return c

    def _add(self, c):
        # addStats (solverBase.nim:75-83)
        self.calls += 1
        self.iterations += c.iterations
        self.iterationsMax = max(self.iterationsMax, c.iterations)
        self.seconds += c.seconds
        self.flops += c.flops
        self.r2 = c.r2
        self.reliableUpdates += c.reliable_updates

    def getStats(self):
        """'its secs Gf/s r2' (solverBase.nim:85-111, single-call form)."""
        s = max(1e-12, self.seconds)
        return f"{self.iterations}  {s:.4g}s  {1e-9 * self.flops / s:.4g}Gf/s  {self.r2:.4g}"


def newSolverParams():
    return SolverParams()


def _solve_xx(s, r, x, m, sp, par):
    c = sp._c()
    call("qexhip_solve_xx", s.h, par, r.h, x.h, float(m), ctypes.byref(c))
    sp._add(c)
    if sp.verbosity > 0:
        print(("solveEE: " if par == EVEN else "solveOO: ") + sp.getStats(), flush=True)
    return c


def solveEE(s, r, x, m, sp):
    """Solve 4(m^2 - D_eo D_oe) r = x on even sites (stagSolve.nim:423-424)."""
    return _solve_xx(s, r, x, m, sp, EVEN)


def solveOO(s, r, x, m, sp):
    """Solve 4(m^2 - D_oe D_eo) r = x on odd sites (stagSolve.nim:426-427)."""
    return _solve_xx(s, r, x, m, sp, ODD)


def solveMulti(s, xs, b, m, sigmas, sp, subset="even"):
    """Multi-shift CG: (4(m^2 - DD) + sigma_k) x_k = b, sigma_0 taken as 0
    (stagSolve.nim:164-229, cgm.nim:68-246)."""
    if len(xs) != len(sigmas):
        raise QexHipError("solveMulti: need one solution field per shift")
    c = sp._c()
    arr = (ctypes.c_void_p * len(xs))(*[f.h.value for f in xs])
    sig = (ctypes.c_double * len(sigmas))(*[float(v) for v in sigmas])
    r2s = (ctypes.c_double * len(sigmas))()
    c.shift_r2 = r2s
    call("qexhip_solve_multishift", s.h, _par(subset), arr, b.h, float(m), sig, len(sigmas), ctypes.byref(c))
    sp._add(c)
    sp.shiftR2 = list(r2s)
    return c


def _vec(ptrs):
    return (ctypes.c_void_p * len(ptrs))(*[p.h.value for p in ptrs])


def solveBatch(ss, xs, bs, masses, sps, subset="even", graph=True, tune=True):
    """Independent FP64 CG solves 4(m_k^2 - DD) x_k = b_k on one parity for
    systems with their own links (lattices may differ), advanced together on
    the GPU, one launch per solver stage for all of them (each system is the
    CG of solveEE/solveOO from x = 0, cg.nim:55-272).  sps: one SolverParams
    per system (r2req, maxits in; iterations and FP64 true residual out)."""
    n = len(ss)
    if not (len(xs) == len(bs) == len(masses) == len(sps) == n):
        raise QexHipError("solveBatch: one field, mass and SolverParams per system")
    cs = (_lib.SolverParamsC * n)(*[sp._c() for sp in sps])
    flags = (_lib.BATCH_GRAPH if graph else 0) | (_lib.BATCH_TUNE if tune else 0)
    call("qexhip_solve_batch", n, _vec(ss), _par(subset), _vec(xs), _vec(bs),
         (ctypes.c_double * n)(*[float(m) for m in masses]), cs, flags)
    for sp, c in zip(sps, cs):
        sp._add(c)
    return sps


def solveMultiBatch(s, xs, bs, m, sigmas, sps, subset="even", graph=True, tune=True, freeze=True):
    """Multi-shift CG for several right-hand sides on one set of links:
    (4(m^2 - DD) + sigma_k) xs[j][k] = bs[j], sigma_0 taken as 0 (cgm.nim:68-246
    per right-hand side); each link load serves up to 8 right-hand sides.
    freeze=False keeps updating shifted systems that already converged, as
    cgm.nim does."""
    nr, ns = len(bs), len(sigmas)
    if len(xs) != nr or any(len(row) != ns for row in xs) or len(sps) != nr:
        raise QexHipError("solveMultiBatch: xs must be nrhs lists of nshift fields, one SolverParams per rhs")
    cs = (_lib.SolverParamsC * nr)(*[sp._c() for sp in sps])
    r2s = [(ctypes.c_double * ns)() for _ in range(nr)]
    for c, a in zip(cs, r2s):
        c.shift_r2 = a
    flags = (_lib.BATCH_GRAPH if graph else 0) | (_lib.BATCH_TUNE if tune else 0) | \
        (0 if freeze else _lib.BATCH_NOFREEZE)
    call("qexhip_solve_multishift_batch", s.h, _par(subset), nr, _vec([f for row in xs for f in row]), _vec(bs),
         float(m), (ctypes.c_double * ns)(*[float(v) for v in sigmas]), ns, cs, flags)
    for sp, c, a in zip(sps, cs, r2s):
        sp._add(c)
        sp.shiftR2 = list(a)
    return sps


def set_autotune(on=True):
    """Launch-shape autotuning (on by default; QEXHIP_BLOCK pins the block size)."""
    call("qexhip_set_autotune", int(bool(on)))


def set_cg_graph(on=True):
    """CG launches: True/1 (default) replays HIP graphs, in the mixed-precision
    CG one graph per reliable-update interval (its iterations, the FP64
    reliable update and the best-iterate copy); 2 replays graphs of 16
    iterations with the reliable updates launched from the host; False/0
    launches kernel by kernel."""
    mode = on if isinstance(on, int) and not isinstance(on, bool) else int(bool(on))
    call("qexhip_set_cg_graph", mode)


def set_cg_cap(cap=-1):
    """Pin the workgroups of CG forms 3 and 4's persistent kernels: form 3 runs
    both stencil kernels on at most cap workgroups (0: one per block), form 4
    its streaming kernel (0: 512); -1 lets the tuner choose."""
    call("qexhip_set_cg_cap", int(cap))


def set_link_cache(lo=None, nbytes=-1):
    """Link bytes the stencils read with cache-allocating loads, so that a
    link field (or its first directions) can stay in the last-level cache
    between sweeps; the rest is read with the non-temporal hint.  lo: a
    Layout (None sets the default of Layouts created afterwards); nbytes: 0
    all non-temporal, -1 (the default) chosen by the tuner."""
    call("qexhip_set_link_cache", lo.h if lo is not None else None, int(nbytes))


def set_cg_block1(block1=0):
    """Pin the block size (64, 128, 256) of the CG iteration's other-parity
    stencil kernel (forms 3 and 4); 0 lets the tuner choose."""
    call("qexhip_set_cg_block1", int(block1))


def set_hop_walk(chunk=-1):
    """Pin the one-parity hop's site walk: 0 canonical order, a multiple of 256
    the t-first walk over chunks of that many sites of a t-slice; -1 lets the
    tuner choose."""
    call("qexhip_set_hop_walk", int(chunk))


def set_hop_two(two=-1):
    """Pin the one-parity hop's stencil loop: 1 both links of a hop per trip,
    0 one link per trip; -1 lets the tuner choose."""
    call("qexhip_set_hop_two", int(two))


def set_hop_split(split=-1):
    """Pin the one-parity hop's threads per output site: 2 or 4 split each
    site's links between that many threads; 0 or 1 one thread per site; -1
    lets the tuner choose."""
    call("qexhip_set_hop_split", int(split))


def set_hop_tmarch(slices=-1):
    """Pin the t-march form of the plain-link one-parity hop: each thread walks
    this many t-slices and keeps its forward t-neighbour in registers for the
    next slice but one; 0 off, -1 lets the tuner choose."""
    call("qexhip_set_hop_tmarch", int(slices))


def set_msb_walk(chunk=-1):
    """Pin the site walk of the batched multi-shift solve's two multi-RHS hops:
    0 canonical order, a multiple of 256 the t-first walk over chunks of that
    many sites of a t-slice; -1 lets the tuner choose (when the solve tunes)."""
    call("qexhip_set_msb_walk", int(chunk))


def set_hop_multi_shape(batches=0, chunk=-1):
    """Pin the multi-RHS hop's vector-load batches per link (1 or 2; 0: tuner)
    and its site walk (0: canonical order, a multiple of 256: t-first walk over
    chunks of that many sites of a t-slice; -1: tuner)."""
    call("qexhip_set_hop_multi_shape", int(batches), int(chunk))


def set_cg_form(form=0):
    """Pin the CG iteration form (2 or 3 kernels per iteration, or 4: form 3
    with the dot-product finish folded into the next iteration's streaming
    kernel); 0 lets the tuner choose."""
    call("qexhip_set_cg_form", int(form))


def autotune_report():
    """{kernel family and shape: {"block", "remap", "form", "cap", "block1", "batches", "chunk",
    "tmarch", "two", "split", "lkcache_mb", "us"}} chosen so far (form: 2-, 3- or 4-kernel CG iteration;
    cap: workgroups of a persistent form, 0 = one pass; block1: block size of
    the 3-kernel iteration's other-parity stencil kernel, 0 = block; batches:
    vector-load batches per link of the multi-RHS hop; chunk: t-first site-walk
    chunk, 0 = canonical order; tmarch: t-slices per thread of the hop's
    t-march form, 0 = off; two: the hop's stencil loop takes both links of a
    hop per trip; split: threads per output site of the hop; lkcache_mb: link
    megabytes read with cache-allocating loads,
    for lattices that leave that budget to the tuner; us: time of one timed
    unit)."""
    buf = ctypes.create_string_buffer(1 << 20)
    call("qexhip_autotune_report", buf, len(buf))
    out = {}
    for line in buf.value.decode().splitlines():
        key, _, rest = line.rpartition(": ")
        f = dict(kv.split("=") for kv in rest.split())
        out[key] = {"block": int(f["block"]), "remap": int(f["remap"]), "form": int(f["form"]),
                    "cap": int(f["cap"]), "block1": int(f.get("block1", 0)),
                    "batches": int(f.get("batches", 1)), "chunk": int(f.get("chunk", 0)),
                    "tmarch": int(f.get("tmarch", 0)), "two": int(f.get("two", 0)), "split": int(f.get("split", 1)),
                    "lkcache_mb": int(f.get("lkcache", 0)), "us": float(f["us"])}
    return out


def _sub_solve(s, x, b, m, sp, r2req, par):
    sub = SolverParams()
    sub.__dict__.update({k: v for k, v in sp.__dict__.items() if not k.startswith("_")})
    sub.resetStats()
    sub.r2req = r2req
    sub.maxits = sp.maxits - sp.iterations
    sub.verbosity = sp.verbosity - 1
    _solve_xx(s, x, b, m, sub, par)
    sp.iterations += sub.iterations
    sp.seconds += sub.seconds
    sp.flops += sub.flops
    sp.reliableUpdates += sub.reliableUpdates


def _solve_recon_r(s, x, b, m, sp, b2e, b2o):
    """solveReconR (stagSolve.nim:238-278): right-preconditioned."""
    lo = s.l
    b2 = b2e + b2o
    r2stop = sp.r2req * b2
    r2stop2 = 0.5 * r2stop
    r2stope = r2stop - b2o if b2o <= r2stop2 else r2stop2
    y = lo.ColorVector()
    if b2e > r2stope:
        _sub_solve(s, y, b, m, sp, r2stope / b2e, EVEN)
        y.axpby(0.0, y, 4.0, "even")
        Ddag(s, x, y, m)
        return
    r2stopo = r2stop - b2e if b2e <= r2stop2 else r2stop2
    if b2o > r2stopo:
        _sub_solve(s, y, b, m, sp, r2stopo / b2o, ODD)
        y.axpby(0.0, y, 4.0, "odd")
        Ddag(s, x, y, m)


def _solve_recon_l(s, x, b, m, sp, b2e, b2o):
    """solveReconL (stagSolve.nim:281-310): D^+ reduce, even solve, odd reconstruct."""
    lo = s.l
    d = lo.ColorVector()
    Ddag(s, d, b, m)
    x.zero()
    d2e = d.norm2("even")
    r2req = 0.99 * sp.r2req * (b2e + b2o) * m * m / d2e
    _sub_solve(s, x, d, m, sp, r2req, EVEN)
    x.axpby(0.0, x, 4.0, "even")
    eoReconstruct(s, x, b, m)


def solve(s, x, b, m, sp, cpuonly=False, sloppySolve=None):
    """Full-lattice D(m) x = b (stagSolve.nim:469-557): even-odd preconditioned
    inner solves (solveInner, :313-326) inside an FP64 true-residual loop.
    With a number in place of sp it is the residual r2req of QEX's short form
    (stagSolve.nim:683-693); cpuonly asks for QEX's own CPU backend, which
    this library does not provide, and is refused."""
    if isinstance(sp, (int, float)):
        res = float(sp)
        sp = SolverParams()
        sp.r2req = res
        sp.verbosity = 1
        sp.subsetName = "even"
        sp.sloppySolve = SloppyType.SloppyNone if sloppySolve is None else sloppySolve
        if cpuonly:
            sp.backend = SolverBackend.sbQex
    if sp.backend != SolverBackend.sbHip:
        raise QexHipError(f"backend {SolverBackend(sp.backend).name} is not provided by this library")
    lo = s.l
    b2 = b.norm2()
    r2stop = sp.r2req * b2
    r = lo.ColorVector()
    if sp.usePrevSoln:
        D(s, r, x, m)
        r.axpby(1.0, b, -1.0)
    else:
        x.zero()
        r.copy_from(b)
    r2e, r2o = r.norm2("even"), r.norm2("odd")
    r2 = r2e + r2o
    y = lo.ColorVector()
    its0 = sp.iterations
    inner = SolverParams()
    inner.__dict__.update({k: v for k, v in sp.__dict__.items() if not k.startswith("_")})
    inner.resetStats()
    while r2 > r2stop:
        inner.maxits = sp.maxits - inner.iterations
        if inner.maxits <= 0:
            break
        inner.r2req = r2stop / r2
        y.zero()
        rs = r2stop / r2 * (r2e + r2o)
        rs2 = 0.5 * rs
        r2stope = rs - r2o if r2o <= rs2 else rs2
        r2stopo = rs - r2e if r2e <= rs2 else rs2
        if r2e <= r2stope or r2o <= r2stopo or m == 0.0:
            _solve_recon_r(s, y, r, m, inner, r2e, r2o)
        else:
            _solve_recon_l(s, y, r, m, inner, r2e, r2o)
        x.axpby(1.0, y, 1.0)
        D(s, r, x, m)
        r.axpby(1.0, b, -1.0)
        r2e, r2o = r.norm2("even"), r.norm2("odd")
        r2 = r2e + r2o
    sp.calls += 1
    sp.iterations = its0 + inner.iterations
    sp.seconds += inner.seconds
    sp.flops += inner.flops
    sp.reliableUpdates += inner.reliableUpdates
    sp.r2 = r2 / b2 if b2 > 0 else 0.0
    return sp


def solveMultiMass(s, xs, b, ms, sp):
    """Multi-mass full-lattice solve (stagSolve.nim:560-668): CGM on one parity
    with shifts 4(m_k^2 - m_0^2), outer FP64 residual loop on m_0, then
    x_k = D^+(m_k) x_k."""
    lo = s.l
    if len(xs) != len(ms):
        raise QexHipError("solveMultiMass: need one solution field per mass")
    b2, b2e, b2o = b.norm2(), b.norm2("even"), b.norm2("odd")
    r = lo.ColorVector().copy_from(b)
    r2 = b2e + b2o
    r2stop = sp.r2req * b2
    shifts = [0.0] + [4.0 * (mk * mk - ms[0] * ms[0]) for mk in ms[1:]]
    for f in xs:
        f.zero()
    ys = [lo.ColorVector() for _ in ms]
    xt = lo.ColorVector()
    its = 0
    while r2 > r2stop:
        inner = SolverParams()
        inner.__dict__.update({k: v for k, v in sp.__dict__.items() if not k.startswith("_")})
        inner.resetStats()
        inner.maxits = sp.maxits - its
        if inner.maxits <= 0:
            break
        req = r2stop
        rs2 = 0.5 * req
        r2stope = req - b2o if b2o <= rs2 else rs2
        r2stopo = req - b2e if b2e <= rs2 else rs2
        if b2e > r2stope:
            inner.r2req, sub = r2stope / b2e, "even"
        elif b2o > r2stopo:
            inner.r2req, sub = r2stopo / b2o, "odd"
        else:
            break
        solveMulti(s, ys, r, ms[0], shifts, inner, subset=sub)
        its += inner.iterations
        for f, y in zip(xs, ys):
            f.axpby(4.0, y, 1.0, sub)
        Ddag(s, xt, xs[0], ms[0])
        D(s, r, xt, ms[0])
        r.axpby(1.0, b, -1.0)
        b2e, b2o = r.norm2("even"), r.norm2("odd")
        r2 = b2e + b2o
    for k, f in enumerate(xs):
        Ddag(s, xt, f, ms[k])
        f.copy_from(xt)
    sp.calls += 1
    sp.iterations += its
    sp.r2 = r2 / b2 if b2 > 0 else 0.0
    return sp


def stagForceSolve(s, psi, phi, mass, sp):
    """psi_e = (m^2 - D_eo D_oe)^-1 phi_e ; psi_o = (2D)_oe psi_e (hmcAction.nim:669-721)."""
    c = sp._c()
    call("qexhip_force_solve", s.h, psi.h, phi.h, float(mass), ctypes.byref(c))
    sp._add(c)
    return c


def fermForce(s, psi, out=None, outer_only=False, device=None):
    """F_mu(x) = TAH(eps(x) psi(x) psi(x+mu)^+ U_mu(x)^+) (hmcAction.nim:723-792, unsmeared).

    Returns a canonical (4, V, 3, 3) complex128 array (torch on `device` if given).
    With outer_only the outer products eps psi psi(x+mu)^+ are returned; for
    HISQ links (8 fields) as [2mu] = 1-hop and [2mu+1] = 3-hop (x+3mu)."""
    V = s.l.nSites
    nk = s.nl if outer_only else 4
    if out is None:
        if device is not None:
            out = torch.zeros((nk, V, 3, 3), dtype=torch.complex128, device=device)
        else:
            out = np.zeros((nk, V, 3, 3), dtype=np.complex128)
    p, dev = _ptr(out)
    call("qexhip_force", s.h, psi.h, p, dev, int(outer_only))
    return out


def toSingle(x):
    """toSingle (stagD.nim:56-86, fieldET.nim:346-352).  For an operator: the
    same operator, since this library keeps per-precision copies of the links
    and picks them by the precision of the fields it is applied to.  For a
    Field: a single-precision copy."""
    if isinstance(x, Staggered):
        return x
    if isinstance(x, Field):
        if torch is None or not torch.cuda.is_available():
            raise QexHipError("toSingle needs the GPU")
        return x.l.ColorVector(prec=FP32).set(x.get(device=torch.device("cuda")))
    raise QexHipError("toSingle: expected a Staggered operator or a Field")


def _outer_acc(s, a, b, ce, co, F):
    if not (torch is not None and isinstance(F, torch.Tensor) and F.is_cuda and F.dtype == torch.complex128
            and F.is_contiguous() and tuple(F.shape) == (s.nl, s.l.nSites, 3, 3)):
        raise QexHipError(f"derivative fields must be a contiguous complex128 CUDA tensor ({s.nl}, V, 3, 3)")
    torch.cuda.synchronize()
    call("qexhip_outer_acc", s.h, a.h, b.h, float(ce), float(co), ctypes.c_void_p(F.data_ptr()))


def stagD2deriv(s, g, c, x):
    """g_k += d/dU_k Re<c, (2D) x>: g_k(y) += c(y) x(y+d_k)^+ - x(y) c(y+d_k)^+ on
    every site (stagD.nim:640-684); g is (nl, V, 3, 3) complex128 on the device."""
    _outer_acc(s, c, x, 1.0, 1.0, g)
    _outer_acc(s, x, c, -1.0, -1.0, g)


def stagDeriv(s, f, x):
    """f_mu(y) += eps(y) x(y) x(y+mu)^+ (eps = +1 even, -1 odd), then f is rephased
    with the antiperiodic-t boundary and the staggered phases (stagD.nim:686-715,
    rephase :110-113); plain links, f (4, V, 3, 3) complex128 on the device."""
    if s.nl != 4:
        raise QexHipError("stagDeriv: plain (4-link) operators only")
    _outer_acc(s, x, x, 1.0, -1.0, f)
    rephase(s, f)


def rephase(s, g):
    """setBC then stagPhase on the plain link array g, in place (stagD.nim:110-113).
    The operator holds its own copy of its links: after rephasing them, pass
    them on with s.update(g)."""
    from . import gauge
    if tuple(g.shape) != (4, s.l.nSites, 3, 3):
        raise QexHipError("rephase: g must be (4, V, 3, 3)")
    gauge.phase_bc(s.l, g)


def time_hop(s, prec=FP64, nrep=100, recon=18):
    """Average ms of one one-parity hop (HIP events on the library stream)."""
    v = ctypes.c_double()
    call("qexhip_time_hop", s.h, prec, recon, nrep, ctypes.byref(v))
    return v.value


def time_hop_multi(s, prec=FP64, nrhs=8, nrep=100):
    """Average ms of one multi-RHS hop over nrhs fields (HIP events)."""
    v = ctypes.c_double()
    call("qexhip_time_hop_multi", s.h, prec, nrhs, nrep, ctypes.byref(v))
    return v.value


def stream_read_gbs(mbytes=2048, nrep=10):
    """Read bandwidth (GB/s) of the library's grid-stride streaming-read kernel:
    the practical ceiling for the bandwidth-bound kernels."""
    v = ctypes.c_double()
    call("qexhip_stream_read", int(mbytes), int(nrep), ctypes.byref(v))
    return v.value


def recon_ok(s):
    """True when the links are phased SU(3) matrices (12-real storage exact)."""
    v = ctypes.c_int()
    call("qexhip_links_recon_ok", s.h, ctypes.byref(v))
    return bool(v.value)
