"""qexhip_stagD2xx_noalloc (the allocation-free variant of qexhip_stagD2xx)
against qexhip_stagD2xx, through the C ABI.

Two calls in a row, the second on a smaller lattice, in FP64 and FP32; the
variant keeps its odd-parity temporary in the library's static device buffer,
which the second call reuses with the first call's contents still in it.  The
hop has no reduction, so each site's result does not depend on the launch
shape: the two entry points agree bit for bit.  Also: the FP64 result against
the oracle's stagD2xx, and the two refusals (links without an FP32 copy, a
lattice larger than the buffer)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 4242


def _call(lib, name, *args):
    rc = getattr(lib, name)(*args)
    return rc, lib.qexhip_last_error().decode()


def _apply(gpu, lib, name, s, par, prec, x, m2):
    lo = s.l
    fx = lo.ColorVector(prec).set(x)
    fr = lo.ColorVector(prec)
    rc, msg = _call(lib, name, s.h, par, fr.h, fx.h, ctypes.c_double(m2))
    assert rc == 0, msg
    return fr.get(sub=par)


@pytest.mark.parametrize("prec_name", ["fp64", "fp32"])
def test_noalloc_matches_stagD2xx_two_calls(oracle, gpu, prec_name):
    from qex_amd import _lib
    O = oracle
    lib = _lib.load()
    prec = gpu.FP64 if prec_name == "fp64" else gpu.FP32
    # first call on 16^4, second on a smaller 8^4 lattice (the buffer keeps the
    # first call's temporary, of the larger lattice, when the second one runs)
    for k, (L, par, m2) in enumerate([((16, 16, 16, 16), 0, 0.01), ((8, 8, 8, 8), 1, 0.0225)]):
        print(f"noalloc {prec_name} call {k}: lattice {L} parity {par}", flush=True)
        U = O.phased(L, O.gauge_random(L, SEED + k))
        x = O.source_gauss(L, 7 + k, 0, par)
        lo = gpu.Layout(L)
        s = gpu.Staggered(lo, U)
        if prec == gpu.FP32:
            # the FP32 link copy is packed by an FP32 operator call
            _apply(gpu, lib, "qexhip_stagD2xx", s, par, prec, x, m2)
        got = _apply(gpu, lib, "qexhip_stagD2xx_noalloc", s, par, prec, x, m2)
        want = _apply(gpu, lib, "qexhip_stagD2xx", s, par, prec, x, m2)
        again = _apply(gpu, lib, "qexhip_stagD2xx_noalloc", s, par, prec, x, m2)
        assert np.array_equal(got, want), float(np.abs(got - want).max())
        assert np.array_equal(again, want)
        if prec == gpu.FP64:
            sl = slice(0, lo.nEven) if par == 0 else slice(lo.nEven, lo.nSites)
            ref = O.stagD2xx(L, U, par, x, m2)
            err = np.linalg.norm(got[sl] - ref[sl]) / np.linalg.norm(ref[sl])
            assert err < 1e-12, err


def test_noalloc_refusals(oracle, gpu):
    import torch
    from qex_amd import _lib, gauge
    lib = _lib.load()
    L = (8, 8, 8, 8)
    lo = gpu.Layout(L)
    s = gpu.Staggered(lo, oracle.phased(L, oracle.gauge_random(L, SEED)))
    fx, fr = lo.ColorVector(gpu.FP32), lo.ColorVector(gpu.FP32)
    rc, msg = _call(lib, "qexhip_stagD2xx_noalloc", s.h, 0, fr.h, fx.h, ctypes.c_double(0.01))
    assert rc != 0 and "not packed" in msg, msg
    # 48^3 x 64: 3 538 944 sites per parity need 170 MB of FP64 scratch
    Lb = (48, 48, 48, 64)
    lob = gpu.Layout(Lb)
    Ub = gauge.random_gauge(lob, SEED)
    gauge.phase_bc(lob, Ub)
    sb = gpu.Staggered(lob, Ub)
    del Ub
    torch.cuda.empty_cache()
    fxb, frb = lob.ColorVector(), lob.ColorVector()
    rc, msg = _call(lib, "qexhip_stagD2xx_noalloc", sb.h, 0, frb.h, fxb.h, ctypes.c_double(0.01))
    assert rc != 0 and "scratch" in msg, msg
