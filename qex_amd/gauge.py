"""Gauge fields on the GPU through the qexhip library: synthetic inputs and MILC
binary gauge files.

Philox4x32-10 keyed by (seed, lexicographic site, stream, component) so the
fields do not depend on the layout; U = exp(i*0.5*H) with H traceless
Hermitian (unit Gaussian entries); complex-Gaussian sources with E|z|^2 = 1;
QEX's setBC + stagPhase (src/gauge/gaugeUtils.nim:133-140,
src/physics/stagD.nim:560-571); HISQ-form links W = 0.8U + 0.05G,
L = -(1/24) U U U (src/gauge/fat7l.nim:145-154).  Arrays are canonical
(QEX single-rank site order) torch complex128 tensors on the GPU.
"""
import ctypes

import torch

from ._lib import call


def _dev(device):
    return torch.device(device if device is not None else "cuda")


def random_gauge(lo, seed, device=None):
    U = torch.zeros((4, lo.nSites, 3, 3), dtype=torch.complex128, device=_dev(device))
    torch.cuda.synchronize()
    call("qexhip_gen_gauge", lo.h, ctypes.c_uint64(seed), ctypes.c_void_p(U.data_ptr()))
    return U


def unit_gauge(lo, device=None):
    U = torch.zeros((4, lo.nSites, 3, 3), dtype=torch.complex128, device=_dev(device))
    for a in range(3):
        U[:, :, a, a] = 1.0
    torch.cuda.synchronize()
    return U


def phase_bc(lo, U):
    """setBC (antiperiodic t) then stagPhase, in place, on plain links."""
    torch.cuda.synchronize()
    call("qexhip_gen_phase_bc", lo.h, ctypes.c_void_p(U.data_ptr()))
    return U


def hisq_links(lo, seed, Uph):
    g8 = torch.zeros((8, lo.nSites, 3, 3), dtype=torch.complex128, device=Uph.device)
    torch.cuda.synchronize()
    call("qexhip_gen_hisq", lo.h, ctypes.c_uint64(seed), ctypes.c_void_p(Uph.data_ptr()),
         ctypes.c_void_p(g8.data_ptr()))
    return g8


def source(lo, seed, stream=0, par=0, device=None):
    x = torch.zeros((lo.nSites, 3), dtype=torch.complex128, device=_dev(device))
    torch.cuda.synchronize()
    call("qexhip_gen_source", lo.h, ctypes.c_uint64(seed), int(stream), int(par), ctypes.c_void_p(x.data_ptr()))
    return x


def milcHeader(path):
    """Header of a MILC binary gauge file (src/new/milcIO.nim:9-25) as a dict."""
    from ._lib import MilcInfoC
    info = MilcInfoC()
    call("qexhip_milc_read_info", str(path).encode(), ctypes.byref(info))
    return _info(info)


def _info(info):
    return {"dims": tuple(info.dims), "timestamp": info.timestamp.decode(errors="replace"), "order": info.order,
            "byteswapped": bool(info.byteswapped), "sum29": info.sum29, "sum31": info.sum31,
            "checksum_ok": bool(info.checksum_ok)}


def readMilc(lo, path, device=None):
    """Read a MILC gauge file (openMilcFileGaugeField, src/new/milcFileField.nim:46-57)
    into canonical (4, V, 3, 3) complex128 links on the GPU; the checksum is verified.
    Returns (U, header dict)."""
    from ._lib import MilcInfoC
    U = torch.zeros((4, lo.nSites, 3, 3), dtype=torch.complex128, device=_dev(device))
    info = MilcInfoC()
    torch.cuda.synchronize()
    call("qexhip_milc_read", lo.h, str(path).encode(), ctypes.c_void_p(U.data_ptr()), ctypes.byref(info))
    return U, _info(info)


def writeMilc(lo, path, U, timestamp=None):
    """Write canonical links U as a MILC gauge file in single precision with its
    checksum (newMilcFileGaugeField, src/new/milcFileField.nim:30-37,58-69)."""
    if not (U.is_cuda and U.dtype == torch.complex128 and U.is_contiguous() and U.shape == (4, lo.nSites, 3, 3)):
        raise ValueError("U must be a contiguous complex128 CUDA tensor of shape (4, nSites, 3, 3)")
    torch.cuda.synchronize()
    call("qexhip_milc_write", lo.h, str(path).encode(), ctypes.c_void_p(U.data_ptr()),
         timestamp.encode() if timestamp else None)


def _qex_shape(lo, innerGeom, nfields, ncomp):
    ig = [int(v) for v in innerGeom]
    if len(ig) != 4:
        raise ValueError("innerGeom needs 4 entries")
    return (ctypes.c_int * 4)(*ig), (nfields, lo.nSites * ncomp * 2)


def fromQexLayout(lo, innerGeom, raw, ncomp):
    """Canonical (nfields, V, ncomp) complex128 fields from QEX's own storage
    of Layout(innerGeom) fields: raw is a float64 CUDA tensor (nfields,
    V*ncomp*2) holding each field's site data as QEX keeps it (the host repack
    of quda/qudaWrapperImpl.nim:204-246, done on the device)."""
    ig, shape = _qex_shape(lo, innerGeom, raw.shape[0], ncomp)
    if not (raw.is_cuda and raw.dtype == torch.float64 and raw.is_contiguous() and tuple(raw.shape) == shape):
        raise ValueError(f"raw must be a contiguous float64 CUDA tensor of shape {shape}")
    out = torch.empty((shape[0], lo.nSites, ncomp), dtype=torch.complex128, device=raw.device)
    torch.cuda.synchronize()
    call("qexhip_from_qex_layout", lo.h, ig, ctypes.c_void_p(raw.data_ptr()), shape[0], ncomp,
         ctypes.c_void_p(out.data_ptr()))
    return out


def toQexLayout(lo, innerGeom, fields):
    """QEX's own storage (nfields, V*ncomp*2) float64 of canonical fields
    (nfields, V, ncomp) complex128."""
    if not (fields.is_cuda and fields.dtype == torch.complex128 and fields.is_contiguous() and fields.dim() == 3
            and fields.shape[1] == lo.nSites):
        raise ValueError("fields must be a contiguous complex128 CUDA tensor (nfields, V, ncomp)")
    nf, _, ncomp = fields.shape
    ig, shape = _qex_shape(lo, innerGeom, nf, ncomp)
    out = torch.empty(shape, dtype=torch.float64, device=fields.device)
    torch.cuda.synchronize()
    call("qexhip_to_qex_layout", lo.h, ig, ctypes.c_void_p(fields.data_ptr()), nf, ncomp,
         ctypes.c_void_p(out.data_ptr()))
    return out
