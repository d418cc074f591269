"""Synthetic inputs generated on the GPU by the qexhip library.

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
