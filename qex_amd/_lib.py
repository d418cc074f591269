"""ctypes binding of the qexhip C ABI (include/qexhip.h).

The product path: every call goes to the HIP library built in-tree
(qex_amd/libqexhip.so).  There is no CPU fallback; loading fails loudly
when the library is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("QEXHIP_LIB") or os.path.join(_HERE, "libqexhip.so")

FP64, FP32, FP16, BF16, Q16 = 0, 1, 2, 3, 4
EVEN, ODD, ALL = 0, 1, 2

# every symbol include/qexhip.h declares
EXPORTS = [
    "qexhip_init", "qexhip_finalize", "qexhip_last_error", "qexhip_sync", "qexhip_stream",
    "qexhip_lattice_create", "qexhip_lattice_destroy",
    "qexhip_links_create", "qexhip_links_update", "qexhip_links_destroy",
    "qexhip_field_create", "qexhip_field_destroy", "qexhip_field_import", "qexhip_field_export",
    "qexhip_field_norm2", "qexhip_field_redot", "qexhip_field_axpby",
    "qexhip_stagD2", "qexhip_stagD", "qexhip_stagD2xx",
    "qexhip_solve_xx", "qexhip_solve_ee_host", "qexhip_solve_oo_host", "qexhip_solve_multishift",
    "qexhip_force_solve", "qexhip_force",
    "qexhip_gen_gauge", "qexhip_gen_source", "qexhip_gen_phase_bc", "qexhip_gen_hisq",
    "qexhip_time_hop", "qexhip_links_recon_ok",
    "qexhip_plaquette", "qexhip_wilson_force", "qexhip_random_momenta", "qexhip_link_update",
    "qexhip_solve_batch", "qexhip_solve_multishift_batch", "qexhip_set_autotune", "qexhip_autotune_report",
    "qexhip_set_cg_form", "qexhip_stagD2_multi", "qexhip_time_hop_multi", "qexhip_stream_read",
    "qexhip_set_cg_graph",
    "qexhip_hisq_coefs_init", "qexhip_make_imp_links", "qexhip_project_u", "qexhip_hisq_smear",
    "qexhip_time_hisq_smear", "qexhip_fat7l_deriv", "qexhip_project_u_deriv", "qexhip_hisq_smear_deriv",
    "qexhip_hisq_force", "qexhip_milc_read_info", "qexhip_milc_read", "qexhip_milc_write",
    "qexhip_scidac_read_info", "qexhip_scidac_read", "qexhip_scidac_create", "qexhip_scidac_append",
    "qexhip_time_scidac", "qexhip_from_qex_layout", "qexhip_to_qex_layout", "qexhip_outer_acc",
    "qexhip_stagD2xx_noalloc", "qexhip_set_hop_multi_shape", "qexhip_set_hop_walk", "qexhip_set_hop_tmarch", "qexhip_set_hop_two", "qexhip_set_hop_split", "qexhip_set_msb_walk", "qexhip_set_cg_block1", "qexhip_set_cg_cap", "qexhip_set_link_cache",
]

# qexhip_solve_batch / qexhip_solve_multishift_batch flags
BATCH_GRAPH, BATCH_TUNE, BATCH_NOFREEZE = 1, 2, 4


class SolverParamsC(ctypes.Structure):
    _fields_ = [
        ("r2req", ctypes.c_double),
        ("maxits", ctypes.c_int),
        ("sloppy", ctypes.c_int),
        ("sloppy_links", ctypes.c_int),
        ("recon", ctypes.c_int),
        ("reliable_delta", ctypes.c_double),
        ("reliable_max_its", ctypes.c_int),
        ("use_prev_soln", ctypes.c_int),
        ("verbosity", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("reliable_updates", ctypes.c_int),
        ("r2", ctypes.c_double),
        ("seconds", ctypes.c_double),
        ("flops", ctypes.c_double),
        ("shift_r2", ctypes.POINTER(ctypes.c_double)),
    ]


class Fat7CoefsC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("one_link", "three_staple", "five_staple", "seven_staple", "lepage")]


class HisqCoefsC(ctypes.Structure):
    _fields_ = [("fat7first", Fat7CoefsC), ("fat7second", Fat7CoefsC), ("naik", ctypes.c_double)]


class MilcInfoC(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int * 4), ("timestamp", ctypes.c_char * 65), ("order", ctypes.c_int),
                ("byteswapped", ctypes.c_int), ("sum29", ctypes.c_uint32), ("sum31", ctypes.c_uint32),
                ("checksum_ok", ctypes.c_int)]


class ScidacInfoC(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int * 4), ("nrecords", ctypes.c_int), ("datatype", ctypes.c_char * 64),
                ("date", ctypes.c_char * 64), ("precision", ctypes.c_char), ("colors", ctypes.c_int),
                ("spins", ctypes.c_int), ("typesize", ctypes.c_int), ("datacount", ctypes.c_int),
                ("ildg", ctypes.c_int), ("has_checksum", ctypes.c_int), ("suma", ctypes.c_uint32),
                ("sumb", ctypes.c_uint32), ("calc_suma", ctypes.c_uint32), ("calc_sumb", ctypes.c_uint32),
                ("checksum_ok", ctypes.c_int), ("filemd_len", ctypes.c_longlong), ("recmd_len", ctypes.c_longlong)]


class QexHipError(RuntimeError):
    pass


_lib = None


def load(path=LIB_PATH):
    """Load the library and declare the ABI (no GPU call is made here)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise QexHipError(f"HIP library not built: {path} (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    vp, ip, dp, I, D = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), \
        ctypes.c_int, ctypes.c_double
    vpp = ctypes.POINTER(ctypes.c_void_p)
    spp = ctypes.POINTER(SolverParamsC)
    U64 = ctypes.c_uint64
    sig = {
        "qexhip_init": [I], "qexhip_finalize": [], "qexhip_sync": [],
        "qexhip_lattice_create": [ip, vpp], "qexhip_lattice_destroy": [vp],
        "qexhip_links_create": [vp, I, vp, I, vpp], "qexhip_links_update": [vp, vp, I],
        "qexhip_links_destroy": [vp],
        "qexhip_field_create": [vp, I, vpp], "qexhip_field_destroy": [vp],
        "qexhip_field_import": [vp, I, vp, I], "qexhip_field_export": [vp, I, vp, I],
        "qexhip_field_norm2": [vp, I, dp], "qexhip_field_redot": [vp, vp, I, dp],
        "qexhip_field_axpby": [vp, vp, I, D, D],
        "qexhip_stagD2": [vp, I, vp, vp, D, D], "qexhip_stagD": [vp, I, vp, vp, D, D, D],
        "qexhip_stagD2xx": [vp, I, vp, vp, D], "qexhip_stagD2xx_noalloc": [vp, I, vp, vp, D],
        "qexhip_solve_xx": [vp, I, vp, vp, D, spp],
        "qexhip_solve_ee_host": [vp, vp, vp, D, spp], "qexhip_solve_oo_host": [vp, vp, vp, D, spp],
        "qexhip_solve_multishift": [vp, I, vpp, vp, D, dp, I, spp],
        "qexhip_force_solve": [vp, vp, vp, D, spp], "qexhip_force": [vp, vp, vp, I, I],
        "qexhip_gen_gauge": [vp, U64, vp], "qexhip_gen_source": [vp, U64, I, I, vp],
        "qexhip_gen_phase_bc": [vp, vp], "qexhip_gen_hisq": [vp, U64, vp, vp],
        "qexhip_time_hop": [vp, I, I, I, dp], "qexhip_links_recon_ok": [vp, ip],
        "qexhip_plaquette": [vp, vp, dp], "qexhip_wilson_force": [vp, vp, D, vp],
        "qexhip_random_momenta": [vp, U64, vp], "qexhip_link_update": [vp, vp, D, vp],
        "qexhip_solve_batch": [I, vpp, I, vpp, vpp, dp, spp, I],
        "qexhip_solve_multishift_batch": [vp, I, I, vpp, vpp, D, dp, I, spp, I],
        "qexhip_set_autotune": [I], "qexhip_autotune_report": [ctypes.c_char_p, I],
        "qexhip_set_cg_form": [I], "qexhip_set_hop_multi_shape": [I, ctypes.c_long], "qexhip_set_hop_walk": [ctypes.c_long], "qexhip_set_hop_tmarch": [I], "qexhip_set_hop_two": [I], "qexhip_set_hop_split": [I], "qexhip_set_msb_walk": [ctypes.c_long], "qexhip_set_cg_block1": [I], "qexhip_set_cg_cap": [I], "qexhip_set_link_cache": [ctypes.c_void_p, ctypes.c_long], "qexhip_set_cg_graph": [I],
        "qexhip_stagD2_multi": [vp, I, I, vpp, vpp, D, D], "qexhip_time_hop_multi": [vp, I, I, I, dp],
        "qexhip_stream_read": [ctypes.c_long, I, dp],
        "qexhip_make_imp_links": [vp, vp, vp, ctypes.POINTER(Fat7CoefsC), vp, vp, D],
        "qexhip_project_u": [ctypes.c_long, vp, vp],
        "qexhip_hisq_smear": [vp, ctypes.POINTER(HisqCoefsC), vp, vp, vp],
        "qexhip_time_hisq_smear": [vp, vp, vp, vp, I, dp],
        "qexhip_fat7l_deriv": [vp, vp, vp, vp, ctypes.POINTER(Fat7CoefsC), vp, vp, vp, D],
        "qexhip_project_u_deriv": [ctypes.c_long, vp, vp, vp, vp],
        "qexhip_hisq_smear_deriv": [vp, ctypes.POINTER(HisqCoefsC), vp, vp, vp, vp],
        "qexhip_hisq_force": [vp, ctypes.POINTER(HisqCoefsC), vp, vp, vp],
        "qexhip_milc_read_info": [ctypes.c_char_p, ctypes.POINTER(MilcInfoC)],
        "qexhip_milc_read": [vp, ctypes.c_char_p, vp, ctypes.POINTER(MilcInfoC)],
        "qexhip_milc_write": [vp, ctypes.c_char_p, vp, ctypes.c_char_p],
        "qexhip_scidac_read_info": [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ScidacInfoC), ctypes.c_char_p,
                                    ctypes.c_longlong, ctypes.c_char_p, ctypes.c_longlong],
        "qexhip_scidac_read": [vp, ctypes.c_char_p, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(ScidacInfoC)],
        "qexhip_scidac_create": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.c_char_p],
        "qexhip_scidac_append": [vp, ctypes.c_char_p, ctypes.c_char_p, vp, ctypes.c_int, ctypes.c_int, ctypes.c_char,
                                 ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ScidacInfoC)],
        "qexhip_time_scidac": [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_char, ctypes.c_int, dp],
        "qexhip_from_qex_layout": [vp, ip, vp, I, I, vp], "qexhip_to_qex_layout": [vp, ip, vp, I, I, vp],
        "qexhip_outer_acc": [vp, vp, vp, D, D, vp],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    lib.qexhip_hisq_coefs_init.restype = None
    lib.qexhip_hisq_coefs_init.argtypes = [ctypes.POINTER(HisqCoefsC)]
    lib.qexhip_last_error.restype = ctypes.c_char_p
    lib.qexhip_last_error.argtypes = []
    lib.qexhip_stream.restype = ctypes.c_void_p
    lib.qexhip_stream.argtypes = []
    _lib = lib
    return lib


def call(name, *args):
    """Call an ABI function; raise QexHipError with the library's message on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.qexhip_last_error().decode(errors="replace")
        raise QexHipError(f"{name}: {msg}")
    return rc


_initialized = False


def init(device=0):
    global _initialized
    if not _initialized:
        call("qexhip_init", device)
        _initialized = True
