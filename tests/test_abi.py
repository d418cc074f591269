"""The C-ABI library loads and exports every symbol include/qexhip.h declares
(no GPU call is made); the host-side mirror of QEX's SolverParams behaves like
solverBase.nim."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "qexhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qexhip_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_matches_python_export_list():
    from qex_amd import _lib
    assert header_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from qex_amd import _lib
    lib = _lib.load()
    for name in header_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (qexhip_\w+)", out))
    assert set(header_functions()) <= exported
    # nothing else with the qexhip_ prefix leaks out
    assert exported == set(header_functions())


def test_library_is_gfx950():
    from qex_amd import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_solver_params_struct_layout():
    from qex_amd._lib import SolverParamsC
    # doubles and ints in declaration order, natural alignment (include/qexhip.h)
    assert ctypes.sizeof(SolverParamsC) == 88
    assert SolverParamsC.reliable_delta.offset == 24 and SolverParamsC.r2.offset == 56
    assert SolverParamsC.shift_r2.offset == 80


def test_solver_params_defaults_like_qex():
    from qex_amd.stag import SolverParams, SolverBackend, SloppyType
    sp = SolverParams()
    # solverBase.nim:49-59
    assert sp.r2req == 1e-6 and sp.maxits == 50000 and sp.verbosity == 1
    assert sp.sloppySolve == SloppyType.SloppyNone and sp.usePrevSoln is False
    assert sp.backend == SolverBackend.sbHip
    c = sp._c()
    assert c.sloppy == 0 and c.sloppy_links == -1 and c.reliable_delta == 0.1 and c.reliable_max_its == 100


def test_other_backends_refused():
    from qex_amd.stag import SolverParams, SolverBackend
    from qex_amd import QexHipError
    sp = SolverParams()
    for be in (SolverBackend.sbQex, SolverBackend.sbQuda, SolverBackend.sbGrid):
        sp.backend = be
        with pytest.raises(QexHipError):
            sp._c()


def test_stats_string_format():
    from qex_amd._lib import SolverParamsC
    from qex_amd.stag import SolverParams
    sp = SolverParams()
    c = SolverParamsC()
    c.iterations, c.seconds, c.flops, c.r2 = 26, 0.5, 2e9, 3.298e-21
    sp._add(c)
    # solverBase.nim:110-111 single-call form: "its  secs  Gf/s  r2"
    assert sp.getStats() == "26  0.5s  4Gf/s  3.298e-21"
    assert sp.calls == 1 and sp.iterationsMax == 26
