"""qex_amd: MI355X-native staggered Dslash / CG / multi-shift CG / fermion force
behind QEX's solver-backend seam.  The compute path is the HIP library
qex_amd/libqexhip.so (C ABI in include/qexhip.h); this package is its host-side
mirror of QEX's Nim API."""
from ._lib import ALL, BF16, EVEN, FP16, FP32, FP64, ODD, QexHipError, load  # noqa: F401
from .stag import (peqDdag, rephase, stagD1, stagD1x, stagD2deriv, stagDb, stagDeriv, toSingle,  # noqa: F401
                   D, Ddag, Field, Layout, SloppyType, SolverBackend, SolverParams,
                   Staggered, eoReconstruct, eoReduce, fermForce, newLayout, newSolverParams, newStag,
                   newStag3, solve, solveBatch, solveEE, solveMulti, solveMultiBatch, solveMultiMass, solveOO,
                   stagD, stagD2, stagD2ee, stagD2oo, stagForceSolve, time_hop, autotune_report, set_autotune,
                   set_cg_form, set_cg_graph, set_hop_multi_shape, set_hop_walk, set_hop_tmarch, set_hop_two, set_hop_split, set_msb_walk, set_cg_block1, set_cg_cap, set_link_cache, stagD2multi, stagDN, time_hop_multi,
                   stream_read_gbs)
from .hisq import (Fat7lCoefs, HisqCoefs, fat7lDeriv, hisqForce, hisqStag, makeImpLinks, newHISQ,  # noqa: F401,E402
                   projectU, projectUderiv, smearGetForce)
from .fileio import (Reader, Writer, getFileLattice, loadGauge, newReader, newWriter,  # noqa: F401,E402
                     saveGauge)
from .cg import CgmState, CgPrecon, CgState, cgSolve, newCgmState, newCgState  # noqa: F401,E402
