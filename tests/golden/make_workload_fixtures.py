"""Write tests/golden/workloads.json: oracle results of the full-size solves.

The C oracle (oracle/qexoracle.c, the CPU restatement of QEX's stagD2xx,
cg.solve, solveXX and cgm.solve) is run once on the inputs the GPU workload
tests use, and its iteration counts, residuals and a strided sample of each
solution are kept as a fixture, so that the GPU tests can compare with the
oracle at sizes where the oracle takes minutes (the whole script runs about
half an hour on 8 host cores):

  plain_32x32x32x32   plain staggered 32^4, m = 0.01, solveEE to r2req 1e-20
                      (QEX stagSolve.nim:330-421 over cg.nim:55-272)
  hisq_32x32x32x64    HISQ-form 32^3 x 64 (W = 0.8U + 0.05G fat links,
                      L = -(1/24) U U U long links), m = 0.005, same solve
  cgm_24x24x24x48     HISQ-form 24^3 x 48, source 0 of the 8-source batch,
                      12 shifts log-spaced over [1e-4, 1] (QEX cgm.nim:68-246)
  batch16             16 plain solves on 8^4, 12^4, 16^4, 16^3 x 32 with
                      masses in [0.01, 0.1] (cg.nim:55-272)

Inputs are the oracle's seeded generators, which the GPU generators reproduce
to 1e-13 (tests/test_gpu_dslash.py::test_generators_match_oracle).

Usage: python tests/golden/make_workload_fixtures.py [name ...]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "workloads.json")
SEED = 20240601
R2REQ = 1e-20
NSAMPLE = 64


def sample_idx(ne):
    """64 even-site indices spread over the parity (first and last included)."""
    return [int(v) for v in np.linspace(0, ne - 1, NSAMPLE).round()]


def pack(x, idx):
    return [[[float(v.real), float(v.imag)] for v in x[i]] for i in idx]


def solve_case(name, L, hisq, m, gseed, sseed):
    U = O.phased(L, O.gauge_random(L, gseed))
    g = O.hisq_links(L, gseed + 1, U) if hisq else U
    del U
    b = O.source_gauss(L, sseed, 0, 0)
    ne = O.volume(L) // 2
    t0 = time.time()
    x, its, tr2 = O.solveXX(L, g, 0, b, m, R2REQ)
    dt = time.time() - t0
    idx = sample_idx(ne)
    return {
        "lattice": list(L), "hisq": hisq, "mass": m, "gauge_seed": gseed,
        "hisq_seed": gseed + 1 if hisq else None, "source_seed": sseed, "source_stream": 0,
        "r2req": R2REQ, "iterations": its, "true_r2": tr2,
        "x_norm2": O.norm2(x[:ne]), "re_b_dot_x": float(np.vdot(b[:ne], x[:ne]).real),
        "sample_sites": idx, "x_sample": pack(x, idx), "oracle_seconds": round(dt, 1),
    }


def cgm_case():
    L = (24, 24, 24, 48)
    gseed = SEED + 7
    U = O.phased(L, O.gauge_random(L, gseed))
    g = O.hisq_links(L, gseed + 1, U)
    del U
    sig = list(np.logspace(-4, 0, 12))
    m = 0.5 * sig[0] ** 0.5
    shifts = [0.0] + [v - sig[0] for v in sig[1:]]
    b = O.source_gauss(L, 900, 0, 0)
    ne = O.volume(L) // 2
    t0 = time.time()
    xs, its, r2 = O.cgm(L, g, 0, b, m, shifts, R2REQ)
    dt = time.time() - t0
    idx = sample_idx(ne)
    b2 = O.norm2(b[:ne])
    res = []
    for k, s in enumerate(shifts):
        r = b - O.stagD2xx(L, g, 0, xs[k], m * m) - s * xs[k]
        res.append(O.norm2(r[:ne]) / b2)
    return {
        "lattice": list(L), "hisq": True, "mass": m, "sigmas": sig, "shifts": shifts,
        "gauge_seed": gseed, "hisq_seed": gseed + 1, "source_seed": 900, "source_stream": 0,
        "r2req": R2REQ, "iterations": its, "true_r2_per_shift": res,
        "x_norm2_per_shift": [O.norm2(xs[k][:ne]) for k in range(len(shifts))],
        "sample_sites": idx, "x_sample_per_shift": [pack(xs[k], idx) for k in range(len(shifts))],
        "oracle_seconds": round(dt, 1),
    }


def batch16_case():
    """16 independent plain solves, 4 each on 8^4, 12^4, 16^4 and 16^3 x 32, with
    distinct seeds and masses uniform in [0.01, 0.1]: each one's oracle CG
    iteration count and true residual (cg.nim:55-272 from x = 0)."""
    lats = [(8, 8, 8, 8)] * 4 + [(12, 12, 12, 12)] * 4 + [(16, 16, 16, 16)] * 4 + [(16, 16, 16, 32)] * 4
    masses = [float(v) for v in np.random.default_rng(SEED).uniform(0.01, 0.1, len(lats))]
    gseeds = [SEED + 100 + k for k in range(len(lats))]
    sseeds = [300 + k for k in range(len(lats))]
    its, r2s = [], []
    t0 = time.time()
    for k, L in enumerate(lats):
        U = O.phased(L, O.gauge_random(L, gseeds[k]))
        b = O.source_gauss(L, sseeds[k], 0, 0)
        x, n, _ = O.cg(L, U, 0, b, masses[k], R2REQ)
        ne = O.volume(L) // 2
        r = b - O.stagD2xx(L, U, 0, x, masses[k] ** 2)
        its.append(n)
        r2s.append(O.norm2(r[:ne]) / O.norm2(b[:ne]))
    return {"lattices": [list(v) for v in lats], "masses": masses, "gauge_seeds": gseeds, "source_seeds": sseeds,
            "r2req": R2REQ, "iterations": its, "true_r2": r2s, "oracle_seconds": round(time.time() - t0, 1)}


CASES = {
    "batch16": batch16_case,
    "plain_32x32x32x32": lambda: solve_case("plain", (32, 32, 32, 32), False, 0.01, SEED, 7),
    "cgm_24x24x24x48": cgm_case,
    "hisq_32x32x32x64": lambda: solve_case("hisq", (32, 32, 32, 64), True, 0.005, SEED, 7),
}


def main(names):
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for name in names or list(CASES):
        print(f"{name} ...", flush=True)
        data[name] = CASES[name]()
        print(f"{name}: {data[name]['iterations']} iterations, {data[name]['oracle_seconds']} s", flush=True)
        data["_generator"] = "tests/golden/make_workload_fixtures.py (C oracle, FP64)"
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])
