"""Writes the small golden fixtures under tests/golden/ from the CPU oracle.

QEX ships no golden vectors for the staggered Dslash, CG or force (its tests
print, not assert, these quantities), so the fixtures are oracle outputs on
seeded synthetic inputs; the oracle itself is pinned by tests/test_oracle.py
(closed forms, free-field spectrum, dense operator, finite differences).
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def main():
    L = (4, 4, 4, 4)
    seed = 31337
    U = O.phased(L, O.gauge_random(L, seed))
    g8 = O.hisq_links(L, seed + 1, U)
    x = O.source_gauss(L, 5, 0, 2)
    out = {"L": np.array(L), "seed": np.array(seed), "x": x}
    for name, g in (("plain", U), ("hisq", g8)):
        r = np.zeros_like(x)
        O.stagD2(L, g, 0, r, x, 0.0, 0.0)
        out[f"{name}_hop_even"] = r
        out[f"{name}_D_m0.1"] = O.D(L, g, x, 0.1)
        out[f"{name}_D2ee_m2_0.01"] = O.stagD2xx(L, g, 0, x, 0.01)
        b = O.source_gauss(L, 6, 0, 0)
        sol, its, tr = O.solveXX(L, g, 0, b, 0.2, 1e-22)
        out[f"{name}_cg_b"] = b
        out[f"{name}_cg_x"] = sol
        out[f"{name}_cg_its"] = np.array(its)
    phi = O.source_gauss(L, 7, 0, 0)
    psi, _ = O.force_solve(L, U, phi, 0.1, 1e-26)
    out["force_phi"] = phi
    out["force_psi"] = psi
    out["force_F"] = O.ferm_force(L, U, psi)
    # multi-shift CG on the plain operator
    sig = np.array([0.0, 0.05, 0.5])
    bm = O.source_gauss(L, 8, 0, 0)
    xs, its, _ = O.cgm(L, U, 0, bm, 0.1, sig, 1e-22)
    out["cgm_sigmas"], out["cgm_b"], out["cgm_x"], out["cgm_its"] = sig, bm, xs, np.array(its)
    # gauge-field pieces of the HMC, on the unphased links
    Uu = O.gauge_random(L, seed)
    out["plaquette"] = np.array(O.plaquette(L, Uu))
    out["wilson_force_b5.6"] = O.wilson_force(L, Uu, 5.6)
    P = O.random_momenta(L, seed + 2)
    out["momenta"] = P
    W = Uu.copy()
    O.link_update(L, P, 0.1, W)
    out["link_update_0.1"] = W
    np.savez_compressed(os.path.join(HERE, "golden_4x4x4x4.npz"), **out)
    # U itself is regenerated from the seed by the generator under test
    print("wrote golden_4x4x4x4.npz")


if __name__ == "__main__":
    main()
