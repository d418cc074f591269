"""Fermion force on a 24^4 plain staggered lattice, m = 0.05, phi = (D^+ eta)_e:
the GPU force against the CPU oracle's (same seeded inputs, independent
solves) and against central finite differences on 64 seeded links (h = 1e-5,
random su(3) directions; the action difference through the exact identity
S(U') - S(U) = -Y'^+ (M' - M) Y with (M' - M) applied through the link
difference)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = (24, 24, 24, 24)
MASS = 0.05


@pytest.fixture(scope="module")
def setup(gpu):
    from qex_amd import gauge, hmc
    lo = gpu.Layout(L)
    U = gauge.random_gauge(lo, 2024)
    Uph = gauge.phase_bc(lo, U.clone())
    s = gpu.Staggered(lo, Uph)
    phi = hmc.pseudofermion(lo, s, MASS, 2025)
    return lo, U, Uph, s, phi


def test_force_matches_oracle_24(oracle, gpu, setup):
    O = oracle
    lo, U, Uph, s, phi = setup
    fphi = lo.ColorVector().set(phi)
    psi = lo.ColorVector()
    sp = gpu.SolverParams()
    sp.r2req, sp.verbosity = 1e-26, 0
    gpu.stagForceSolve(s, psi, fphi, MASS, sp)
    F = gpu.fermForce(s, psi)
    # oracle: its own solve and force from the same links and pseudofermion
    g = Uph.cpu().numpy()
    psi_o, _ = O.force_solve(L, g, phi.cpu().numpy(), MASS, 1e-26)
    F_o = O.ferm_force(L, g, psi_o)
    assert np.linalg.norm(psi.get() - psi_o) / np.linalg.norm(psi_o) < 1e-10
    assert np.linalg.norm(F - F_o) / np.linalg.norm(F_o) < 1e-10
    # and the force kernel alone on the same psi
    F_same = O.ferm_force(L, g, psi.get())
    assert np.abs(F - F_same).max() <= 1e-12 * np.abs(F_same).max()


def test_force_finite_differences_64_links_24(gpu, setup):
    """Plain central differences at h = 1e-5 are truncation-limited (O(h^2),
    relative errors up to ~2e-8 on links whose directional derivative is small
    against the link's force); relative to the link's force magnitude the
    agreement is ~1e-10.  With the O(h^2) term removed (Richardson, h and 2h)
    every link agrees to 1e-8 relative."""
    from qex_amd import hmc
    lo, U, Uph, s, phi = setup
    det = []
    wf, wg, sec = hmc.force_fd_check(lo, Uph, phi, MASS, nlinks=64, seed=7, h=1e-5, beta=5.6, U=U, details=det)
    normed = max(abs(fd - an) / nf for _, _, fd, an, nf in det)
    print(f"24^4 FD h=1e-5, 64 links: fermion worst rel {wf:.3e} (force-normalized {normed:.3e}), "
          f"gauge worst rel {wg:.3e}, {sec:.1f} s")
    assert wf < 1e-7 and wg < 1e-7 and normed < 1e-9
    det = []
    wf, wg, sec = hmc.force_fd_check(lo, Uph, phi, MASS, nlinks=64, seed=7, h=1e-5, beta=5.6, U=U, details=det,
                                     richardson=True)
    normed = max(abs(fd - an) / nf for _, _, fd, an, nf in det)
    print(f"24^4 FD Richardson h=1e-5,2e-5, 64 links: fermion worst rel {wf:.3e} (force-normalized "
          f"{normed:.3e}), gauge worst rel {wg:.3e}, {sec:.1f} s")
    assert wf < 1e-8 and wg < 1e-8
