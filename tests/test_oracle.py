"""Pins the CPU oracle (oracle/qexoracle.c) against known answers and against
what the mathematics fixes, chosen so that a dropped term, a wrong sign, a
wrong index or a transposed operand fails one of them:

 - Philox4x32-10 Random123 known-answer vectors;
 - closed-form point-source norms (QEX stagD.nim:735-752: 8.04 and 2.01);
 - free-field spectrum: H^2 psi_p = -4 sum_mu (sin p_mu + c3 sin 3p_mu)^2 psi_p
   for plane waves with antiperiodic t (plain c3 = 0, HISQ form c3 = -1/24);
 - dense 4^4 operator: anti-Hermiticity of the hop, eps-symmetry, spectrum
   m +- i*lambda, and dense LU solves against CG / solveXX;
 - gauge covariance of the hop (plain and HISQ form) and of the force;
 - multi-shift CG: each shifted system's own residual and agreement with
   independent single-shift CG;
 - the fermion force against central finite differences of
   S = phi^+ (m^2 - D_eo D_oe)^-1 phi along random su(3) directions.
"""
import numpy as np
import pytest
import scipy.linalg

SEED = 4242


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ------------------------------------------------------------------- basics
def test_philox_known_answers(oracle):
    # Random123 kat_vectors for philox4x32_10
    assert oracle.philox((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert oracle.philox((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2) == (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)
    assert oracle.philox((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)) == \
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_layout_bijection_even_first(oracle):
    L = (4, 6, 2, 8)
    V = oracle.volume(L)
    C = oracle.all_coords(L)
    seen = set()
    for idx in range(V):
        c = oracle.coords(L, idx)
        assert tuple(C[idx]) == c
        assert oracle.index(L, c) == idx
        assert ((sum(c) & 1) == 1) == (idx >= V // 2)
        seen.add(c)
    assert len(seen) == V


def test_gauge_links_are_su3(oracle):
    L = (4, 4, 4, 4)
    U = oracle.gauge_random(L, SEED)
    eye = np.eye(3)
    assert np.abs(U @ np.conj(np.swapaxes(U, -1, -2)) - eye).max() < 1e-14
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-14
    # U = exp(i*0.5*H) with H ~ unit Gaussian: |tr U|/3 well away from 1
    assert 0.2 < np.abs(np.trace(U, axis1=-2, axis2=-1)).mean() / 3 < 0.9


def test_su3_exp_matches_scipy(oracle):
    rng = np.random.default_rng(1)
    A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    assert np.abs(oracle.su3_exp(A) - scipy.linalg.expm(A)).max() < 1e-13
    H = A + A.conj().T
    assert np.abs(oracle.su3_exp_i(H, 0.5) - scipy.linalg.expm(0.5j * H)).max() < 1e-13


def test_project_tah(oracle):
    rng = np.random.default_rng(2)
    X = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    R = oracle.project_tah(X)
    assert np.abs(R + R.conj().T).max() < 1e-15
    assert abs(np.trace(R)) < 1e-15
    want = 0.5 * (X - X.conj().T)
    want -= np.trace(want) / 3 * np.eye(3)
    assert np.abs(R - want).max() < 1e-15


def test_phases_and_bc(oracle):
    L = (4, 4, 4, 6)
    U = oracle.gauge_unit(L)
    W = oracle.phased(L, U)
    C = oracle.all_coords(L)
    x, y, z, t = C.T
    eta = [(-1.0) ** t, (-1.0) ** (x + t), (-1.0) ** (x + y + t), np.ones(len(x))]
    eta[3] = np.where(t == L[3] - 1, -1.0, 1.0)
    for mu in range(4):
        assert np.allclose(W[mu, :, 0, 0], eta[mu])
        assert np.allclose(W[mu, :, 1, 1], eta[mu])


# ------------------------------------------------------- closed-form norms
def test_point_source_norms(oracle):
    """stagD.nim:735-752 on unit links with phases and BC, delta at site 0."""
    L = (4, 4, 4, 4)
    U = oracle.phased(L, oracle.gauge_unit(L))
    x = np.zeros((256, 3), complex)
    x[0, 0] = 1.0
    m = 0.1
    r = np.zeros_like(x)
    oracle.stagD2(L, U, 2, r, x, 0.0, m / 0.5)  # stagDb: m/(0.5 sc), no 1/2
    assert abs(oracle.norm2(r) - ((2 * m) ** 2 + 8)) < 1e-14
    assert abs(oracle.norm2(oracle.D(L, U, x, m)) - (m * m + 8 * 0.25)) < 1e-14


def test_point_source_norm_naik(oracle):
    L = (8, 8, 8, 8)
    Uph = oracle.phased(L, oracle.gauge_unit(L))
    g8 = np.empty((8,) + Uph.shape[1:], complex)
    g8[0::2] = Uph
    g8[1::2] = long_links_numpy(L, Uph)
    x = np.zeros((oracle.volume(L), 3), complex)
    x[5, 1] = 1.0
    m = 0.3
    want = m * m + 0.25 * (8 + 8 / 24.0 ** 2)
    assert abs(oracle.norm2(oracle.D(L, g8, x, m)) - want) < 1e-13


def long_links_numpy(L, Uph):
    """-(1/24) U_mu(x) U_mu(x+mu) U_mu(x+2mu) computed independently in numpy."""
    C = _coords_np(L)
    out = np.empty_like(Uph)
    for mu in range(4):
        c1 = C.copy(); c1[:, mu] = (c1[:, mu] + 1) % L[mu]
        c2 = C.copy(); c2[:, mu] = (c2[:, mu] + 2) % L[mu]
        i1, i2 = _index_np(L, c1), _index_np(L, c2)
        out[mu] = (-1 / 24) * Uph[mu] @ Uph[mu][i1] @ Uph[mu][i2]
    return out


def _coords_np(L):
    from oracle import oracle as O
    return O.all_coords(L)


def _index_np(L, C):
    V = int(np.prod(L))
    lex = C[:, 0] + L[0] * (C[:, 1] + L[1] * (C[:, 2] + L[2] * C[:, 3]))
    p = C.sum(1) & 1
    return (lex + p * V) // 2


def test_hisq_links_long_link_matches_numpy(oracle):
    L = (4, 4, 4, 6)
    Uph = oracle.phased(L, oracle.gauge_random(L, SEED))
    g8 = oracle.hisq_links(L, SEED + 1, Uph)
    assert np.abs(g8[1::2] - long_links_numpy(L, Uph)).max() < 1e-15
    # fat link noise: W - 0.8 U = 0.05 G with E|G_ij|^2 = 1
    G = (g8[0::2] - 0.8 * Uph) / 0.05
    assert 0.9 < np.mean(np.abs(G) ** 2) < 1.1


# -------------------------------------------------------------- free field
def _plane_wave(L, n, col):
    C = _coords_np(L)
    p = [2 * np.pi * n[0] / L[0], 2 * np.pi * n[1] / L[1], 2 * np.pi * n[2] / L[2],
         np.pi * (2 * n[3] + 1) / L[3]]
    ph = np.exp(1j * (C @ np.array(p)))
    psi = np.zeros((len(C), 3), complex)
    psi[:, col] = ph
    return psi, p


@pytest.mark.parametrize("hisq", [False, True])
@pytest.mark.parametrize("n", [(0, 0, 0, 0), (1, 0, 2, 1), (3, 1, 1, 2), (2, 3, 0, 0)])
def test_free_field_spectrum(oracle, hisq, n):
    L = (8, 6, 8, 6)
    Uph = oracle.phased(L, oracle.gauge_unit(L))
    if hisq:
        g = np.empty((8,) + Uph.shape[1:], complex)
        g[0::2] = Uph
        g[1::2] = long_links_numpy(L, Uph)
        c3 = -1.0 / 24
    else:
        g, c3 = Uph, 0.0
    psi, p = _plane_wave(L, n, 1)
    # D psi via stagD with m=0: D = (2D)/2; D^2 psi = -sum (sin p + c3 sin 3p)^2 psi
    Dp = oracle.D(L, g, psi, 0.0)
    DDp = oracle.D(L, g, Dp, 0.0)
    lam = sum((np.sin(q) + c3 * np.sin(3 * q)) ** 2 for q in p)
    assert np.abs(DDp + lam * psi).max() < 1e-13


@pytest.mark.parametrize("hisq", [False, True])
def test_free_field_cg_solution(oracle, hisq):
    """Even-projected plane waves are eigenvectors of 4(m^2 - D_eo D_oe)."""
    L = (8, 6, 8, 6)
    Uph = oracle.phased(L, oracle.gauge_unit(L))
    if hisq:
        g = np.empty((8,) + Uph.shape[1:], complex)
        g[0::2] = Uph
        g[1::2] = long_links_numpy(L, Uph)
        c3 = -1.0 / 24
    else:
        g, c3 = Uph, 0.0
    psi, p = _plane_wave(L, (1, 2, 3, 1), 2)
    V = oracle.volume(L)
    b = psi.copy()
    b[V // 2:] = 0
    m = 0.2
    lam = sum((np.sin(q) + c3 * np.sin(3 * q)) ** 2 for q in p)
    x, its, tr = oracle.solveXX(L, g, 0, b, m, 1e-26)
    assert rel(x[: V // 2], b[: V // 2] / (4 * (m * m + lam))) < 1e-12


# ---------------------------------------------------------- dense operator
def _dense(oracle, L, g, m):
    V = oracle.volume(L)
    n = 3 * V
    M = np.zeros((n, n), complex)
    for j in range(n):
        e = np.zeros((V, 3), complex)
        e.flat[j] = 1.0
        M[:, j] = oracle.D(L, g, e, m).ravel()
    return M


@pytest.fixture(scope="module")
def dense44(oracle):
    L = (4, 4, 4, 4)
    U = oracle.phased(L, oracle.gauge_random(L, SEED))
    return L, U, _dense(oracle, L, U, 0.0)


def test_dense_hop_antihermitian_and_eps(oracle, dense44):
    L, U, K = dense44
    assert np.abs(K + K.conj().T).max() < 1e-15
    V = oracle.volume(L)
    eps = np.repeat(np.where(np.arange(V) < V // 2, 1.0, -1.0), 3)
    assert np.abs(eps[:, None] * K * eps[None, :] + K).max() < 1e-15
    # D(m) = m + K has spectrum m +- i lambda
    ev = np.linalg.eigvals(K + 0.3 * np.eye(len(K)))
    assert np.abs(ev.real - 0.3).max() < 1e-12
    im = np.sort(ev.imag)
    assert np.abs(im + im[::-1]).max() < 1e-12


def test_dense_matches_stagD2xx_and_solvers(oracle, dense44):
    L, U, K = dense44
    V = oracle.volume(L)
    ne = 3 * V // 2
    m = 0.15
    # A_ee = 4(m^2 - K_eo K_oe) with K = D (hop/2)
    A = 4 * (m * m * np.eye(ne) - K[:ne, ne:] @ K[ne:, :ne])
    b = oracle.source_gauss(L, 3, 0, 0)
    Ax = oracle.stagD2xx(L, U, 0, b, m * m)
    assert rel(Ax[: V // 2].ravel(), A @ b[: V // 2].ravel()) < 1e-14
    want = np.linalg.solve(A, b[: V // 2].ravel())
    x, its, tr = oracle.solveXX(L, U, 0, b, m, 1e-24)
    assert rel(x[: V // 2].ravel(), want) < 1e-11
    assert tr <= 1e-24
    x2, its2, r2 = oracle.cg(L, U, 0, b, m, 1e-24)
    assert rel(x2[: V // 2].ravel(), want) < 1e-11
    # odd parity: A_oo = 4(m^2 - K_oe K_eo)
    Ao = 4 * (m * m * np.eye(ne) - K[ne:, :ne] @ K[:ne, ne:])
    bo = oracle.source_gauss(L, 4, 0, 1)
    xo, _, _ = oracle.solveXX(L, U, 1, bo, m, 1e-24)
    assert rel(xo[V // 2:].ravel(), np.linalg.solve(Ao, bo[V // 2:].ravel())) < 1e-11


def test_dense_stagD_variants(oracle, dense44):
    L, U, K = dense44
    x = oracle.source_gauss(L, 8, 1, 2)
    r0 = oracle.source_gauss(L, 9, 1, 2)
    V = oracle.volume(L)
    Kx = (K @ x.ravel()).reshape(V, 3)
    r = r0.copy()
    oracle.stagD(L, U, 1, r, x, 0.7, -1.0, 0.25)
    want = r0.copy()
    want[V // 2:] = 0.25 * r0[V // 2:] + 0.7 * x[V // 2:] - Kx[V // 2:]
    assert rel(r, want) < 1e-14
    assert rel(oracle.Ddag(L, U, x, 0.7), 0.7 * x - Kx) < 1e-14


# ---------------------------------------------------------- gauge covariance
def _random_su3_field(oracle, L, seed):
    G = oracle.gauge_random(L, seed)
    return G[0]


def _transform(L, U, g, hops):
    """U'_k(x) = g(x) U_k(x) g(x + l_k mu_k)^+"""
    C = _coords_np(L)
    out = np.empty_like(U)
    for k in range(U.shape[0]):
        mu, ln = hops[k]
        c1 = C.copy(); c1[:, mu] = (c1[:, mu] + ln) % L[mu]
        out[k] = g @ U[k] @ np.conj(np.swapaxes(g[_index_np(L, c1)], -1, -2))
    return out


@pytest.mark.parametrize("hisq", [False, True])
def test_gauge_covariance(oracle, hisq):
    L = (8, 4, 4, 8)
    U = oracle.phased(L, oracle.gauge_random(L, SEED))
    if hisq:
        U = oracle.hisq_links(L, SEED + 1, U)
        hops = [(k // 2, 3 if k & 1 else 1) for k in range(8)]
    else:
        hops = [(k, 1) for k in range(4)]
    g = _random_su3_field(oracle, L, 99)
    Ug = _transform(L, U, g, hops)
    x = oracle.source_gauss(L, 1, 0, 2)
    gx = np.einsum("vab,vb->va", g, x)
    lhs = oracle.D(L, Ug, gx, 0.4)
    rhs = np.einsum("vab,vb->va", g, oracle.D(L, U, x, 0.4))
    assert rel(lhs, rhs) < 1e-13


# -------------------------------------------------------------- multi-shift
def test_cgm_residuals_and_single_shift_agreement(oracle):
    L = (4, 4, 4, 8)
    U = oracle.phased(L, oracle.gauge_random(L, SEED))
    b = oracle.source_gauss(L, 2, 0, 0)
    m = 0.1
    sig = [0.0, 1e-3, 0.05, 0.4, 2.0]
    xs, its, r2 = oracle.cgm(L, U, 0, b, m, sig, 1e-22)
    V = oracle.volume(L)
    ne = V // 2
    for k, s in enumerate(sig):
        r = b - oracle.stagD2xx(L, U, 0, xs[k], m * m) - s * xs[k]
        assert np.sqrt(oracle.norm2(r[:ne]) / oracle.norm2(b)) < 1e-10
        # (A + s) = 4((m^2 + s/4) - DD): single-shift CG with that mass
        xk, _, _ = oracle.solveXX(L, U, 0, b, np.sqrt(m * m + s / 4), 1e-24)
        assert rel(xs[k][:ne], xk[:ne]) < 1e-9


# -------------------------------------------------------------------- force
def _action(oracle, L, U, phi, m):
    """S = phi^+ (m^2 - D_eo D_oe)^-1 phi and Y = (m^2 - D_eo D_oe)^-1 phi (even)."""
    y, its, tr = oracle.solveXX(L, U, 0, phi, m, 1e-28)
    Y = 4 * y
    V = oracle.volume(L)
    return float(np.vdot(phi[: V // 2], Y[: V // 2]).real), Y


def _mapply(oracle, L, U, Y, m):
    """(m^2 - D_eo D_oe) Y on even sites."""
    return 0.25 * oracle.stagD2xx(L, U, 0, Y, m * m)


def _rand_su3_alg(rng):
    A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    X = 0.5 * (A - A.conj().T)
    return X - np.trace(X) / 3 * np.eye(3)


@pytest.mark.parametrize("L,m", [((4, 4, 4, 4), 0.2), ((4, 4, 6, 4), 0.05)])
def test_force_finite_differences(oracle, L, m):
    """dS/dw along U -> exp(w X) U equals -1/2 Re tr(X F), F = fermForce output.

    S differences use the exact local identity S(U') - S(U) = -Y'^+ (M' - M) Y,
    so solver residuals do not limit the comparison."""
    U = oracle.phased(L, oracle.gauge_random(L, SEED))
    phi = oracle.source_gauss(L, 5, 0, 0)
    V = oracle.volume(L)
    ne = V // 2
    psi, _ = oracle.force_solve(L, U, phi, m, 1e-28)
    F = oracle.ferm_force(L, U, psi)
    S0, Y0 = _action(oracle, L, U, phi, m)
    rng = np.random.default_rng(7)
    h = 1e-5
    worst = 0.0
    for trial in range(6):
        mu = int(rng.integers(4))
        site = int(rng.integers(V))
        X = _rand_su3_alg(rng)
        analytic = -0.5 * np.trace(X @ F[mu, site]).real
        dS = []
        for sgn in (1, -1):
            Up = U.copy()
            Up[mu, site] = scipy.linalg.expm(sgn * h * X) @ U[mu, site]
            _, Yp = _action(oracle, L, Up, phi, m)
            diffM = _mapply(oracle, L, Up, Y0, m) - _mapply(oracle, L, U, Y0, m)
            dS.append(-np.vdot(Yp[:ne], diffM[:ne]).real)
        fd = (dS[0] - dS[1]) / (2 * h)
        worst = max(worst, abs(fd - analytic) / max(abs(analytic), 1e-3))
    assert worst < 1e-8


def test_force_gauge_covariance(oracle):
    L = (4, 4, 4, 4)
    U = oracle.phased(L, oracle.gauge_random(L, SEED))
    psi = oracle.source_gauss(L, 6, 0, 2)
    g = _random_su3_field(oracle, L, 55)
    Ug = _transform(L, U, g, [(k, 1) for k in range(4)])
    gpsi = np.einsum("vab,vb->va", g, psi)
    F = oracle.ferm_force(L, U, psi)
    Fg = oracle.ferm_force(L, Ug, gpsi)
    want = g[None] @ F @ np.conj(np.swapaxes(g, -1, -2))[None]
    assert rel(Fg, want) < 1e-13
    # F is traceless anti-Hermitian
    assert np.abs(F + np.conj(np.swapaxes(F, -1, -2))).max() < 1e-13
    assert np.abs(np.trace(F, axis1=-2, axis2=-1)).max() < 1e-13


# ------------------------------------------------------------ HMC pieces
def _staple_np(oracle, L, U, mu, site):
    """Sum of the six staples at (mu, site), computed independently in numpy."""
    C = _coords_np(L)
    c = C[site]

    def idx(d):
        y = (c + np.array(d)) % np.array(L)
        return int(_index_np(L, y[None])[0])

    A = np.zeros((3, 3), complex)
    e = np.eye(4, dtype=int)
    for nu in range(4):
        if nu == mu:
            continue
        dag = lambda m: m.conj().T
        A += U[nu, idx(e[mu])] @ dag(U[mu, idx(e[nu])]) @ dag(U[nu, site])
        A += dag(U[nu, idx(e[mu] - e[nu])]) @ dag(U[mu, idx(-e[nu])]) @ U[nu, idx(-e[nu])]
    return A


def test_plaquette_closed_forms(oracle):
    L = (4, 4, 4, 6)
    assert oracle.plaquette(L, oracle.gauge_unit(L)) == 18 * oracle.volume(L)
    U = oracle.gauge_random(L, SEED)
    g = _random_su3_field(oracle, L, 3)
    Ug = _transform(L, U, g, [(k, 1) for k in range(4)])
    assert abs(oracle.plaquette(L, Ug) - oracle.plaquette(L, U)) < 1e-10


def test_wilson_force_finite_differences(oracle):
    """dS/dw = -1/2 Re tr(X F) for S = -(beta/3) sum Re tr U_P; the action change
    is evaluated locally through the staple sum (exact up to rounding)."""
    L = (4, 4, 4, 4)
    beta = 5.6
    U = oracle.gauge_random(L, SEED)
    F = oracle.wilson_force(L, U, beta)
    assert np.abs(F + np.conj(np.swapaxes(F, -1, -2))).max() < 1e-13
    rng = np.random.default_rng(11)
    h = 1e-5
    for trial in range(8):
        mu, site = int(rng.integers(4)), int(rng.integers(oracle.volume(L)))
        X = _rand_su3_alg(rng)
        A = _staple_np(oracle, L, U, mu, site)
        S = lambda W: -(beta / 3) * np.trace(W @ A).real
        fd = (S(scipy.linalg.expm(h * X) @ U[mu, site]) - S(scipy.linalg.expm(-h * X) @ U[mu, site])) / (2 * h)
        an = -0.5 * np.trace(X @ F[mu, site]).real
        assert abs(fd - an) <= 1e-8 * max(1.0, abs(an))
    # and the staple route agrees with the whole-lattice plaquette difference
    mu, site = 2, 77
    X = _rand_su3_alg(rng)
    Up = U.copy()
    Up[mu, site] = scipy.linalg.expm(1e-3 * X) @ U[mu, site]
    dfull = -(beta / 3) * (oracle.plaquette(L, Up) - oracle.plaquette(L, U))
    A = _staple_np(oracle, L, U, mu, site)
    dloc = -(beta / 3) * (np.trace(Up[mu, site] @ A).real - np.trace(U[mu, site] @ A).real)
    assert abs(dfull - dloc) < 1e-9


def test_momenta_and_link_update(oracle):
    L = (4, 4, 4, 4)
    P = oracle.random_momenta(L, 9)
    assert np.abs(P + np.conj(np.swapaxes(P, -1, -2))).max() == 0.0
    assert np.abs(np.trace(P, axis1=-2, axis2=-1)).max() < 1e-14
    # density exp(tr P^2/2): 8 real degrees of freedom of unit variance per link
    kin = -0.5 * np.trace(P @ P, axis1=-2, axis2=-1).real
    assert abs(kin.mean() - 4.0) < 0.2
    U = oracle.gauge_random(L, 3)
    W = U.copy()
    oracle.link_update(L, P, 0.05, W)
    for k in (0, 100, 1000):
        mu, s = divmod(k, oracle.volume(L))
        assert np.abs(W[mu, s] - scipy.linalg.expm(0.05 * P[mu, s]) @ U[mu, s]).max() < 1e-14
