"""Pins the oracle's HISQ link construction (qo_make_imp_links, qo_project_u,
qo_hisq_smear; QEX gauge/fat7l.nim:71-159, maths/matrixFunctions.nim:337-352,
physics/hisqLinks.nim:10-51) with checks that do not restate its algorithm:

 - unit links: every path contributes the identity, so the fat link is
   (oneLink + 6 threeStaple + 24 fiveStaple + 48 sevenStaple + 6 lepage) * 1,
   the count QEX's own fat7l test checks (gauge/fat7l.nim:183-217);
 - phased, BC-applied unit links: every plaquette is -1, so staples flip sign;
   the HISQ first level gives the links back, the second level 9/8 of them,
   and the long link is -1/24 of the three-link product (the Naik-improved
   free-field operator 9/8 sin p - 1/24 sin 3p);
 - brute force: an explicit sum over the 90 link paths per direction, each
   written as a step sequence, on a random SU(3) field of unequal extents;
 - projectU against numpy's SVD polar factor;
 - gauge covariance of the whole HISQ smearing under random SU(3) transforms.
"""
import itertools

import numpy as np

SEED = 777


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def site_index(L, C):
    V = int(np.prod(L))
    lex = C[:, 0] + L[0] * (C[:, 1] + L[1] * (C[:, 2] + L[2] * C[:, 3]))
    p = C.sum(axis=1) & 1
    return (lex + p * V) // 2


def dag(M):
    return np.conj(np.swapaxes(M, -1, -2))


def path_field(oracle, U, L, steps):
    """Product of links along `steps` from every site: +(d+1) is a forward
    step in direction d (multiply U_d(y), then y += d), -(d+1) a backward one
    (y -= d, then multiply U_d(y)^+)."""
    C = oracle.all_coords(L).copy()
    V = C.shape[0]
    M = np.broadcast_to(np.eye(3, dtype=np.complex128), (V, 3, 3)).copy()
    for s in steps:
        d = abs(s) - 1
        if s > 0:
            M = M @ U[d][site_index(L, C)]
            C[:, d] = (C[:, d] + 1) % L[d]
        else:
            C[:, d] = (C[:, d] - 1) % L[d]
            M = M @ dag(U[d][site_index(L, C)])
    return M


def fat_by_paths(oracle, U, L, coef):
    c1, c3, c5, c7, cL = coef
    out = np.empty_like(U)
    for mu in range(4):
        m = mu + 1
        others = [d for d in range(4) if d != mu]
        acc = (c1 - 6.0 * cL) * U[mu]
        for nu in others:
            for a in (nu + 1, -(nu + 1)):
                if c3:
                    acc = acc + c3 * path_field(oracle, U, L, [a, m, -a])
                if cL:
                    for b in (nu + 1, -(nu + 1)):
                        acc = acc + cL * path_field(oracle, U, L, [a, b, m, -b, -a])
        if c5:
            for d1, d2 in itertools.permutations(others, 2):
                for s1, s2 in itertools.product((1, -1), repeat=2):
                    a, b = s1 * (d1 + 1), s2 * (d2 + 1)
                    acc = acc + c5 * path_field(oracle, U, L, [a, b, m, -b, -a])
        if c7:
            for d1, d2, d3 in itertools.permutations(others, 3):
                for s1, s2, s3 in itertools.product((1, -1), repeat=3):
                    a, b, c = s1 * (d1 + 1), s2 * (d2 + 1), s3 * (d3 + 1)
                    acc = acc + c7 * path_field(oracle, U, L, [a, b, c, m, -c, -b, -a])
        out[mu] = acc
    return out


TERMS = {
    "oneLink": (1.0, 0, 0, 0, 0),
    "threeStaple": (0, 1.0, 0, 0, 0),
    "fiveStaple": (0, 0, 1.0, 0, 0),
    "sevenStaple": (0, 0, 0, 1.0, 0),
    "lepage": (0, 0, 0, 0, 1.0),
    "all": (0.3, -0.7, 0.45, -0.2, 0.15),
}


def test_fat7_unit_links_path_counts(oracle):
    L = (4, 4, 4, 4)
    U = oracle.gauge_unit(L)
    for name, c in TERMS.items():
        fl, _ = oracle.make_imp_links(L, U, oracle.fat7_coefs(*c))
        expect = c[0] + 6 * c[1] + 24 * c[2] + 48 * c[3] + 6 * c[4]
        assert np.abs(fl - expect * U).max() < 1e-13, name


def test_hisq_on_phased_unit_links(oracle):
    L = (4, 4, 4, 6)
    Uph = oracle.phased(L, oracle.gauge_unit(L))
    first, second, naik = oracle.hisq_coefs()
    assert abs(first[0] - 1 / 8) < 1e-16 and abs(second[0] - 1.0) < 1e-16 and abs(second[4] + 1 / 8) < 1e-16
    # every plaquette of the phased links is -1: each staple level flips sign
    v1, _ = oracle.make_imp_links(L, Uph, first)
    assert np.abs(v1 - Uph).max() < 1e-14          # 1/8 + 6/16 + 24/64 + 48/384 = 1
    fl, ll = oracle.hisq_smear(L, Uph)
    assert np.abs(fl - 9.0 / 8.0 * Uph).max() < 1e-13
    C = oracle.all_coords(L)
    for mu in range(4):
        prod = Uph[mu].copy()
        for k in (1, 2):
            Ck = C.copy()
            Ck[:, mu] = (Ck[:, mu] + k) % L[mu]
            prod = prod @ Uph[mu][site_index(L, Ck)]
        assert np.abs(ll[mu] - naik * prod).max() < 1e-15
    # the long link of phased unit links is eta_mu(x) times the BC sign of its path
    assert np.allclose(np.abs(ll), np.abs(naik) * np.abs(Uph))


def test_fat7_matches_explicit_path_sum(oracle):
    L = (4, 6, 4, 8)
    U = oracle.gauge_random(L, SEED)
    for name, c in list(TERMS.items()) + [("hisq second", tuple(oracle.hisq_coefs()[1]))]:
        fl, _ = oracle.make_imp_links(L, U, oracle.fat7_coefs(*c))
        ref = fat_by_paths(oracle, U, L, c)
        assert rel(fl, ref) < 1e-13, (name, rel(fl, ref))


def test_naik_long_links_explicit_path(oracle):
    L = (4, 6, 4, 8)
    U = oracle.gauge_random(L, SEED + 1)
    W = oracle.gauge_random(L, SEED + 2)
    fl, ll = oracle.make_imp_links(L, U, oracle.fat7_coefs(1.0), naik=0.37, gl=W)
    assert rel(fl, U) < 1e-15
    for mu in range(4):
        ref = 0.37 * path_field(oracle, W, L, [mu + 1] * 3)
        assert rel(ll[mu], ref) < 1e-14


def test_project_u_is_the_polar_factor(oracle):
    rng = np.random.default_rng(SEED)
    n = 400
    X = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    X[:100] = np.eye(3) + 0.2 * X[:100]  # well conditioned
    X[100:164] = 2.5 * oracle.gauge_random((2, 2, 2, 2), SEED).reshape(-1, 3, 3)  # scaled SU(3): equal eigenvalues
    u = oracle.project_u(X)
    W, s, Vh = np.linalg.svd(X)
    polar = W @ Vh
    assert np.abs(u - polar).max() < 1e-12
    # forming x^+ x squares the condition number: rounding grows with it
    assert np.abs(dag(u[:164]) @ u[:164] - np.eye(3)).max() < 1e-13
    assert np.abs(dag(u) @ u - np.eye(3)).max() < 1e-11
    H = dag(u) @ X                       # X = u H with H Hermitian positive
    assert np.abs(H - dag(H)).max() < 1e-12
    assert np.linalg.eigvalsh(0.5 * (H + dag(H))).min() > 0
    # exactly unitary input is returned unchanged
    U = oracle.gauge_random((2, 2, 2, 2), SEED + 3).reshape(-1, 3, 3)
    assert np.abs(oracle.project_u(U) - U).max() < 1e-13


def test_hisq_smear_gauge_covariance(oracle):
    L = (4, 4, 4, 6)
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 4))
    G = oracle.gauge_random(L, SEED + 5)[0]          # one SU(3) matrix per site
    C = oracle.all_coords(L)

    def fwd(mu, k):
        Ck = C.copy()
        Ck[:, mu] = (Ck[:, mu] + k) % L[mu]
        return site_index(L, Ck)

    Ug = np.stack([G @ U[mu] @ dag(G[fwd(mu, 1)]) for mu in range(4)])
    fl, ll = oracle.hisq_smear(L, U)
    flg, llg = oracle.hisq_smear(L, Ug)
    # on these rough links some first-level sums have condition numbers of a
    # few hundred, which projectU squares: rounding reaches ~1e-12
    for mu in range(4):
        assert rel(flg[mu], G @ fl[mu] @ dag(G[fwd(mu, 1)])) < 1e-10
        assert rel(llg[mu], G @ ll[mu] @ dag(G[fwd(mu, 3)])) < 1e-10


# --------------------------------------------------------------- chain rule
# A fat7 level is a polynomial of degree <= 7 in the links, so the 9-point
# central difference (exact through degree 8) gives its directional
# derivative exactly up to rounding; projectU and the full smearing are
# smooth and the same stencil with a small step leaves O(h^8) truncation.
W9 = np.array([1 / 280, -4 / 105, 1 / 5, -4 / 5, 0, 4 / 5, -1 / 5, 4 / 105, -1 / 280])


def redot(a, b):
    return float(np.real(np.sum(np.conj(a) * b)))


def _crand(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def fd9(f, h):
    return float(np.dot(W9, [f(k * h) for k in range(-4, 5)]) / h)


def test_fat7_deriv_against_exact_differences(oracle):
    L = (4, 4, 4, 6)
    V = oracle.volume(L)
    rng = np.random.default_rng(SEED)
    U = oracle.gauge_random(L, SEED + 6)
    C, dU = _crand(rng, 4, V, 3, 3), _crand(rng, 4, V, 3, 3)
    for name, c in list(TERMS.items()) + [("hisq second", tuple(oracle.hisq_coefs()[1]))]:
        coef = oracle.fat7_coefs(*c)
        fd = fd9(lambda t: redot(C, oracle.make_imp_links(L, U + t * dU, coef)[0]), 1e-2)
        d, _ = oracle.fat7l_deriv(L, U, C, coef)
        assert abs(redot(d, dU) - fd) < 1e-12 * abs(fd), name
    # Naik long links from a separate field
    W = oracle.gauge_random(L, SEED + 7)
    Cl = _crand(rng, 4, V, 3, 3)
    fd = fd9(lambda t: redot(Cl, oracle.make_imp_links(L, U, oracle.fat7_coefs(1.0), naik=-1 / 24,
                                                       gl=W + t * dU)[1]), 1e-2)
    _, lld = oracle.fat7l_deriv(L, U, np.zeros_like(U), oracle.fat7_coefs(1.0), naik=-1 / 24, gl=W, llmid=Cl)
    assert abs(redot(lld, dU) - fd) < 1e-12 * abs(fd)
    # one link at a time: the derivative's entries, not only a projection of them
    for mu, site in [(0, 3), (3, V - 1), (2, V // 2)]:
        E = np.zeros_like(U)
        E[mu, site] = _crand(rng, 3, 3)
        coef = oracle.fat7_coefs(*TERMS["all"])
        fd = fd9(lambda t: redot(C, oracle.make_imp_links(L, U + t * E, coef)[0]), 1e-2)
        d, _ = oracle.fat7l_deriv(L, U, C, coef)
        assert abs(redot(d, E) - fd) < 1e-11 * abs(fd)


def test_project_u_deriv_against_differences(oracle):
    rng = np.random.default_rng(SEED + 1)
    X = np.eye(3) + 0.3 * _crand(rng, 64, 3, 3)
    C, dX = _crand(rng, 64, 3, 3), _crand(rng, 64, 3, 3)
    fd = fd9(lambda t: redot(C, oracle.project_u(X + t * dX)), 3e-4)
    r = oracle.project_u_deriv(oracle.project_u(X), X, C)
    assert abs(redot(r, dX) - fd) < 1e-10 * abs(fd)


def test_hisq_smear_deriv_against_differences(oracle):
    L = (4, 4, 4, 6)
    V = oracle.volume(L)
    rng = np.random.default_rng(SEED + 2)
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 8))
    Cf, Cl, dU = _crand(rng, 4, V, 3, 3), _crand(rng, 4, V, 3, 3), _crand(rng, 4, V, 3, 3)

    def S(t):
        fl, ll = oracle.hisq_smear(L, U + t * dU)
        return redot(Cf, fl) + redot(Cl, ll)

    # the rough first-level links include condition numbers ~100, so the
    # projection is strongly curved there: a small step keeps the O(h^8) term down
    fd = fd9(S, 3e-5)
    assert abs(redot(oracle.hisq_smear_deriv(L, U, Cf, Cl), dU) - fd) < 1e-9 * abs(fd)


def _g8(oracle, L, U):
    fl, ll = oracle.hisq_smear(L, U)
    return np.ascontiguousarray(np.stack([t for mu in range(4) for t in (fl[mu], ll[mu])]))


def test_hisq_fermion_force_against_differences(oracle):
    """dS/dU of S = phi^+ (m^2 - D_eo D_oe)^-1 phi with the HISQ operator built
    from smear(U): the outer products 1/2 eps psi psi(x+mu)^+ and
    1/2 eps psi psi(x+3mu)^+ chained through both smearing levels."""
    L, m = (4, 4, 4, 6), 0.2
    V = oracle.volume(L)
    rng = np.random.default_rng(SEED + 3)
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 9))
    phi = oracle.source_gauss(L, 5, 0, 0)

    def S(t):
        y, _, _ = oracle.solveXX(L, _g8(oracle, L, U + t * dU), 0, phi, m, 1e-30)
        return float(np.vdot(phi[: V // 2], 4 * y[: V // 2]).real)

    psi, _ = oracle.force_solve(L, _g8(oracle, L, U), phi, m, 1e-30)
    G8 = oracle.hisq_outer(L, psi)
    GU = oracle.hisq_smear_deriv(L, U, 0.5 * G8[0::2], 0.5 * G8[1::2])
    dU = _crand(rng, 4, V, 3, 3)
    fd = fd9(S, 3e-5)       # a first-level link has condition number ~430 here
    assert abs(redot(GU, dU) - fd) < 1e-9 * abs(fd)


def test_trivial_smearing_reproduces_the_plain_force(oracle):
    """With fl = projectU(U) = U and no Naik term the chain rule through the
    projection changes only the Hermitian part, which TAH removes: TAH(2 dS/dU U^+)
    equals the plain fermion force."""
    L = (4, 4, 4, 4)
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 10))
    psi = oracle.source_gauss(L, 6, 0, 2)
    one = oracle.fat7_coefs(1.0)
    G8 = oracle.hisq_outer(L, psi)
    GU = oracle.hisq_smear_deriv(L, U, 0.5 * G8[0::2], 0.5 * G8[1::2], coefs=(one, one, 0.0))
    T = 2 * GU @ dag(U)
    F = 0.5 * (T - dag(T))
    F -= np.trace(F, axis1=-2, axis2=-1)[..., None, None] / 3 * np.eye(3)
    assert rel(F, oracle.ferm_force(L, U, psi)) < 1e-12

