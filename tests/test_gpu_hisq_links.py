"""HISQ link construction on the GPU (qexhip_make_imp_links, qexhip_project_u,
qexhip_hisq_smear) against the CPU oracle, plus closed forms at full size:

 - every fat7 term alone and combined, and the HISQ second level with Naik
   long links from a separate field, equal the oracle's to 1e-12;
 - projectU equals the oracle's and is unitary;
 - the full HISQ smearing equals the oracle's, and the Dslash built from the
   GPU-smeared links equals the oracle's Dslash on the oracle's links;
 - 32^3 x 64 phased unit links give fat = 9/8 U and long = -1/24 U U U
   exactly (each staple level flips sign on them), and projectU maps every
   link of a random 32^3 x 64 field to a unitary matrix.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 913


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


TERMS = [
    (1.0, 0, 0, 0, 0),
    (0, 1.0, 0, 0, 0),
    (0, 0, 1.0, 0, 0),
    (0, 0, 0, 1.0, 0),
    (0, 0, 0, 0, 1.0),
    (0.3, -0.7, 0.45, -0.2, 0.15),
]


@pytest.mark.parametrize("L", [(8, 8, 8, 8), (4, 6, 4, 8)])
def test_fat7_terms_match_oracle(oracle, gpu, L):
    import torch
    from qex_amd import hisq
    print(f"fat7 terms L={L}", flush=True)
    lo = gpu.Layout(L)
    U = oracle.gauge_random(L, SEED)
    Ug = _cuda(U)
    for c in TERMS:
        fl = torch.empty_like(Ug)
        hisq.makeImpLinks(lo, fl, Ug, hisq.Fat7lCoefs(*c))
        want, _ = oracle.make_imp_links(L, U, oracle.fat7_coefs(*c))
        assert rel(fl.cpu().numpy(), want) < 1e-12, c


def test_fat7_with_naik_from_other_field(oracle, gpu):
    import torch
    from qex_amd import hisq
    L = (8, 8, 8, 8)
    lo = gpu.Layout(L)
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 1))
    W = oracle.gauge_random(L, SEED + 2)
    second = oracle.hisq_coefs()[1]
    fl, ll = torch.empty_like(_cuda(U)), torch.empty_like(_cuda(U))
    hisq.makeImpLinks(lo, fl, _cuda(U), hisq.Fat7lCoefs(*second), ll, _cuda(W), -1.0 / 24)
    wf, wl = oracle.make_imp_links(L, U, second, naik=-1.0 / 24, gl=W)
    assert rel(fl.cpu().numpy(), wf) < 1e-12
    assert rel(ll.cpu().numpy(), wl) < 1e-13
    # without gfLong the long links come from gf
    hisq.makeImpLinks(lo, fl, _cuda(U), hisq.Fat7lCoefs(*second), ll, None, 0.5)
    _, wl2 = oracle.make_imp_links(L, U, second, naik=0.5)
    assert rel(ll.cpu().numpy(), wl2) < 1e-13


def test_project_u_matches_oracle(oracle, gpu):
    from qex_amd import hisq
    rng = np.random.default_rng(SEED)
    X = rng.normal(size=(4, 2000, 3, 3)) + 1j * rng.normal(size=(4, 2000, 3, 3))
    X[0] = np.eye(3) + 0.2 * X[0]
    u = hisq.projectU(_cuda(X)).cpu().numpy()
    want = oracle.project_u(X)
    assert np.abs(u[0] - want[0]).max() < 1e-14
    assert rel(u, want) < 1e-12
    uu = np.conj(np.swapaxes(u, -1, -2)) @ u
    assert np.abs(uu[0] - np.eye(3)).max() < 1e-13


@pytest.mark.parametrize("L", [(8, 8, 8, 8), (4, 4, 6, 8)])
def test_hisq_smear_matches_oracle(oracle, gpu, L):
    from qex_amd import hisq
    print(f"hisq smear L={L}", flush=True)
    lo = gpu.Layout(L)
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 3))
    fl, ll = hisq.newHISQ().smear(lo, _cuda(U))
    wf, wl = oracle.hisq_smear(L, U)
    # rough random links: first-level sums with condition numbers of a few
    # hundred, squared by projectU; rounding differences reach ~1e-13
    assert rel(fl.cpu().numpy(), wf) < 1e-12
    assert rel(ll.cpu().numpy(), wl) < 1e-12
    # the library's defaults are QEX's
    c = hisq.default_coefs_c()
    first, second, naik = oracle.hisq_coefs()
    assert abs(c.naik - naik) == 0.0
    assert [c.fat7first.one_link, c.fat7first.three_staple, c.fat7first.five_staple, c.fat7first.seven_staple,
            c.fat7first.lepage] == list(first)
    assert [c.fat7second.one_link, c.fat7second.three_staple, c.fat7second.five_staple,
            c.fat7second.seven_staple, c.fat7second.lepage] == list(second)


def test_dslash_on_smeared_links(oracle, gpu):
    from qex_amd import hisq
    L = (8, 8, 8, 8)
    lo = gpu.Layout(L)
    U = oracle.phased(L, oracle.gauge_random(L, SEED + 4))
    s, fl, ll = hisq.hisqStag(lo, _cuda(U))
    wf, wl = oracle.hisq_smear(L, U)
    g8 = np.ascontiguousarray(np.stack([t for mu in range(4) for t in (wf[mu], wl[mu])]))
    x = oracle.source_gauss(L, 5, 0, 0)
    fx = lo.ColorVector().set(x)
    fr = lo.ColorVector()
    gpu.stagD2ee(s.se, s.so, fr, s.g, fx, 0.05)
    got = fr.get(sub=0)
    want = oracle.stagD2xx(L, g8, 0, x, 0.05)
    assert rel(got, want) < 1e-12


def test_full_size_closed_forms(oracle, gpu):
    import torch
    from qex_amd import gauge, hisq
    L = (32, 32, 32, 64)
    print(f"closed forms L={L}", flush=True)
    lo = gpu.Layout(L)
    Uph = gauge.phase_bc(lo, gauge.unit_gauge(lo))
    fl, ll = hisq.newHISQ().smear(lo, Uph)
    assert torch.abs(fl - (9.0 / 8.0) * Uph).max().item() < 1e-13
    # long link of phased unit links: -1/24 times the product of the three signs
    # along the path, i.e. -1/24 U_mu(x) U_mu(x+mu) U_mu(x+2mu)
    C = torch.from_numpy(oracle.all_coords(L)).cuda()
    V = lo.nSites
    for mu in range(4):
        prod = Uph[mu].clone()
        for k in (1, 2):
            Ck = C.clone()
            Ck[:, mu] = (Ck[:, mu] + k) % L[mu]
            lex = Ck[:, 0] + L[0] * (Ck[:, 1] + L[1] * (Ck[:, 2] + L[2] * Ck[:, 3]))
            idx = (lex + (Ck.sum(1) & 1) * V) // 2
            prod = prod @ Uph[mu][idx]
        assert torch.abs(ll[mu] + prod / 24.0).max().item() < 1e-15
    del fl, ll
    U = gauge.random_gauge(lo, SEED)
    W = hisq.projectU((0.8 * U + 0.3 * torch.roll(U, 1, dims=1)).contiguous())
    Wd = W.transpose(-1, -2).conj().resolve_conj().contiguous()
    err = torch.abs(Wd @ W - torch.eye(3, dtype=W.dtype, device=W.device)).max().item()
    assert err < 1e-11
