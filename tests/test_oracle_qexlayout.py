"""Pins the oracle's restatement of QEX's SIMD field layout (layoutSetupQ and
layoutIndexImpl, src/layout/qlayout.nim:10-150) with checks that do not
restate it: layoutCoordQ (qlayout.nim:154-213, a separate code path that QEX
itself checks against the index) inverts it on every site; the map is a
bijection onto [0, V) with the even sites first; innerGeom = 1 reduces it to
the canonical order every other part of the library uses; the storage
round trip is exact; and the layouts QEX refuses are refused."""
import numpy as np
import pytest

CASES = [
    ((4, 4, 4, 4), (1, 1, 1, 1)),
    ((4, 4, 4, 4), (1, 1, 1, 2)),
    ((4, 4, 4, 4), (1, 1, 2, 2)),
    ((8, 4, 4, 4), (2, 1, 1, 2)),
    ((4, 4, 6, 6), (1, 1, 1, 2)),   # odd outer extent: the innerCb shift is active
    ((6, 4, 2, 2), (2, 1, 1, 1)),   # odd outer extent in x
    ((4, 4, 4, 8), (2, 2, 2, 2)),
]


@pytest.mark.parametrize("L,ig", CASES)
def test_index_is_a_bijection_inverted_by_layout_coord(oracle, L, ig):
    q = oracle.qex_layout_setup(L, ig)
    V = oracle.volume(L)
    C = oracle.all_coords(L)
    idx = np.array([oracle.qex_layout_index(q, C[s]) for s in range(V)])
    assert sorted(idx.tolist()) == list(range(V))
    for s in range(V):
        assert oracle.qex_layout_coord(q, int(idx[s])) == tuple(int(v) for v in C[s])
    # even sites first: their outer slots come before every odd site's
    even = (C.sum(axis=1) & 1) == 0
    assert (idx[even] // q["nSitesInner"]).max() < q["nEvenOuter"] <= (idx[~even] // q["nSitesInner"]).min()


def test_inner_geom_one_is_the_canonical_order(oracle):
    L = (4, 6, 2, 8)
    q = oracle.qex_layout_setup(L, (1, 1, 1, 1))
    C = oracle.all_coords(L)
    for s in range(oracle.volume(L)):
        assert oracle.qex_layout_index(q, C[s]) == oracle.index(L, tuple(int(v) for v in C[s])) == s


def test_innercb_choice(oracle):
    q = oracle.qex_layout_setup((4, 4, 6, 6), (1, 1, 1, 2))
    assert q["innerCb"] == 3 and q["innerCbDir"] == 2 and q["og"] == [4, 4, 6, 3]
    q = oracle.qex_layout_setup((4, 4, 4, 4), (1, 1, 2, 2))
    assert q["innerCb"] == 0 and q["innerCbDir"] == 0 and q["nSitesInner"] == 4


def test_refused_layouts(oracle):
    with pytest.raises(AssertionError, match="divide"):
        oracle.qex_layout_setup((4, 4, 4, 6), (1, 1, 1, 4))
    with pytest.raises(AssertionError, match="above 2"):
        oracle.qex_layout_setup((8, 4, 4, 4), (4, 1, 1, 1))
    with pytest.raises(AssertionError, match="cb choice"):
        oracle.qex_layout_setup((4, 4, 4, 6), (1, 1, 1, 2))      # og = (4,4,4,3): innerCb 2 is even
    with pytest.raises(AssertionError, match="2's"):
        oracle.qex_layout_setup((2, 2, 2, 6), (2, 2, 2, 2))      # og = (1,1,1,3): no even outer extent


@pytest.mark.parametrize("L,ig", [((4, 4, 6, 6), (1, 1, 1, 2)), ((8, 4, 4, 4), (2, 1, 1, 2))])
def test_storage_roundtrip(oracle, L, ig):
    V = oracle.volume(L)
    rng = np.random.default_rng(5)
    f = rng.normal(size=(2, V, 9)) + 1j * rng.normal(size=(2, V, 9))
    raw = oracle.to_qex_layout(L, ig, f)
    q = oracle.qex_layout_setup(L, ig)
    assert raw.shape == (2, q["nSitesOuter"] * 9 * 2 * q["nSitesInner"])
    assert np.array_equal(oracle.from_qex_layout(L, ig, raw, 9), f)
    # the site at the origin is outer slot 0, lane 0: its first real part leads the buffer
    assert raw[0, 0] == f[0, 0, 0].real and raw[0, q["nSitesInner"]] == f[0, 0, 0].imag
