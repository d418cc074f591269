"""Pins the oracle's MILC gauge-file pieces (src/new/milcIO.nim,
src/new/milcFileField.nim): the checksum against answers worked out by hand
(a rotate-direction, modulus or offset mistake fails one of them), and the
numpy writer/reader against each other in both byte orders.  The reference
ships no MILC sample file, so no external file pins them."""
import numpy as np

SEED = 777


# ------------------------------------------------------------ MILC gauge files
def test_milc_checksum_known_answers(oracle):
    # 30 ones: sum29 sets bits 0..28 once and bit 0 again at k = 29; sum31 sets bits 0..29
    assert oracle.milc_checksum(np.ones(30, dtype=np.uint32)) == (0x1FFFFFFE, 0x3FFFFFFF)
    # the top bit rotates into bit 0 at k = 1: a rotation, not a shift
    assert oracle.milc_checksum([0x80000000, 0x80000000]) == (0x80000001, 0x80000001)
    # the word index counts from the offset
    assert oracle.milc_checksum([1], offset=29) == (1, 1 << 29)
    assert oracle.milc_checksum([], offset=0) == (0, 0)


def test_milc_file_roundtrip_and_byte_order(oracle, tmp_path):
    L = (4, 6, 2, 4)
    U = oracle.gauge_random(L, SEED)
    s = oracle.milc_write(str(tmp_path / "a.milc"), L, U)
    L2, U2, stored, computed = oracle.milc_read(str(tmp_path / "a.milc"))
    assert L2 == L and stored == computed == s
    assert np.array_equal(U2, U.astype(np.complex64).astype(np.complex128))
    oracle.milc_write(str(tmp_path / "b.milc"), L, U, byteswap=True)
    _, U3, stored3, computed3 = oracle.milc_read(str(tmp_path / "b.milc"))
    assert stored3 == computed3 == s and np.array_equal(U3, U2)
    assert (tmp_path / "a.milc").stat().st_size == 96 + 288 * oracle.volume(L)
