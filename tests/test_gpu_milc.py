"""MILC binary gauge files through the C ABI (qexhip_milc_read_info,
qexhip_milc_read, qexhip_milc_write) against the oracle's numpy reader and
writer: files the oracle writes (either byte order) read back to the oracle's
values exactly, files the library writes read back in the oracle with the same
checksum, a flipped bit fails the checksum, and a lattice of the wrong shape
is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 404


def test_read_oracle_files(oracle, gpu, tmp_path):
    from qex_amd import gauge
    L = (8, 6, 4, 8)
    U = oracle.gauge_random(L, SEED)
    lo = gpu.Layout(L)
    for swap in (False, True):
        path = tmp_path / f"cfg{int(swap)}.milc"
        s = oracle.milc_write(str(path), L, U, stamp=b"Tue Jan 02 03:04:05 2024", byteswap=swap)
        hdr = gauge.milcHeader(path)
        assert hdr["dims"] == L and hdr["byteswapped"] == swap and (hdr["sum29"], hdr["sum31"]) == s
        assert hdr["timestamp"].startswith("Tue Jan 02 03:04:05 2024")
        Ug, info = gauge.readMilc(lo, path)
        assert info["checksum_ok"]
        assert np.array_equal(Ug.cpu().numpy(), U.astype(np.complex64).astype(np.complex128))


def test_write_then_read(oracle, gpu, tmp_path):
    import torch
    from qex_amd import gauge
    L = (4, 8, 6, 4)
    lo = gpu.Layout(L)
    U = gauge.random_gauge(lo, SEED)
    path = tmp_path / "w.milc"
    gauge.writeMilc(lo, path, U, timestamp="qexhip")
    L2, Uo, stored, computed = oracle.milc_read(str(path))
    assert L2 == L and stored == computed
    U32 = U.cpu().numpy().astype(np.complex64).astype(np.complex128)
    assert np.array_equal(Uo, U32)
    U2, info = gauge.readMilc(lo, path)
    assert info["checksum_ok"] and torch.equal(U2.cpu(), torch.from_numpy(U32))


def test_corrupt_and_mismatched_files(oracle, gpu, tmp_path):
    from qex_amd import gauge, QexHipError
    L = (4, 4, 4, 4)
    lo = gpu.Layout(L)
    path = tmp_path / "c.milc"
    oracle.milc_write(str(path), L, oracle.gauge_random(L, SEED + 1))
    raw = bytearray(path.read_bytes())
    raw[96 + 1000] ^= 0x10
    path.write_bytes(bytes(raw))
    with pytest.raises(QexHipError, match="checksum"):
        gauge.readMilc(lo, path)
    with pytest.raises(QexHipError, match="dimensions"):
        gauge.readMilc(gpu.Layout((4, 4, 4, 6)), path)
    (tmp_path / "x.milc").write_bytes(b"\0" * 200)
    with pytest.raises(QexHipError, match="not a MILC"):
        gauge.milcHeader(tmp_path / "x.milc")


@pytest.mark.parametrize("L", [(2, 2, 2, 2), (32, 32, 32, 32)])
def test_roundtrip_sizes_and_checksum(oracle, gpu, tmp_path, L):
    """2^4: a partial 64-site block; 32^4: six t-slice chunks, the last one
    ragged.  The header's checksum equals the oracle's over the file's words."""
    import torch
    from qex_amd import gauge
    print(f"milc round trip L={L}", flush=True)
    lo = gpu.Layout(L)
    U = gauge.random_gauge(lo, SEED + 7)
    path = tmp_path / "big.milc"
    gauge.writeMilc(lo, path, U, timestamp="sizes")
    hdr = gauge.milcHeader(path)
    words = np.fromfile(path, dtype=np.uint32, offset=96)
    assert words.size == 72 * lo.nSites
    assert (hdr["sum29"], hdr["sum31"]) == oracle.milc_checksum(words)
    del words
    G, info = gauge.readMilc(lo, path)
    assert info["checksum_ok"] and torch.equal(G, U.to(torch.complex64).to(torch.complex128))
    path.unlink()
