"""SciDAC / ILDG files through the C ABI (qexhip_scidac_*), mirroring QEX's
tests/base/tgaugeio.nim and tfieldio.nim:

 - a gauge field saved in double precision loads back bit for bit, and in
   single precision loads back as the float32 rounding of the field, with the
   plaquette following (tgaugeio.nim:20-51);
 - the bytes the library writes equal the oracle's file byte for byte (same
   metadata and date), so the device checksum equals the oracle's;
 - files the oracle writes, ILDG-named records and pure ILDG files included,
   read back exactly; records of several kinds in one file read back in turn
   with their metadata (tfieldio.nim:27-96);
 - a flipped bit fails the checksum; a lattice, datacount or element-count
   mismatch is refused;
 - at 32^4 (eleven 64 MiB chunks, the last one ragged) the FP64 round trip is
   exact and the checksums computed on writing and reading agree; a 2^4
   lattice exercises a partial 64-site block."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 5150
DATE = "Thu Dec  3 12:00:00 2026 UTC"


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_gauge_roundtrip_both_precisions(oracle, gpu, tmp_path):
    import torch
    from qex_amd import fileio, gauge, hmc
    L = (8, 8, 8, 8)
    print(f"scidac gauge round trip L={L}", flush=True)
    lo = gpu.Layout(L)
    U = gauge.random_gauge(lo, SEED)
    p = hmc.plaquette(lo, U)
    fn = tmp_path / "tmplat.lime"
    assert fileio.saveGauge(lo, U, fn) == 0
    G = torch.zeros_like(U)
    assert fileio.loadGauge(lo, G, fn) == 0
    assert torch.equal(G, U) and hmc.plaquette(lo, G) == p
    assert fileio.saveGauge(lo, U, fn, "F") == 0
    G.zero_()
    rd = fileio.newReader(lo, fn)
    assert rd.fileMetadata == fileio.defFileMd and rd.latsize == list(L)
    info = rd.read(G)
    assert info["checksum_ok"] and info["precision"] == "F" and info["typesize"] == 72
    assert rd.recordMetadata == fileio.defRecMd
    U32 = U.to(torch.complex64).to(torch.complex128)
    assert torch.equal(G, U32)
    assert abs(hmc.plaquette(lo, G) - hmc.plaquette(lo, U32)) <= 1e-12 * abs(p)


@pytest.mark.parametrize("prec,ildg", [("D", False), ("F", False), ("D", True)])
def test_bytes_equal_oracle_file(oracle, gpu, tmp_path, prec, ildg):
    L = (4, 6, 2, 8)
    V = oracle.volume(L)
    lo = gpu.Layout(L)
    U = oracle.gauge_random(L, SEED + 1)
    fn = tmp_path / "g.lime"
    from qex_amd import fileio
    w = fileio.newWriter(lo, fn, "file md")
    info = w.write(_cuda(U), "record md", prec, ildg=ildg, date=DATE)
    want = oracle.scidac_file(L, "file md", [(U.reshape(4, V, 9), prec, "record md", DATE, ildg)])
    assert fn.read_bytes() == want
    sb = 4 * 9 * (8 if prec == "F" else 16)
    binary = oracle.scidac_binary(L, U.reshape(4, V, 9), prec)
    assert (info["suma"], info["sumb"]) == oracle.scidac_checksum(binary, sb)


def test_read_oracle_files_and_records(oracle, gpu, tmp_path):
    import torch
    from qex_amd import fileio
    L = (4, 4, 6, 8)
    V = oracle.volume(L)
    lo = gpu.Layout(L)
    U = oracle.gauge_random(L, SEED + 2).reshape(4, V, 9)
    x = oracle.source_gauss(L, 3, 0, 2).reshape(1, V, 3)
    y = oracle.source_gauss(L, 4, 0, 2).reshape(V, 3)[:, :1].reshape(1, V, 1).copy()
    fn = tmp_path / "multi.lime"
    fn.write_bytes(oracle.scidac_file(L, "test filemd", [
        (U, "D", "gauge md", DATE, True), (x, "F", "vector md", DATE, False), (y, "D", "complex md", DATE, False)]))
    rd = fileio.newReader(lo, fn)
    assert rd.nrecords == 3 and rd.fileMetadata == "test filemd"
    G = torch.zeros((4, V, 3, 3), dtype=torch.complex128, device="cuda")
    info = rd.read(G)
    assert info["ildg"] and info["checksum_ok"] and rd.recordMetadata == "gauge md" and rd.recordDate == DATE
    assert np.array_equal(G.cpu().numpy().reshape(4, V, 9), U)
    f = lo.ColorVector()
    rd.read(f)
    assert rd.recordMetadata == "vector md"
    assert np.array_equal(f.get(), x[0].astype(np.complex64).astype(np.complex128))
    c = torch.zeros(V, dtype=torch.complex128, device="cuda")
    info = rd.read(c)
    assert info["datatype"] == "QDP_D_Complex" and np.array_equal(c.cpu().numpy(), y[0, :, 0])
    with pytest.raises(fileio.QexHipError, match="no field record"):
        rd.read(c)
    # a pure ILDG file: no SciDAC records, dimensions and precision from ildg-format
    binary = oracle.scidac_binary(L, U, "F")
    raw = (oracle.lime_record("ildg-format", oracle.ildg_format(L, "F").encode(), True, False)
           + oracle.lime_record("ildg-binary-data", binary, False, False)
           + oracle.lime_record("ildg-data-lfn", b"lfn://test", False, True))
    (tmp_path / "pure.ildg").write_bytes(raw)
    assert fileio.getFileLattice(str(tmp_path / "pure.ildg")) == list(L)
    G.zero_()
    info = fileio.newReader(lo, tmp_path / "pure.ildg").read(G)
    assert not info["has_checksum"] and info["precision"] == "F"
    assert np.array_equal(G.cpu().numpy().reshape(4, V, 9), U.astype(np.complex64).astype(np.complex128))


def test_write_several_records_then_read(oracle, gpu, tmp_path):
    import torch
    from qex_amd import fileio, gauge
    L = (4, 4, 4, 8)
    lo = gpu.Layout(L)
    V = lo.nSites
    U = gauge.random_gauge(lo, SEED + 3)
    fs = [lo.ColorVector().set(oracle.source_gauss(L, 7 + k, 0, 2)) for k in range(2)]
    fn = tmp_path / "tmpfield.lime"
    w = fileio.newWriter(lo, fn, "test filemd")
    w.write(fs, "two vectors", "D", date=DATE)
    w.write([U[mu] for mu in range(4)], "gauge as four matrices", "F", date=DATE)
    raw = fn.read_bytes()
    got, _, user, stored, computed = oracle.scidac_read(raw, L, 2, 3, 0)
    assert user == "two vectors" and stored == computed
    assert np.array_equal(got, np.stack([f.get() for f in fs]))
    got, _, user, stored, computed = oracle.scidac_read(raw, L, 4, 9, 1)
    assert stored == computed
    assert np.array_equal(got, U.cpu().numpy().reshape(4, V, 9).astype(np.complex64).astype(np.complex128))
    rd = fileio.newReader(lo, fn)
    gs = [lo.ColorVector(), lo.ColorVector()]
    rd.read(gs)
    for a, b in zip(gs, fs):
        assert np.array_equal(a.get(), b.get())
    parts = [torch.zeros((V, 3, 3), dtype=torch.complex128, device="cuda") for _ in range(4)]
    rd.read(parts)
    assert torch.equal(torch.stack(parts), U.to(torch.complex64).to(torch.complex128))


def test_corrupt_and_mismatched(oracle, gpu, tmp_path):
    import torch
    from qex_amd import fileio, QexHipError
    L = (4, 4, 4, 4)
    V = oracle.volume(L)
    lo = gpu.Layout(L)
    U = oracle.gauge_random(L, SEED + 4).reshape(4, V, 9)
    fn = tmp_path / "c.lime"
    raw = bytearray(oracle.scidac_file(L, "f", [(U, "D", "r", DATE, False)]))
    G = torch.zeros((4, V, 3, 3), dtype=torch.complex128, device="cuda")
    fn.write_bytes(bytes(raw))
    with pytest.raises(QexHipError, match="lattice"):
        fileio.newReader(gpu.Layout((4, 4, 4, 6)), fn).read(
            torch.zeros((4, 384, 3, 3), dtype=torch.complex128, device="cuda"))
    with pytest.raises(QexHipError, match="count"):
        fileio.newReader(lo, fn).read(G[:2])
    with pytest.raises(QexHipError, match="elements"):
        fileio.newReader(lo, fn).read(torch.zeros((4, V, 3), dtype=torch.complex128, device="cuda"))
    pos = bytes(raw).find(b"scidac-binary-data") - 16 + 144 + 5000
    raw[pos] ^= 0x04
    fn.write_bytes(bytes(raw))
    with pytest.raises(QexHipError, match="checksum"):
        fileio.loadGauge(lo, G, fn)
    (tmp_path / "junk").write_bytes(b"\1" * 300)
    with pytest.raises(QexHipError, match="LIME"):
        fileio.getFileLattice(str(tmp_path / "junk"))
    (tmp_path / "short").write_bytes(bytes(raw[:len(raw) // 2]))
    with pytest.raises(QexHipError, match="truncated"):
        fileio.getFileLattice(str(tmp_path / "short"))


@pytest.mark.parametrize("L", [(2, 2, 2, 2), (32, 32, 32, 32)])
def test_roundtrip_sizes(oracle, gpu, tmp_path, L):
    import torch
    from qex_amd import fileio, gauge
    print(f"scidac round trip L={L}", flush=True)
    lo = gpu.Layout(L)
    U = gauge.random_gauge(lo, SEED + 5)
    fn = tmp_path / "big.lime"
    w = fileio.newWriter(lo, fn, "")
    winfo = w.write(U, "", "D")
    G = torch.zeros_like(U)
    rinfo = fileio.newReader(lo, fn).read(G)
    assert torch.equal(G, U)
    assert (rinfo["calc_suma"], rinfo["calc_sumb"]) == (winfo["suma"], winfo["sumb"]) == (rinfo["suma"],
                                                                                         rinfo["sumb"])
    if lo.nSites <= 4096:
        raw = fn.read_bytes()
        _, _, _, stored, computed = oracle.scidac_read(raw, L, 4, 9, 0)
        assert stored == computed == (winfo["suma"], winfo["sumb"])
    fn.unlink()


def test_kernel_timing_entry(gpu):
    from qex_amd import fileio, gauge
    lo = gpu.Layout((8, 8, 8, 8))
    U = gauge.random_gauge(lo, SEED + 6)
    for prec in ("D", "F"):
        pk, up = fileio.time_io_kernels(lo, U, prec, nrep=2)
        assert pk > 0 and up > 0
