"""Pins the oracle's SciDAC / ILDG file pieces.  QEX does this work in the
qiolite package (qex.nimble:35, not vendored), so the anchors are:

 - CRC-32: the reference's own known answer (src/io/crc32.nim:113-116,
   0x414FA339) and zlib's crc32, the textbook routine, on random bytes;
 - the record checksum: recomputed here from zlib.crc32 and explicit
   rotations, site by site, with a rank offset (a wrong modulus, rotation
   direction or rank fails it);
 - LIME headers: bytes assembled by hand from the published layout;
 - the file: written and read back (FP64 exactly, FP32 as numpy's rounding),
   with each site's bytes found where lexicographic order puts them.

The reference ships no SciDAC or ILDG sample file, so no external file pins
the XML records; their text follows QIO's, as QEX's Reader expects."""
import struct
import zlib

import numpy as np

SEED = 4242


def rotl(v, k):
    return ((v << k) | (v >> (32 - k))) & 0xFFFFFFFF if k else v


def test_crc32_known_answers(oracle):
    assert oracle.crc32(b"The quick brown fox jumps over the lazy dog") == 0x414FA339
    assert oracle.crc32(b"") == 0
    rng = np.random.default_rng(SEED)
    for n in (1, 3, 8, 1000):
        b = rng.bytes(n)
        assert oracle.crc32(b) == zlib.crc32(b)


def test_scidac_checksum_against_zlib(oracle):
    rng = np.random.default_rng(SEED + 1)
    site_bytes, nsites, rank0 = 24, 70, 5
    data = rng.bytes(site_bytes * nsites)
    a = b = 0
    for k in range(nsites):
        c = zlib.crc32(data[k * site_bytes:(k + 1) * site_bytes])
        r = rank0 + k
        a ^= rotl(c, r % 29)
        b ^= rotl(c, r % 31)
    assert oracle.scidac_checksum(data, site_bytes, rank0) == (a, b)
    # a site's contribution depends on its rank: swapping two sites changes the sums
    sw = data[site_bytes:2 * site_bytes] + data[:site_bytes] + data[2 * site_bytes:]
    assert oracle.scidac_checksum(sw, site_bytes, rank0) != (a, b)


def test_lime_header_layout(oracle):
    rec = oracle.lime_record("abc", b"12345", True, False)
    want = bytes.fromhex("456789ab" "0001" "8000" "0000000000000005") + b"abc" + b"\0" * 125 + b"12345" + b"\0" * 3
    assert rec == want
    rec = oracle.lime_record("x", b"", False, True)
    assert rec[:16] == bytes.fromhex("456789ab" "0001" "4000" "0000000000000000") and len(rec) == 144
    recs = oracle.lime_records(oracle.lime_record("a", b"x" * 8, True, False) + oracle.lime_record("b", b"yy", False, True))
    assert [(t, mb, me, d) for t, mb, me, d in recs] == [("a", True, False, b"x" * 8), ("b", False, True, b"yy")]


def test_scidac_file_roundtrip_and_site_order(oracle):
    L = (4, 6, 2, 4)
    V = oracle.volume(L)
    U = np.ascontiguousarray(oracle.gauge_random(L, SEED).reshape(4, V, 9))
    raw = oracle.scidac_file(L, "file md", [(U, "D", "record md", "today", False),
                                            (U, "F", "second", "today", True)])
    got, priv, user, stored, computed = oracle.scidac_read(raw, L, 4, 9, 0)
    assert np.array_equal(got, U) and user == "record md" and stored == computed
    assert "<datatype>QDP_D3_ColorMatrix</datatype>" in priv and "<typesize>144</typesize>" in priv
    got, priv, user, stored, computed = oracle.scidac_read(raw, L, 4, 9, 1)
    assert np.array_equal(got, U.astype(np.complex64).astype(np.complex128)) and stored == computed
    # site (x,y,z,t) = (1,2,1,3), link mu = 2, element (0,1): big-endian doubles at
    # ((lex*4 + mu)*9 + 1)*16 bytes into the first binary record
    recs = oracle.lime_records(raw)
    binary = [d for t, _, _, d in recs if t == "scidac-binary-data"][0]
    x = (1, 2, 1, 3)
    lex = x[0] + L[0] * (x[1] + L[1] * (x[2] + L[2] * x[3]))
    idx = oracle.index(L, x)
    off = ((lex * 4 + 2) * 9 + 1) * 16
    re, im = struct.unpack(">dd", binary[off:off + 16])
    assert complex(re, im) == U[2, idx, 1]
    types = [t for t, _, _, _ in recs]
    assert types == ["scidac-private-file-xml", "scidac-file-xml",
                     "scidac-private-record-xml", "scidac-record-xml", "scidac-binary-data", "scidac-checksum",
                     "scidac-private-record-xml", "scidac-record-xml", "ildg-format", "ildg-binary-data",
                     "scidac-checksum"]
