"""The SciDAC/ILDG file scanner of the C library (qexhip_scidac_read_info, host
code only, no GPU needed) on valid, malformed and fuzzed files: valid files
describe themselves as the oracle wrote them; every malformed file is refused
with an error naming the problem, never a crash; metadata copies respect the
caller's buffer size."""
import ctypes
import struct

import numpy as np
import pytest

SEED = 9090
L = (4, 4, 2, 6)


def _file(oracle, records=None, filemd="file md"):
    V = oracle.volume(L)
    U = oracle.gauge_random(L, SEED).reshape(4, V, 9)
    if records is None:
        records = [(U, "D", "record md", "Thu Dec  3 12:00:00 2026 UTC", False),
                   (U[:1, :, :3].copy(), "F", "vector md", "d2", False)]
    return oracle.scidac_file(L, filemd, records)


def _info(path, record=0):
    from qex_amd import fileio
    return fileio.fileInfo(str(path), record)


def test_valid_file_description(oracle, tmp_path):
    raw = _file(oracle)
    p = tmp_path / "ok.lime"
    p.write_bytes(raw)
    info, fmd, rmd = _info(p)
    assert info["dims"] == L and info["nrecords"] == 2 and fmd == "file md" and rmd == "record md"
    assert info["datatype"] == "QDP_D3_ColorMatrix" and info["typesize"] == 144 and info["datacount"] == 4
    assert info["date"] == "Thu Dec  3 12:00:00 2026 UTC" and info["has_checksum"]
    V = oracle.volume(L)
    U = oracle.gauge_random(L, SEED).reshape(4, V, 9)
    assert (info["suma"], info["sumb"]) == oracle.scidac_checksum(oracle.scidac_binary(L, U, "D"), 576)
    info, _, rmd = _info(p, 1)
    assert rmd == "vector md" and info["datatype"] == "QDP_F3_ColorVector" and info["typesize"] == 24
    assert info["precision"] == "F" and info["datacount"] == 1


def test_metadata_buffer_capacity(oracle, tmp_path):
    from qex_amd._lib import ScidacInfoC, call
    p = tmp_path / "ok.lime"
    p.write_bytes(_file(oracle, filemd="0123456789"))
    info = ScidacInfoC()
    fmd = ctypes.create_string_buffer(b"\xff" * 8)
    call("qexhip_scidac_read_info", str(p).encode(), 0, ctypes.byref(info), fmd, 4, None, 0)
    assert fmd.raw[:4] == b"012\0" and fmd.raw[4:8] == b"\xff" * 4 and info.filemd_len == 10


def _refused(path, match, record=0):
    from qex_amd import QexHipError
    with pytest.raises(QexHipError, match=match):
        _info(path, record)


def test_malformed_files_are_refused(oracle, tmp_path):
    raw = bytearray(_file(oracle))
    p = tmp_path / "bad.lime"
    cases = []
    cases.append((b"", "dimensions"))                                # nothing at all
    cases.append((b"\0" * 200, "not a LIME"))                        # bad magic
    cases.append((bytes(raw[:100]), "truncated record header"))      # cut inside the first header
    cases.append((bytes(raw[:-200]), "truncated"))                   # cut inside the last record
    big = bytearray(raw)
    big[8:16] = struct.pack(">Q", 1 << 62)                           # absurd record length
    cases.append((bytes(big), "truncated"))
    bad_mid = bytearray(raw)
    second = 144 + ((struct.unpack(">Q", raw[8:16])[0] + 7) // 8) * 8
    bad_mid[second:second + 4] = b"\0\0\0\0"                          # second header's magic
    cases.append((bytes(bad_mid), "corrupt record header"))
    for data, match in cases:
        p.write_bytes(bytes(data))
        _refused(p, match)
    # XML content: no dims, a 3-d lattice, a bad precision
    for xml, match in [("<scidacFile><spacetime>4</spacetime></scidacFile>", "dims"),
                       ("<scidacFile><spacetime>3</spacetime><dims>4 4 4 </dims></scidacFile>", "4-dimensional"),
                       ("<scidacFile><spacetime>4</spacetime><dims>4 x 4 4</dims></scidacFile>", "dims")]:
        f = (oracle.lime_record("scidac-private-file-xml", xml.encode(), True, False)
             + oracle.lime_record("scidac-file-xml", b"", False, True))
        p.write_bytes(f)
        _refused(p, match)
    priv = oracle.scidac_private_record(4, 9, "D", "now").replace("<precision>D</precision>",
                                                                  "<precision>Q</precision>")
    f = (oracle.lime_record("scidac-private-file-xml", oracle.scidac_private_file(L).encode(), True, False)
         + oracle.lime_record("scidac-file-xml", b"", False, True)
         + oracle.lime_record("scidac-private-record-xml", priv.encode(), True, False)
         + oracle.lime_record("scidac-binary-data", b"\0" * (oracle.volume(L) * 576), False, True))
    p.write_bytes(f)
    _refused(p, "precision")
    # binary length that does not match lattice x datacount x typesize
    f = (oracle.lime_record("scidac-private-file-xml", oracle.scidac_private_file(L).encode(), True, False)
         + oracle.lime_record("scidac-file-xml", b"", False, True)
         + oracle.lime_record("scidac-private-record-xml", oracle.scidac_private_record(4, 9, "D", "now").encode(),
                              True, False)
         + oracle.lime_record("scidac-binary-data", b"\0" * 1000, False, True))
    p.write_bytes(f)
    _refused(p, "length")
    # ILDG record of another field type or precision
    for fmt, match in [(oracle.ildg_format(L, "D").replace("su3gauge", "su2gauge"), "su3gauge"),
                       (oracle.ildg_format(L, "D").replace("<precision>64", "<precision>16"), "32 or 64")]:
        f = (oracle.lime_record("ildg-format", fmt.encode(), True, False)
             + oracle.lime_record("ildg-binary-data", b"\0" * (oracle.volume(L) * 576), False, True))
        p.write_bytes(f)
        _refused(p, match)
    # record index out of range
    p.write_bytes(bytes(raw))
    _refused(p, "out of range", record=2)
    _refused(p, "out of range", record=-1)


def test_fuzzed_headers_and_xml_never_crash(oracle, tmp_path):
    """Random byte changes in the header and XML region of a valid file: each
    one either still parses or is refused with an error."""
    from qex_amd import QexHipError
    raw = _file(oracle)
    # offsets of everything but the binary payloads
    spans, pos = [], 0
    while pos < len(raw):
        n = struct.unpack(">Q", raw[pos + 8:pos + 16])[0]
        t = raw[pos + 16:pos + 144].split(b"\0")[0]
        spans.append((pos, pos + 144 + (0 if t.endswith(b"binary-data") else n)))
        pos += 144 + ((n + 7) // 8) * 8
    rng = np.random.default_rng(SEED)
    p = tmp_path / "fuzz.lime"
    refused = parsed = 0
    for trial in range(300):
        b = bytearray(raw)
        for _ in range(int(rng.integers(1, 4))):
            a, e = spans[int(rng.integers(len(spans)))]
            b[int(rng.integers(a, e))] = int(rng.integers(256))
        p.write_bytes(bytes(b))
        try:
            _info(p, int(rng.integers(0, 2)))
            parsed += 1
        except QexHipError:
            refused += 1
    assert parsed + refused == 300 and refused > 0
