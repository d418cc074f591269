"""SciDAC / ILDG lattice files: QEX's Reader and Writer (src/io/readerQiolite.nim,
src/io/writerQiolite.nim) and its loadGauge / saveGauge (src/gauge/gaugeUtils.nim:
96-131), over the C ABI (qexhip_scidac_*).  The device converts between the
library's canonical complex128 arrays and the file's big-endian lexicographic
bytes and computes the SciDAC checksum in the same pass; the host moves the
bytes through pinned buffers.

Fields are passed as they live in this library:

 - a gauge field: a (4, V, 3, 3) complex128 CUDA tensor in canonical site
   order (or a list of four (V, 3, 3) tensors, QEX's array[4, ColorMatrix]);
 - a color-vector Field (Layout.ColorVector()) or a (V, 3) tensor;
 - a (V,) complex tensor; or a list of up to four fields of one kind.

Single-precision records ("F") are widened to FP64 on reading; writing "F"
rounds each number to float32 (nearest even), as QEX's writer does when the
record precision is below the field's (writerQiolite.nim:83-100)."""
import ctypes
import os

import torch

from ._lib import QexHipError, ScidacInfoC, call

defFileMd = '<?xml version="1.0"?>\n<note>generated by QEX</note>\n'   # gaugeUtils.nim:118
defRecMd = '<?xml version="1.0"?>\n<note>gauge configuration</note>\n'  # gaugeUtils.nim:119


def _info_dict(info):
    # file text may hold any bytes: decode without failing
    return {"dims": tuple(info.dims), "nrecords": info.nrecords, "datatype": info.datatype.decode(errors="replace"),
            "date": info.date.decode(errors="replace"), "precision": info.precision.decode(errors="replace"),
            "colors": info.colors,
            "spins": info.spins, "typesize": info.typesize, "datacount": info.datacount, "ildg": bool(info.ildg),
            "has_checksum": bool(info.has_checksum), "suma": info.suma, "sumb": info.sumb,
            "calc_suma": info.calc_suma, "calc_sumb": info.calc_sumb, "checksum_ok": bool(info.checksum_ok)}


def fileInfo(fn, record=0):
    """(info dict, file metadata, record metadata) of field record `record`."""
    info = ScidacInfoC()
    call("qexhip_scidac_read_info", str(fn).encode(), int(record), ctypes.byref(info), None, 0, None, 0)
    fmd = ctypes.create_string_buffer(int(info.filemd_len) + 1)
    rmd = ctypes.create_string_buffer(int(info.recmd_len) + 1)
    call("qexhip_scidac_read_info", str(fn).encode(), int(record), ctypes.byref(info), fmd, len(fmd), rmd, len(rmd))
    return _info_dict(info), fmd.value.decode(errors="replace"), rmd.value.decode(errors="replace")


def getFileLattice(fn):
    """Lattice extents of a SciDAC/ILDG file (readerQiolite.nim:11-17); None if
    the file does not exist."""
    if not os.path.exists(fn):
        return None
    return list(fileInfo(fn)[0]["dims"])


class _Canon:
    """Canonical [nobj][V][nelem] complex128 CUDA view of the fields passed in,
    with the copy back after a read."""

    def __init__(self, lo, v):
        from .stag import Field
        V = lo.nSites
        self.lo, self.v, self.back = lo, v, None
        items = list(v) if isinstance(v, (list, tuple)) else None
        if isinstance(v, Field) or (items and isinstance(items[0], Field)):
            fs = items if items else [v]
            self.nobj, self.nelem = len(fs), 3
            self.t = torch.stack([f.get(device=torch.device("cuda")) for f in fs]).contiguous()
            self.back = lambda: [f.set(self.t[k]) for k, f in enumerate(fs)]
        elif items is not None:
            for t in items:
                _check_tensor(t)
            self.t = torch.stack(items).contiguous()
            if self.t.dim() < 2 or self.t.shape[1] != V:
                raise ValueError(f"field shape {tuple(items[0].shape)} does not match the lattice ({V} sites)")
            self.nobj = len(items)
            self.nelem = _nelem(self.t.shape[2:], V)
            self.back = lambda: [t.copy_(self.t[k]) for k, t in enumerate(items)]
        else:
            _check_tensor(v)
            shp = tuple(v.shape)
            if len(shp) >= 1 and shp[0] == V:
                self.nobj, self.nelem = 1, _nelem(shp[1:], V)
            elif len(shp) >= 2 and shp[1] == V:
                self.nobj, self.nelem = shp[0], _nelem(shp[2:], V)
            else:
                raise ValueError(f"field shape {shp} does not match the lattice ({V} sites)")
            if v.is_contiguous():
                self.t = v
            else:
                self.t = v.contiguous()
                self.back = lambda: v.copy_(self.t)
        if not 1 <= self.nobj <= 4:
            raise ValueError("1 to 4 fields per record")

    def ptr(self):
        return ctypes.c_void_p(self.t.data_ptr())


def _check_tensor(t):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.complex128):
        raise ValueError("fields must be complex128 CUDA tensors or ColorVector Fields")


def _nelem(inner, V):
    inner = tuple(inner)
    if inner == (3, 3):
        return 9
    if inner == (3,):
        return 3
    if inner == ():
        return 1
    raise ValueError(f"per-site shape {inner} is not (3, 3), (3,) or ()")


class Reader:
    """Reader of a SciDAC/ILDG file (readerQiolite.nim:19-51).  `read` moves to
    the next field record before reading unless `nextRecord` already did, as
    the reference does (readerQiolite.nim:95-98)."""

    def __init__(self, lo, fn):
        if not os.path.exists(fn):
            raise FileNotFoundError(fn)
        self.layout = lo
        self.fileName = str(fn)
        self.status = 0
        self.verb = 0
        info, self.fileMetadata, _ = fileInfo(fn, 0)
        self.latsize = list(info["dims"])
        self.nrecords = info["nrecords"]
        self.record = 0
        self.atNextRecord = True

    def _info(self):
        return fileInfo(self.fileName, self.record)

    @property
    def recordMetadata(self):
        return self._info()[2]

    @property
    def recordDate(self):
        return self._info()[0]["date"]

    @property
    def recordInfo(self):
        return self._info()[0]

    def nextRecord(self):
        self.record += 1
        self.atNextRecord = True

    def read(self, v):
        """Read the current field record into v (gauge tensor, Field, or list)."""
        if not self.atNextRecord:
            self.nextRecord()
        if self.record >= self.nrecords:
            raise QexHipError(f"{self.fileName}: no field record {self.record} (file holds {self.nrecords})")
        c = _Canon(self.layout, v)
        info = ScidacInfoC()
        torch.cuda.synchronize()
        call("qexhip_scidac_read", self.layout.h, self.fileName.encode(), int(self.record), c.ptr(), int(c.nobj),
             int(c.nelem), ctypes.byref(info))
        if c.back:
            c.back()
        self.atNextRecord = False
        self.lastInfo = _info_dict(info)
        return self.lastInfo

    def close(self):
        pass


def newReader(lo, fn):
    return Reader(lo, fn)


class Writer:
    """Writer of a SciDAC file (writerQiolite.nim:11-42): creating it writes the
    file record with metadata md; every `write` appends one field record."""

    def __init__(self, lo, fn, md):
        self.layout = lo
        self.fileName = str(fn)
        self.fileMetadata = md
        self.latsize = list(lo.physGeom)
        self.status = 0
        self.verb = 0
        call("qexhip_scidac_create", self.fileName.encode(), (ctypes.c_int * 4)(*lo.physGeom),
             (md or "").encode())

    def write(self, v, md="", prec="", ildg=False, date=None):
        """Append v as a record of precision prec ("" = the fields' own, FP64;
        "F"; "D") with record metadata md (writerQiolite.nim:83-160)."""
        if prec not in ("", "F", "D"):
            raise QexHipError(f"unknown precision {prec!r}")
        c = _Canon(self.layout, v)
        info = ScidacInfoC()
        torch.cuda.synchronize()
        call("qexhip_scidac_append", self.layout.h, self.fileName.encode(), (md or "").encode(), c.ptr(),
             int(c.nobj), int(c.nelem), (prec or "D").encode(), 1 if ildg else 0,
             date.encode() if date else None, ctypes.byref(info))
        self.lastInfo = _info_dict(info)
        return self.lastInfo

    def close(self):
        pass


def newWriter(lo, fn, md):
    return Writer(lo, fn, md)


def loadGauge(lo, g, fn):
    """Read a gauge configuration into g (gaugeUtils.nim:96-116); returns 0."""
    rd = newReader(lo, fn)
    rd.read(g)
    rd.close()
    return rd.status


def saveGauge(lo, g, fn, prec="", filemd=defFileMd, recordmd=defRecMd, ildg=False):
    """Write g as a SciDAC gauge file (gaugeUtils.nim:120-131); returns 0."""
    wr = newWriter(lo, fn, filemd)
    wr.write(g, recordmd, prec, ildg=ildg)
    wr.close()
    return wr.status


def time_io_kernels(lo, v, prec="D", nrep=5):
    """Device milliseconds per whole-lattice launch of the write-side (pack)
    and read-side (unpack) conversion + checksum kernels for fields v."""
    c = _Canon(lo, v)
    ms = (ctypes.c_double * 2)()
    torch.cuda.synchronize()
    call("qexhip_time_scidac", lo.h, c.ptr(), int(c.nobj), int(c.nelem), prec.encode(), int(nrep), ms)
    return ms[0], ms[1]
