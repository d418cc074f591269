// SciDAC and ILDG lattice files, the format of QEX's loadGauge/saveGauge and of
// its Reader/Writer (src/gauge/gaugeUtils.nim:96-131, src/io/readerQiolite.nim,
// src/io/writerQiolite.nim; QEX does the container work in the qiolite package,
// qex.nimble:35, which follows USQCD's QIO and c-lime).
//
// Container (LIME): records of a 144-byte big-endian header {uint32 magic
// 0x456789ab, uint16 version 1, uint16 flags (0x8000 begins a message, 0x4000
// ends one), uint64 data length, char type[128]}, then the data, zero-padded
// to a multiple of 8 bytes.  A SciDAC file is a file message
// {scidac-private-file-xml, scidac-file-xml} followed by one message per field
// record {scidac-private-record-xml, scidac-record-xml, [ildg-format],
// scidac-binary-data | ildg-binary-data, scidac-checksum}.  Binary data are
// big-endian, sites in lexicographic order (x fastest), each site holding
// `datacount` objects of `typesize` bytes (a 3x3 complex matrix row-major).
// The checksum: every site's file bytes get their own CRC-32 c, and with r the
// site's lexicographic rank, suma ^= rotl(c, r mod 29), sumb ^= rotl(c, r mod 31).
//
// Division of work: the host scans and writes the LIME headers and XML; the
// device converts between the library's canonical complex128 arrays
// [nobj][site][nelem] and file bytes and computes the per-site CRCs and the
// rank-rotated XOR partials in the same pass (one wave per 64 sites, the
// site's words staged through LDS so that file bytes move coalesced).  The
// binary record streams through two pinned host buffers in whole t-slices of
// at most 64 MiB, so disk transfers of one chunk (pread/pwrite from up to 8
// host threads) overlap the device work on the next.
#include <stdio.h>
#include <string.h>
#include <fcntl.h>
#include <time.h>
#include <unistd.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "internal.h"
#include "iopipe.h"
#include "su3.h"

namespace qh {

namespace {

constexpr uint32_t LIME_MAGIC = 0x456789abu;
constexpr int LIME_HEADER = 144;
using namespace iop;  // SITES_PER_BLOCK (64 sites, one per lane of a wave), Pipe, Chunking, pio

// ---------------------------------------------------------------------------
// device side
//
// A block holds 64 consecutive sites and Q = 4 (or 2) neighbouring lanes per
// site, so that the lanes of a site load and store neighbouring complex
// numbers of that site together.  Lane (s, k) converts site s's
// complex numbers q = k, k+Q, ... into the site's row of the LDS tile, then
// runs the CRC register over words [k W/Q, (k+1) W/Q) of that row with
// slicing-by-4 tables (four independent lookups per 32-bit word).  Its
// register r_k is moved to the end of the site by the GF(2)-linear map Z^(Q-1-k),
// Z = the register update over W/Q*4 zero bytes (the same combination as
// src/io/crc32.nim:78-91 zeroPadCrc32): crc = ~(Z^3 r_0 ^ Z^2 r_1 ^ Z r_2 ^ r_3)
// with r_0 started from ~0 and the others from 0.  Z^1..Z^3 come in as 32
// columns each, built on the host.

constexpr int MAXQ = 4;

// T[0] = byte table of the reflected CRC-32 polynomial 0xedb88320
// (src/io/crc32.nim:8-14); T[j][i] = (T[j-1][i] >> 8) ^ T[0][T[j-1][i] & 0xff]
__device__ __forceinline__ void crc_tables(uint32_t *T) {
  for (int e = threadIdx.x; e < 256; e += blockDim.x) {
    uint32_t r = (uint32_t)e;
    for (int j = 0; j < 8; j++) r = (r & 1) ? (r >> 1) ^ 0xedb88320u : r >> 1;
    T[e] = r;
  }
  __syncthreads();
  for (int j = 1; j < 4; j++) {
    for (int e = threadIdx.x; e < 256; e += blockDim.x) {
      const uint32_t p = T[(j - 1) * 256 + e];
      T[j * 256 + e] = (p >> 8) ^ T[p & 0xff];
    }
    __syncthreads();
  }
}

// four file bytes (w's bytes in memory order) through the CRC register
__device__ __forceinline__ uint32_t crc_word(const uint32_t *T, uint32_t c, uint32_t w) {
  c ^= w;
  return T[768 + (c & 0xff)] ^ T[512 + ((c >> 8) & 0xff)] ^ T[256 + ((c >> 16) & 0xff)] ^ T[c >> 24];
}

// M (32 columns) applied to c over GF(2)
__device__ __forceinline__ uint32_t gf2_apply(const uint32_t *M, uint32_t c) {
  uint32_t o = 0;
#pragma unroll
  for (int i = 0; i < 32; i++) o ^= M[i] & (0u - ((c >> i) & 1u));
  return o;
}

__device__ __forceinline__ long canon_of_lex(const Geom &g, long lex) {
  int c[4];
  long r = lex;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    c[d] = (int)(r % g.L[d]);
    r /= g.L[d];
  }
  return canon_index(g, c);
}

// LDS: 4 CRC tables, the Z^1..Z^3 columns, the tile (64 rows of W+1 words,
// an odd stride so lanes of one wave hit distinct banks), Q x 64 registers
struct Smem {
  uint32_t *T, *Z, *tile, *comb;
  int W, ld;
  __device__ Smem(uint32_t *lds, int w) : W(w), ld(w + 1) {
    T = lds;
    Z = lds + 1024;
    tile = Z + 96;
    comb = tile + SITES_PER_BLOCK * ld;
  }
};

__device__ __forceinline__ void load_shift(uint32_t *Z, const uint32_t *zmat) {
  for (int i = threadIdx.x; i < 96; i += blockDim.x) Z[i] = zmat[i];
}

// CRC of lane (s, k)'s part of row `row`, moved to the end of the site
__device__ __forceinline__ uint32_t part_crc(const Smem &m, const uint32_t *row, int k, int Q) {
  const int wq = m.W / Q;
  uint32_t r = k == 0 ? 0xffffffffu : 0u;
  for (int w = k * wq; w < (k + 1) * wq; w++) r = crc_word(m.T, r, row[w]);
  const int sh = Q - 1 - k;
  return sh > 0 ? gf2_apply(m.Z + 32 * (sh - 1), r) : r;
}

// wave 0: combine the Q registers of every site, rotate by the site's rank,
// XOR over the block into part[2*blockIdx.x ..]
__device__ __forceinline__ void block_checksum(const Smem &m, int Q, long lex_first, long nv, uint32_t *part) {
  if (threadIdx.x >= SITES_PER_BLOCK) return;
  const int s = threadIdx.x;
  uint32_t a = 0, b = 0;
  if (s < nv) {
    uint32_t c = 0;
    for (int k = 0; k < Q; k++) c ^= m.comb[k * SITES_PER_BLOCK + s];
    c = ~c;
    const long lex = lex_first + s;
    a = rotl32(c, (int)(lex % 29));
    b = rotl32(c, (int)(lex % 31));
  }
  a = wave_xor(a);
  b = wave_xor(b);
  if (s == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// canonical [nobj][V][nelem] complex128 -> file words of the sites
// lex0 .. lex0+n-1 (n counted from the chunk start), and the block's
// (suma, sumb) partial; blockDim = 64 Q
template <bool SINGLE>
__global__ void __launch_bounds__(SITES_PER_BLOCK *MAXQ)
    k_scidac_pack(Geom g, const double2 *src, int nobj, int nelem, long lex0, long n, uint32_t *out, uint32_t *part,
                  const uint32_t *zmat) {
  extern __shared__ uint32_t lds[];
  constexpr int WPC = SINGLE ? 2 : 4;  // words per complex number
  const int N = nobj * nelem;
  const Smem m(lds, N * WPC);
  const int Q = blockDim.x / SITES_PER_BLOCK;
  const int s = threadIdx.x / Q, k = threadIdx.x % Q;  // a site's Q lanes side by side in one wave
  load_shift(m.Z, zmat);
  crc_tables(m.T);
  const long V = 2 * g.Vh;
  const long s0 = (long)blockIdx.x * SITES_PER_BLOCK;
  const long nv = n - s0 < SITES_PER_BLOCK ? n - s0 : SITES_PER_BLOCK;
  uint32_t *row = m.tile + s * m.ld;
  if (s < nv) {
    const long idx = canon_of_lex(g, lex0 + s0 + s);
    for (int q = k; q < N; q += Q) {
      const int o = q / nelem, e = q - o * nelem;
      const double2 z = src[((long)o * V + idx) * nelem + e];
      if (SINGLE) {
        row[2 * q] = __builtin_bswap32(__float_as_uint((float)z.x));
        row[2 * q + 1] = __builtin_bswap32(__float_as_uint((float)z.y));
      } else {
        const uint64_t br = __builtin_bswap64((uint64_t)__double_as_longlong(z.x));
        const uint64_t bi = __builtin_bswap64((uint64_t)__double_as_longlong(z.y));
        row[4 * q] = (uint32_t)br;
        row[4 * q + 1] = (uint32_t)(br >> 32);
        row[4 * q + 2] = (uint32_t)bi;
        row[4 * q + 3] = (uint32_t)(bi >> 32);
      }
    }
  }
  __syncthreads();
  m.comb[k * SITES_PER_BLOCK + s] = s < nv ? part_crc(m, row, k, Q) : 0u;
  __syncthreads();
  block_checksum(m, Q, lex0 + s0, nv, part);
  const int total = (int)nv * m.W;
  uint32_t *dst = out + s0 * m.W;
  for (int w = threadIdx.x; w < total; w += blockDim.x) {
    const int r = w / m.W;
    dst[w] = m.tile[r * m.ld + (w - r * m.W)];
  }
}

// file words -> canonical complex128, with the same checksum partials
template <bool SINGLE>
__global__ void __launch_bounds__(SITES_PER_BLOCK *MAXQ)
    k_scidac_unpack(Geom g, const uint32_t *in, int nobj, int nelem, long lex0, long n, double2 *dst, uint32_t *part,
                    const uint32_t *zmat) {
  extern __shared__ uint32_t lds[];
  constexpr int WPC = SINGLE ? 2 : 4;
  const int N = nobj * nelem;
  const Smem m(lds, N * WPC);
  const int Q = blockDim.x / SITES_PER_BLOCK;
  const int s = threadIdx.x / Q, k = threadIdx.x % Q;  // a site's Q lanes side by side in one wave
  load_shift(m.Z, zmat);
  const long s0 = (long)blockIdx.x * SITES_PER_BLOCK;
  const long nv = n - s0 < SITES_PER_BLOCK ? n - s0 : SITES_PER_BLOCK;
  const int total = (int)nv * m.W;
  const uint32_t *src = in + s0 * m.W;
  for (int w = threadIdx.x; w < total; w += blockDim.x) {
    const int r = w / m.W;
    m.tile[r * m.ld + (w - r * m.W)] = src[w];
  }
  crc_tables(m.T);  // ends with a barrier, after which the tile is complete
  const long V = 2 * g.Vh;
  const uint32_t *row = m.tile + s * m.ld;
  if (s < nv) {
    const long idx = canon_of_lex(g, lex0 + s0 + s);
    for (int q = k; q < N; q += Q) {
      const int o = q / nelem, e = q - o * nelem;
      double2 z;
      if (SINGLE) {
        z.x = (double)__uint_as_float(__builtin_bswap32(row[2 * q]));
        z.y = (double)__uint_as_float(__builtin_bswap32(row[2 * q + 1]));
      } else {
        z.x = __longlong_as_double((long long)__builtin_bswap64(((uint64_t)row[4 * q + 1] << 32) | row[4 * q]));
        z.y = __longlong_as_double((long long)__builtin_bswap64(((uint64_t)row[4 * q + 3] << 32) | row[4 * q + 2]));
      }
      dst[((long)o * V + idx) * nelem + e] = z;
    }
  }
  m.comb[k * SITES_PER_BLOCK + s] = s < nv ? part_crc(m, row, k, Q) : 0u;
  __syncthreads();
  block_checksum(m, Q, lex0 + s0, nv, part);
}

// ---------------------------------------------------------------------------
// host side: files, LIME records, XML

struct File {
  FILE *f = nullptr;
  File(const char *path, const char *mode) : f(fopen(path, mode)) {
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  }
  ~File() {
    if (f) fclose(f);
  }
  File(const File &) = delete;
  File &operator=(const File &) = delete;
};

void put_be(unsigned char *p, uint64_t v, int nbytes) {
  for (int i = nbytes - 1; i >= 0; i--, v >>= 8) p[i] = (unsigned char)(v & 0xff);
}

uint64_t get_be(const unsigned char *p, int nbytes) {
  uint64_t v = 0;
  for (int i = 0; i < nbytes; i++) v = (v << 8) | p[i];
  return v;
}

void fwrite_all(FILE *f, const void *p, size_t n) {
  if (n && fwrite(p, 1, n, f) != n) throw std::runtime_error("SciDAC write: write failed");
}

void lime_header(FILE *f, const char *type, uint64_t len, bool mb, bool me) {
  unsigned char h[LIME_HEADER];
  memset(h, 0, sizeof h);
  put_be(h, LIME_MAGIC, 4);
  put_be(h + 4, 1, 2);
  put_be(h + 6, (mb ? 0x8000u : 0u) | (me ? 0x4000u : 0u), 2);
  put_be(h + 8, len, 8);
  strncpy((char *)h + 16, type, 128);
  fwrite_all(f, h, sizeof h);
}

void lime_pad(FILE *f, uint64_t len) {
  static const unsigned char zero[8] = {0};
  fwrite_all(f, zero, (8 - len % 8) % 8);
}

void lime_record(FILE *f, const char *type, const std::string &data, bool mb, bool me) {
  lime_header(f, type, data.size(), mb, me);
  fwrite_all(f, data.data(), data.size());
  lime_pad(f, data.size());
}

struct LimeRec {
  std::string type;
  long long off = 0, len = 0;  // data offset and length in the file
  bool mb = false, me = false;
};

std::vector<LimeRec> lime_scan(FILE *f) {
  if (fseeko(f, 0, SEEK_END) != 0) throw std::runtime_error("SciDAC read: seek failed");
  const long long size = ftello(f);
  std::vector<LimeRec> recs;
  long long pos = 0;
  while (pos < size) {
    unsigned char h[LIME_HEADER];
    if (fseeko(f, pos, SEEK_SET) != 0 || fread(h, 1, sizeof h, f) != sizeof h)
      throw std::runtime_error("LIME file: truncated record header");
    if (get_be(h, 4) != LIME_MAGIC)
      throw std::runtime_error(recs.empty() ? "not a LIME (SciDAC/ILDG) file" : "LIME file: corrupt record header");
    LimeRec r;
    const unsigned flags = (unsigned)get_be(h + 6, 2);
    r.mb = (flags & 0x8000u) != 0;
    r.me = (flags & 0x4000u) != 0;
    const uint64_t len = get_be(h + 8, 8);
    r.type.assign((const char *)h + 16, strnlen((const char *)h + 16, 128));
    r.off = pos + LIME_HEADER;
    if (len > (uint64_t)(size - r.off)) throw std::runtime_error("LIME file: truncated record '" + r.type + "'");
    r.len = (long long)len;
    recs.push_back(r);
    pos = r.off + r.len + (8 - r.len % 8) % 8;
  }
  return recs;
}

std::string read_string(FILE *f, const LimeRec &r) {
  if (r.len > (1ll << 30)) throw std::runtime_error("LIME file: metadata record too large");
  std::string s((size_t)r.len, '\0');
  if (fseeko(f, r.off, SEEK_SET) != 0 || (r.len && fread(&s[0], 1, (size_t)r.len, f) != (size_t)r.len))
    throw std::runtime_error("LIME file: short read");
  s.resize(strnlen(s.data(), s.size()));  // writers may include the C string's NUL
  return s;
}

bool xml_tag(const std::string &doc, const char *tag, std::string *val) {
  const std::string open = std::string("<") + tag + ">", close = std::string("</") + tag + ">";
  const size_t a = doc.find(open);
  if (a == std::string::npos) return false;
  const size_t b = doc.find(close, a + open.size());
  if (b == std::string::npos) return false;
  *val = doc.substr(a + open.size(), b - a - open.size());
  return true;
}

long xml_long(const std::string &doc, const char *tag, const char *what) {
  std::string v;
  if (!xml_tag(doc, tag, &v)) throw std::runtime_error(std::string(what) + ": no <" + tag + ">");
  char *end = nullptr;
  const long x = strtol(v.c_str(), &end, 10);
  if (end == v.c_str()) throw std::runtime_error(std::string(what) + ": bad <" + tag + ">");
  return x;
}

const char *XML_DECL = "<?xml version=\"1.0\" encoding=\"UTF-8\"?>";

struct SciRecord {
  std::string priv, user, ildg, date, datatype;
  LimeRec bin;
  bool has_ck = false;
  uint32_t suma = 0, sumb = 0;
  char prec = 'D';
  int colors = 0, spins = 0, typesize = 0, datacount = 0;
  bool ildg_bin = false;
};

struct SciFile {
  int dims[4] = {0, 0, 0, 0};
  std::string filemd;
  std::vector<SciRecord> recs;
};

void finish_record(SciRecord &r) {
  if (!r.priv.empty()) {
    const char *w = "scidac-private-record-xml";
    xml_tag(r.priv, "datatype", &r.datatype);
    xml_tag(r.priv, "date", &r.date);
    std::string p;
    if (!xml_tag(r.priv, "precision", &p) || (p != "F" && p != "D"))
      throw std::runtime_error("SciDAC record: precision must be F or D");
    r.prec = p[0];
    r.colors = (int)xml_long(r.priv, "colors", w);
    r.spins = (int)xml_long(r.priv, "spins", w);
    r.typesize = (int)xml_long(r.priv, "typesize", w);
    r.datacount = (int)xml_long(r.priv, "datacount", w);
  } else if (!r.ildg.empty()) {  // ILDG-only record: an SU(3) gauge field, four links per site
    std::string field;
    if (!xml_tag(r.ildg, "field", &field) || field != "su3gauge")
      throw std::runtime_error("ILDG record: only su3gauge fields are supported");
    const long p = xml_long(r.ildg, "precision", "ildg-format");
    if (p != 32 && p != 64) throw std::runtime_error("ILDG record: precision must be 32 or 64");
    r.prec = p == 32 ? 'F' : 'D';
    r.datatype = p == 32 ? "QDP_F3_ColorMatrix" : "QDP_D3_ColorMatrix";
    r.colors = 3;
    r.spins = 1;
    r.typesize = p == 32 ? 72 : 144;
    r.datacount = 4;
  } else {
    throw std::runtime_error("SciDAC file: binary record without a record description");
  }
  if (r.typesize <= 0 || r.datacount <= 0) throw std::runtime_error("SciDAC record: bad typesize or datacount");
}

SciFile scidac_scan(FILE *f) {
  const std::vector<LimeRec> lime = lime_scan(f);
  SciFile sf;
  bool have_dims = false;
  std::string ildg_first;
  size_t i = 0;
  while (i < lime.size()) {
    // one LIME message: from a record with the begin flag up to one with the end flag
    size_t j = i;
    while (j + 1 < lime.size() && !lime[j].me && !lime[j + 1].mb) j++;
    SciRecord r;
    bool has_bin = false;
    for (size_t k = i; k <= j; k++) {
      const LimeRec &l = lime[k];
      if (l.type == "scidac-private-file-xml") {
        const std::string x = read_string(f, l);
        if (xml_long(x, "spacetime", "scidac-private-file-xml") != 4)
          throw std::runtime_error("SciDAC file: only 4-dimensional lattices are supported");
        std::string d;
        if (!xml_tag(x, "dims", &d) ||
            sscanf(d.c_str(), "%d %d %d %d", &sf.dims[0], &sf.dims[1], &sf.dims[2], &sf.dims[3]) != 4)
          throw std::runtime_error("SciDAC file: bad <dims>");
        have_dims = true;
      } else if (l.type == "scidac-file-xml") {
        sf.filemd = read_string(f, l);
      } else if (l.type == "scidac-private-record-xml") {
        r.priv = read_string(f, l);
      } else if (l.type == "scidac-record-xml") {
        r.user = read_string(f, l);
      } else if (l.type == "ildg-format") {
        r.ildg = read_string(f, l);
        if (ildg_first.empty()) ildg_first = r.ildg;
      } else if (l.type == "scidac-binary-data" || l.type == "ildg-binary-data") {
        if (has_bin) throw std::runtime_error("SciDAC file: two binary records in one message");
        r.bin = l;
        r.ildg_bin = l.type == "ildg-binary-data";
        has_bin = true;
      } else if (l.type == "scidac-checksum") {
        const std::string x = read_string(f, l);
        std::string a, b;
        if (!xml_tag(x, "suma", &a) || !xml_tag(x, "sumb", &b))
          throw std::runtime_error("SciDAC file: bad scidac-checksum record");
        r.suma = (uint32_t)strtoul(a.c_str(), nullptr, 16);
        r.sumb = (uint32_t)strtoul(b.c_str(), nullptr, 16);
        r.has_ck = true;
      }  // other record types (ildg-data-lfn, ...) carry nothing needed here
    }
    if (has_bin) {
      finish_record(r);
      sf.recs.push_back(r);
    }
    i = j + 1;
  }
  if (!have_dims) {
    if (ildg_first.empty()) throw std::runtime_error("SciDAC file: no lattice dimensions");
    const char *tags[4] = {"lx", "ly", "lz", "lt"};
    for (int d = 0; d < 4; d++) sf.dims[d] = (int)xml_long(ildg_first, tags[d], "ildg-format");
  }
  long long V = 1;
  for (int d = 0; d < 4; d++) {
    if (sf.dims[d] <= 0) throw std::runtime_error("SciDAC file: bad lattice dimensions");
    V *= sf.dims[d];
  }
  for (const SciRecord &r : sf.recs)
    if (r.bin.len != V * r.datacount * (long long)r.typesize)
      throw std::runtime_error("SciDAC file: binary record length does not match lattice, typesize and datacount");
  return sf;
}

void info_of(const SciFile &sf, int record, ScidacInfo *info) {
  for (int d = 0; d < 4; d++) info->dims[d] = sf.dims[d];
  info->nrecords = (int)sf.recs.size();
  info->filemd = sf.filemd;
  if (record < 0 || record >= (int)sf.recs.size()) return;
  const SciRecord &r = sf.recs[record];
  info->recmd = r.user;
  info->datatype = r.datatype;
  info->date = r.date;
  info->prec = r.prec;
  info->colors = r.colors;
  info->spins = r.spins;
  info->typesize = r.typesize;
  info->datacount = r.datacount;
  info->ildg = r.ildg_bin ? 1 : 0;
  info->has_checksum = r.has_ck ? 1 : 0;
  info->suma = r.suma;
  info->sumb = r.sumb;
}

std::string type_name(int nelem, char prec) {
  switch (nelem) {
    case 9: return std::string("QDP_") + prec + "3_ColorMatrix";
    case 3: return std::string("QDP_") + prec + "3_ColorVector";
    default: return std::string("QDP_") + prec + "_Complex";
  }
}

std::string now_utc() {
  const time_t t = time(nullptr);
  struct tm tmv;
  gmtime_r(&t, &tmv);
  char buf[64];
  strftime(buf, sizeof buf, "%a %b %e %H:%M:%S %Y UTC", &tmv);
  return buf;
}

// launch shape of the pack/unpack kernels: Q lanes per site (W/Q words each)
struct Kern {
  int W, Q;
  dim3 block;
  size_t lds;
  Kern(int nobj, int nelem, bool single) {
    W = nobj * nelem * (single ? 2 : 4);
    Q = W % 4 == 0 ? 4 : 2;
    block = dim3(SITES_PER_BLOCK * Q);
    lds = (1024 + 96 + (size_t)SITES_PER_BLOCK * (W + 1) + (size_t)SITES_PER_BLOCK * Q) * sizeof(uint32_t);
  }
};

// Z^1..Z^3 as 32 columns each, Z = the CRC register update over W/Q*4 zero
// bytes (column i = the register 1 << i after those bytes), to dev (96 words)
void upload_shift(uint32_t *dev, const Kern &k, hipStream_t st) {
  uint32_t T0[256];
  for (uint32_t e = 0; e < 256; e++) {
    uint32_t r = e;
    for (int j = 0; j < 8; j++) r = (r & 1) ? (r >> 1) ^ 0xedb88320u : r >> 1;
    T0[e] = r;
  }
  const int bytes = k.W / k.Q * 4;
  uint32_t M[96];
  for (int j = 1; j <= 3; j++)
    for (int i = 0; i < 32; i++) {
      uint32_t c = 1u << i;
      for (int b = 0; b < j * bytes; b++) c = T0[c & 0xff] ^ (c >> 8);
      M[(j - 1) * 32 + i] = c;
    }
  if (hipMemcpyAsync(dev, M, sizeof M, hipMemcpyHostToDevice, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    throw std::runtime_error("SciDAC I/O: copy to device failed");
}

void launch_pack(bool single, const Kern &k, unsigned grid, hipStream_t st, const Geom &g, const double *src,
                 int nobj, int nelem, long lex0, long n, uint32_t *out, uint32_t *part, const uint32_t *zmat) {
  if (single)
    hipLaunchKernelGGL(k_scidac_pack<true>, dim3(grid), k.block, k.lds, st, g, (const double2 *)src, nobj, nelem, lex0,
                       n, out, part, zmat);
  else
    hipLaunchKernelGGL(k_scidac_pack<false>, dim3(grid), k.block, k.lds, st, g, (const double2 *)src, nobj, nelem,
                       lex0, n, out, part, zmat);
}

void launch_unpack(bool single, const Kern &k, unsigned grid, hipStream_t st, const Geom &g, const uint32_t *in,
                   int nobj, int nelem, long lex0, long n, double *dst, uint32_t *part, const uint32_t *zmat) {
  if (single)
    hipLaunchKernelGGL(k_scidac_unpack<true>, dim3(grid), k.block, k.lds, st, g, in, nobj, nelem, lex0, n,
                       (double2 *)dst, part, zmat);
  else
    hipLaunchKernelGGL(k_scidac_unpack<false>, dim3(grid), k.block, k.lds, st, g, in, nobj, nelem, lex0, n,
                       (double2 *)dst, part, zmat);
}

void check_shape(int nobj, int nelem) {
  if (nelem != 1 && nelem != 3 && nelem != 9) throw std::runtime_error("SciDAC I/O: nelem must be 1, 3 or 9");
  if (nobj < 1 || nobj > 4) throw std::runtime_error("SciDAC I/O: nobj must be 1 to 4");
}

}  // namespace

void scidac_info(const char *path, int record, ScidacInfo *info) {
  File f(path, "rb");
  const SciFile sf = scidac_scan(f.f);
  if (record < 0 || record >= (int)sf.recs.size()) {
    if (!(record == 0 && sf.recs.empty())) throw std::runtime_error("SciDAC file: record index out of range");
  }
  info_of(sf, record, info);
}

void scidac_read(const Geom &g, const char *path, int record, double *dst, int nobj, int nelem, ScidacInfo *info,
                 hipStream_t st) {
  check_shape(nobj, nelem);
  File f(path, "rb");
  const SciFile sf = scidac_scan(f.f);
  if (record < 0 || record >= (int)sf.recs.size()) throw std::runtime_error("SciDAC file: record index out of range");
  info_of(sf, record, info);
  for (int d = 0; d < 4; d++)
    if (sf.dims[d] != g.L[d]) throw std::runtime_error("SciDAC file: field lattice differs from the file's lattice");
  const SciRecord &r = sf.recs[record];
  if (r.datacount != nobj) throw std::runtime_error("SciDAC file: field count differs from the record's datacount");
  const bool single = r.prec == 'F';
  if (r.typesize != nelem * 2 * (single ? 4 : 8))
    throw std::runtime_error("SciDAC file: field data elements differ from the record's typesize");
  const Chunking ch(g, (size_t)nobj * r.typesize);
  Pipe p(st, (size_t)ch.per * ch.site_bytes, ch.nblocks, 96, "SciDAC I/O");
  const int fd = fileno(f.f);
  const Kern kern(nobj, nelem, single);
  upload_shift(p.aux, kern, st);
  long blk = 0;
  for (long c = 0; c < ch.nchunks; c++) {
    const int b = (int)(c & 1);
    const long n = ch.sites(g, c);
    const size_t bytes = (size_t)n * ch.site_bytes;
    if (c >= 2 && hipEventSynchronize(p.ev[b]) != hipSuccess) throw std::runtime_error("SciDAC read: device error");
    pio(false, fd, (char *)p.host[b], bytes, (off_t)(r.bin.off + (long long)(c * ch.per) * (long long)ch.site_bytes),
        "SciDAC read");
    if (hipMemcpyAsync(p.dev[b], p.host[b], bytes, hipMemcpyHostToDevice, st) != hipSuccess)
      throw std::runtime_error("SciDAC read: copy to device failed");
    const unsigned grid = (unsigned)((n + SITES_PER_BLOCK - 1) / SITES_PER_BLOCK);
    launch_unpack(single, kern, grid, st, g, p.dev[b], nobj, nelem, c * ch.per, n, dst, p.part + 2 * blk, p.aux);
    if (hipGetLastError() != hipSuccess || hipEventRecord(p.ev[b], st) != hipSuccess)
      throw std::runtime_error("SciDAC read: kernel launch failed");
    blk += grid;
  }
  if (hipStreamSynchronize(st) != hipSuccess) throw std::runtime_error("SciDAC read: device error");
  p.checksum(ch.nblocks, &info->calc_suma, &info->calc_sumb, "SciDAC read");
  info->checksum_ok = r.has_ck ? (info->calc_suma == r.suma && info->calc_sumb == r.sumb) : 1;
  if (!info->checksum_ok) throw std::runtime_error("SciDAC file: checksum mismatch");
}

void scidac_create(const char *path, const int dims[4], const char *filemd) {
  for (int d = 0; d < 4; d++)
    if (dims[d] <= 0) throw std::runtime_error("SciDAC write: bad lattice dimensions");
  char priv[512];
  snprintf(priv, sizeof priv,
           "%s<scidacFile><version>1.1</version><spacetime>4</spacetime><dims>%d %d %d %d </dims>"
           "<volfmt>0</volfmt></scidacFile>",
           XML_DECL, dims[0], dims[1], dims[2], dims[3]);
  File f(path, "wb");
  lime_record(f.f, "scidac-private-file-xml", priv, true, false);
  lime_record(f.f, "scidac-file-xml", filemd ? filemd : "", false, true);
  if (fflush(f.f) != 0) throw std::runtime_error("SciDAC write: write failed");
}

void scidac_append(const Geom &g, const char *path, const char *recmd, const double *src, int nobj, int nelem,
                   char prec, int ildg, const char *date, ScidacInfo *info, hipStream_t st) {
  check_shape(nobj, nelem);
  if (prec != 'F' && prec != 'D') throw std::runtime_error("SciDAC write: precision must be F or D");
  if (ildg && !(nobj == 4 && nelem == 9)) throw std::runtime_error("SciDAC write: ILDG records hold gauge fields only");
  File f(path, "r+b");
  const SciFile sf = scidac_scan(f.f);
  for (int d = 0; d < 4; d++)
    if (sf.dims[d] != g.L[d]) throw std::runtime_error("SciDAC write: field lattice differs from the file's lattice");
  const bool single = prec == 'F';
  const int typesize = nelem * 2 * (single ? 4 : 8);
  const std::string when = date ? std::string(date) : now_utc();
  const std::string dtype = type_name(nelem, prec);
  char priv[1024];
  snprintf(priv, sizeof priv,
           "%s<scidacRecord><version>1.1</version><date>%s</date><recordtype>0</recordtype>"
           "<datatype>%s</datatype><precision>%c</precision><colors>%d</colors><spins>1</spins>"
           "<typesize>%d</typesize><datacount>%d</datacount></scidacRecord>",
           XML_DECL, when.c_str(), dtype.c_str(), prec, nelem > 1 ? 3 : 1, typesize, nobj);
  if (fseeko(f.f, 0, SEEK_END) != 0) throw std::runtime_error("SciDAC write: seek failed");
  lime_record(f.f, "scidac-private-record-xml", priv, true, false);
  lime_record(f.f, "scidac-record-xml", recmd ? recmd : "", false, false);
  if (ildg) {
    char fmt[768];
    snprintf(fmt, sizeof fmt,
             "%s<ildgFormat xmlns=\"http://www.lqcd.org/ildg\" xmlns:xsi=\"http://www.w3.org/2001/XMLSchema-instance\" "
             "xsi:schemaLocation=\"http://www.lqcd.org/ildg http://www.lqcd.org/ildg/filefmt.xsd\">"
             "<version>1.0</version><field>su3gauge</field><precision>%d</precision>"
             "<lx>%d</lx><ly>%d</ly><lz>%d</lz><lt>%d</lt></ildgFormat>",
             XML_DECL, single ? 32 : 64, g.L[0], g.L[1], g.L[2], g.L[3]);
    lime_record(f.f, "ildg-format", fmt, false, false);
  }
  const long V = 2 * g.Vh;
  const Chunking ch(g, (size_t)nobj * typesize);
  const uint64_t total = (uint64_t)V * ch.site_bytes;
  lime_header(f.f, ildg ? "ildg-binary-data" : "scidac-binary-data", total, false, false);
  if (fflush(f.f) != 0) throw std::runtime_error("SciDAC write: write failed");
  const off_t base = ftello(f.f);
  const int fd = fileno(f.f);
  (void)posix_fallocate(fd, base, (off_t)total);  // reserve the record's blocks; a refusal is harmless
  Pipe p(st, (size_t)ch.per * ch.site_bytes, ch.nblocks, 96, "SciDAC I/O");
  const Kern kern(nobj, nelem, single);
  upload_shift(p.aux, kern, st);
  std::vector<size_t> pending(ch.nchunks, 0);
  long blk = 0;
  auto drain = [&](long c) {  // write chunk c's bytes once its copy to the host is done
    const int b = (int)(c & 1);
    if (hipEventSynchronize(p.ev[b]) != hipSuccess) throw std::runtime_error("SciDAC write: device error");
    pio(true, fd, (char *)p.host[b], pending[c], base + (off_t)(c * ch.per) * (off_t)ch.site_bytes, "SciDAC write");
  };
  for (long c = 0; c < ch.nchunks; c++) {
    const int b = (int)(c & 1);
    if (c >= 2) drain(c - 2);
    const long n = ch.sites(g, c);
    pending[c] = (size_t)n * ch.site_bytes;
    const unsigned grid = (unsigned)((n + SITES_PER_BLOCK - 1) / SITES_PER_BLOCK);
    launch_pack(single, kern, grid, st, g, src, nobj, nelem, c * ch.per, n, p.dev[b], p.part + 2 * blk, p.aux);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(p.host[b], p.dev[b], pending[c], hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipEventRecord(p.ev[b], st) != hipSuccess)
      throw std::runtime_error("SciDAC write: conversion failed");
    blk += grid;
  }
  for (long c = ch.nchunks >= 2 ? ch.nchunks - 2 : 0; c < ch.nchunks; c++) drain(c);
  if (fseeko(f.f, base + (off_t)total, SEEK_SET) != 0) throw std::runtime_error("SciDAC write: seek failed");
  lime_pad(f.f, total);
  uint32_t a, b;
  p.checksum(ch.nblocks, &a, &b, "SciDAC write");
  char ck[256];
  snprintf(ck, sizeof ck,
           "%s<scidacChecksum><version>1.0</version><suma>%x</suma><sumb>%x</sumb></scidacChecksum>", XML_DECL, a,
           b);
  lime_record(f.f, "scidac-checksum", ck, false, true);
  if (fflush(f.f) != 0) throw std::runtime_error("SciDAC write: write failed");
  if (info) {
    for (int d = 0; d < 4; d++) info->dims[d] = g.L[d];
    info->nrecords = (int)sf.recs.size() + 1;
    info->filemd = sf.filemd;
    info->recmd = recmd ? recmd : "";
    info->datatype = dtype;
    info->date = when;
    info->prec = prec;
    info->colors = nelem > 1 ? 3 : 1;
    info->spins = 1;
    info->typesize = typesize;
    info->datacount = nobj;
    info->ildg = ildg ? 1 : 0;
    info->has_checksum = 1;
    info->suma = info->calc_suma = a;
    info->sumb = info->calc_sumb = b;
    info->checksum_ok = 1;
  }
}

void scidac_time(const Geom &g, const double *src, int nobj, int nelem, char prec, int nrep, double ms[2],
                 hipStream_t st) {
  check_shape(nobj, nelem);
  if (prec != 'F' && prec != 'D') throw std::runtime_error("SciDAC timing: precision must be F or D");
  if (nrep < 1) throw std::runtime_error("SciDAC timing: nrep must be positive");
  const bool single = prec == 'F';
  const long V = 2 * g.Vh;
  const size_t site_bytes = (size_t)nobj * nelem * 2 * (single ? 4 : 8);
  const unsigned grid = (unsigned)((V + SITES_PER_BLOCK - 1) / SITES_PER_BLOCK);
  const Kern kern(nobj, nelem, single);
  struct Bufs {
    uint32_t *file = nullptr, *part = nullptr, *zmat = nullptr;
    double *canon = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    hipStream_t st;
    ~Bufs() {
      (void)hipStreamSynchronize(st);
      if (file) (void)hipFree(file);
      if (part) (void)hipFree(part);
      if (zmat) (void)hipFree(zmat);
      if (canon) (void)hipFree(canon);
      for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    }
  } b;
  b.st = st;
  if (hipMalloc((void **)&b.file, (size_t)V * site_bytes) != hipSuccess ||
      hipMalloc((void **)&b.part, (size_t)grid * 2 * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void **)&b.zmat, 96 * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void **)&b.canon, (size_t)nobj * V * nelem * sizeof(double2)) != hipSuccess)
    throw std::runtime_error("SciDAC timing: out of device memory");
  for (auto &e : b.ev)
    if (hipEventCreate(&e) != hipSuccess) throw std::runtime_error("SciDAC timing: event creation failed");
  upload_shift(b.zmat, kern, st);
  auto pack = [&]() { launch_pack(single, kern, grid, st, g, src, nobj, nelem, 0L, V, b.file, b.part, b.zmat); };
  auto unpack = [&]() { launch_unpack(single, kern, grid, st, g, b.file, nobj, nelem, 0L, V, b.canon, b.part, b.zmat); };
  pack();
  unpack();
  if (hipEventRecord(b.ev[0], st) != hipSuccess) throw std::runtime_error("SciDAC timing: event failed");
  for (int r = 0; r < nrep; r++) pack();
  if (hipEventRecord(b.ev[1], st) != hipSuccess) throw std::runtime_error("SciDAC timing: event failed");
  for (int r = 0; r < nrep; r++) unpack();
  if (hipEventRecord(b.ev[2], st) != hipSuccess || hipGetLastError() != hipSuccess ||
      hipEventSynchronize(b.ev[2]) != hipSuccess)
    throw std::runtime_error("SciDAC timing: kernels failed");
  float t0 = 0, t1 = 0;
  (void)hipEventElapsedTime(&t0, b.ev[0], b.ev[1]);
  (void)hipEventElapsedTime(&t1, b.ev[1], b.ev[2]);
  ms[0] = t0 / nrep;
  ms[1] = t1 / nrep;
}

}  // namespace qh
