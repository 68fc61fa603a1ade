// Device Parquet scan: the footer and page headers are parsed on the host
// (Thrift compact protocol), the page bodies (definition levels, dictionary
// indices, PLAIN values) are decoded on the device into device columns.
//
// It takes the place of Table::open / read_row_group / prune
// (crates/minilake-storage/src/table.rs:59-240) and read_column_chunk /
// convert_stats (crates/minilake-storage/src/parquet_reader.rs:210-358), whose
// page decoding the reference delegates to the `parquet` crate (60.0.0, the
// Apache Parquet format). Supported: flat schemas (REQUIRED / OPTIONAL
// columns), UNCOMPRESSED and SNAPPY chunks, data pages v1 and v2, the PLAIN,
// dictionary (PLAIN_DICTIONARY / RLE_DICTIONARY), RLE (booleans),
// DELTA_BINARY_PACKED, DELTA_LENGTH_BYTE_ARRAY, DELTA_BYTE_ARRAY and
// BYTE_STREAM_SPLIT encodings, RLE definition levels. Byte arrays (strings and
// BYTE_ARRAY decimals) are assembled on the host; integers of DELTA pages are
// unpacked and prefix-summed on the device. SNAPPY is the only codec, as in
// the reference, which builds the crate with the "snap" feature alone
// (Cargo.toml:30).
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "kernels.h"

namespace mlg {
namespace pq {

// Parquet enums (parquet.thrift)
enum PhysType { T_BOOLEAN = 0, T_INT32 = 1, T_INT64 = 2, T_INT96 = 3, T_FLOAT = 4, T_DOUBLE = 5, T_BYTE_ARRAY = 6, T_FLBA = 7 };
enum Conv { C_BOOL, C_INT32, C_DATE, C_INT64, C_FLOAT, C_DOUBLE, C_UTF8, C_DEC_I32, C_DEC_I64, C_DEC_BYTES };
enum { CT_UTF8 = 0, CT_DECIMAL = 5, CT_DATE = 6 };
enum {
  ENC_PLAIN = 0,
  ENC_PLAIN_DICTIONARY = 2,
  ENC_RLE = 3,
  ENC_DELTA_BINARY_PACKED = 5,
  ENC_DELTA_LENGTH_BYTE_ARRAY = 6,
  ENC_DELTA_BYTE_ARRAY = 7,
  ENC_RLE_DICTIONARY = 8,
  ENC_BYTE_STREAM_SPLIT = 9
};
enum { PAGE_DATA = 0, PAGE_INDEX = 1, PAGE_DICTIONARY = 2, PAGE_DATA_V2 = 3 };

[[noreturn]] static void corrupt(const char* what) { fail(ML_ERR_PARQUET, std::string("corrupt parquet: ") + what); }

// Thrift compact protocol
struct TReader {
  const uint8_t* p;
  size_t n;
  size_t pos = 0;
  uint8_t byte() {
    if (pos >= n) corrupt("metadata ends early");
    return p[pos++];
  }
  uint64_t varint() {
    uint64_t v = 0;
    for (int s = 0; s < 64; s += 7) {
      const uint8_t b = byte();
      v |= (uint64_t)(b & 0x7f) << s;
      if (!(b & 0x80)) return v;
    }
    corrupt("varint too long");
  }
  int64_t zz() {
    const uint64_t v = varint();
    return (int64_t)(v >> 1) ^ -(int64_t)(v & 1);
  }
  std::string_view binary() {
    const uint64_t len = varint();
    if (len > n - pos) corrupt("binary field past the end");
    std::string_view s(reinterpret_cast<const char*>(p) + pos, (size_t)len);
    pos += (size_t)len;
    return s;
  }
  void list(int* et, uint64_t* size) {
    const uint8_t h = byte();
    *et = h & 0x0f;
    uint64_t s = h >> 4;
    if (s == 15) s = varint();
    *size = s;
  }
  // next field of the current struct; false at its STOP byte
  bool field(int* last, int* id, int* type) {
    const uint8_t h = byte();
    if (h == 0) return false;
    *type = h & 0x0f;
    const int d = h >> 4;
    *id = d ? *last + d : (int)zz();
    *last = *id;
    return true;
  }
  int64_t integer(int type) {
    if (type == 3) return (int8_t)byte();
    if (type == 4 || type == 5 || type == 6) return zz();
    corrupt("integer field of another type");
  }
  void skip(int type, int depth = 0) {
    if (depth > 32) corrupt("nesting too deep");
    switch (type) {
      case 1: case 2: break;  // boolean field: the value is the type
      case 3: byte(); break;
      case 4: case 5: case 6: varint(); break;
      case 7:
        if (n - pos < 8) corrupt("double past the end");
        pos += 8;
        break;
      case 8: binary(); break;
      case 9: case 10: {
        int et;
        uint64_t sz;
        list(&et, &sz);
        for (uint64_t i = 0; i < sz; i++) {
          if (et == 1 || et == 2) byte();
          else skip(et, depth + 1);
        }
        break;
      }
      case 11: {
        const uint64_t sz = varint();
        if (sz) {
          const uint8_t kv = byte();
          for (uint64_t i = 0; i < sz; i++) {
            skip(kv >> 4, depth + 1);
            skip(kv & 0x0f, depth + 1);
          }
        }
        break;
      }
      case 12: {
        int last = 0, id, t;
        while (field(&last, &id, &t)) skip(t, depth + 1);
        break;
      }
      default: corrupt("unknown thrift type");
    }
  }
};

struct SchemaElem {
  int type = -1, type_length = 0, repetition = 0, num_children = 0, converted = -1, scale = 0;
  std::string name;
};

struct Stats {
  bool has_min = false, has_max = false, has_null_count = false;
  bool min_exact = true, max_exact = true;
  std::string min, max;
  int64_t null_count = 0;
};

struct ChunkMeta {
  int type = -1, codec = 0;
  int64_t num_values = 0, total_compressed = 0, data_page_offset = -1, dict_page_offset = -1;
  bool has_stats = false;
  Stats stats;
};

struct RowGroupMeta {
  int64_t num_rows = 0, compressed = 0;
  std::vector<ChunkMeta> cols;
};

static void parse_stats(TReader& r, Stats& s) {
  int last = 0, id, t;
  while (r.field(&last, &id, &t)) {
    if (id == 3 && t == 6) {
      s.null_count = r.zz();
      s.has_null_count = true;
    } else if (id == 5 && t == 8) {
      s.max = std::string(r.binary());
      s.has_max = true;
    } else if (id == 6 && t == 8) {
      s.min = std::string(r.binary());
      s.has_min = true;
    } else if (id == 7 && (t == 1 || t == 2)) {
      s.max_exact = t == 1;
    } else if (id == 8 && (t == 1 || t == 2)) {
      s.min_exact = t == 1;
    } else {
      r.skip(t);  // the deprecated min / max (1, 2) are not used
    }
  }
}

static void parse_column_meta(TReader& r, ChunkMeta& c) {
  int last = 0, id, t;
  while (r.field(&last, &id, &t)) {
    switch (id) {
      case 1: c.type = (int)r.integer(t); break;
      case 4: c.codec = (int)r.integer(t); break;
      case 5: c.num_values = r.integer(t); break;
      case 7: c.total_compressed = r.integer(t); break;
      case 9: c.data_page_offset = r.integer(t); break;
      case 11: c.dict_page_offset = r.integer(t); break;
      case 12:
        if (t != 12) corrupt("statistics");
        parse_stats(r, c.stats);
        c.has_stats = true;
        break;
      default: r.skip(t);
    }
  }
}

static void parse_row_group(TReader& r, RowGroupMeta& g) {
  int last = 0, id, t;
  while (r.field(&last, &id, &t)) {
    if (id == 1 && t == 9) {
      int et;
      uint64_t sz;
      r.list(&et, &sz);
      for (uint64_t i = 0; i < sz; i++) {  // ColumnChunk
        ChunkMeta c;
        int l2 = 0, id2, t2;
        bool has_meta = false;
        while (r.field(&l2, &id2, &t2)) {
          if (id2 == 1 && t2 == 8) fail(ML_ERR_UNSUPPORTED, "parquet column chunks in other files");
          if (id2 == 3 && t2 == 12) {
            parse_column_meta(r, c);
            has_meta = true;
          } else {
            r.skip(t2);
          }
        }
        if (!has_meta) corrupt("column chunk without metadata");
        g.cols.push_back(std::move(c));
      }
    } else if (id == 3) {
      g.num_rows = r.integer(t);
    } else if (id == 6) {
      g.compressed = r.integer(t);
    } else {
      r.skip(t);
    }
  }
}

struct PageHeader {
  int type = -1;
  int64_t uncompressed = 0;
  int64_t compressed = 0;
  int64_t num_values = 0;
  int encoding = -1, def_encoding = -1;
  // DataPageHeaderV2: rows, and the level bytes stored uncompressed before the values
  int64_t num_rows = -1, def_len = 0, rep_len = 0;
  bool values_compressed = true;
};

static void parse_page_header(TReader& r, PageHeader& h) {
  int last = 0, id, t;
  while (r.field(&last, &id, &t)) {
    if (id == 1) {
      h.type = (int)r.integer(t);
    } else if (id == 2) {
      h.uncompressed = r.integer(t);
    } else if (id == 3) {
      h.compressed = r.integer(t);
    } else if ((id == 5 || id == 7) && t == 12) {  // DataPageHeader / DictionaryPageHeader
      int l2 = 0, id2, t2;
      while (r.field(&l2, &id2, &t2)) {
        if (id2 == 1) h.num_values = r.integer(t2);
        else if (id2 == 2) h.encoding = (int)r.integer(t2);
        else if (id2 == 3 && id == 5) h.def_encoding = (int)r.integer(t2);
        else r.skip(t2);
      }
    } else if (id == 8 && t == 12) {  // DataPageHeaderV2
      int l2 = 0, id2, t2;
      while (r.field(&l2, &id2, &t2)) {
        if (id2 == 1) h.num_values = r.integer(t2);
        else if (id2 == 3) h.num_rows = r.integer(t2);
        else if (id2 == 4) h.encoding = (int)r.integer(t2);
        else if (id2 == 5) h.def_len = r.integer(t2);
        else if (id2 == 6) h.rep_len = r.integer(t2);
        else if (id2 == 7 && (t2 == 1 || t2 == 2)) h.values_compressed = t2 == 1;
        else r.skip(t2);
      }
    } else {
      r.skip(t);
    }
  }
}

// Snappy, from the format's published description: the uncompressed length
// as a little-endian varint, then elements: a literal (tag & 3 == 0; length
// - 1 in the tag's upper 6 bits, or in the 1-4 bytes after it when those bits
// are 60-63) or a back-reference of 1-, 2- or 4-byte offset (tag & 3 = 1, 2,
// 3), copied forward byte by byte so a reference may overlap its output.
static void snappy_decompress(const uint8_t* in, size_t n, size_t expect, std::vector<uint8_t>& out) {
  size_t i = 0;
  uint64_t len = 0;
  for (int shift = 0;; shift += 7) {
    if (i >= n || shift > 28) corrupt("snappy length");
    const uint8_t b = in[i++];
    len |= (uint64_t)(b & 0x7f) << shift;
    if (!(b & 0x80)) break;
  }
  if (len != expect) corrupt("snappy length differs from the page header");
  const size_t base = out.size();
  out.resize(base + (size_t)len);
  uint8_t* o = out.data() + base;
  size_t op = 0;
  while (i < n) {
    const uint8_t tag = in[i++];
    size_t l, off;
    if ((tag & 3) == 0) {
      l = tag >> 2;
      if (l >= 60) {
        const size_t nb = l - 59;
        if (i + nb > n) corrupt("snappy literal length");
        l = 0;
        for (size_t k = 0; k < nb; k++) l |= (size_t)in[i + k] << (8 * k);
        i += nb;
      }
      l += 1;
      if (l > n - i || l > len - op) corrupt("snappy literal past the end");
      std::memcpy(o + op, in + i, l);
      i += l;
      op += l;
      continue;
    }
    if ((tag & 3) == 1) {
      if (i >= n) corrupt("snappy copy");
      l = 4 + ((tag >> 2) & 7);
      off = ((size_t)(tag >> 5) << 8) | in[i++];
    } else if ((tag & 3) == 2) {
      if (n - i < 2) corrupt("snappy copy");
      l = (size_t)(tag >> 2) + 1;
      off = (size_t)in[i] | ((size_t)in[i + 1] << 8);
      i += 2;
    } else {
      if (n - i < 4) corrupt("snappy copy");
      l = (size_t)(tag >> 2) + 1;
      off = (size_t)in[i] | ((size_t)in[i + 1] << 8) | ((size_t)in[i + 2] << 16) | ((size_t)in[i + 3] << 24);
      i += 4;
    }
    if (off == 0 || off > op || l > len - op) corrupt("snappy back-reference");
    for (size_t k = 0; k < l; k++) o[op + k] = o[op + k - off];
    op += l;
  }
  if (op != len) corrupt("snappy output short");
}

// A page of a column chunk: its header and where its (uncompressed) body lies.
struct PageRef {
  PageHeader h;
  size_t body, end;
};

// How a data page's non-NULL values are stored.
enum PageKind : int32_t {
  PK_PLAIN = 0,  // PLAIN bytes at values_off (strings: host-assigned codes from plain_base)
  PK_DICT = 1,   // dictionary indices, expanded into idx
  PK_BITS = 2,   // RLE booleans, expanded into idx
  PK_SPLIT = 3,  // BYTE_STREAM_SPLIT: byte k of value j at values_off + k * stride + j
  PK_DELTA = 4   // DELTA_BINARY_PACKED, decoded into the delta array
};

// Page table entry of a column chunk (device): rows [first_row, first_row + rows).
struct PageDesc {
  int64_t first_row;
  int64_t values_off;  // PLAIN / BYTE_STREAM_SPLIT values (bytes into the chunk buffer)
  int64_t plain_base;  // strings: index of the page's first value in the host-assigned codes
  int64_t stride;      // BYTE_STREAM_SPLIT: values in the page's stream
  int32_t rows;
  int32_t kind;
};

// A piece of a DELTA_BINARY_PACKED stream (device): up to 64 numbers of one
// miniblock, or (bw < 0) the page's first value. Value i of a piece is the
// page's first value + the deltas of the page's earlier pieces + its own first
// i + 1 deltas, each min_delta + an unpacked number, wrapping at 64 bits (the
// low 32 bits are the INT32 result).
struct DeltaRec {
  int64_t out;         // output position of the piece's first value
  int64_t bit;         // bit offset of its first packed number in the chunk
  uint64_t min_delta;
  uint64_t first;      // the page's first value
  int64_t head;        // index of the page's first-value record
  int32_t count;
  int32_t bw;          // 0..64; -1: the page's first value
};

// A run of the RLE / bit-packed hybrid encoding, placed at out_start.
struct Run {
  int64_t out_start;
  int64_t count;
  int64_t byte_off;  // bit-packed values (bytes into the chunk buffer)
  uint32_t value;    // RLE value
  int32_t bw;
  int32_t packed;
  int32_t _pad;
};

// Parse hybrid runs in [pos, end) holding up to `want` values, appended with
// output positions out_base + j.
static void hybrid_runs(const uint8_t* chunk, size_t pos, size_t end, int bw, int64_t want, int64_t out_base,
                        std::vector<Run>& runs) {
  TReader r{chunk, end, pos};
  int64_t got = 0;
  const int vbytes = (bw + 7) / 8;
  while (got < want && r.pos < end) {
    const uint64_t h = r.varint();
    Run run{};
    run.out_start = out_base + got;
    run.bw = bw;
    if (h & 1) {  // bit-packed: (h >> 1) groups of 8 values
      const uint64_t groups = h >> 1;
      const uint64_t bytes = groups * (uint64_t)bw;
      if (bytes > end - r.pos) corrupt("bit-packed run past the page");
      run.packed = 1;
      run.byte_off = (int64_t)r.pos;
      run.count = (int64_t)std::min<uint64_t>(groups * 8, (uint64_t)(want - got));
      r.pos += (size_t)bytes;
    } else {
      run.count = (int64_t)std::min<uint64_t>(h >> 1, (uint64_t)(want - got));
      uint32_t v = 0;
      for (int b = 0; b < vbytes; b++) v |= (uint32_t)r.byte() << (8 * b);
      run.value = v;
    }
    if (run.count > 0) runs.push_back(run);
    got += run.count;
  }
}

static uint64_t host_unpack(const uint8_t* chunk, size_t end, int64_t bit, int bw) {
  if (bw == 0) return 0;
  const size_t b0 = (size_t)(bit >> 3);
  const int sh = (int)(bit & 7);
  unsigned __int128 x = 0;
  for (int k = 0; k < 9; k++)
    if (b0 + (size_t)k < end) x |= (unsigned __int128)chunk[b0 + (size_t)k] << (8 * k);
  const uint64_t v = (uint64_t)(x >> sh);
  return bw == 64 ? v : v & ((1ull << bw) - 1);
}

// DELTA_BINARY_PACKED, from the format's published description: a header
// <block size> <miniblocks per block> <total count> <first value (zigzag)>,
// then blocks of <min delta (zigzag)> <a bit width byte per miniblock>
// <miniblocks, bit-packed LSB first>; each value is the previous one + min
// delta + an unpacked number, wrapping at the type's width. Miniblocks the
// last block does not need take no bytes. Walks the stream in [pos, end),
// which may hold at most max_total values: first(v) for the first value, then
// mini(bit offset, count, bit width, min delta) for each miniblock used.
// Returns the position after the last miniblock read.
template <class First, class Mini>
static size_t delta_walk(const uint8_t* chunk, size_t pos, size_t end, int bits, int64_t max_total, int64_t* total_out,
                         First first, Mini mini) {
  TReader r{chunk, end, pos};
  const uint64_t block = r.varint(), nmb = r.varint(), total = r.varint();
  const int64_t v0 = r.zz();
  if (block == 0 || block > (1u << 30) || block % 128 || nmb == 0 || block % nmb || (block / nmb) % 32)
    corrupt("delta header");
  if (total > (uint64_t)max_total) corrupt("delta value count");
  *total_out = (int64_t)total;
  if (total == 0) return r.pos;
  first((uint64_t)v0);
  const uint64_t per = block / nmb;
  uint64_t left = total - 1;
  while (left > 0) {
    const uint64_t md = (uint64_t)r.zz();
    if (nmb > end - r.pos) corrupt("delta bit widths past the page");
    const size_t widths = r.pos;
    r.pos += (size_t)nmb;
    for (uint64_t i = 0; i < nmb && left > 0; i++) {
      const int bw = chunk[widths + i];
      if (bw > bits) corrupt("delta bit width");
      const uint64_t bytes = per * (uint64_t)bw / 8;
      if (bytes > end - r.pos) corrupt("delta miniblock past the page");
      const uint64_t n = std::min(per, left);
      mini((int64_t)r.pos * 8, (int64_t)n, bw, md);
      r.pos += (size_t)bytes;
      left -= n;
    }
  }
  return r.pos;
}

// A DELTA_BINARY_PACKED stream of 32-bit numbers (byte array lengths), decoded
// on the host as signed values.
static size_t delta_host32(const uint8_t* chunk, size_t pos, size_t end, int64_t max_total, std::vector<int64_t>& out) {
  out.clear();
  uint64_t v = 0;
  int64_t total = 0;
  return delta_walk(chunk, pos, end, 32, max_total, &total,
                    [&](uint64_t f) {
                      v = f;
                      out.push_back((int32_t)(uint32_t)v);
                    },
                    [&](int64_t bit, int64_t n, int bw, uint64_t md) {
                      for (int64_t q = 0; q < n; q++) {
                        v += md + host_unpack(chunk, end, bit + q * bw, bw);
                        out.push_back((int32_t)(uint32_t)v);
                      }
                    });
}

// DELTA_LENGTH_BYTE_ARRAY: the lengths as DELTA_BINARY_PACKED, then the bytes
// back to back. DELTA_BYTE_ARRAY (prefixed): prefix lengths as
// DELTA_BINARY_PACKED, then the suffixes as DELTA_LENGTH_BYTE_ARRAY; value k
// is the first prefix[k] bytes of value k - 1, then suffix k. Exactly `want`
// values, passed to emit(view, the view stays valid while the chunk does).
template <class F>
static void delta_byte_arrays(const uint8_t* chunk, size_t pos, size_t end, bool prefixed, int64_t want, F emit) {
  std::vector<int64_t> pre, lens;
  if (prefixed) pos = delta_host32(chunk, pos, end, want, pre);
  pos = delta_host32(chunk, pos, end, want, lens);
  if ((int64_t)lens.size() != want || (prefixed && pre.size() != lens.size())) corrupt("delta value count");
  std::string prev;
  for (size_t k = 0; k < lens.size(); k++) {
    const int64_t l = lens[k];
    if (l < 0 || (uint64_t)l > end - pos) corrupt("delta length");
    const char* v = reinterpret_cast<const char*>(chunk) + pos;
    if (prefixed) {
      if (pre[k] < 0 || (size_t)pre[k] > prev.size()) corrupt("delta prefix");
      prev.resize((size_t)pre[k]);
      prev.append(v, (size_t)l);
      emit(std::string_view(prev), false);
    } else {
      emit(std::string_view(v, (size_t)l), true);
    }
    pos += (size_t)l;
  }
}

// ---------------------------------------------------------------- kernels
__device__ __forceinline__ uint32_t packed_value(const uint8_t* chunk, int64_t len, int64_t off, int64_t j, int bw) {
  if (bw == 0) return 0;
  const int64_t bit = j * bw;
  const int64_t b0 = off + (bit >> 3);
  uint64_t x = 0;
  for (int k = 0; k < 5; k++)
    if (b0 + k < len) x |= (uint64_t)chunk[b0 + k] << (8 * k);
  return (uint32_t)((x >> (bit & 7)) & ((bw >= 32) ? 0xFFFFFFFFull : ((1ull << bw) - 1)));
}

// One thread per decoded value: its run by binary search over the run starts.
__global__ void expand_runs_kernel(const Run* runs, const int64_t* cum, int nruns, int64_t total, const uint8_t* chunk,
                                   int64_t len, int max_def, uint8_t* valid, uint32_t* flag, uint32_t* idx) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    int lo = 0, hi = nruns - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (cum[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const Run& R = runs[lo];
    const int64_t j = t - cum[lo];
    const uint32_t v = R.packed ? packed_value(chunk, len, R.byte_off, j, R.bw) : R.value;
    const int64_t o = R.out_start + j;
    if (idx) {
      idx[o] = v;
    } else {
      const uint8_t ok = v == (uint32_t)max_def;
      valid[o] = ok;
      flag[o] = ok;
    }
  }
}

__device__ __forceinline__ uint64_t unpack64(const uint8_t* chunk, int64_t len, int64_t bit, int bw) {
  if (bw == 0) return 0;
  const int64_t b0 = bit >> 3;
  const int sh = (int)(bit & 7);
  uint64_t lo = 0;
  for (int k = 0; k < 8; k++)
    if (b0 + k < len) lo |= (uint64_t)chunk[b0 + k] << (8 * k);
  uint64_t v = lo >> sh;
  if (sh + bw > 64 && b0 + 8 < len) v |= (uint64_t)chunk[b0 + 8] << (64 - sh);  // sh >= 1 here
  return bw == 64 ? v : v & ((1ull << bw) - 1);
}

// The sum of each piece's deltas (0 for a page's first value).
__global__ void delta_sum_kernel(const DeltaRec* recs, int64_t n, const uint8_t* chunk, int64_t len, uint64_t* sums) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += stride) {
    const DeltaRec R = recs[m];
    uint64_t sum = 0;
    if (R.bw >= 0) {
      sum = (uint64_t)R.count * R.min_delta;
      for (int i = 0; i < R.count; i++) sum += unpack64(chunk, len, R.bit + (int64_t)i * R.bw, R.bw);
    }
    sums[m] = sum;
  }
}

// Exclusive prefix sums of n 64-bit numbers, wrapping; one workgroup of 1024
// threads walks the input 1024 numbers at a time.
__global__ __launch_bounds__(1024) void scan_u64_one_block(const uint64_t* in, int64_t n, uint64_t* out) {
  __shared__ uint64_t wsum[16];
  const int t = (int)threadIdx.x, lane = t & 63, w = t >> 6;
  uint64_t carry = 0;
  for (int64_t base = 0; base < n; base += 1024) {
    const int64_t i = base + t;
    const uint64_t x = i < n ? in[i] : 0;
    uint64_t v = x;
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t y = __shfl_up(v, d, 64);
      if (lane >= d) v += y;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    uint64_t before = carry, total = 0;
    for (int k = 0; k < 16; k++) {
      if (k < w) before += wsum[k];
      total += wsum[k];
    }
    if (i < n) out[i] = before + v - x;
    carry += total;
    __syncthreads();
  }
}

// Each piece's values into out (one thread per piece, at most 64 values).
__global__ void delta_decode_kernel(const DeltaRec* recs, int64_t n, const uint8_t* chunk, int64_t len,
                                    const uint64_t* pre, uint64_t* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += stride) {
    const DeltaRec R = recs[m];
    if (R.bw < 0) {
      out[R.out] = R.first;
      continue;
    }
    uint64_t v = R.first + (pre[m] - pre[R.head]);
    for (int i = 0; i < R.count; i++) {
      v += R.min_delta + unpack64(chunk, len, R.bit + (int64_t)i * R.bw, R.bw);
      out[R.out + i] = v;
    }
  }
}

struct Gather {
  int conv;
  int width;           // physical value bytes (decimal bytes: the FLBA length; 16 for BYTE_ARRAY)
  double div;          // DECIMAL: 10^scale
  const uint8_t* chunk;
  int64_t len;
  const uint8_t* valid;  // null: REQUIRED
  const uint32_t* rank;  // exclusive scan of valid
  const uint32_t* idx;
  const uint64_t* delta;  // DELTA_BINARY_PACKED values, at the positions of idx
  const PageDesc* pages;
  int npages;
  int64_t dict_off;    // dictionary values (bytes into the chunk)
  int64_t dict_n;
  const uint32_t* plain_codes;
  int64_t plain_n;
  const uint32_t* remap;
  int64_t nrows;
  void* out;
};

__device__ __forceinline__ uint64_t f64_bits(double d) {
  uint64_t b;
  __builtin_memcpy(&b, &d, 8);
  return b;
}

// An integer of the column's physical type (INT32 in the low 32 bits) in the
// column's conversion.
__device__ __forceinline__ uint64_t convert_int(const Gather& G, uint64_t x) {
  switch (G.conv) {
    case C_INT32: case C_DATE: return (uint32_t)x;
    case C_DEC_I32: return f64_bits((double)(int32_t)(uint32_t)x / G.div);
    case C_DEC_I64: return f64_bits((double)(int64_t)x / G.div);
    default: return x;
  }
}

__global__ void gather_kernel(Gather G) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < G.nrows; r += stride) {
    uint64_t bits = 0;
    uint32_t code = 0;
    if (!G.valid || G.valid[r]) {
      int lo = 0, hi = G.npages - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (G.pages[mid].first_row <= r) lo = mid;
        else hi = mid - 1;
      }
      const PageDesc& P = G.pages[lo];
      const int64_t j = G.rank ? (int64_t)(G.rank[r] - G.rank[P.first_row]) : r - P.first_row;
      if (G.conv == C_UTF8) {
        if (P.kind == PK_DICT) {
          const uint32_t k = G.idx[P.first_row + j];
          code = k < (uint64_t)G.dict_n ? G.remap[k] : 0u;
        } else {
          code = P.plain_base + j < G.plain_n ? G.plain_codes[P.plain_base + j] : 0u;
        }
      } else if (P.kind == PK_BITS) {
        bits = G.idx[P.first_row + j] & 1u;
      } else if (P.kind == PK_DELTA) {
        bits = convert_int(G, G.delta[P.first_row + j]);
      } else if (G.conv == C_BOOL) {
        const int64_t bit = P.values_off * 8 + j;
        bits = (bit >> 3) < G.len ? (G.chunk[bit >> 3] >> (bit & 7)) & 1 : 0;
      } else {
        // the value's bytes: base + k * step for k < width
        int64_t base = -1, step = 1;
        if (P.kind == PK_DICT) {
          const uint32_t k = G.idx[P.first_row + j];
          if (k < (uint64_t)G.dict_n) base = G.dict_off + (int64_t)k * G.width;
        } else if (P.kind == PK_SPLIT) {
          if (j < P.stride) base = P.values_off + j;
          step = P.stride;
        } else {
          base = P.values_off + j * G.width;
        }
        const int nb = G.conv == C_DEC_BYTES ? (G.width < 16 ? G.width : 16) : G.width;
        if (base >= 0 && nb > 0 && base + (int64_t)(nb - 1) * step < G.len) {
          const uint8_t* v = G.chunk + base;
          if (G.conv == C_DEC_BYTES) {  // be_i128 of the first 16 bytes at most
            unsigned __int128 x = (v[0] & 0x80) ? ~(unsigned __int128)0 : 0;
            for (int k = 0; k < nb; k++) x = (x << 8) | v[(int64_t)k * step];
            bits = f64_bits((double)(__int128)x / G.div);
          } else {
            uint64_t x = 0;
            for (int k = 0; k < nb; k++) x |= (uint64_t)v[(int64_t)k * step] << (8 * k);
            if (G.conv == C_FLOAT) {
              const uint32_t u = (uint32_t)x;
              float f;
              __builtin_memcpy(&f, &u, 4);
              bits = f64_bits((double)f);
            } else {
              bits = convert_int(G, x);
            }
          }
        }
      }
    }
    switch (G.conv) {
      case C_UTF8: ((uint32_t*)G.out)[r] = code; break;
      case C_BOOL: ((uint8_t*)G.out)[r] = (uint8_t)bits; break;
      case C_INT32: case C_DATE: ((uint32_t*)G.out)[r] = (uint32_t)bits; break;
      default: ((uint64_t*)G.out)[r] = bits; break;
    }
  }
}

__global__ void u8_to_u32_kernel(const uint8_t* in, int64_t n, uint32_t* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

static unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

}  // namespace pq
}  // namespace mlg

using namespace mlg;
using namespace mlg::pq;

struct ml_parquet {
  ml_ctx* ctx = nullptr;
  const uint8_t* file = nullptr;  // the caller's bytes
  int64_t len = 0;
  std::vector<SchemaElem> leaves;
  std::vector<int> conv, max_def, ml_type;
  std::vector<RowGroupMeta> groups;
};

namespace {

template <class F>
ml_status run_guarded(ml_ctx* ctx, F f) {
  try {
    f();
    if (ctx) ctx->last_error.clear();
    return ML_OK;
  } catch (const Error& e) {
    if (ctx) ctx->last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->last_error = "host allocation failed";
    return ML_ERR_RESOURCES_EXHAUSTED;
  } catch (const std::exception& e) {
    if (ctx) ctx->last_error = e.what();
    return ML_ERR_INTERNAL;
  }
}

// Well-formed UTF-8 (no overlong forms, surrogates or code points past U+10FFFF).
bool valid_utf8(const std::string& b) {
  size_t k = 0;
  const size_t n = b.size();
  while (k < n) {
    const uint8_t c0 = (uint8_t)b[k];
    if (c0 < 0x80) {
      k++;
      continue;
    }
    int extra;
    uint8_t lo = 0x80, hi = 0xBF;
    if (c0 >= 0xC2 && c0 <= 0xDF) extra = 1;
    else if (c0 >= 0xE0 && c0 <= 0xEF) {
      extra = 2;
      if (c0 == 0xE0) lo = 0xA0;
      if (c0 == 0xED) hi = 0x9F;
    } else if (c0 >= 0xF0 && c0 <= 0xF4) {
      extra = 3;
      if (c0 == 0xF0) lo = 0x90;
      if (c0 == 0xF4) hi = 0x8F;
    } else {
      return false;
    }
    if (k + (size_t)extra >= n) return false;
    const uint8_t c1 = (uint8_t)b[k + 1];
    if (c1 < lo || c1 > hi) return false;
    for (int e = 2; e <= extra; e++)
      if (((uint8_t)b[k + (size_t)e] >> 6) != 2) return false;
    k += (size_t)extra + 1;
  }
  return true;
}

// Conversion::for_column (parquet_reader.rs:55-80)
int conversion_of(const SchemaElem& e) {
  switch (e.type) {
    case T_BOOLEAN: return C_BOOL;
    case T_INT32: return e.converted == CT_DATE ? C_DATE : (e.converted == CT_DECIMAL ? C_DEC_I32 : C_INT32);
    case T_INT64: return e.converted == CT_DECIMAL ? C_DEC_I64 : C_INT64;
    case T_FLOAT: return C_FLOAT;
    case T_DOUBLE: return C_DOUBLE;
    case T_BYTE_ARRAY: return e.converted == CT_DECIMAL ? C_DEC_BYTES : C_UTF8;
    case T_FLBA:
      if (e.converted == CT_DECIMAL) return C_DEC_BYTES;
      break;
    default: break;
  }
  fail(ML_ERR_UNSUPPORTED, "parquet column '" + e.name + "' has an unsupported type");
}

int ml_type_of(int conv) {
  switch (conv) {
    case C_BOOL: return ML_TYPE_BOOLEAN;
    case C_INT32: return ML_TYPE_INT32;
    case C_DATE: return ML_TYPE_DATE;
    case C_INT64: return ML_TYPE_INT64;
    case C_UTF8: return ML_TYPE_UTF8;
    default: return ML_TYPE_FLOAT64;
  }
}

// 10f64.powi(s): square-and-multiply, as compiler-rt's __powidf2
double pow10i(int s) {
  const bool recip = s < 0;
  int b = recip ? -s : s;
  double a = 10.0, r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

}  // namespace

extern "C" {

ml_status ml_parquet_open(ml_ctx* ctx, const void* data, int64_t len, ml_parquet** out) {
  if (out) *out = nullptr;
  return run_guarded(ctx, [&] {
    if (!out || (!data && len > 0)) fail(ML_ERR_INTERNAL, "null argument");
    const uint8_t* f = static_cast<const uint8_t*>(data);
    if (len < 12 || std::memcmp(f, "PAR1", 4) != 0 || std::memcmp(f + len - 4, "PAR1", 4) != 0)
      fail(ML_ERR_PARQUET, "not a parquet file");
    uint32_t flen;
    std::memcpy(&flen, f + len - 8, 4);
    if ((int64_t)flen > len - 12) fail(ML_ERR_PARQUET, "footer length past the file");
    std::unique_ptr<ml_parquet> P(new ml_parquet());
    P->ctx = ctx;
    P->file = f;
    P->len = len;
    TReader r{f + len - 8 - flen, flen, 0};
    std::vector<SchemaElem> schema;
    int last = 0, id, t;
    while (r.field(&last, &id, &t)) {
      if (id == 2 && t == 9) {
        int et;
        uint64_t sz;
        r.list(&et, &sz);
        for (uint64_t i = 0; i < sz; i++) {
          SchemaElem e;
          int l2 = 0, id2, t2;
          while (r.field(&l2, &id2, &t2)) {
            switch (id2) {
              case 1: e.type = (int)r.integer(t2); break;
              case 2: e.type_length = (int)r.integer(t2); break;
              case 3: e.repetition = (int)r.integer(t2); break;
              case 4: e.name = std::string(r.binary()); break;
              case 5: e.num_children = (int)r.integer(t2); break;
              case 6: e.converted = (int)r.integer(t2); break;
              case 7: e.scale = (int)r.integer(t2); break;
              default: r.skip(t2);
            }
          }
          schema.push_back(std::move(e));
        }
      } else if (id == 4 && t == 9) {
        int et;
        uint64_t sz;
        r.list(&et, &sz);
        for (uint64_t i = 0; i < sz; i++) {
          RowGroupMeta g;
          parse_row_group(r, g);
          P->groups.push_back(std::move(g));
        }
      } else {
        r.skip(t);
      }
    }
    if (schema.empty()) corrupt("no schema");
    // a flat schema: the root, then one leaf per column (Table::open, table.rs:75-84)
    if ((size_t)schema[0].num_children + 1 != schema.size()) fail(ML_ERR_UNSUPPORTED, "nested parquet schema");
    for (size_t i = 1; i < schema.size(); i++) {
      const SchemaElem& e = schema[i];
      if (e.num_children > 0 || e.repetition == 2) fail(ML_ERR_UNSUPPORTED, "nested parquet column '" + e.name + "'");
      const int cv = conversion_of(e);
      if (cv == C_DEC_BYTES && e.type == T_FLBA && e.type_length < 0) corrupt("fixed-length byte array length");
      P->leaves.push_back(e);
      P->conv.push_back(cv);
      P->max_def.push_back(e.repetition == 1 ? 1 : 0);
      P->ml_type.push_back(ml_type_of(cv));
    }
    for (auto& g : P->groups)
      if (g.cols.size() != P->leaves.size()) corrupt("row group column count");
    *out = P.release();
  });
}

int32_t ml_parquet_num_columns(const ml_parquet* p) { return (int32_t)p->leaves.size(); }
int32_t ml_parquet_column_type(const ml_parquet* p, int32_t c) {
  return c >= 0 && c < (int32_t)p->ml_type.size() ? p->ml_type[(size_t)c] : -1;
}
const char* ml_parquet_column_name(const ml_parquet* p, int32_t c) {
  return c >= 0 && c < (int32_t)p->leaves.size() ? p->leaves[(size_t)c].name.c_str() : nullptr;
}
int32_t ml_parquet_column_nullable(const ml_parquet* p, int32_t c) {
  return c >= 0 && c < (int32_t)p->max_def.size() ? p->max_def[(size_t)c] : 0;
}
int32_t ml_parquet_num_row_groups(const ml_parquet* p) { return (int32_t)p->groups.size(); }
int64_t ml_parquet_row_group_rows(const ml_parquet* p, int32_t rg) {
  return rg >= 0 && rg < (int32_t)p->groups.size() ? p->groups[(size_t)rg].num_rows : -1;
}

// convert_stats (parquet_reader.rs:290-358): only exact bounds; no bounds for
// booleans and byte-array decimals; strings only when valid UTF-8.
ml_status ml_parquet_column_stats(const ml_parquet* p, int32_t rg, int32_t c, ml_column_stats* out) {
  return run_guarded(p->ctx, [&] {
    if (rg < 0 || rg >= (int32_t)p->groups.size() || c < 0 || c >= (int32_t)p->leaves.size())
      fail(ML_ERR_PLAN, "row group or column out of range");
    *out = ml_column_stats{};
    out->type = p->ml_type[(size_t)c];
    const ChunkMeta& m = p->groups[(size_t)rg].cols[(size_t)c];
    if (!m.has_stats) return;
    const Stats& s = m.stats;
    out->has_null_count = s.has_null_count;
    out->null_count = s.null_count;
    const int cv = p->conv[(size_t)c];
    auto set = [&](const std::string& b, bool has, bool exact, int32_t* has_out, int64_t* i, double* f,
                   const char** str, int64_t* slen) {
      if (!has || !exact) return;
      switch (cv) {
        case C_INT32: case C_DATE: case C_DEC_I32: {
          if (b.size() != 4 || m.type != T_INT32) return;
          int32_t x;
          std::memcpy(&x, b.data(), 4);
          if (cv == C_DEC_I32) *f = (double)x / pow10i(p->leaves[(size_t)c].scale);
          else *i = x;
          break;
        }
        case C_INT64: case C_DEC_I64: {
          if (b.size() != 8 || m.type != T_INT64) return;
          int64_t x;
          std::memcpy(&x, b.data(), 8);
          if (cv == C_DEC_I64) *f = (double)x / pow10i(p->leaves[(size_t)c].scale);
          else *i = x;
          break;
        }
        case C_DOUBLE: {
          if (b.size() != 8) return;
          std::memcpy(f, b.data(), 8);
          break;
        }
        case C_FLOAT: {
          if (b.size() != 4) return;
          float x;
          std::memcpy(&x, b.data(), 4);
          *f = x;
          break;
        }
        case C_UTF8: {
          if (!valid_utf8(b)) return;  // std::str::from_utf8
          *str = b.data();
          *slen = (int64_t)b.size();
          break;
        }
        default: return;
      }
      *has_out = 1;
    };
    set(s.min, s.has_min, s.min_exact, &out->has_min, &out->min_i64, &out->min_f64, &out->min_str, &out->min_len);
    set(s.max, s.has_max, s.max_exact, &out->has_max, &out->max_i64, &out->max_f64, &out->max_str, &out->max_len);
  });
}

ml_status ml_parquet_read_row_group(ml_parquet* p, int32_t rg, const int32_t* projection, int32_t nproj,
                                    ml_result** out) {
  if (out) *out = nullptr;
  ml_ctx* ctx = p->ctx;
  return run_guarded(ctx, [&] {
    if (rg < 0 || rg >= (int32_t)p->groups.size()) fail(ML_ERR_PLAN, "row group out of range");
    const RowGroupMeta& G = p->groups[(size_t)rg];
    const int64_t N = G.num_rows;
    hipStream_t s = ctx->stream;
    std::unique_ptr<ml_result> res(new ml_result());
    res->ctx = ctx;
    res->device = ctx->device;
    res->num_rows = N;
    std::vector<Buffer> keep;  // scratch that must outlive the launches (freed after the final sync)
    std::deque<std::vector<uint8_t>> host_keep;
    for (int32_t q = 0; q < nproj; q++) {
      const int c = projection[q];
      if (c < 0 || c >= (int)p->leaves.size()) fail(ML_ERR_PLAN, "projected column out of range");
      const ChunkMeta& m = G.cols[(size_t)c];
      const SchemaElem& e = p->leaves[(size_t)c];
      const int cv = p->conv[(size_t)c];
      const int maxd = p->max_def[(size_t)c];
      if (m.codec != 0 && m.codec != 1)
        fail(ML_ERR_UNSUPPORTED, "parquet column chunk codec " + std::to_string(m.codec) + " (only UNCOMPRESSED and SNAPPY)");
      const int64_t start = m.dict_page_offset >= 0 && m.dict_page_offset < m.data_page_offset ? m.dict_page_offset
                                                                                               : m.data_page_offset;
      if (start < 4 || m.total_compressed < 0 || start + m.total_compressed > p->len - 8) corrupt("column chunk range");
      const uint8_t* raw = p->file + start;
      const int64_t rawlen = m.total_compressed;
      // the pages; a SNAPPY chunk's bodies are decompressed into one host image
      // (a v2 page keeps its level bytes, stored uncompressed, in front of its values)
      std::vector<PageRef> refs;
      std::vector<uint8_t> image;
      for (size_t pos = 0; pos < (size_t)rawlen;) {
        PageRef pr;
        TReader r{raw, (size_t)rawlen, pos};
        parse_page_header(r, pr.h);
        if (pr.h.compressed < 0 || (int64_t)r.pos + pr.h.compressed > rawlen) corrupt("page past the column chunk");
        const int64_t levels = pr.h.type == PAGE_DATA_V2 ? pr.h.def_len + pr.h.rep_len : 0;
        if (pr.h.def_len < 0 || pr.h.rep_len < 0 || levels > pr.h.compressed) corrupt("v2 level lengths");
        if (m.codec == 1) {
          if (pr.h.uncompressed < levels) corrupt("page size");
          pr.body = image.size();
          image.insert(image.end(), raw + r.pos, raw + r.pos + levels);
          const uint8_t* vals = raw + r.pos + levels;
          const size_t vlen = (size_t)(pr.h.compressed - levels);
          if (pr.h.type != PAGE_DATA_V2 || pr.h.values_compressed)
            snappy_decompress(vals, vlen, (size_t)(pr.h.uncompressed - levels), image);
          else
            image.insert(image.end(), vals, vals + vlen);
          pr.end = image.size();
        } else {
          pr.body = r.pos;
          pr.end = r.pos + (size_t)pr.h.compressed;
        }
        pos = r.pos + (size_t)pr.h.compressed;
        refs.push_back(pr);
      }
      const int64_t clen = m.codec == 1 ? (int64_t)image.size() : rawlen;
      if (m.codec == 1) host_keep.push_back(std::move(image));  // uploaded below; kept until the final synchronize
      const uint8_t* chunk = m.codec == 1 ? host_keep.back().data() : raw;
      // page headers (host)
      std::vector<PageDesc> pages;
      std::vector<Run> def_runs, idx_runs;
      std::vector<DeltaRec> deltas;
      std::vector<uint32_t> plain_codes, remap;
      std::vector<uint8_t> extra;  // values decoded on the host, fixed width, placed after the chunk's bytes
      std::deque<std::string> decoded;  // strings assembled on the host (stable keys for code_of)
      auto dict = std::make_shared<HostDict>();
      std::unordered_map<std::string_view, uint32_t> code_of;
      auto intern = [&](std::string_view v, bool stable) {
        auto it = code_of.find(v);
        if (it != code_of.end()) return it->second;
        if (!stable) {
          decoded.emplace_back(v);
          v = decoded.back();
        }
        const uint32_t k = (uint32_t)dict->entries.size();
        dict->entries.emplace_back(v);
        code_of.emplace(v, k);
        return k;
      };
      // BYTE_ARRAY decimals: each value as the 16 big-endian bytes with the same be_i128
      auto put16 = [&](std::string_view v) {
        uint8_t o[16];
        const size_t L = v.size();
        if (L >= 16) {
          std::memcpy(o, v.data(), 16);
        } else {
          std::memset(o, (L && ((uint8_t)v[0] & 0x80)) ? 0xFF : 0, 16 - L);
          std::memcpy(o + 16 - L, v.data(), L);
        }
        extra.insert(extra.end(), o, o + 16);
      };
      const bool byte_dec = cv == C_DEC_BYTES && m.type == T_BYTE_ARRAY;
      // length-prefixed PLAIN byte arrays in [v, end)
      auto plain_arrays = [&](size_t v, size_t end, int64_t limit, auto&& emit) {
        for (int64_t k = 0; k < limit && v < end; k++) {
          if (end - v < 4) corrupt("byte array value");
          uint32_t l;
          std::memcpy(&l, chunk + v, 4);
          v += 4;
          if (l > end - v) corrupt("byte array past the page");
          emit(std::string_view(reinterpret_cast<const char*>(chunk) + v, l));
          v += l;
        }
      };
      // values held by the runs appended since `before`
      auto run_values = [&](size_t before) {
        int64_t got = 0;
        for (size_t k = before; k < idx_runs.size(); k++) got += idx_runs[k].count;
        return got;
      };
      int64_t dict_off = 0, dict_n = 0, row = 0;
      bool have_dict = false;
      const int width = byte_dec ? 16
                        : cv == C_DEC_BYTES ? e.type_length
                        : (m.type == T_INT64 || m.type == T_DOUBLE) ? 8 : (m.type == T_BOOLEAN ? 0 : 4);
      for (const PageRef& pr : refs) {
        if (row >= N) break;
        const PageHeader& h = pr.h;
        const size_t body = pr.body, end = pr.end;
        if (h.type == PAGE_DICTIONARY) {
          if (h.encoding != ENC_PLAIN && h.encoding != ENC_PLAIN_DICTIONARY) fail(ML_ERR_UNSUPPORTED, "dictionary page encoding");
          if (cv == C_BOOL) fail(ML_ERR_UNSUPPORTED, "dictionary-encoded booleans");
          dict_off = (int64_t)body;
          dict_n = h.num_values;
          have_dict = true;
          if (dict_n < 0) corrupt("dictionary size");
          if (cv == C_UTF8 || byte_dec) {
            dict_off = clen + (int64_t)extra.size();
            int64_t got = 0;
            plain_arrays(body, end, dict_n, [&](std::string_view v) {
              if (byte_dec) put16(v);
              else remap.push_back(intern(v, true));
              got++;
            });
            if (got != dict_n) corrupt("dictionary page");
          } else if ((int64_t)dict_n * width > (int64_t)(end - body)) {
            corrupt("dictionary page too short");
          }
        } else if (h.type == PAGE_DATA || h.type == PAGE_DATA_V2) {
          const bool v2 = h.type == PAGE_DATA_V2;
          const int64_t rows = v2 ? h.num_rows : h.num_values;
          if (v2 && h.num_values != h.num_rows) fail(ML_ERR_UNSUPPORTED, "parquet page with repeated values");
          if (rows < 0 || row + rows > N) fail(ML_ERR_PARQUET, "expected " + std::to_string(N) + " rows, decoded more");
          size_t vpos = body;
          int64_t nvalid = rows;
          if (v2) vpos += (size_t)h.rep_len;
          if (maxd > 0) {  // definition levels: RLE / bit-packed hybrid of bit width 1 (v1: after a u32 length)
            size_t dl;
            if (v2) {
              dl = (size_t)h.def_len;
            } else {
              if (h.def_encoding != ENC_RLE) fail(ML_ERR_UNSUPPORTED, "definition level encoding");
              if (end - vpos < 4) corrupt("definition levels");
              uint32_t l;
              std::memcpy(&l, chunk + vpos, 4);
              vpos += 4;
              dl = l;
            }
            if (dl > end - vpos) corrupt("definition levels past the page");
            const size_t before = def_runs.size();
            hybrid_runs(chunk, vpos, vpos + dl, 1, rows, row, def_runs);
            int64_t got = 0;
            nvalid = 0;
            for (size_t k = before; k < def_runs.size(); k++) {
              const Run& R = def_runs[k];
              got += R.count;
              if (!R.packed) {
                nvalid += R.value == (uint32_t)maxd ? R.count : 0;
              } else {
                for (int64_t q = 0; q < R.count; q++) nvalid += (chunk[R.byte_off + (q >> 3)] >> (q & 7)) & 1;
              }
            }
            if (got != rows) corrupt("definition levels short");
            vpos += dl;
          } else if (v2) {
            vpos += (size_t)h.def_len;
          }
          if (vpos > end) corrupt("page levels past the page");
          PageDesc pd{};
          pd.first_row = row;
          pd.rows = (int32_t)rows;
          const int enc = h.encoding;
          const bool ints = m.type == T_INT32 || m.type == T_INT64;
          const bool fixed = cv != C_BOOL && cv != C_UTF8 && !byte_dec;
          if (enc == ENC_PLAIN) {
            pd.values_off = (int64_t)vpos;
            if (cv == C_UTF8) {
              pd.plain_base = (int64_t)plain_codes.size();
              plain_arrays(vpos, end, INT64_MAX, [&](std::string_view v) { plain_codes.push_back(intern(v, true)); });
            } else if (byte_dec) {
              pd.values_off = clen + (int64_t)extra.size();
              plain_arrays(vpos, end, nvalid, put16);
              if (pd.values_off + nvalid * 16 != clen + (int64_t)extra.size()) corrupt("byte array values short");
            }
          } else if (enc == ENC_RLE_DICTIONARY || enc == ENC_PLAIN_DICTIONARY) {
            if (!have_dict) corrupt("dictionary indices without a dictionary page");
            pd.kind = PK_DICT;
            if (end - vpos < 1) corrupt("dictionary indices");
            const int bw = chunk[vpos];
            if (bw > 32) corrupt("dictionary index bit width");
            const size_t before = idx_runs.size();
            hybrid_runs(chunk, vpos + 1, end, bw, rows, row, idx_runs);
            if (run_values(before) < nvalid) corrupt("dictionary indices short");
          } else if (enc == ENC_RLE && cv == C_BOOL) {  // u32 length, then the hybrid of bit width 1
            pd.kind = PK_BITS;
            if (end - vpos < 4) corrupt("boolean RLE length");
            uint32_t l;
            std::memcpy(&l, chunk + vpos, 4);
            if (l > end - vpos - 4) corrupt("boolean RLE past the page");
            const size_t before = idx_runs.size();
            hybrid_runs(chunk, vpos + 4, vpos + 4 + l, 1, nvalid, row, idx_runs);
            if (run_values(before) < nvalid) corrupt("boolean values short");
          } else if (enc == ENC_DELTA_BINARY_PACKED && ints) {
            pd.kind = PK_DELTA;
            const int64_t head = (int64_t)deltas.size();
            int64_t total = 0, at = row;
            delta_walk(chunk, vpos, end, m.type == T_INT32 ? 32 : 64, nvalid, &total,
                       [&](uint64_t f) { deltas.push_back(DeltaRec{at++, 0, 0, f, head, 1, -1}); },
                       [&](int64_t bit, int64_t n, int bw, uint64_t md) {
                         for (int64_t q = 0; q < n; q += 64) {
                           const int32_t c = (int32_t)std::min<int64_t>(64, n - q);
                           deltas.push_back(DeltaRec{at, bit + q * bw, md, deltas[(size_t)head].first, head, c, bw});
                           at += c;
                         }
                       });
            if (total != nvalid) corrupt("delta value count");
          } else if ((enc == ENC_DELTA_LENGTH_BYTE_ARRAY && m.type == T_BYTE_ARRAY) ||
                     (enc == ENC_DELTA_BYTE_ARRAY && (m.type == T_BYTE_ARRAY || m.type == T_FLBA))) {
            pd.plain_base = (int64_t)plain_codes.size();
            pd.values_off = clen + (int64_t)extra.size();
            delta_byte_arrays(chunk, vpos, end, enc == ENC_DELTA_BYTE_ARRAY, nvalid, [&](std::string_view v, bool stable) {
              if (cv == C_UTF8) {
                plain_codes.push_back(intern(v, stable));
              } else if (byte_dec) {
                put16(v);
              } else {
                if ((int64_t)v.size() != width) corrupt("fixed-length value of another length");
                extra.insert(extra.end(), v.begin(), v.end());
              }
            });
          } else if (enc == ENC_BYTE_STREAM_SPLIT && fixed) {
            pd.kind = PK_SPLIT;
            pd.values_off = (int64_t)vpos;
            if (width <= 0 || (end - vpos) % (size_t)width) corrupt("byte stream split size");
            pd.stride = (int64_t)((end - vpos) / (size_t)width);
            if (pd.stride < nvalid) corrupt("byte stream split count");
          } else {
            fail(ML_ERR_UNSUPPORTED, "parquet data page encoding " + std::to_string(enc));
          }
          pages.push_back(pd);
          row += rows;
        }
      }
      if (row != N) fail(ML_ERR_PARQUET, "expected " + std::to_string(N) + " rows, decoded " + std::to_string(row));
      // device: chunk bytes, tables, runs
      const int64_t blen = clen + (int64_t)extra.size();
      Buffer dchunk(ctx->arena, (size_t)std::max<int64_t>(blen, 1));
      hip_check(hipMemcpyAsync(dchunk.ptr, chunk, (size_t)clen, hipMemcpyHostToDevice, s), "upload column chunk");
      if (!extra.empty()) {
        host_keep.push_back(std::move(extra));
        hip_check(hipMemcpyAsync(dchunk.as<uint8_t>() + clen, host_keep.back().data(), host_keep.back().size(),
                                 hipMemcpyHostToDevice, s),
                  "upload decoded values");
      }
      auto put = [&](const void* src, size_t bytes) {  // the host copy lives until the final synchronize
        Buffer b(ctx->arena, std::max<size_t>(bytes, 8));
        host_keep.emplace_back(static_cast<const uint8_t*>(src), static_cast<const uint8_t*>(src) + bytes);
        if (bytes) hip_check(hipMemcpyAsync(b.ptr, host_keep.back().data(), bytes, hipMemcpyHostToDevice, s), "upload");
        return b;
      };
      Buffer dpages = put(pages.data(), pages.size() * sizeof(PageDesc));
      ResultColumn rc;
      rc.type = p->ml_type[(size_t)c];
      rc.data = Buffer(ctx->arena, (size_t)std::max<int64_t>(N, 1) * (size_t)type_width(rc.type));
      Gather g{};
      g.conv = cv;
      g.width = width;
      g.div = (cv == C_DEC_I32 || cv == C_DEC_I64 || cv == C_DEC_BYTES) ? pow10i(e.scale) : 1.0;
      g.chunk = dchunk.as<uint8_t>();
      g.len = blen;
      g.pages = (const PageDesc*)dpages.ptr;
      g.npages = (int)pages.size();
      g.dict_off = dict_off;
      g.dict_n = cv == C_UTF8 ? (int64_t)remap.size() : dict_n;
      g.nrows = N;
      g.out = rc.data.ptr;
      auto expand = [&](const std::vector<Run>& runs, uint8_t* valid, uint32_t* flag, uint32_t* idx) {
        if (runs.empty()) return;
        std::vector<int64_t> cum(runs.size() + 1, 0);
        for (size_t k = 0; k < runs.size(); k++) cum[k + 1] = cum[k] + runs[k].count;
        Buffer druns = put(runs.data(), runs.size() * sizeof(Run));
        Buffer dcum = put(cum.data(), cum.size() * 8);
        expand_runs_kernel<<<grid_of(cum.back()), 256, 0, s>>>((const Run*)druns.ptr, dcum.as<int64_t>(),
                                                               (int)runs.size(), cum.back(), g.chunk, clen, maxd,
                                                               valid, flag, idx);
        keep.push_back(std::move(druns));
        keep.push_back(std::move(dcum));
      };
      if (maxd > 0 && N > 0) {
        rc.valid = Buffer(ctx->arena, (size_t)N);
        Buffer flag(ctx->arena, (size_t)N * 4);
        expand(def_runs, rc.valid.as<uint8_t>(), flag.as<uint32_t>(), nullptr);
        Buffer rank(ctx->arena, (size_t)(N + 1) * 4);
        Buffer sums(ctx->arena, (size_t)(scan_u32_blocks(N) + 1) * 4);
        launch_scan_u32(flag.as<uint32_t>(), N, rank.as<uint32_t>(), sums.as<uint32_t>(), s);
        g.valid = rc.valid.as<uint8_t>();
        g.rank = rank.as<uint32_t>();
        keep.push_back(std::move(flag));
        keep.push_back(std::move(rank));
        keep.push_back(std::move(sums));
      }
      Buffer didx;
      if (!idx_runs.empty()) {
        didx = Buffer(ctx->arena, (size_t)N * 4);
        expand(idx_runs, nullptr, nullptr, didx.as<uint32_t>());
        g.idx = didx.as<uint32_t>();
      }
      Buffer ddelta;
      if (!deltas.empty()) {  // DELTA_BINARY_PACKED: piece sums, their prefix sums, then the values
        const int64_t nrec = (int64_t)deltas.size();
        ddelta = Buffer(ctx->arena, (size_t)N * 8);
        Buffer drecs = put(deltas.data(), deltas.size() * sizeof(DeltaRec));
        Buffer dsums(ctx->arena, (size_t)nrec * 8), dpre(ctx->arena, (size_t)nrec * 8);
        const DeltaRec* recs = (const DeltaRec*)drecs.ptr;
        delta_sum_kernel<<<grid_of(nrec), 256, 0, s>>>(recs, nrec, g.chunk, clen, dsums.as<uint64_t>());
        scan_u64_one_block<<<1, 1024, 0, s>>>(dsums.as<uint64_t>(), nrec, dpre.as<uint64_t>());
        delta_decode_kernel<<<grid_of(nrec), 256, 0, s>>>(recs, nrec, g.chunk, clen, dpre.as<uint64_t>(),
                                                            ddelta.as<uint64_t>());
        g.delta = ddelta.as<uint64_t>();
        keep.push_back(std::move(drecs));
        keep.push_back(std::move(dsums));
        keep.push_back(std::move(dpre));
      }
      Buffer dplain = put(plain_codes.data(), plain_codes.size() * 4);
      Buffer dremap = put(remap.data(), remap.size() * 4);
      g.plain_codes = dplain.as<uint32_t>();
      g.plain_n = (int64_t)plain_codes.size();
      g.remap = dremap.as<uint32_t>();
      if (N > 0) gather_kernel<<<grid_of(N), 256, 0, s>>>(g);
      if (cv == C_UTF8) {
        dict->finish_view();
        rc.dict = dict;
      }
      keep.push_back(std::move(dchunk));
      keep.push_back(std::move(dpages));
      keep.push_back(std::move(didx));
      keep.push_back(std::move(ddelta));
      keep.push_back(std::move(dplain));
      keep.push_back(std::move(dremap));
      res->cols.push_back(std::move(rc));
    }
    res->finish_views();
    hip_check(hipStreamSynchronize(s), "synchronize");
    *out = res.release();
  });
}

// PrunePredicate::may_match (stats.rs:53-70) for each row group, keeping the
// ones whose statistics do not prove that no row matches (Table::prune,
// table.rs:165-172). Comparison as ScalarValue::compare (scalar.rs:115-131).
ml_status ml_parquet_prune(const ml_parquet* p, const ml_prune_predicate* preds, int32_t npreds, int32_t* out_rgs,
                           int32_t* out_n) {
  return run_guarded(p->ctx, [&] {
    for (int32_t k = 0; k < npreds; k++)
      if (preds[k].column < 0 || preds[k].column >= (int32_t)p->leaves.size()) fail(ML_ERR_PLAN, "prune column out of range");
    // -2: no order (a missing bound, incomparable types or NaN)
    auto cmp = [&](const ml_column_stats& st, bool use_min, const ml_prune_predicate& v) -> int {
      if (use_min ? !st.has_min : !st.has_max) return -2;
      const int bt = st.type, vt = v.type;
      auto ord = [](auto a, auto b) { return a < b ? -1 : (a > b ? 1 : 0); };
      if (bt == ML_TYPE_UTF8 || vt == ML_TYPE_UTF8) {
        if (bt != vt) return -2;
        const std::string_view a(use_min ? st.min_str : st.max_str, (size_t)(use_min ? st.min_len : st.max_len));
        const std::string_view b(v.str ? v.str : "", (size_t)v.str_len);
        const int c = a.compare(b);
        return c < 0 ? -1 : (c > 0 ? 1 : 0);
      }
      const bool bint = bt == ML_TYPE_INT32 || bt == ML_TYPE_INT64, vint = vt == ML_TYPE_INT32 || vt == ML_TYPE_INT64;
      const int64_t bi = use_min ? st.min_i64 : st.max_i64;
      if ((bt == ML_TYPE_DATE && vt == ML_TYPE_DATE) || (bint && vint)) return ord(bi, v.i64);
      const double bf = bt == ML_TYPE_FLOAT64 ? (use_min ? st.min_f64 : st.max_f64) : (double)bi;
      const double vf = vt == ML_TYPE_FLOAT64 ? v.f64 : (double)v.i64;
      if (bf != bf || vf != vf) return -2;
      return ord(bf, vf);
    };
    int32_t n = 0;
    for (int32_t rg = 0; rg < (int32_t)p->groups.size(); rg++) {
      bool keep = true;
      for (int32_t k = 0; k < npreds && keep; k++) {
        const ml_prune_predicate& v = preds[k];
        ml_column_stats st;
        if (ml_parquet_column_stats(p, rg, v.column, &st) != ML_OK) fail(ML_ERR_INTERNAL, "statistics");
        const int lo = cmp(st, true, v), hi = cmp(st, false, v);
        switch (v.op) {
          case ML_PRUNE_EQ: keep = lo != 1 && hi != -1; break;
          case ML_PRUNE_LT: keep = !(lo == 1 || lo == 0); break;
          case ML_PRUNE_LTEQ: keep = lo != 1; break;
          case ML_PRUNE_GT: keep = !(hi == -1 || hi == 0); break;
          case ML_PRUNE_GTEQ: keep = hi != -1; break;
          default: fail(ML_ERR_PLAN, "bad prune operator");
        }
      }
      if (keep) out_rgs[n++] = rg;
    }
    *out_n = n;
  });
}

void ml_parquet_free(ml_parquet* p) { delete p; }

}  // extern "C"
