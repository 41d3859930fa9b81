// segment.cpp — thinDB segment files (.dat) read on the host, their integer
// and float column blocks staged into HBM.
//
// readSegment (src/storage/segment_reader.zig:1350-1455) parses the 32-byte
// header, the 8-byte trailer and the footer (per row group: offset, length,
// row count, per-column Stats, per-column block offset and length; then the
// string-table directory, readStringTables :1459-1512).  A column block is a
// 12-byte header (compression, flags, encoding, reserved, uncompressed and
// stored sizes; format.zig:203-231) and its stored bytes; readBlockRaw /
// decompressBlockInto (:1577-1672) undo zstd or LZ4 and, for an int_streams
// block, restore its rewritten integer arrays (restoreIntStreams :1700-1716,
// src/storage/int_streams.zig:1-40, :155-210, :370-475).  The payload that
// leaves this file is exactly what thinDB's block cache holds, and
// tdbg_column_stage_encoded expands its raw/FOR/RLE body in HBM.
//
// The file image is the caller's memory (a read or an mmap of the .dat
// file); every offset and length read from it is checked against its size
// before use.
#include <dlfcn.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "runtime.h"

using tdbg::rt::fail;

namespace {

constexpr uint32_t kHeaderSize = 32;        // format.header_size
constexpr uint32_t kTrailerSize = 8;        // format.footer_trailer_size
constexpr uint32_t kBlockHeaderSize = 12;   // format.column_block_header_size
constexpr uint32_t kStatsSize = 56;         // i128 min, max, sum + u64 null_count
constexpr uint32_t kStringParts = 4;        // format.StringBlockPart values
constexpr uint32_t kEntrySize = 36;         // int_streams.entry_size
constexpr uint16_t kVersion = 17;           // format.segment_version
constexpr uint8_t kShuffledBit = 0x80;      // int_streams.shuffled_bit
enum : uint8_t { C_NONE = 0, C_ZSTD = 1, C_LZ4 = 2 };

uint16_t rd16(const uint8_t* p) { uint16_t v; std::memcpy(&v, p, 2); return v; }
uint32_t rd32(const uint8_t* p) { uint32_t v; std::memcpy(&v, p, 4); return v; }
uint64_t rd64(const uint8_t* p) { uint64_t v; std::memcpy(&v, p, 8); return v; }
__int128 rd128(const uint8_t* p) { __int128 v; std::memcpy(&v, p, 16); return v; }

struct BlockStats {
  __int128 min, max, sum;
  uint64_t null_count;
};

struct RowGroupMeta {
  uint64_t offset;
  uint32_t length, row_count;
  std::vector<BlockStats> stats;
  std::vector<uint64_t> col_offsets;
  std::vector<uint32_t> col_lengths;
};

// The codecs the reference links (compression.zig:9-25), opened from the
// system's shared libraries on first use.
struct Codecs {
  size_t (*zstd_decompress)(void*, size_t, const void*, size_t) = nullptr;
  unsigned (*zstd_is_error)(size_t) = nullptr;
  int (*lz4_decompress_safe)(const char*, char*, int, int) = nullptr;
};

const Codecs& codecs() {
  static Codecs c;
  static std::once_flag once;
  std::call_once(once, [] {
    if (void* z = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL)) {
      c.zstd_decompress = reinterpret_cast<size_t (*)(void*, size_t, const void*, size_t)>(dlsym(z, "ZSTD_decompress"));
      c.zstd_is_error = reinterpret_cast<unsigned (*)(size_t)>(dlsym(z, "ZSTD_isError"));
    }
    if (void* l = dlopen("liblz4.so.1", RTLD_NOW | RTLD_LOCAL))
      c.lz4_decompress_safe = reinterpret_cast<int (*)(const char*, char*, int, int)>(dlsym(l, "LZ4_decompress_safe"));
  });
  return c;
}

// One codec frame into exactly `out_len` bytes (compression.decompressInto /
// lz4DecompressInto: a frame that does not fill the payload is corrupt).
int32_t decompress_into(uint8_t kind, const uint8_t* in, uint64_t in_len, uint8_t* out, uint64_t out_len) {
  if (kind == C_NONE) {
    if (in_len != out_len) return fail(TDBG_ERR_INVALID, "stored block of %llu bytes for a %llu-byte payload",
                                       (unsigned long long)in_len, (unsigned long long)out_len);
    if (out_len) std::memmove(out, in, out_len);
    return TDBG_OK;
  }
  const Codecs& c = codecs();
  if (kind == C_ZSTD) {
    if (!c.zstd_decompress || !c.zstd_is_error) return fail(TDBG_ERR_UNSUPPORTED, "libzstd.so.1 not available");
    const size_t n = c.zstd_decompress(out, out_len, in, in_len);
    if (c.zstd_is_error(n) || n != out_len) return fail(TDBG_ERR_INVALID, "zstd frame does not decode to %llu bytes",
                                                        (unsigned long long)out_len);
    return TDBG_OK;
  }
  if (!c.lz4_decompress_safe) return fail(TDBG_ERR_UNSUPPORTED, "liblz4.so.1 not available");
  if (in_len > 0x7fffffffull || out_len > 0x7fffffffull) return fail(TDBG_ERR_INVALID, "LZ4 block beyond 2 GiB");
  const int n = c.lz4_decompress_safe(reinterpret_cast<const char*>(in), reinterpret_cast<char*>(out), (int)in_len,
                                      (int)out_len);
  if (n < 0 || (uint64_t)n != out_len) return fail(TDBG_ERR_INVALID, "LZ4 block does not decode to %llu bytes",
                                                   (unsigned long long)out_len);
  return TDBG_OK;
}

// ---- int_streams (src/storage/int_streams.zig) ------------------------------
enum : uint8_t { M_PLAIN = 0, M_FRAME = 1, M_DELTA = 2 };

struct StreamEntry {
  uint32_t offset, count;
  uint8_t width, method, stored_width;
  bool shuffled;
  uint64_t base, factor, first;
};

uint64_t stored_count(const StreamEntry& e) { return e.method == M_DELTA ? (uint64_t)e.count - 1 : e.count; }

bool width_ok(uint8_t w, bool allow_zero) { return (allow_zero && w == 0) || w == 1 || w == 2 || w == 3 || w == 4 || w == 8; }

// readEntry (:370-399).
int32_t read_entry(const uint8_t* in, StreamEntry* e) {
  e->offset = rd32(in);
  e->count = rd32(in + 4);
  e->width = in[8];
  e->method = in[9] & (uint8_t)~kShuffledBit;
  e->shuffled = (in[9] & kShuffledBit) != 0;
  e->stored_width = in[10];
  e->base = rd64(in + 12);
  e->factor = rd64(in + 20);
  e->first = rd64(in + 28);
  if (e->method > M_DELTA || !width_ok(e->width, false) || !width_ok(e->stored_width, true) ||
      e->stored_width > e->width || (e->method == M_PLAIN && e->stored_width != e->width) ||
      (e->method == M_DELTA && e->count == 0))
    return fail(TDBG_ERR_INVALID, "corrupt int_streams entry");
  return TDBG_OK;
}

uint64_t load_le(const uint8_t* p, uint32_t w) {
  uint64_t v = 0;
  std::memcpy(&v, p, w);
  return v;
}

// The directory heading a block's stored bytes (readDirectory :155-171):
// its entries and the length of the stored payload in the frame after it.
int32_t read_directory(const uint8_t* stored, uint64_t stored_len, uint64_t payload_len,
                       std::vector<StreamEntry>* entries, uint64_t* dir_bytes, uint64_t* frame_payload) {
  if (stored_len < 4) return fail(TDBG_ERR_INVALID, "int_streams block without a directory");
  const uint32_t n = rd32(stored);
  if (n > (stored_len - 4) / kEntrySize) return fail(TDBG_ERR_INVALID, "int_streams directory overruns its block");
  entries->resize(n);
  uint64_t shed = 0;
  for (uint32_t i = 0; i < n; i++) {
    StreamEntry& e = (*entries)[i];
    if (int32_t rc = read_entry(stored + 4 + (uint64_t)i * kEntrySize, &e)) return rc;
    shed += (uint64_t)e.count * e.width - stored_count(e) * e.stored_width;
  }
  if (shed > payload_len) return fail(TDBG_ERR_INVALID, "int_streams directory sheds more than its payload");
  *dir_bytes = 4 + (uint64_t)n * kEntrySize;
  *frame_payload = payload_len - shed;
  return TDBG_OK;
}

// Rebuilds the payload in `buf` (len bytes) from the stored payload at its
// tail, front to back in place (restore :173-210, decode :401-430).  Every
// array is stored no wider than it is restored, so value i is written only
// over bytes the loop has already read; byte-plane arrays are copied out
// first because each plane spans the whole array.
int32_t restore(const std::vector<StreamEntry>& entries, uint64_t stored_payload, uint8_t* buf, uint64_t len) {
  if (stored_payload > len) return fail(TDBG_ERR_INVALID, "corrupt int_streams payload");
  uint64_t src = len - stored_payload, dst = 0;
  std::vector<uint8_t> planes;
  for (const StreamEntry& e : entries) {
    const uint64_t out_len = (uint64_t)e.count * e.width;
    if (e.offset < dst || e.offset + out_len > len) return fail(TDBG_ERR_INVALID, "corrupt int_streams array");
    const uint64_t gap = e.offset - dst;
    if (src + gap > len) return fail(TDBG_ERR_INVALID, "corrupt int_streams gap");
    std::memmove(buf + dst, buf + src, gap);
    src += gap;
    dst = e.offset;
    const uint64_t n = stored_count(e);
    const uint64_t array_len = n * e.stored_width;
    if (src + array_len > len) return fail(TDBG_ERR_INVALID, "corrupt int_streams stored array");
    const uint8_t* in = buf + src;
    if (e.shuffled && array_len) {
      planes.assign(in, in + array_len);
      in = planes.data();
    }
    uint8_t* out = buf + dst;
    uint64_t prev = e.first, k = 0;
    if (e.method == M_DELTA) {
      std::memcpy(out, &prev, e.width);  // value_0 = first
      k = 1;
    }
    for (uint64_t i = 0; i < n; i++, k++) {
      uint64_t s = 0;
      if (e.stored_width && e.shuffled) {
        for (uint32_t b = 0; b < e.stored_width; b++) s |= (uint64_t)in[b * n + i] << (8 * b);
      } else if (e.stored_width) {
        s = load_le(in + i * e.stored_width, e.stored_width);
      }
      uint64_t v = s;
      if (e.method == M_FRAME) v = e.base + s * e.factor;
      else if (e.method == M_DELTA) v = prev = prev + (e.base + s * e.factor);
      std::memcpy(out + k * e.width, &v, e.width);  // truncated to `width` bytes, little-endian
    }
    src += array_len;
    dst += out_len;
  }
  if (src != dst) return fail(TDBG_ERR_INVALID, "int_streams payload does not end where it was stored");
  return TDBG_OK;
}

}  // namespace

struct tdbg_segment {
  const uint8_t* bytes = nullptr;
  uint64_t size = 0;
  uint64_t segment_id = 0, row_count = 0, fingerprint = 0;
  uint32_t n_cols = 0;
  std::vector<RowGroupMeta> rgs;
};

namespace {

// The block frame of (rg, col): its 12-byte header and stored bytes, inside
// the file image.
int32_t block_frame(const tdbg_segment* seg, uint32_t rg, uint32_t col, const uint8_t** frame, uint64_t* len) {
  if (!seg) return fail(TDBG_ERR_INVALID, "null segment");
  if (rg >= seg->rgs.size() || col >= seg->n_cols)
    return fail(TDBG_ERR_INVALID, "block (%u, %u) outside the segment", rg, col);
  const RowGroupMeta& m = seg->rgs[rg];
  const uint64_t off = m.col_offsets[col], l = m.col_lengths[col];
  if (l < kBlockHeaderSize || off > seg->size || l > seg->size - off)
    return fail(TDBG_ERR_INVALID, "block (%u, %u) outside the file", rg, col);
  const uint8_t* f = seg->bytes + off;
  if ((uint64_t)kBlockHeaderSize + rd32(f + 8) != l)  // checkFrame (:1553-1562)
    return fail(TDBG_ERR_INVALID, "block (%u, %u) header does not size its frame", rg, col);
  if (f[0] > C_LZ4) return fail(TDBG_ERR_INVALID, "block (%u, %u) compression %u", rg, col, f[0]);
  *frame = f;
  *len = l;
  return TDBG_OK;
}

// readBlockRaw / decompressBlockInto: the block's payload into `out`
// (exactly its header's uncompressed size).
int32_t read_payload(const uint8_t* frame, uint64_t frame_len, uint8_t* out, uint64_t out_len) {
  const uint8_t kind = frame[0];
  if (rd32(frame + 4) != out_len) return fail(TDBG_ERR_INVALID, "payload buffer of the wrong size");
  const uint8_t* body = frame + kBlockHeaderSize;
  const uint64_t body_len = frame_len - kBlockHeaderSize;
  if (!(frame[1] & 2u)) return decompress_into(kind, body, body_len, out, out_len);
  std::vector<StreamEntry> entries;
  uint64_t dir = 0, stored = 0;
  if (int32_t rc = read_directory(body, body_len, out_len, &entries, &dir, &stored)) return rc;
  if (int32_t rc = decompress_into(kind, body + dir, body_len - dir, out + (out_len - stored), stored)) return rc;
  return restore(entries, stored, out, out_len);
}

}  // namespace

extern "C" {

int32_t tdbg_segment_open(const void* bytes, uint64_t n_bytes, uint32_t n_columns, tdbg_segment** out) {
  if (!bytes || !out) return fail(TDBG_ERR_INVALID, "null argument");
  *out = nullptr;
  if (n_columns == 0 || n_columns > 65536) return fail(TDBG_ERR_INVALID, "a segment of %u columns", n_columns);
  const uint8_t* b = static_cast<const uint8_t*>(bytes);
  if (n_bytes < kHeaderSize + kTrailerSize) return fail(TDBG_ERR_INVALID, "segment too small");
  if (std::memcmp(b, "tDBS", 4) != 0) return fail(TDBG_ERR_INVALID, "bad segment magic");
  if (rd16(b + 4) != kVersion) return fail(TDBG_ERR_UNSUPPORTED, "segment version %u (reads %u)", rd16(b + 4), kVersion);
  const uint8_t* trailer = b + n_bytes - kTrailerSize;
  if (std::memcmp(trailer + 4, "tDBS", 4) != 0) return fail(TDBG_ERR_INVALID, "bad footer magic");
  const uint64_t footer_size = rd32(trailer);
  if (footer_size < kTrailerSize || footer_size > n_bytes - kHeaderSize) return fail(TDBG_ERR_INVALID, "corrupt footer size");
  const uint8_t* f = b + n_bytes - footer_size;
  const uint64_t flen = footer_size - kTrailerSize;  // up to the trailer
  if (flen < 4) return fail(TDBG_ERR_INVALID, "corrupt footer");
  const uint64_t n_rg = rd32(f);
  const uint64_t per_rg = 16 + (uint64_t)n_columns * (kStatsSize + 8 + 4);
  if (4 + n_rg * per_rg + 4 > flen) return fail(TDBG_ERR_INVALID, "footer shorter than its row groups");
  auto seg = std::make_unique<tdbg_segment>();
  seg->bytes = b;
  seg->size = n_bytes;
  seg->fingerprint = rd64(b + 8);
  seg->segment_id = rd64(b + 16);
  seg->row_count = rd64(b + 24);
  seg->n_cols = n_columns;
  seg->rgs.resize(n_rg);
  uint64_t off = 4, rows = 0;
  for (RowGroupMeta& m : seg->rgs) {
    m.offset = rd64(f + off);
    m.length = rd32(f + off + 8);
    m.row_count = rd32(f + off + 12);
    off += 16;
    m.stats.resize(n_columns);
    for (BlockStats& s : m.stats) {
      s.min = rd128(f + off);
      s.max = rd128(f + off + 16);
      s.sum = rd128(f + off + 32);
      s.null_count = rd64(f + off + 48);
      off += kStatsSize;
    }
    m.col_offsets.resize(n_columns);
    for (uint64_t& o : m.col_offsets) o = rd64(f + off), off += 8;
    m.col_lengths.resize(n_columns);
    for (uint32_t& l : m.col_lengths) l = rd32(f + off), off += 4;
    rows += m.row_count;
  }
  // the string-table directory (readStringTables): walked for its size only,
  // since string columns are not read here
  const uint64_t n_tables = rd32(f + off);
  off += 4;
  if (n_tables > n_columns) return fail(TDBG_ERR_INVALID, "corrupt string-table directory");
  for (uint64_t t = 0; t < n_tables; t++) {
    if (off + 24 > flen) return fail(TDBG_ERR_INVALID, "corrupt string-table directory");
    const uint64_t pages = rd32(f + off + 12);
    off += 24;
    const uint64_t rest = pages * 16 + n_rg * kStringParts * 4 + 12;
    if (rest > flen - off) return fail(TDBG_ERR_INVALID, "corrupt string-table directory");
    off += rest;
  }
  if (off != flen) return fail(TDBG_ERR_INVALID, "footer of %llu bytes, directory ends at %llu",
                               (unsigned long long)flen, (unsigned long long)off);
  if (rows != seg->row_count) return fail(TDBG_ERR_INVALID, "row groups hold %llu rows, the header %llu",
                                          (unsigned long long)rows, (unsigned long long)seg->row_count);
  *out = seg.release();
  return TDBG_OK;
}

void tdbg_segment_close(tdbg_segment* seg) { delete seg; }

int32_t tdbg_segment_get_info(const tdbg_segment* seg, tdbg_segment_info* out) {
  if (!seg || !out) return fail(TDBG_ERR_INVALID, "null argument");
  out->segment_id = seg->segment_id;
  out->row_count = seg->row_count;
  out->schema_fingerprint = seg->fingerprint;
  out->n_row_groups = (uint32_t)seg->rgs.size();
  out->n_columns = seg->n_cols;
  return TDBG_OK;
}

int32_t tdbg_segment_block_info(const tdbg_segment* seg, uint32_t rg, uint32_t col, tdbg_segment_block* out) {
  if (!out) return fail(TDBG_ERR_INVALID, "null argument");
  const uint8_t* fr;
  uint64_t len;
  if (int32_t rc = block_frame(seg, rg, col, &fr, &len)) return rc;
  const RowGroupMeta& m = seg->rgs[rg];
  const BlockStats& s = m.stats[col];
  std::memset(out, 0, sizeof(*out));
  out->row_count = m.row_count;
  out->payload_bytes = rd32(fr + 4);
  out->stored_bytes = rd32(fr + 8);
  out->compression = fr[0];
  out->has_nulls = fr[1] & 1u;
  out->int_streams = (fr[1] >> 1) & 1u;
  out->encoding = fr[2];
  out->null_count = s.null_count;
  out->min_lo = (uint64_t)s.min;
  out->min_hi = (uint64_t)((unsigned __int128)s.min >> 64);
  out->max_lo = (uint64_t)s.max;
  out->max_hi = (uint64_t)((unsigned __int128)s.max >> 64);
  out->sum_lo = (uint64_t)s.sum;
  out->sum_hi = (uint64_t)((unsigned __int128)s.sum >> 64);
  return TDBG_OK;
}

int32_t tdbg_segment_read_block(const tdbg_segment* seg, uint32_t rg, uint32_t col, void* out, uint64_t capacity) {
  if (!out) return fail(TDBG_ERR_INVALID, "null argument");
  const uint8_t* fr;
  uint64_t len;
  if (int32_t rc = block_frame(seg, rg, col, &fr, &len)) return rc;
  const uint64_t payload = rd32(fr + 4);
  if (capacity < payload) return fail(TDBG_ERR_INVALID, "buffer of %llu bytes for a %llu-byte payload",
                                      (unsigned long long)capacity, (unsigned long long)payload);
  return read_payload(fr, len, static_cast<uint8_t*>(out), payload);
}

int32_t tdbg_segment_stage(tdbg_context* ctx, const tdbg_segment* seg, uint64_t table_uid, uint32_t rg, uint32_t col,
                           uint32_t type) {
  if (!ctx) return fail(TDBG_ERR_INVALID, "null context");
  const uint8_t* fr;
  uint64_t len;
  if (int32_t rc = block_frame(seg, rg, col, &fr, &len)) return rc;
  const RowGroupMeta& m = seg->rgs[rg];
  if (m.row_count == 0) return fail(TDBG_ERR_INVALID, "empty row group %u", rg);
  const uint64_t payload = rd32(fr + 4);
  std::vector<uint8_t> buf;
  const uint8_t* p = fr + kBlockHeaderSize;  // an uncompressed, unrewritten block stages in place
  if (fr[0] != C_NONE || (fr[1] & 2u)) {
    buf.resize(payload);
    if (int32_t rc = read_payload(fr, len, buf.data(), payload)) return rc;
    p = buf.data();
  } else if (len - kBlockHeaderSize != payload) {
    return fail(TDBG_ERR_INVALID, "stored block of %llu bytes for a %llu-byte payload",
                (unsigned long long)(len - kBlockHeaderSize), (unsigned long long)payload);
  }
  // the footer's zone map of an integer-family block (sign-extended to i128,
  // format.zig:233-239); float extents are order keys and an all-NULL
  // block's are placeholders, so those blocks derive theirs from the values
  const BlockStats& s = m.stats[col];
  tdbg_block_stats st{};
  const __int128 lo = INT64_MIN, hi = INT64_MAX;
  if (tdbg::rt::is_int_type(type) && s.null_count < m.row_count && s.min >= lo && s.max <= hi && s.min <= s.max) {
    st.min_i = (int64_t)s.min;
    st.max_i = (int64_t)s.max;
    st.has_minmax = 1;
  }
  const tdbg_column_key key{table_uid, seg->segment_id, rg, col};
  return tdbg_column_stage_encoded(ctx, &key, type, fr[2], p, payload, m.row_count, fr[1] & 1u,
                                   st.has_minmax ? &st : nullptr);
}

}  // extern "C"
