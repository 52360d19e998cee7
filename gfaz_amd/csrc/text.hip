// P- and W-line step text on gfx950: formatting decoded traversals as the step
// lists of GFA P/W lines, and parsing those step lists back into signed node IDs.
//
// Formatting restates src/io/gfa_write_utils.cpp:238-290
// (format_path_line_numeric: "P\t" name "\t" steps "\t" overlap "\n", steps
// "12+,13-" with |id| in decimal and the orientation after it;
// format_walk_line_numeric: "W\t" six fields "\t" steps "\n", steps ">12<13").
// The caller supplies each line's bytes before and after the steps (its head and
// tail), as the reference builds them from names and fields; the GPU writes the
// steps and places the heads and tails. Lines are back to back.
//
// Parsing restates the numeric-name fast path of src/io/gfa_parser.cpp:676-790
// with resolve_node_id -> parse_uint32 (:69-91, :1210-1219): a P token is the
// text between commas, its last byte the orientation ('-' = reverse, anything
// else forward) and the rest the name; a W token starts at '<' or '>' and runs to
// the next one (bytes before the first are skipped). A name's value is its
// leading decimal digits; no leading digit or a value past 2^32 - 1 is an error
// (the reference throws), an empty name or the value 0 drops the step.
//
// Both directions work in tiles: one pass sizes every tile (bytes of its steps,
// or its kept tokens), a scan places the tiles, and a second pass writes. Output
// bytes are staged per tile in LDS and stored as whole 32-bit words where the
// destination allows.
#include "codec.hpp"
#include "prims.hpp"
#include "text.hpp"

#include <algorithm>
#include <cstring>

namespace gfz {

namespace {

constexpr uint32_t kFmtItems = 8;                    // steps per thread
constexpr uint32_t kFmtTile = kBlock * kFmtItems;    // steps per tile
constexpr uint32_t kMaxStepBytes = 12;               // ',' + 10 digits + orientation
constexpr uint32_t kParseItems = 32;                 // text bytes per thread
constexpr uint32_t kParseTile = kBlock * kParseItems;  // 8 KB
constexpr uint32_t kParseSub = 2048;                   // bytes whose tokens are listed at once
constexpr uint32_t kParseSubs = kParseTile / kParseSub;

__device__ inline uint32_t ndigits(uint32_t v) {
  uint32_t d = 1;
  uint32_t p = 10;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    d += v >= p ? 1u : 0u;
    p = p <= 0xFFFFFFFFu / 10 ? p * 10 : 0xFFFFFFFFu;
  }
  return d;
}

// Bytes of step i: P "[,]digits(+|-)", W "(<|>)digits".
__device__ inline uint32_t step_bytes(int kind, int32_t v, bool first) {
  const uint32_t d = ndigits(uabs(v));
  return kind == kTextP ? d + 1 + (first ? 0u : 1u) : d + 1;
}

// Writes step text at dst (dst[0 .. step_bytes)).
__device__ inline void put_step(int kind, int32_t v, bool first, uint8_t *dst) {
  const uint32_t a = uabs(v);
  const uint32_t d = ndigits(a);
  uint32_t o = 0;
  if (kind == kTextP) {
    if (!first)
      dst[o++] = ',';
  } else {
    dst[o++] = v < 0 ? '<' : '>';
  }
  uint32_t x = a;
  for (uint32_t k = d; k > 0; --k) {
    dst[o + k - 1] = (uint8_t)('0' + x % 10u);
    x /= 10u;
  }
  o += d;
  if (kind == kTextP)
    dst[o] = v < 0 ? '-' : '+';
}

__device__ inline bool bit_at(const uint32_t *heads, uint64_t i) {
  return (heads[i >> 5] >> (i & 31)) & 1u;
}

// Sum of the step bytes of each tile.
__global__ __launch_bounds__(kBlock) void k_fmt_tile_bytes(int kind, const int32_t *__restrict__ x,
                                                           uint64_t n,
                                                           const uint32_t *__restrict__ heads,
                                                           uint32_t *__restrict__ tile_bytes) {
  __shared__ uint32_t lds[kBlock / kWave + 1];
  const uint64_t base = (uint64_t)blockIdx.x * kFmtTile + (uint64_t)threadIdx.x * kFmtItems;
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t j = 0; j < kFmtItems; ++j) {
    const uint64_t i = base + j;
    if (i < n)
      sum += step_bytes(kind, x[i], bit_at(heads, i));
  }
  uint32_t total;
  block_exclusive_scan<kBlock>(sum, lds, total);
  if (threadIdx.x == 0)
    tile_bytes[blockIdx.x] = total;
}

// tpos[t] = byte position (in the steps-only stream) of traversal t's first step;
// tpos[n_trav] = the total. One thread per traversal, summing from its tile start.
__global__ void k_fmt_trav_pos(int kind, const int32_t *__restrict__ x, uint64_t n,
                               const uint32_t *__restrict__ heads,
                               const uint64_t *__restrict__ tstart, uint64_t n_trav,
                               const uint64_t *__restrict__ tile_off,
                               uint64_t *__restrict__ tpos) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_trav)
    return;
  const uint64_t i = tstart[t];
  const uint64_t tile = i / kFmtTile;
  uint64_t p = tile_off[tile];
  for (uint64_t k = tile * kFmtTile; k < i && k < n; ++k)
    p += step_bytes(kind, x[k], bit_at(heads, k));
  tpos[t] = p;
}

// Steps of one tile formatted into LDS, then copied to their lines: a tile's
// steps are contiguous in the output except where a traversal ends inside it
// (its tail and the next line's head come in between).
__global__ __launch_bounds__(kBlock) void k_fmt_write(
    int kind, const int32_t *__restrict__ x, uint64_t n, const uint32_t *__restrict__ heads,
    const uint64_t *__restrict__ tstart, uint64_t n_trav, const uint64_t *__restrict__ tile_off,
    const uint64_t *__restrict__ tpos, const uint64_t *__restrict__ line_off,
    const uint64_t *__restrict__ head_len, uint8_t *__restrict__ out) {
  __shared__ uint8_t buf[kFmtTile * kMaxStepBytes];
  __shared__ uint32_t lds[kBlock / kWave + 1];
  __shared__ uint32_t s_total;
  const uint64_t tile0 = (uint64_t)blockIdx.x * kFmtTile;
  const uint64_t base = tile0 + (uint64_t)threadIdx.x * kFmtItems;
  int32_t v[kFmtItems];
  uint32_t first = 0, sum = 0;
#pragma unroll
  for (uint32_t j = 0; j < kFmtItems; ++j) {
    const uint64_t i = base + j;
    v[j] = i < n ? x[i] : 0;
    const bool f = i < n && bit_at(heads, i);
    first |= f ? 1u << j : 0u;
    sum += i < n ? step_bytes(kind, v[j], f) : 0u;
  }
  uint32_t total;
  uint32_t o = block_exclusive_scan<kBlock>(sum, lds, total);
#pragma unroll
  for (uint32_t j = 0; j < kFmtItems; ++j) {
    if (base + j < n) {
      const bool f = (first >> j) & 1;
      put_step(kind, v[j], f, buf + o);
      o += step_bytes(kind, v[j], f);
    }
  }
  if (threadIdx.x == 0)
    s_total = total;
  __syncthreads();
  // Copy out, one traversal piece at a time: steps [a, b) of the tile belong to
  // traversal t, whose steps start at output byte line_off[t] + head_len[t].
  const uint64_t tile_end = tile0 + kFmtTile < n ? tile0 + kFmtTile : n;
  // First traversal with a step in the tile: last t with tstart[t] <= tile0 that
  // has steps at or after tile0 (empty traversals are skipped by the search).
  uint64_t lo = 0, hi = n_trav;  // tstart[n_trav] = n > tile0
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) / 2;
    if (tstart[mid] <= tile0)
      lo = mid;
    else
      hi = mid;
  }
  uint64_t t = lo;
  const uint64_t tile_pos = tile_off[blockIdx.x];
  uint64_t a = tile0;
  while (a < tile_end) {
    while (tstart[t + 1] <= a)
      ++t;
    const uint64_t b = tstart[t + 1] < tile_end ? tstart[t + 1] : tile_end;
    // Stream positions (bytes of all earlier steps) of the piece's ends: a piece
    // starts at the tile or at a traversal start, and ends at the tile's end or at
    // the next traversal start; LDS holds the tile from stream position tile_pos.
    const uint64_t sa = a == tile0 ? tile_pos : tpos[t];
    const uint64_t sb = b == tile_end ? tile_pos + s_total : tpos[t + 1];
    const uint32_t la = (uint32_t)(sa - tile_pos), lb = (uint32_t)(sb - tile_pos);
    // Output bytes [d, d + (lb - la)) where d = line start + head + (sa - tpos[t]).
    const uint64_t d = line_off[t] + head_len[t] + (sa - tpos[t]);
    const uint32_t len = lb - la;
    // Byte edges up to 4-byte alignment of the destination, whole words between.
    const uint32_t head_b = (uint32_t)((4 - (d & 3)) & 3) < len ? (uint32_t)((4 - (d & 3)) & 3) : len;
    const uint32_t words = (len - head_b) / 4;
    const uint32_t tail_b = len - head_b - words * 4;
    for (uint32_t k = threadIdx.x; k < head_b; k += kBlock)
      out[d + k] = buf[la + k];
    uint32_t *dw = reinterpret_cast<uint32_t *>(out + d + head_b);
    for (uint32_t w = threadIdx.x; w < words; w += kBlock) {
      const uint32_t s = la + head_b + 4 * w;
      dw[w] = (uint32_t)buf[s] | ((uint32_t)buf[s + 1] << 8) | ((uint32_t)buf[s + 2] << 16) |
              ((uint32_t)buf[s + 3] << 24);
    }
    for (uint32_t k = threadIdx.x; k < tail_b; k += kBlock) {
      const uint32_t s = la + head_b + 4 * words + k;
      out[d + head_b + 4 * words + k] = buf[s];
    }
    a = b;
  }
}

// Heads and tails of every line: one block per traversal.
__global__ void k_fmt_heads(const uint8_t *__restrict__ heads_txt,
                            const uint64_t *__restrict__ head_off,
                            const uint8_t *__restrict__ tails_txt,
                            const uint64_t *__restrict__ tail_off,
                            const uint64_t *__restrict__ line_off,
                            const uint64_t *__restrict__ tpos, uint8_t *__restrict__ out) {
  const uint64_t t = blockIdx.x;
  const uint64_t hl = head_off[t + 1] - head_off[t];
  const uint64_t tl = tail_off[t + 1] - tail_off[t];
  const uint64_t d = line_off[t];
  for (uint64_t k = threadIdx.x; k < hl; k += blockDim.x)
    out[d + k] = heads_txt[head_off[t] + k];
  const uint64_t d2 = d + hl + (tpos[t + 1] - tpos[t]);
  for (uint64_t k = threadIdx.x; k < tl; k += blockDim.x)
    out[d2 + k] = tails_txt[tail_off[t] + k];
}

// ---- parsing ---------------------------------------------------------------

// Byte-wise tests on a 32-bit word of text, four bits out (bit k = byte k).
// zero_bytes is exact (no borrow crosses bytes): 0x80 in each zero byte.
__device__ inline uint32_t zero_bytes(uint32_t x) {
  return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
}
__device__ inline uint32_t byte_bits(uint32_t z) {  // 0x80 per byte -> 4 bits
  return (((z >> 7) * 0x00204081u) >> 21) & 0xFu;
}
// Token separators of a word: ',' for P; '<' or '>' (0x3C, 0x3E) for W.
__device__ inline uint32_t sep_bits(int kind, uint32_t v) {
  return kind == kTextP ? byte_bits(zero_bytes(v ^ 0x2C2C2C2Cu))
                        : byte_bits(zero_bytes((v | 0x02020202u) ^ 0x3E3E3E3Eu));
}

// Bytes of the text as a kernel reads them: from an LDS window [w0, w0 + wn) of
// the tile being parsed where it holds them, else from global memory.
struct TextReader {
  const uint8_t *g;
  const uint8_t *win;
  uint64_t w0, wn;
  __device__ uint8_t operator()(uint64_t p) const { return p - w0 < wn ? win[p - w0] : g[p]; }
};

// Where a token's field ends: inside the tile's window [w0, w0 + wn] a bit per
// byte marks the ends of the fields with a byte in the tile (no field's end
// lies strictly inside another field; a field may end where the next begins),
// past it the field holding p is searched for.
struct FieldEnds {
  const uint32_t *bits;
  uint64_t w0, wn;
  const uint64_t *fb, *fe;
  uint64_t n;
  __device__ bool at(uint64_t q) const {
    const uint64_t o = q - w0;
    return (bits[o >> 5] >> (o & 31)) & 1u;
  }
  __device__ uint64_t search(uint64_t p) const {  // fe of the last field with fb <= p
    uint64_t lo = 0, hi = n;
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) / 2;
      if (fb[mid] <= p)
        lo = mid;
      else
        hi = mid;
    }
    return fe[lo];
  }
};

// The token starting at byte p (a token start inside a field). For P the token
// runs to the next ',' or the field's end, its last byte is the orientation and
// the bytes before it the name; for W it runs from the '<' / '>' at p to the
// next one or the field's end, the name after p. One forward pass keeps the
// state of the name's leading digit run (value, digits, overflow); for P the
// state before the last byte is the name's. Returns 0 (a dropped step: empty
// name or value 0) or 1 (a step: *id signed); sets *err to 1 (no digits) or 2
// (overflow) on a malformed name, as parse_uint32 throws
// (src/io/gfa_parser.cpp:69-91).
__device__ inline int parse_token(int kind, const TextReader &txt, const FieldEnds &fends,
                                  uint64_t p, int32_t *id, uint32_t *err) {
  const bool pk = kind == kTextP;
  // Fast path: a token ending within 12 bytes inside the window. Its 16 bytes
  // come from five LDS words (byte-aligned by alignbyte), its end is the first
  // separator or field end after p, and the name's leading digit run is summed
  // in a fixed loop (at most 11 digits: exact in 64 bits, so overflow is a
  // compare); every other token takes the byte loop below, with the same result.
  const uint64_t o = p - fends.w0;
  if (o + 16 <= fends.wn) {
    const uint32_t *w32 = reinterpret_cast<const uint32_t *>(txt.win) + (o >> 2);
    const uint32_t sh = (uint32_t)(o & 3);
    uint32_t raw[5];
#pragma unroll
    for (uint32_t k = 0; k < 5; ++k)
      raw[k] = w32[k];
    uint32_t d[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      d[k] = __builtin_amdgcn_alignbyte(raw[k + 1], raw[k], sh);
    uint32_t sep = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      sep |= sep_bits(kind, d[k]) << (4 * k);
    const uint64_t eb2 = (uint64_t)fends.bits[(o >> 5) + 1] << 32 | fends.bits[o >> 5];
    const uint32_t ends = (uint32_t)(eb2 >> (o & 31)) & 0xFFFFu;
    const uint32_t stop = (sep | ends) & ~1u;  // the token's own first byte never ends it
    const uint32_t len = stop ? (uint32_t)__builtin_ctz(stop) : 32u;
    if (len <= 12) {
      const uint32_t name0 = pk ? 0u : 1u;   // the name: bytes [name0, name0 + name_len)
      const uint32_t name_len = len - 1;
      if (!name_len)
        return 0;
      uint64_t val = 0;
      uint32_t ndig = 0;
      bool run = true;
#pragma unroll
      for (uint32_t i = 0; i < 11; ++i) {
        const uint32_t j = name0 + i;
        const uint32_t dg = ((d[j >> 2] >> (8 * (j & 3))) & 0xFFu) - '0';
        run = run && i < name_len && dg < 10u;
        val = run ? val * 10 + dg : val;
        ndig += run ? 1u : 0u;
      }
      if (val > 0xFFFFFFFFull) {
        atomicOr(err, 2u);
        return 0;
      }
      if (!ndig) {
        atomicOr(err, 1u);
        return 0;
      }
      if (val == 0)
        return 0;
      const uint32_t rb = pk ? len - 1 : 0u;  // the orientation byte
      const bool rev = ((d[rb >> 2] >> (8 * (rb & 3))) & 0xFFu) == (pk ? '-' : '<');
      *id = (int32_t)(rev ? 0u - (uint32_t)val : (uint32_t)val);
      return 1;
    }
  }
  const uint8_t c0 = txt(p);
  uint32_t val = 0, ndig = 0;
  bool ovf = false, run = true;
  uint32_t bval = 0, bdig = 0;  // P: the state before the byte just consumed
  bool bovf = false;
  uint64_t q = pk ? p : p + 1;
  uint64_t e = ~0ull;  // the field's end once the token leaves the window
  for (;; ++q) {
    if (q - fends.w0 <= fends.wn) {
      if (q > p && fends.at(q))  // at p itself: the end of an adjacent field before
        break;
    } else {
      if (e == ~0ull)
        e = fends.search(p);
      if (q >= e)
        break;
    }
    const uint8_t c = txt(q);
    if (pk ? c == ',' : (c == '<' || c == '>'))
      break;
    bval = val;
    bdig = ndig;
    bovf = ovf;
    if (run) {
      if (c >= '0' && c <= '9') {
        const uint32_t dg = c - '0';
        if (!ovf) {
          if (val > 0xFFFFFFFFu / 10 || (val == 0xFFFFFFFFu / 10 && dg > 0xFFFFFFFFu % 10))
            ovf = true;
          else
            val = val * 10 + dg;
        }
        ++ndig;
      } else {
        run = false;
      }
    }
  }
  bool rev;
  uint64_t name_len;
  if (pk) {  // token [p, q), q > p; orientation txt[q - 1], name [p, q - 1)
    rev = txt(q - 1) == '-';
    name_len = q - 1 - p;
    val = bval;
    ndig = bdig;
    ovf = bovf;
  } else {
    rev = c0 == '<';
    name_len = q - p - 1;
  }
  if (!name_len)
    return 0;
  if (ovf) {
    atomicOr(err, 2u);
    return 0;
  }
  if (!ndig) {
    atomicOr(err, 1u);
    return 0;
  }
  if (val == 0)
    return 0;
  *id = (int32_t)(rev ? 0u - val : val);  // the reference's uint32 -> NodeId conversion
  return 1;
}

// Field cursor of one thread: the field holding (or next after) its position.
struct FieldCursor {
  const uint64_t *fb, *fe;
  uint64_t n, l, b, e, nb;  // field l = [b, e); nb = the next field's begin
  // Last field in [lo, hi) beginning at or before p (lo if none does): the
  // fields that can hold a byte of a tile are few, found per tile beforehand.
  __device__ void seek(uint64_t p, uint64_t lo, uint64_t hi) {
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) / 2;
      if (fb[mid] <= p)
        lo = mid;
      else
        hi = mid;
    }
    l = lo;
    b = fb[l];
    e = fe[l];
    nb = l + 1 < n ? fb[l + 1] : ~0ull;
  }
  __device__ void advance(uint64_t p) {
    while (nb <= p) {
      ++l;
      b = nb;
      e = fe[l];
      nb = l + 1 < n ? fb[l + 1] : ~0ull;
    }
  }
};

// Stages text bytes [w0, w0 + kParseTile + kParseHalo) of the block's tile into LDS
// (16-byte loads where the text allows).
constexpr uint32_t kParseHalo = 64;
constexpr uint32_t kParseWin = kParseTile + kParseHalo;
__device__ inline uint64_t stage_window(const uint8_t *txt, uint64_t n_bytes, uint64_t w0,
                                        uint8_t *win) {
  const uint64_t wn = w0 + kParseWin < n_bytes ? kParseWin : n_bytes - w0;
  const bool aligned = ((uintptr_t)(txt + w0) & 15) == 0;
  if (aligned) {
    const uint64_t nv = wn / 16;
    const uint4 *src = reinterpret_cast<const uint4 *>(txt + w0);
    uint4 *dst = reinterpret_cast<uint4 *>(win);
    for (uint64_t i = threadIdx.x; i < nv; i += kBlock)
      dst[i] = src[i];
    for (uint64_t i = nv * 16 + threadIdx.x; i < wn; i += kBlock)
      win[i] = txt[w0 + i];
  } else {
    for (uint64_t i = threadIdx.x; i < wn; i += kBlock)
      win[i] = txt[w0 + i];
  }
  __syncthreads();
  return wn;
}

// The tile's tokens, in byte order. Classification: thread t looks at bytes
// t, t + kBlock, ... (a wave at 64 consecutive bytes per step, so its starts are
// one ballot, stored as two bitmap words), recording the field of each start;
// then, a kParseSub stretch at a time, the bitmap's set bits are listed by a
// block scan and parsed a lane per token (a token is several bytes, most bytes
// start none). A block takes an 8 KB tile so that the dependent global reads
// that open it (its first field, the field search, the window) are paid once
// per 8 KB; the token list and the parsed IDs stay at one stretch's size.
struct TokenLds {
  __align__(16) uint8_t win[kParseWin];
  uint32_t start_bits[kParseTile / 32];
  uint32_t end_bits[kParseWin / 32 + 1];  // field ends at window offsets 0 .. kParseWin
  uint16_t list[kParseSub];               // one stretch's token starts (tile offsets)
  uint32_t scan[kBlock / kWave + 1];
};

// first[t] = the first field beginning at or after tile t's first byte
// (t = 0 .. n_tiles): the fields that can hold a byte of tile t are
// first[t] - 1 .. first[t + 1] - 1.
__global__ void k_parse_first(const uint64_t *__restrict__ fb, uint64_t n_lines,
                              uint64_t n_tiles, uint64_t *__restrict__ first) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t <= n_tiles;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t at = t * kParseTile;
    uint64_t lo = 0, hi = n_lines;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) / 2;
      if (fb[mid] < at)
        lo = mid + 1;
      else
        hi = mid;
    }
    first[t] = lo;
  }
}

// Token starts of the tile: thread t classifies the 32 bytes [32 t, 32 t + 32)
// from two 16-byte LDS reads (start_bits word t). The fields meeting the chunk
// (usually one) give its in-field and field-begin masks; a P start is an in-field
// byte other than ',' that begins its field or follows a ','; a W start is an
// in-field '<' or '>'. Every field with a byte in the tile has its end marked in
// end_bits when the window holds it (by the thread holding its last byte, or by
// the tile's last thread for a field that runs past the tile).
__device__ inline void tile_classify(int kind, const uint8_t *__restrict__ txt,
                                     uint64_t n_bytes, const uint64_t *__restrict__ fb,
                                     const uint64_t *__restrict__ fe, uint64_t n_lines,
                                     const uint64_t *__restrict__ first, uint64_t w0,
                                     TokenLds &sm, TextReader &rd, FieldEnds &fends) {
  const uint64_t f0 = first[blockIdx.x], f1 = first[blockIdx.x + 1];
  for (uint32_t i = threadIdx.x; i < kParseWin / 32 + 1; i += kBlock)
    sm.end_bits[i] = 0;
  rd = TextReader{txt, sm.win, w0, stage_window(txt, n_bytes, w0, sm.win)};  // syncs
  fends = FieldEnds{sm.end_bits, w0, rd.wn, fb, fe, n_lines};
  static_assert(kParseTile == 32 * kBlock, "one 32-byte chunk per thread");
  const uint32_t r0 = threadIdx.x * 32;
  const uint64_t p0 = w0 + r0, p1 = p0 + 32;
  const bool last_chunk = threadIdx.x == kBlock - 1;
  uint32_t in = 0, beg = 0;
  if (p0 < n_bytes) {
    FieldCursor fc{fb, fe, n_lines, 0, 0, 0, 0};
    fc.seek(p0, f0 ? f0 - 1 : 0, f1 > f0 ? f1 : f0 + 1 < n_lines ? f0 + 1 : n_lines);
    for (;;) {
      if (fc.b < p1 && fc.e > p0 && fc.e > fc.b) {
        const uint32_t lo = fc.b > p0 ? (uint32_t)(fc.b - p0) : 0u;
        const uint32_t hi = fc.e < p1 ? (uint32_t)(fc.e - p0) : 32u;
        in |= (hi == 32 ? ~0u : (1u << hi) - 1u) & ~((1u << lo) - 1u);
        if (fc.b >= p0)
          beg |= 1u << lo;
        if ((fc.e <= p1 || last_chunk) && fc.e - w0 <= rd.wn) {
          const uint64_t o = fc.e - w0;
          atomicOr(&sm.end_bits[o >> 5], 1u << (o & 31));
        }
      }
      if (fc.nb >= p1)
        break;
      fc.advance(fc.nb);
    }
  }
  uint32_t word = 0;
  if (in) {
    const uint4 *w4 = reinterpret_cast<const uint4 *>(sm.win + r0);  // 16-byte aligned
    const uint4 a = w4[0], b = w4[1];
    const uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t sep = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k)
      sep |= sep_bits(kind, v[k]) << (4 * k);
    if (kind == kTextP) {
      const uint32_t prev = (p0 > 0 && rd(p0 - 1) == ',') ? 1u : 0u;
      word = in & ~sep & ((sep << 1) | prev | beg);
    } else {
      word = in & sep;
    }
  }
  sm.start_bits[threadIdx.x] = word;
  __syncthreads();
}

// Lists the token starts of stretch s of the tile (tile offsets, byte order);
// returns their count. The caller syncs before the next stretch refills the list.
__device__ inline uint32_t sub_tokens(TokenLds &sm, uint32_t s) {
  constexpr uint32_t kWords = kParseSub / 32;
  static_assert(kWords <= kBlock, "one bitmap word per thread");
  const uint32_t word = threadIdx.x < kWords ? sm.start_bits[s * kWords + threadIdx.x] : 0u;
  uint32_t total;
  uint32_t at = block_exclusive_scan<kBlock>((uint32_t)__popc(word), sm.scan, total);
  for (uint32_t bits = word; bits; bits &= bits - 1)
    sm.list[at++] = (uint16_t)(s * kParseSub + threadIdx.x * 32 + __builtin_ctz(bits));
  __syncthreads();
  return total;
}

// Kept steps per tile of text bytes, and for every field that starts inside the
// tile the kept steps of the tile before its first byte (partial[l]).
__global__ __launch_bounds__(kBlock) void k_parse_count(int kind, const uint8_t *__restrict__ txt,
                                                        uint64_t n_bytes,
                                                        const uint64_t *__restrict__ fb,
                                                        const uint64_t *__restrict__ fe,
                                                        uint64_t n_lines,
                                                        const uint64_t *__restrict__ first,
                                                        uint32_t *__restrict__ tile_count,
                                                        uint32_t *__restrict__ partial,
                                                        uint32_t *__restrict__ err) {
  __shared__ TokenLds sm;
  __shared__ uint32_t keep_bits[kParseTile / 32];  // kept-token starts, one bit per byte
  const uint64_t w0 = (uint64_t)blockIdx.x * kParseTile;
  for (uint32_t i = threadIdx.x; i < kParseTile / 32; i += kBlock)
    keep_bits[i] = 0;
  TextReader rd;
  FieldEnds fends;
  tile_classify(kind, txt, n_bytes, fb, fe, n_lines, first, w0, sm, rd, fends);  // syncs
  uint32_t c = 0;
  for (uint32_t s = 0; s < kParseSubs; ++s) {
    const uint32_t n_tok = sub_tokens(sm, s);  // syncs
    for (uint32_t i = threadIdx.x; i < n_tok; i += kBlock) {
      const uint32_t r = sm.list[i];
      int32_t id;
      if (parse_token(kind, rd, fends, w0 + r, &id, err)) {
        ++c;
        atomicOr(&keep_bits[r >> 5], 1u << (r & 31));
      }
    }
    __syncthreads();  // the next stretch refills the list
  }
  uint32_t total;
  block_exclusive_scan<kBlock>(c, sm.scan, total);  // its barriers publish keep_bits
  if (threadIdx.x == 0)
    tile_count[blockIdx.x] = total;
  // fields starting in the tile: first[t] .. first[t + 1] - 1
  for (uint64_t l = first[blockIdx.x] + threadIdx.x; l < first[blockIdx.x + 1]; l += kBlock) {
    const uint32_t off = (uint32_t)(fb[l] - w0);  // < kParseTile
    uint32_t n = 0;
    for (uint32_t w = 0; w < off / 32; ++w)
      n += (uint32_t)__popc(keep_bits[w]);
    if (off % 32)
      n += (uint32_t)__popc(keep_bits[off / 32] & ((1u << (off % 32)) - 1));
    partial[l] = n;
  }
}

// Kept steps before each field start (lpos[n_lines]: all of them): the tile's
// prefix plus the field's partial count; a field starting at the text's end past
// every tile has them all.
__global__ void k_parse_line_pos(const uint64_t *__restrict__ fb, uint64_t n_lines,
                                 uint64_t n_tiles, const uint64_t *__restrict__ tile_pos,
                                 const uint32_t *__restrict__ partial,
                                 uint64_t *__restrict__ lpos) {
  const uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l > n_lines)
    return;
  if (l == n_lines) {
    lpos[l] = tile_pos[n_tiles];
    return;
  }
  const uint64_t t = fb[l] / kParseTile;
  lpos[l] = t < n_tiles ? tile_pos[t] + partial[l] : tile_pos[n_tiles];
}

// Steps in byte order: 2 KB of the tile at a time, tokens are parsed a lane each
// (token i of the stretch by thread i % kBlock), a ballot per wave records the
// kept ones as bits in list order, one wave turns the 64 bitmap words into
// prefix counts, and each lane stores its kept ID at the tile's position plus its
// rank (consecutive lanes hold consecutive tokens, so the stores are contiguous).
constexpr int32_t kDropped = 0;  // no kept step has ID 0
__global__ __launch_bounds__(kBlock) void k_parse_write(int kind, const uint8_t *__restrict__ txt,
                                                        uint64_t n_bytes,
                                                        const uint64_t *__restrict__ fb,
                                                        const uint64_t *__restrict__ fe,
                                                        uint64_t n_lines,
                                                        const uint64_t *__restrict__ first,
                                                        const uint64_t *__restrict__ tile_pos,
                                                        int32_t *__restrict__ nodes,
                                                        uint32_t *__restrict__ err) {
  constexpr uint32_t kWords = kParseSub / 32;      // kept bits of a stretch
  constexpr uint32_t kPer = kParseSub / kBlock;    // tokens per thread and stretch
  static_assert(kWords == kWave, "one wave scans a stretch's kept words");
  __shared__ TokenLds sm;
  __shared__ uint32_t kbits[kWords];
  __shared__ uint32_t kpre[kWords + 1];
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const uint64_t w0 = (uint64_t)blockIdx.x * kParseTile;
  TextReader rd;
  FieldEnds fends;
  tile_classify(kind, txt, n_bytes, fb, fe, n_lines, first, w0, sm, rd, fends);  // syncs
  uint64_t base = tile_pos[blockIdx.x];
  for (uint32_t s = 0; s < kParseSubs; ++s) {
    const uint32_t n_tok = sub_tokens(sm, s);  // syncs
    int32_t id[kPer];
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
      const uint32_t i = k * kBlock + threadIdx.x;
      id[k] = kDropped;
      if (i < n_tok && !parse_token(kind, rd, fends, w0 + sm.list[i], &id[k], err))
        id[k] = kDropped;
      const uint64_t m = __ballot(id[k] != kDropped);
      if (lane == 0) {
        const uint32_t w = (k * kBlock + wid * kWave) / 32;
        kbits[w] = (uint32_t)m;
        kbits[w + 1] = (uint32_t)(m >> 32);
      }
    }
    __syncthreads();
    if (wid == 0) {
      const uint32_t c = (uint32_t)__popc(kbits[lane]);
      uint32_t incl = c;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d);
        if (lane >= (uint32_t)d)
          incl += y;
      }
      kpre[lane] = incl - c;
      if (lane == kWave - 1)
        kpre[kWave] = incl;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
      const uint32_t i = k * kBlock + threadIdx.x;
      if (id[k] != kDropped)
        nodes[base + kpre[i >> 5] + (uint32_t)__popc(kbits[i >> 5] & ((1u << (i & 31)) - 1u))] =
            id[k];
    }
    base += kpre[kWords];
    __syncthreads();  // the next stretch refills the list and the kept bits
  }
}

uint32_t grid1(uint64_t n, uint32_t per) {
  return (uint32_t)std::max<uint64_t>(1, ceil_div(n, (uint64_t)per));
}

}  // namespace

void format_plan(Ctx &ctx, int kind, const int32_t *d_nodes, uint64_t n,
                 const uint32_t *lengths, uint64_t n_trav, const uint64_t *head_off,
                 const uint64_t *tail_off, FormatPlan &plan) {
  if (kind != kTextP && kind != kTextW)
    throw Error(kErrInvalid, "line kind must be P or W");
  plan = FormatPlan();
  plan.kind = kind;
  plan.nodes = d_nodes;
  plan.n = n;
  plan.n_trav = n_trav;
  std::vector<uint64_t> tstart(n_trav + 1, 0), starts;
  for (uint64_t t = 0; t < n_trav; ++t) {
    if (lengths[t])
      starts.push_back(tstart[t]);
    tstart[t + 1] = tstart[t] + lengths[t];
  }
  if (tstart[n_trav] != n)
    throw Error(kErrInvalid, "sum of traversal lengths (" + std::to_string(tstart[n_trav]) +
                                 ") != number of steps (" + std::to_string(n) + ")");
  if (head_off[0] != 0 || tail_off[0] != 0)
    throw Error(kErrInvalid, "head and tail offsets must start at 0");
  for (uint64_t t = 0; t < n_trav; ++t)
    if (head_off[t + 1] < head_off[t] || tail_off[t + 1] < tail_off[t])
      throw Error(kErrInvalid, "head and tail offsets must be non-decreasing");
  Pool &pool = ctx.pool();
  plan.n_tiles = ceil_div(std::max<uint64_t>(n, 1), kFmtTile);
  const uint64_t words = ceil_div(std::max<uint64_t>(n, 1), 32);
  plan.heads = DBuf<uint32_t>(pool, words);
  plan.tstart = DBuf<uint64_t>(pool, n_trav + 1);
  plan.tile_off = DBuf<uint64_t>(pool, plan.n_tiles + 1);
  plan.tpos = DBuf<uint64_t>(pool, n_trav + 1);
  DBuf<uint64_t> d_starts(pool, std::max<size_t>(starts.size(), 1));
  DBuf<uint32_t> tile_bytes(pool, plan.n_tiles);
  fill_u32(ctx, plan.heads.get(), 0, words);
  GFZ_HIP(hipMemcpyAsync(plan.tstart.get(), tstart.data(), (n_trav + 1) * 8,
                         hipMemcpyHostToDevice, ctx.stream()));
  if (!starts.empty()) {
    GFZ_HIP(hipMemcpyAsync(d_starts.get(), starts.data(), starts.size() * 8,
                           hipMemcpyHostToDevice, ctx.stream()));
    set_head_bits(ctx, d_starts.get(), starts.size(), plan.heads.get(), n);
  }
  if (n) {
    hipLaunchKernelGGL(k_fmt_tile_bytes, dim3((uint32_t)plan.n_tiles), dim3(kBlock), 0,
                       ctx.stream(), kind, d_nodes, n, plan.heads.get(), tile_bytes.get());
    GFZ_CHECK_LAUNCH();
  }
  exclusive_scan(ctx, tile_bytes.get(), plan.tile_off.get(), n ? plan.n_tiles : 0);
  hipLaunchKernelGGL(k_fmt_trav_pos, dim3(grid1(n_trav + 1, kBlock)), dim3(kBlock), 0,
                     ctx.stream(), kind, d_nodes, n, plan.heads.get(), plan.tstart.get(), n_trav,
                     plan.tile_off.get(), plan.tpos.get());
  GFZ_CHECK_LAUNCH();
  std::vector<uint64_t> h_tpos(n_trav + 1);
  GFZ_HIP(hipMemcpyAsync(h_tpos.data(), plan.tpos.get(), (n_trav + 1) * 8, hipMemcpyDeviceToHost,
                         ctx.stream()));
  ctx.sync();  // also: the pageable tstart/starts copies are done
  plan.line_off.assign(n_trav + 1, 0);
  plan.head_len.resize(n_trav);
  for (uint64_t t = 0; t < n_trav; ++t) {
    plan.head_len[t] = head_off[t + 1] - head_off[t];
    plan.line_off[t + 1] = plan.line_off[t] + plan.head_len[t] + (h_tpos[t + 1] - h_tpos[t]) +
                           (tail_off[t + 1] - tail_off[t]);
  }
  plan.total = plan.line_off[n_trav];
}

void format_write(Ctx &ctx, const FormatPlan &plan, const uint8_t *heads_txt,
                  const uint64_t *head_off, const uint8_t *tails_txt, const uint64_t *tail_off,
                  uint8_t *d_out, uint64_t out_cap) {
  const uint64_t n_trav = plan.n_trav, n = plan.n;
  if (plan.total > out_cap)
    throw Error(kErrInvalid, "output capacity " + std::to_string(out_cap) + " < text size " +
                                 std::to_string(plan.total));
  if (!n_trav)
    return;
  Pool &pool = ctx.pool();
  DBuf<uint64_t> d_line(pool, n_trav + 1), d_hlen(pool, n_trav), d_hoff(pool, n_trav + 1),
      d_toff(pool, n_trav + 1);
  const uint64_t hbytes = head_off[n_trav], tbytes = tail_off[n_trav];
  DBuf<uint8_t> d_htxt(pool, std::max<uint64_t>(hbytes, 1)),
      d_ttxt(pool, std::max<uint64_t>(tbytes, 1));
  GFZ_HIP(hipMemcpyAsync(d_line.get(), plan.line_off.data(), (n_trav + 1) * 8,
                         hipMemcpyHostToDevice, ctx.stream()));
  GFZ_HIP(hipMemcpyAsync(d_hlen.get(), plan.head_len.data(), n_trav * 8, hipMemcpyHostToDevice,
                         ctx.stream()));
  GFZ_HIP(hipMemcpyAsync(d_hoff.get(), head_off, (n_trav + 1) * 8, hipMemcpyHostToDevice,
                         ctx.stream()));
  GFZ_HIP(hipMemcpyAsync(d_toff.get(), tail_off, (n_trav + 1) * 8, hipMemcpyHostToDevice,
                         ctx.stream()));
  if (hbytes)
    GFZ_HIP(hipMemcpyAsync(d_htxt.get(), heads_txt, hbytes, hipMemcpyHostToDevice, ctx.stream()));
  if (tbytes)
    GFZ_HIP(hipMemcpyAsync(d_ttxt.get(), tails_txt, tbytes, hipMemcpyHostToDevice, ctx.stream()));
  ctx.timed("format", [&] {
    if (n) {
      hipLaunchKernelGGL(k_fmt_write, dim3((uint32_t)plan.n_tiles), dim3(kBlock), 0, ctx.stream(),
                         plan.kind, plan.nodes, n, plan.heads.get(), plan.tstart.get(), n_trav,
                         plan.tile_off.get(), plan.tpos.get(), d_line.get(), d_hlen.get(), d_out);
      GFZ_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_fmt_heads, dim3((uint32_t)n_trav), dim3(64), 0, ctx.stream(),
                       d_htxt.get(), d_hoff.get(), d_ttxt.get(), d_toff.get(), d_line.get(),
                       plan.tpos.get(), d_out);
    GFZ_CHECK_LAUNCH();
  });
  ctx.add_bytes("format", 4.0 * n + (double)plan.total);
  ctx.sync();  // the pageable line tables go out of scope
}

void parse_plan(Ctx &ctx, int kind, const uint8_t *d_txt, uint64_t n_bytes,
                const uint64_t *field_begin, const uint64_t *field_end, uint64_t n_lines,
                uint32_t *lengths, ParsePlan &plan) {
  if (kind != kTextP && kind != kTextW)
    throw Error(kErrInvalid, "line kind must be P or W");
  for (uint64_t l = 0; l < n_lines; ++l)
    if (field_end[l] < field_begin[l] || field_end[l] > n_bytes ||
        (l + 1 < n_lines && field_begin[l + 1] < field_end[l]))
      throw Error(kErrInvalid, "step fields must be ordered, disjoint and inside the text");
  plan = ParsePlan();
  plan.kind = kind;
  plan.txt = d_txt;
  plan.n_bytes = n_bytes;
  plan.n_lines = n_lines;
  Pool &pool = ctx.pool();
  plan.n_tiles = ceil_div(std::max<uint64_t>(n_bytes, 1), kParseTile);
  plan.fb = DBuf<uint64_t>(pool, std::max<uint64_t>(n_lines, 1));
  plan.fe = DBuf<uint64_t>(pool, std::max<uint64_t>(n_lines, 1));
  plan.tile_pos = DBuf<uint64_t>(pool, plan.n_tiles + 1);
  plan.first = DBuf<uint64_t>(pool, plan.n_tiles + 1);
  DBuf<uint64_t> lpos(pool, n_lines + 1);
  DBuf<uint32_t> tile_cnt(pool, plan.n_tiles), err(pool, 1);
  DBuf<uint32_t> partial(pool, std::max<uint64_t>(n_lines, 1));
  fill_u32(ctx, err.get(), 0, 1);
  if (n_lines) {
    GFZ_HIP(hipMemcpyAsync(plan.fb.get(), field_begin, n_lines * 8, hipMemcpyHostToDevice,
                           ctx.stream()));
    GFZ_HIP(hipMemcpyAsync(plan.fe.get(), field_end, n_lines * 8, hipMemcpyHostToDevice,
                           ctx.stream()));
  }
  const bool any = n_bytes && n_lines;
  if (any) {
    hipLaunchKernelGGL(k_parse_first, dim3(grid1(plan.n_tiles + 1, kBlock)), dim3(kBlock), 0,
                       ctx.stream(), plan.fb.get(), n_lines, plan.n_tiles, plan.first.get());
    GFZ_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_parse_count, dim3((uint32_t)plan.n_tiles), dim3(kBlock), 0, ctx.stream(),
                       kind, d_txt, n_bytes, plan.fb.get(), plan.fe.get(), n_lines,
                       plan.first.get(), tile_cnt.get(), partial.get(), err.get());
    GFZ_CHECK_LAUNCH();
  }
  exclusive_scan(ctx, tile_cnt.get(), plan.tile_pos.get(), any ? plan.n_tiles : 0);
  if (any) {
    hipLaunchKernelGGL(k_parse_line_pos, dim3(grid1(n_lines + 1, kBlock)), dim3(kBlock), 0,
                       ctx.stream(), plan.fb.get(), n_lines, plan.n_tiles, plan.tile_pos.get(),
                       partial.get(), lpos.get());
    GFZ_CHECK_LAUNCH();
  }
  std::vector<uint64_t> h_lpos(n_lines + 1, 0);
  if (any)
    GFZ_HIP(hipMemcpyAsync(h_lpos.data(), lpos.get(), (n_lines + 1) * 8, hipMemcpyDeviceToHost,
                           ctx.stream()));
  ctx.stage_scalars(err.get(), 1);
  ctx.sync();
  const uint32_t e = (uint32_t)ctx.scalars()[0];
  if (e & 2u)
    throw Error(kErrInvalid, "parse_uint32: overflow");
  if (e & 1u)
    throw Error(kErrInvalid, "parse_uint32: no digits");
  plan.total = h_lpos[n_lines];
  for (uint64_t l = 0; l < n_lines; ++l) {
    const uint64_t c = h_lpos[l + 1] - h_lpos[l];
    if (c > 0xFFFFFFFFull)
      throw Error(kErrInvalid, "a line holds more than 2^32 - 1 steps");
    if (lengths)
      lengths[l] = (uint32_t)c;
  }
}

void parse_write(Ctx &ctx, const ParsePlan &plan, int32_t *d_nodes, uint64_t nodes_cap) {
  if (plan.total > nodes_cap)
    throw Error(kErrInvalid, "node capacity " + std::to_string(nodes_cap) + " < steps " +
                                 std::to_string(plan.total));
  if (plan.n_bytes && plan.n_lines) {
    DBuf<uint32_t> err(ctx.pool(), 1);  // every malformed name was reported by the plan
    ctx.timed("parse", [&] {
      hipLaunchKernelGGL(k_parse_write, dim3((uint32_t)plan.n_tiles), dim3(kBlock), 0,
                         ctx.stream(), plan.kind, plan.txt, plan.n_bytes, plan.fb.get(),
                         plan.fe.get(), plan.n_lines, plan.first.get(), plan.tile_pos.get(),
                         d_nodes, err.get());
    });
    GFZ_CHECK_LAUNCH();
  }
  ctx.add_bytes("parse", (double)plan.n_bytes + 4.0 * plan.total);
  ctx.sync();
}

}  // namespace gfz
