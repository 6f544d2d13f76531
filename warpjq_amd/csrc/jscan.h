// Per-line JSON evaluation shared by the gfx950 kernels (compiled by hipcc
// with __device__) and by a host build used only by the test-suite to pin the
// exact same source against the oracle on the CPU.
//
// Semantics restated from the reference's CPU engine, which is the oracle the
// GPU must match byte for byte:
//   validation          crates/warpjq-core/src/json.rs:213-421
//   lookup              crates/warpjq-core/src/json.rs:444-591
//   decoded strings     crates/warpjq-core/src/json.rs:594-654
//   compare / truthy    crates/warpjq-core/src/json.rs:63-65, :662-693
//   line outcome order  crates/warpjq-core/src/exec/cpu.rs:299-371
//   blank lines         crates/warpjq-core/src/chunk.rs:432-438
//   CSV cells           crates/warpjq-core/src/output.rs:45-81
//
// Design: one lane walks one line once. The walk validates the whole value
// and, in the same pass, captures every needed path of up to WJQ_FUSED_STEPS
// steps ("fused" paths), tracking for each open container level which paths
// are still live (a 16-bit mask per level). Duplicate keys overwrite earlier
// captures, which is jq's last-wins rule; an intermediate value of the wrong
// kind records a type error; null records Missing. Longer paths are resolved
// afterwards by a sequential lookup over the already-validated line.
#pragma once

#include <stdint.h>

#include <type_traits>

#include "../../include/warpjq_kernels.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define WJQ_HD __host__ __device__ __forceinline__
#define WJQ_HDI __host__ __device__
#else
#define WJQ_HD inline
#define WJQ_HDI inline
#endif

namespace wjqd {

enum { JK_MISSING = 0, JK_NULL = 1, JK_BOOL = 2, JK_NUM = 3, JK_STR = 4, JK_ARR = 5, JK_OBJ = 6 };
enum { ST_OK = 0, ST_INVALID = 1, ST_TYPE_ERROR = 2, ST_FALLBACK = 3, ST_BLANK = 4 };

// Paths of at most this many steps are captured during the validating walk.
#define WJQ_FUSED_STEPS 3
// Depth of the container bit-stack kept in registers; deeper levels need a
// reader with a stack of its own (the wave-wide phase), else the line is
// declined (FALLBACK) there and handed to that phase.
#define WJQ_MAX_DEPTH 64
#define WJQ_END 0x0Au  // '\n', the line terminator; also returned past the chunk

// A captured value. `off` is chunk-relative; Missing is kind 0.
struct SlotV {
  uint32_t off;
  uint32_t len;
  uint32_t kind;
};

// Per needed slot: the path it resolves, pre-digested at context creation.
struct SlotPath {
  uint32_t nsteps;
  uint32_t step_off;  // into steps[]
  uint32_t fused;     // 1 when nsteps <= WJQ_FUSED_STEPS and indices < 0xFFFF
  uint32_t klen;      // fused slot whose first step is a key of at most 16 bytes
                      // with no '"', backslash or control byte: its length; else ~0
  uint32_t kw[4];     // that key's bytes, little-endian words, zero past klen
};

// The device view of a lowered program (warpjq_program plus derived masks).
struct DProg {
  const warpjq_step *steps;
  const warpjq_cmp *cmps;
  const warpjq_cond_op *cond;
  const uint8_t *blob;
  const SlotPath *spath;       // [n_slots]
  const uint32_t *slot_of_path;
  const uint32_t *field_slot;  // [n_fields] slot of each projected field
  const uint32_t *prefix_off;
  const uint32_t *prefix_len;
  uint32_t n_slots;
  uint32_t n_cond;
  uint32_t fused_mask;   // slots captured by the walk
  uint32_t key_lens0;    // bit n: a fused slot's first step is a key of n bytes (bit 31: 31+ bytes, or no key)
  uint32_t plain_keys;   // slots with SlotPath::klen set
  uint32_t filter_mask;  // slots the filter reads
  uint32_t post_mask;    // every other needed slot
  uint32_t has_filter;
  uint32_t output_kind;
  uint32_t output_slot;  // OUT_PATH
  uint32_t agg_kind;
  uint32_t agg_slot;     // 0xFFFFFFFF for count
  uint32_t group_slot;   // 0xFFFFFFFF when ungrouped
  uint32_t n_fields;
  uint32_t suffix_off;
  uint32_t suffix_len;
  uint32_t csv_mode;
};

// ---------------------------------------------------------------------------
// byte classes
// ---------------------------------------------------------------------------

WJQ_HD bool is_digit(uint32_t c) { return c - '0' < 10u; }
WJQ_HD bool is_hexd(uint32_t c) { return c - '0' < 10u || (c | 0x20u) - 'a' < 6u; }
WJQ_HD uint32_t hexval(uint32_t c) { return c <= '9' ? c - '0' : (c | 0x20u) - 'a' + 10; }
WJQ_HD bool is_jws(uint32_t c) { return c == ' ' || c == '\t' || c == '\r'; }

// Bit 7 of each byte set when that byte may be '"', '\\' or < 0x20. Exact
// for the lowest flagged byte; bytes above a true hit can be flagged falsely
// (borrow), which callers tolerate by re-checking the byte.
WJQ_HD uint32_t str_special(uint32_t x) {
  uint32_t q = x ^ 0x22222222u, b = x ^ 0x5C5C5C5Cu;
  uint32_t m = ((q - 0x01010101u) & ~q) | ((b - 0x01010101u) & ~b) | ((x - 0x20202020u) & ~x);
  return m & 0x80808080u;
}

// 16 bytes of the chunk, little endian dwords (the unit readers fetch).
struct Blk16 {
  uint32_t x, y, z, w;
};

// High bit of each of a block's 16 bytes as a 16-bit mask (bit k = byte k);
// every byte of a..d holds 0x80 or 0x00. On gfx950 four v_dot4_u32_u8.
WJQ_HD uint32_t movemask16(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t lo = __builtin_amdgcn_udot4(b, 0x80402010u, __builtin_amdgcn_udot4(a, 0x08040201u, 0u, false), false);
  uint32_t hi = __builtin_amdgcn_udot4(d, 0x80402010u, __builtin_amdgcn_udot4(c, 0x08040201u, 0u, false), false);
  return (lo >> 7) | ((hi >> 7) << 8);
#else
  uint32_t w[4] = {a, b, c, d}, m = 0;
  for (int q = 0; q < 4; q++)
    for (int k = 0; k < 4; k++) m |= ((w[q] >> (8 * k + 7)) & 1u) << (4 * q + k);
  return m;
#endif
}

// Bit 7 of each byte set when that byte is not an ASCII digit (exact).
WJQ_HD uint32_t non_digit(uint32_t x) {
  uint32_t ge30 = ((x | 0x80808080u) - 0x30303030u) & ~x;  // b < 0x80 and b >= '0'
  uint32_t gt39 = (x & 0x7F7F7F7Fu) + 0x46464646u;          // (b & 0x7F) > '9'
  return ~(ge30 & ~gt39) & 0x80808080u;
}

WJQ_HD int ctz32(uint32_t x) { return __builtin_ctz(x); }
WJQ_HD int ctz64(uint64_t x) { return __builtin_ctzll(x); }

// Bytes sh..sh+3 of the 8 bytes hi:lo, little-endian (sh < 4).
WJQ_HD uint32_t funnel_bytes(uint32_t hi, uint32_t lo, uint32_t sh) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbyte(hi, lo, sh);
#else
  return sh ? (lo >> (8 * sh)) | (hi << (32 - 8 * sh)) : lo;
#endif
}
WJQ_HD int clz64(uint64_t x) { return __builtin_clzll(x); }

// ---------------------------------------------------------------------------
// scanners. `R` provides byte(i), dword(k) (bytes 4k..4k+3, little endian)
// and block(b) (bytes 16b..16b+15); all read WJQ_END for positions at or
// past the chunk end.
// ---------------------------------------------------------------------------

template <class R>
WJQ_HD uint32_t skip_ws(R &rd, uint32_t i) {
  while (is_jws(rd.byte(i))) i++;
  return i;
}

// First position >= j flagged by str_special (16 bytes per step): the first
// '"', '\\' or byte < 0x20 at or after j, or a byte before it that
// str_special flags falsely after a hit earlier in the same dword (callers
// re-check the byte). Every line ends in '\n', so it exists.
// A reader whose every lane reads the same positions in step can search
// 64 blocks at once (next_special_wide).
template <class R, class = void>
struct wide_reader : std::false_type {};
template <class R>
struct wide_reader<R, std::void_t<decltype(&R::next_special_wide)>> : std::true_type {};

template <class R>
WJQ_HD uint32_t next_special(R &rd, uint32_t j) {
  if constexpr (wide_reader<R>::value) return rd.next_special_wide(j);
  // two blocks per step: the second read does not wait for the first
  for (;;) {
    const Blk16 v = rd.block(j >> 4), u = rd.block((j >> 4) + 1u);
    const uint32_t m = (movemask16(str_special(v.x), str_special(v.y), str_special(v.z), str_special(v.w)) |
                        (movemask16(str_special(u.x), str_special(u.y), str_special(u.z), str_special(u.w)) << 16)) &
                       (0xFFFFFFFFu << (j & 15u));
    if (m) return (j & ~15u) + (uint32_t)ctz32(m);
    j = (j | 15u) + 17u;
  }
}

// First position >= j whose byte is not an ASCII digit.
template <class R>
WJQ_HD uint32_t digits_end(R &rd, uint32_t j) {
  for (;;) {  // two blocks per step, as next_special
    const Blk16 v = rd.block(j >> 4), u = rd.block((j >> 4) + 1u);
    const uint32_t m = (movemask16(non_digit(v.x), non_digit(v.y), non_digit(v.z), non_digit(v.w)) |
                        (movemask16(non_digit(u.x), non_digit(u.y), non_digit(u.z), non_digit(u.w)) << 16)) &
                       (0xFFFFFFFFu << (j & 15u));
    if (m) return (j & ~15u) + (uint32_t)ctz32(m);
    j = (j | 15u) + 17u;
  }
}

// `i` is the opening quote. On success returns true with `i` one past the
// closing quote; `esc` is set when the body holds a backslash.
template <class R>
WJQ_HD bool scan_string(R &rd, uint32_t &i, bool &esc) {
  uint32_t j = i + 1;
  for (;;) {
    uint32_t pos = next_special(rd, j);
    uint32_t c = rd.byte(pos);
    if (c == '"') {
      i = pos + 1;
      return true;
    }
    if (c != '\\') {
      if (c < 0x20) return false;  // raw control, or the line ended inside
      j = pos + 1;                 // a byte str_special flagged by borrow
      continue;
    }
    uint32_t e = rd.byte(pos + 1);
    if (e == '"' || e == '\\' || e == '/' || e == 'b' || e == 'f' || e == 'n' || e == 'r' || e == 't') {
      j = pos + 2;
    } else if (e == 'u') {
      if (!is_hexd(rd.byte(pos + 2)) || !is_hexd(rd.byte(pos + 3)) || !is_hexd(rd.byte(pos + 4)) ||
          !is_hexd(rd.byte(pos + 5)))
        return false;
      j = pos + 6;
    } else {
      return false;
    }
    esc = true;
  }
}

// json.rs:370-410. `i` is '-' or a digit; returns one past the number.
template <class R>
WJQ_HD bool scan_number(R &rd, uint32_t &i) {
  uint32_t j = i;
  if (rd.byte(j) == '-') j++;
  const uint32_t is = j;
  const uint32_t first = rd.byte(j);
  j = digits_end(rd, j);
  if (j == is) return false;
  if (first == '0' && j - is > 1) return false;
  uint32_t c = rd.byte(j);
  if (c == '.') {
    const uint32_t fs = j + 1;
    j = digits_end(rd, fs);
    if (j == fs) return false;
    c = rd.byte(j);
  }
  if ((c | 0x20u) == 'e') {
    j++;
    c = rd.byte(j);
    if (c == '+' || c == '-') j++;
    const uint32_t es = j;
    j = digits_end(rd, es);
    if (j == es) return false;
  }
  i = j;
  return true;
}

template <class R>
WJQ_HD bool scan_lit(R &rd, uint32_t &i, uint32_t c0) {
  if (c0 == 't') {
    if (rd.byte(i + 1) != 'r' || rd.byte(i + 2) != 'u' || rd.byte(i + 3) != 'e') return false;
    i += 4;
  } else if (c0 == 'f') {
    if (rd.byte(i + 1) != 'a' || rd.byte(i + 2) != 'l' || rd.byte(i + 3) != 's' || rd.byte(i + 4) != 'e')
      return false;
    i += 5;
  } else {
    if (rd.byte(i + 1) != 'u' || rd.byte(i + 2) != 'l' || rd.byte(i + 3) != 'l') return false;
    i += 4;
  }
  return true;
}

WJQ_HD uint32_t kind_of(uint32_t c) {
  switch (c) {
    case '{': return JK_OBJ;
    case '[': return JK_ARR;
    case '"': return JK_STR;
    case 't':
    case 'f': return JK_BOOL;
    case 'n': return JK_NULL;
    default: return JK_NUM;
  }
}

// ---------------------------------------------------------------------------
// decoded string iteration (json.rs:594-654): escapes decoded, surrogate
// pairs combined, lone surrogates replaced by U+FFFD, all as UTF-8 bytes.
// ---------------------------------------------------------------------------

template <class R>
struct DecIter {
  R *rd;
  uint32_t p, end;
  uint32_t pend;  // pending UTF-8 bytes, low byte first
  uint32_t npend;

  WJQ_HD DecIter(R &r, uint32_t a, uint32_t b) : rd(&r), p(a), end(b), pend(0), npend(0) {}

  WJQ_HD uint32_t hex4(uint32_t q) {
    return (hexval(rd->byte(q)) << 12) | (hexval(rd->byte(q + 1)) << 8) | (hexval(rd->byte(q + 2)) << 4) |
           hexval(rd->byte(q + 3));
  }

  // Next decoded byte, or -1 at the end. Assumes a validated body.
  WJQ_HD int next() {
    if (npend) {
      uint32_t b = pend & 0xFF;
      pend >>= 8;
      npend--;
      return (int)b;
    }
    if (p >= end) return -1;
    uint32_t c = rd->byte(p);
    if (c != '\\') {
      p++;
      return (int)c;
    }
    uint32_t e = rd->byte(p + 1);
    p += 2;
    switch (e) {
      case 'b': return 0x08;
      case 'f': return 0x0C;
      case 'n': return '\n';
      case 'r': return '\r';
      case 't': return '\t';
      case 'u': break;
      default: return (int)e;  // " \ /
    }
    uint32_t cp = hex4(p);
    p += 4;
    if (cp >= 0xD800 && cp < 0xDC00 && p + 6 <= end && rd->byte(p) == '\\' && rd->byte(p + 1) == 'u') {
      uint32_t lo = hex4(p + 2);
      if (lo >= 0xDC00 && lo < 0xE000) {
        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
        p += 6;
      }
    }
    if (cp >= 0xD800 && cp < 0xE000) cp = 0xFFFD;
    if (cp < 0x80) return (int)cp;
    uint32_t b0;
    if (cp < 0x800) {
      b0 = 0xC0 | (cp >> 6);
      pend = 0x80 | (cp & 0x3F);
      npend = 1;
    } else if (cp < 0x10000) {
      b0 = 0xE0 | (cp >> 12);
      pend = (0x80 | ((cp >> 6) & 0x3F)) | ((0x80 | (cp & 0x3F)) << 8);
      npend = 2;
    } else {
      b0 = 0xF0 | (cp >> 18);
      pend = (0x80 | ((cp >> 12) & 0x3F)) | ((0x80 | ((cp >> 6) & 0x3F)) << 8) | ((0x80 | (cp & 0x3F)) << 16);
      npend = 3;
    }
    return (int)b0;
  }
};

// FNV-1a 32 over decoded bytes of the body [a, b) (query/mod.rs:15-22).
template <class R>
WJQ_HD uint32_t decoded_hash(R &rd, uint32_t a, uint32_t b, bool esc, uint32_t *dlen) {
  uint32_t h = 0x811c9dc5u, n = 0;
  if (!esc) {
    for (uint32_t q = a; q < b; q++) {
      h = (h ^ rd.byte(q)) * 0x01000193u;
    }
    n = b - a;
  } else {
    DecIter<R> it(rd, a, b);
    for (int c; (c = it.next()) >= 0; n++) h = (h ^ (uint32_t)c) * 0x01000193u;
  }
  if (dlen) *dlen = n;
  return h;
}

// Three-way compare of the decoded body [a, b) with bytes lit[0, n).
template <class R, class LP>
WJQ_HD int decoded_cmp(R &rd, uint32_t a, uint32_t b, bool esc, LP lit, uint32_t n) {
  if (!esc) {
    uint32_t m = b - a;
    uint32_t k = m < n ? m : n;
    for (uint32_t q = 0; q < k; q++) {
      uint32_t x = rd.byte(a + q), y = lit[q];
      if (x != y) return x < y ? -1 : 1;
    }
    return m < n ? -1 : (m > n ? 1 : 0);
  }
  DecIter<R> it(rd, a, b);
  uint32_t q = 0;
  for (;;) {
    int c = it.next();
    if (c < 0) return q < n ? -1 : 0;
    if (q >= n) return 1;
    uint32_t y = lit[q++];
    if ((uint32_t)c != y) return (uint32_t)c < y ? -1 : 1;
  }
}

// Does the string body [a, b) contain a backslash?
template <class R>
WJQ_HD bool body_has_escape(R &rd, uint32_t a, uint32_t b) {
  for (uint32_t q = a; q < b; q++)
    if (rd.byte(q) == '\\') return true;
  return false;
}

// ---------------------------------------------------------------------------
// number value (Slot::as_f64, json.rs:73-78): correctly rounded decimal ->
// binary64. Clinger's exact fast path, then Eisel-Lemire with a 128-bit
// power-of-five table (Lemire 2021), then for inputs with more than 19
// significant digits the two-candidate check. Anything still ambiguous is
// reported as not decided (the caller declines the line).
// ---------------------------------------------------------------------------

#ifndef WJQ_TABLE_QUAL
#if defined(__HIPCC__)
#define WJQ_TABLE_QUAL __device__ __constant__ static const
#else
#define WJQ_TABLE_QUAL static const
#endif
#endif
#include "pow5_table.h"

WJQ_HD void mul128(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo) {
  unsigned __int128 p = (unsigned __int128)a * b;
  hi = (uint64_t)(p >> 64);
  lo = (uint64_t)p;
}

// w * 10^q for w != 0 as binary64 bits; false when the product is too close
// to a rounding boundary to decide from 128 bits.
WJQ_HD bool eisel_lemire(uint64_t w, int q, uint64_t &bits) {
  if (q < WJQ_POW5_MIN_Q) {
    bits = 0;
    return true;
  }
  if (q > WJQ_POW5_MAX_Q) {
    bits = 0x7FF0000000000000ull;
    return true;
  }
  int lz = clz64(w);
  w <<= lz;
  const uint64_t *t = &wjq_pow5_128[2 * (q - WJQ_POW5_MIN_Q)];
  uint64_t hi, lo;
  mul128(w, t[0], hi, lo);
  if ((hi & 0x1FF) == 0x1FF) {
    uint64_t hi2, lo2;
    mul128(w, t[1], hi2, lo2);
    lo += hi2;
    if (lo < hi2) hi++;
  }
  if (lo == ~0ull && (q < -27 || q > 55)) return false;
  uint64_t upper = hi >> 63;
  uint64_t m = hi >> (upper + 9);
  int p2 = (int)(((217706 * (int64_t)q) >> 16) + 63 + (int64_t)upper - lz + 1023);
  if (p2 <= 0) {
    if (-p2 + 1 >= 64) {
      bits = 0;
      return true;
    }
    m >>= -p2 + 1;
    m += m & 1;
    m >>= 1;
    bits = m;  // subnormal, or the smallest normal when rounding carried
    return true;
  }
  if (lo <= 1 && q >= -4 && q <= 23 && (m & 3) == 1) {
    if ((m << (upper + 9)) == hi) m &= ~1ull;
  }
  m += m & 1;
  m >>= 1;
  if (m >= (2ull << 52)) {
    m = 1ull << 52;
    p2++;
  }
  m &= ~(1ull << 52);
  if (p2 >= 0x7FF) {
    bits = 0x7FF0000000000000ull;
    return true;
  }
  bits = ((uint64_t)p2 << 52) | m;
  return true;
}

WJQ_HD double bits_to_f64(uint64_t b) {
  union {
    uint64_t u;
    double d;
  } x;
  x.u = b;
  return x.d;
}
WJQ_HD uint64_t f64_to_bits(double d) {
  union {
    uint64_t u;
    double d;
  } x;
  x.d = d;
  return x.u;
}

// ---------------------------------------------------------------------------
// exact decimal -> binary64 for the numbers the 128-bit path cannot settle:
// the simple decimal conversion (N. Tao, "The Simple Decimal Conversion
// Algorithm", 2020; the method behind Go's strconv and Rust's dec2flt slow
// path, whose correctly rounded result Rust's f64::from_str gives,
// json.rs:73-78). The decimal is held as up to WJQ_DEC_DIGITS significant
// digits (enough to decide any binary64: 767 + 1) plus a flag for nonzero
// digits dropped past them, and scaled by powers of two until it lies in
// [0.5, 1); its 53-bit mantissa is then rounded half to even. `B` is a byte
// buffer of WJQ_DEC_BUF bytes (a wave's LDS on the device, all lanes
// computing the same number in step).
// ---------------------------------------------------------------------------

#define WJQ_DEC_DIGITS 800
#define WJQ_DEC_BUF (WJQ_DEC_DIGITS + 32)

template <class B>
struct Decimal {
  B d;          // digit values 0..9
  uint32_t nd;  // significant digits held
  int32_t dp;   // value = 0.d[0]d[1]... x 10^dp
  bool trunc;   // nonzero digits were dropped past WJQ_DEC_DIGITS

  WJQ_HD void trim() {
    while (nd > 0 && d[nd - 1] == 0) nd--;
    if (nd == 0) dp = 0;
  }
  // x 2^-k, 1 <= k <= 60: long division from the most significant digit
  WJQ_HD void shr(uint32_t k) {
    uint32_t r = 0, w = 0;
    uint64_t n = 0;
    while ((n >> k) == 0) {
      if (r >= nd) {
        if (n == 0) {
          nd = 0;
          dp = 0;
          return;
        }
        while ((n >> k) == 0) {
          n *= 10;
          r++;
        }
        break;
      }
      n = n * 10 + d[r];
      r++;
    }
    dp -= (int32_t)r - 1;
    const uint64_t mask = (1ull << k) - 1;
    while (r < nd) {
      const uint32_t q = (uint32_t)(n >> k);
      n = ((n & mask) * 10) + d[r];
      r++;
      d[w++] = (uint8_t)q;
    }
    while (n > 0) {
      const uint32_t q = (uint32_t)(n >> k);
      n = (n & mask) * 10;
      if (w < WJQ_DEC_DIGITS) d[w++] = (uint8_t)q;
      else if (q) trunc = true;
    }
    nd = w;
    trim();
  }
  // x 2^k, 1 <= k <= 60: from the least significant digit, into positions
  // shifted right by an upper bound on the new leading digits, then moved back
  WJQ_HD void shl(uint32_t k) {
    const uint32_t D = ((k * 1233u) >> 12) + 1u;  // > k log10(2)
    int32_t rd_ = (int32_t)nd - 1;
    int32_t wr = (int32_t)nd - 1 + (int32_t)D;
    uint64_t n = 0;
    while (rd_ >= 0) {
      n += (uint64_t)d[rd_] << k;
      const uint64_t q = n / 10;
      d[wr] = (uint8_t)(n - q * 10);
      n = q;
      wr--;
      rd_--;
    }
    while (n > 0) {
      const uint64_t q = n / 10;
      d[wr] = (uint8_t)(n - q * 10);
      n = q;
      wr--;
    }
    const uint32_t s = (uint32_t)(wr + 1);
    uint32_t total = nd + D - s;
    for (uint32_t i = 0; i < total; i++) d[i] = d[i + s];
    dp += (int32_t)(D - s);
    if (total > WJQ_DEC_DIGITS) {
      for (uint32_t i = WJQ_DEC_DIGITS; i < total; i++)
        if (d[i]) trunc = true;
      total = WJQ_DEC_DIGITS;
    }
    nd = total;
    trim();
  }
  // nearest integer, ties to even (a dropped nonzero tail breaks a tie up)
  WJQ_HD uint64_t round_int() const {
    if (nd == 0 || dp < 0) return 0;
    if (dp > 18) return ~0ull;
    uint64_t n = 0;
    for (int32_t i = 0; i < dp; i++) n = n * 10 + ((uint32_t)i < nd ? d[i] : 0u);
    bool up = false;
    if ((uint32_t)dp < nd) {
      const uint32_t c = d[dp];
      up = c >= 5;
      if (c == 5 && (uint32_t)dp + 1 == nd) up = trunc || (n & 1);
    }
    return n + (up ? 1 : 0);
  }
};

// floor(n log2(10)) for n < 19: a right (left) shift by it keeps a decimal
// point of n (-n) from overshooting [0.5, 1).
WJQ_HD uint32_t dec_shift_for(uint32_t n) {
  const uint8_t t[19] = {0, 3, 6, 9, 13, 16, 19, 23, 26, 29, 33, 36, 39, 43, 46, 49, 53, 56, 59};
  return n < 19 ? t[n] : 60u;
}

// Token [a, b) is a validated JSON number; its correctly rounded binary64
// bits (sign included).
template <class R, class B>
WJQ_HD uint64_t slow_decimal(R &rd, uint32_t a, uint32_t b, B buf) {
  Decimal<B> x{buf, 0, 0, false};
  uint32_t i = a;
  const bool neg = rd.byte(i) == '-';
  if (neg) i++;
  auto push = [&](uint32_t c) {
    if (x.nd < WJQ_DEC_DIGITS) x.d[x.nd++] = (uint8_t)(c - '0');
    else if (c != '0') x.trunc = true;
  };
  for (; i < b; i++) {
    const uint32_t c = rd.byte(i);
    if (!is_digit(c)) break;
    if (x.nd == 0 && c == '0') continue;
    push(c);
    x.dp++;
  }
  if (i < b && rd.byte(i) == '.') {
    for (i++; i < b; i++) {
      const uint32_t c = rd.byte(i);
      if (!is_digit(c)) break;
      if (x.nd == 0 && c == '0') x.dp--;
      else push(c);
    }
  }
  int64_t e = 0;
  if (i < b && (rd.byte(i) | 0x20u) == 'e') {
    i++;
    bool eneg = false;
    if (rd.byte(i) == '+') i++;
    else if (rd.byte(i) == '-') {
      eneg = true;
      i++;
    }
    for (; i < b; i++)
      if (e < 100000) e = e * 10 + (rd.byte(i) - '0');
    if (eneg) e = -e;
  }
  const uint64_t sign = neg ? (1ull << 63) : 0ull;
  x.trim();
  if (x.nd == 0) return sign;
  const int64_t dp = (int64_t)x.dp + e;
  if (dp >= 310) return sign | 0x7FF0000000000000ull;
  if (dp <= -326) return sign;
  x.dp = (int32_t)dp;
  int32_t exp2 = 0;
  while (x.dp > 0) {
    const uint32_t n = dec_shift_for((uint32_t)x.dp);
    x.shr(n);
    exp2 += (int32_t)n;
  }
  while (x.dp < 0 || (x.dp == 0 && x.d[0] < 5)) {
    uint32_t n;
    if (x.dp == 0) n = x.d[0] >= 2 ? 1u : 2u;
    else n = dec_shift_for((uint32_t)(-x.dp));
    x.shl(n);
    exp2 -= (int32_t)n;
  }
  // value = V x 2^exp2 with V in [0.5, 1), i.e. (2V) x 2^E
  int32_t E = exp2 - 1;
  if (E < -1022) {
    uint32_t n = (uint32_t)(-1022 - E);
    while (n > 0) {
      const uint32_t k = n > 60 ? 60u : n;
      x.shr(k);
      n -= k;
    }
    E = -1022;
  }
  if (E > 1023) return sign | 0x7FF0000000000000ull;
  x.shl(53);
  uint64_t m = x.round_int();
  if (m >= (1ull << 53)) {
    x.shr(1);
    E++;
    m = x.round_int();
    if (E > 1023) return sign | 0x7FF0000000000000ull;
  }
  uint64_t biased = 0;
  if (m >= (1ull << 52)) {
    biased = (uint64_t)(E + 1023);
    m &= (1ull << 52) - 1;
  }
  return sign | (biased << 52) | m;
}

// Readers that carry a WJQ_DEC_BUF-byte digit buffer (dec_buf()) decide
// every number; the others leave the rare hard ones undecided.
template <class R, class = void>
struct has_dec_buf : std::false_type {};
template <class R>
struct has_dec_buf<R, std::void_t<decltype(&R::dec_buf)>> : std::true_type {};

struct NumVal {
  double v;
  int64_t iv;    // exact integer value when is_int
  bool is_int;   // integer token with |value| <= 2^53
};

template <class R, class = void>
struct has_dword_reads : std::false_type {};
template <class R>
struct has_dword_reads<R, std::void_t<decltype(&R::dword)>> : std::true_type {};

// Value of the 4 digits in bytes 0-3 of w (byte 0 the most significant).
WJQ_HD uint32_t digits4(uint32_t w) {
  const uint32_t p = (w & 0x00FF00FFu) * 10u + ((w >> 8) & 0x00FF00FFu);
  return (p & 0xFFFFu) * 100u + (p >> 16);
}

// Token [a, b) is a validated JSON number. Returns false when undecided.
template <class R>
WJQ_HD bool parse_number(R &rd, uint32_t a, uint32_t b, NumVal &out) {
  if constexpr (has_dword_reads<R>::value && !wide_reader<R>::value) {
    // 1 to 8 bytes, all digits: the integer from two words (the value a
    // digit loop would give: leading zeros add nothing, below 2^53)
    const uint32_t n = b - a;
    if (n - 1u < 8u) {
      const uint32_t k0 = a >> 2, sh = a & 3u;
      const uint32_t d0 = rd.dword(k0), d1 = rd.dword(k0 + 1u), d2 = rd.dword(k0 + 2u);
      const uint32_t mlo = n >= 4u ? ~0u : (1u << (8u * n)) - 1u;
      const uint32_t mhi = n <= 4u ? 0u : n == 8u ? ~0u : (1u << (8u * (n - 4u))) - 1u;
      const uint32_t lo = funnel_bytes(d1, d0, sh) & mlo, hi = funnel_bytes(d2, d1, sh) & mhi;
      const uint32_t vlo = lo ^ (0x30303030u & mlo), vhi = hi ^ (0x30303030u & mhi);
      const uint32_t big = ((((vlo & 0x7F7F7F7Fu) + 0x76767676u) | vlo) | (((vhi & 0x7F7F7F7Fu) + 0x76767676u) | vhi)) &
                           0x80808080u;
      if (!big) {
        // right-align to 8 digits: shift left by 8 - n bytes
        const uint32_t s8 = 8u * (8u - n);
        uint32_t l2, h2;
        if (s8 >= 32u) {
          h2 = vlo << (s8 - 32u);
          l2 = 0;
        } else {
          h2 = s8 ? (vhi << s8) | (vlo >> (32u - s8)) : vhi;
          l2 = vlo << s8;
        }
        const uint32_t v = digits4(l2) * 10000u + digits4(h2);
        out.v = (double)v;
        out.iv = (int64_t)v;
        out.is_int = true;
        return true;
      }
    }
  }
  uint32_t i = a;
  bool neg = rd.byte(i) == '-';
  if (neg) i++;
  uint64_t w = 0;
  int nd = 0;          // significant digits taken into w
  int dropped = 0;     // significant digits beyond 19 (integer part)
  int frac_taken = 0;  // fraction digits consumed into w
  bool trunc = false;
  bool any_frac = false, any_exp = false;
  // integer part
  for (; i < b; i++) {
    uint32_t c = rd.byte(i);
    if (!is_digit(c)) break;
    if (nd == 0 && c == '0') continue;
    if (nd < 19) {
      w = w * 10 + (c - '0');
      nd++;
    } else {
      dropped++;
      if (c != '0') trunc = true;
    }
  }
  if (i < b && rd.byte(i) == '.') {
    any_frac = true;
    i++;
    for (; i < b; i++) {
      uint32_t c = rd.byte(i);
      if (!is_digit(c)) break;
      if (nd == 0 && c == '0') {
        frac_taken++;
        continue;
      }
      if (nd < 19) {
        w = w * 10 + (c - '0');
        nd++;
        frac_taken++;
      } else if (c != '0') {
        trunc = true;
      }
    }
  }
  int64_t e = 0;
  if (i < b && (rd.byte(i) == 'e' || rd.byte(i) == 'E')) {
    any_exp = true;
    i++;
    bool eneg = false;
    if (rd.byte(i) == '+') i++;
    else if (rd.byte(i) == '-') {
      eneg = true;
      i++;
    }
    for (; i < b; i++) {
      uint32_t c = rd.byte(i);
      if (e < 100000) e = e * 10 + (c - '0');
    }
    if (eneg) e = -e;
  }
  int64_t q64 = e + dropped - frac_taken;
  if (q64 < -100000) q64 = -100000;
  if (q64 > 100000) q64 = 100000;
  int q = (int)q64;
  out.is_int = false;
  out.iv = 0;
  if (w == 0) {
    out.v = neg ? -0.0 : 0.0;
    if (!any_frac && !any_exp) {
      out.is_int = true;
      out.iv = 0;
    }
    return true;
  }
  if (!any_frac && !any_exp && !trunc && dropped == 0 && w <= (1ull << 53)) {
    out.is_int = true;
    out.iv = neg ? -(int64_t)w : (int64_t)w;
  }
  if (!trunc && q >= -22 && q <= 22 && w <= (1ull << 53)) {
    const double p10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                            1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    double d = (double)w;
    d = q < 0 ? d / p10[-q] : d * p10[q];
    out.v = neg ? -d : d;
    return true;
  }
  uint64_t bits;
  bool ok = eisel_lemire(w, q, bits);
  if (ok && trunc) {
    uint64_t bits2;
    ok = eisel_lemire(w + 1, q, bits2) && bits2 == bits;
  }
  if (!ok) {
    // a reader with a digit buffer (the wave-wide phase) settles it exactly
    if constexpr (has_dec_buf<R>::value) {
      out.v = bits_to_f64(slow_decimal(rd, a, b, rd.dec_buf()));
      return true;
    }
    return false;
  }
  double d = bits_to_f64(bits);
  out.v = neg ? -d : d;
  return true;
}

// ---------------------------------------------------------------------------
// the walk
// ---------------------------------------------------------------------------

WJQ_HD uint32_t lvl_get(uint64_t packed, uint32_t L) { return (uint32_t)(packed >> (16 * L)) & 0xFFFFu; }
WJQ_HD uint64_t lvl_set(uint64_t packed, uint32_t L, uint32_t v) {
  uint64_t sh = 16ull * L;
  return (packed & ~(0xFFFFull << sh)) | ((uint64_t)(v & 0xFFFFu) << sh);
}

// Does the step `st` name the object member whose raw key body is [a, b)?
// Raw bodies (no backslash) compare by length, then bytes.
template <class R, class PP, class ST>
WJQ_HD bool key_step_match(R &rd, const PP &P, const ST &st, uint32_t a, uint32_t b, bool esc) {
  if (st.kind != WARPJQ_STEP_KEY) return false;
  if (!esc) {
    if (st.key_len != b - a) return false;
    const uint32_t off = st.key_off;
    for (uint32_t q = 0; q < b - a; q++)
      if (rd.byte(a + q) != P.blob[off + q]) return false;
    return true;
  }
  return decoded_cmp(rd, a, b, esc, P.blob + st.key_off, st.key_len) == 0;
}

// Slots whose path step `L` is the key [ks, ke) (among `A`), and those for
// which it is the last step.
template <class R, class PP>
WJQ_HD void key_hits(R &rd, const PP &P, uint32_t A, uint32_t L, uint32_t ks, uint32_t ke, bool esc,
                     uint32_t &hit, uint32_t &hit_fin) {
  hit = 0;
  hit_fin = 0;
  while (A) {
    const uint32_t s = (uint32_t)ctz32(A);
    A &= A - 1;
    const auto &sp = P.spath[s];
    if (key_step_match(rd, P, P.steps[sp.step_off + L], ks, ke, esc)) {
      hit |= 1u << s;
      if (sp.nsteps == L + 1) hit_fin |= 1u << s;
    }
  }
}

// Slots whose path step `L` is array index `k` (among `A`).
template <class PP>
WJQ_HD void index_hits(const PP &P, uint32_t A, uint32_t L, uint32_t k, uint32_t &hit, uint32_t &hit_fin) {
  hit = 0;
  hit_fin = 0;
  while (A) {
    const uint32_t s = (uint32_t)ctz32(A);
    A &= A - 1;
    const auto &sp = P.spath[s];
    const auto &st = P.steps[sp.step_off + L];
    if (st.kind == WARPJQ_STEP_INDEX && st.index == k) {
      hit |= 1u << s;
      if (sp.nsteps == L + 1) hit_fin |= 1u << s;
    }
  }
}

// A value of `kind` starts at `i` for the slots in `hit`: final ones record
// it, intermediate ones check that it is the container their next step
// needs (null: Missing; anything else: a type error). Returns the slots that
// descend into it.
template <int NS, class PP>
WJQ_HD uint32_t value_hits(const PP &P, uint32_t hit, uint32_t hit_fin, uint32_t kind, uint32_t i, uint32_t d,
                           SlotV (&sl)[NS], uint32_t &te) {
  uint32_t next_act = 0;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    if (!((hit >> s) & 1)) continue;
    if ((hit_fin >> s) & 1) {
      sl[s].off = i;
      sl[s].kind = kind;
      sl[s].len = 0;
      te &= ~(1u << s);
    } else {
      const auto &st = P.steps[P.spath[s].step_off + d];
      const uint32_t want = st.kind == WARPJQ_STEP_INDEX ? (uint32_t)JK_ARR : (uint32_t)JK_OBJ;
      sl[s].kind = JK_MISSING;
      if (kind == JK_NULL) {
        te &= ~(1u << s);
      } else if (kind == want) {
        te &= ~(1u << s);
        next_act |= 1u << s;
      } else {
        te |= 1u << s;
      }
    }
  }
  return next_act;
}

template <int NS>
WJQ_HD void close_hits(uint32_t F, uint32_t i, SlotV (&sl)[NS]) {
#pragma unroll
  for (int s = 0; s < NS; s++)
    if ((F >> s) & 1) sl[s].len = i - sl[s].off;
}

// Readers that hold container levels past WJQ_MAX_DEPTH (deep_cap(),
// deep_set(level, is_array), deep_get(level)) let the walk go deeper; with
// the others a deeper line is declined.
template <class R, class = void>
struct has_deep_stack : std::false_type {};
template <class R>
struct has_deep_stack<R, std::void_t<decltype(&R::deep_cap)>> : std::true_type {};

// Validates the value starting at `i` and the whitespace after it up to the
// end of the line; captures fused slots. Returns ST_OK / ST_INVALID /
// ST_FALLBACK. On ST_OK `end` is the line terminator's position.
//
// One loop iteration per object member (key, ':', value, then ',' or the
// close) or per array element, so a flat object costs one iteration per
// member. Containers open by falling into the VALUE state.
template <int NS, class R, class PP>
WJQ_HD int walk(R &rd, uint32_t i, const PP &P, SlotV (&sl)[NS], uint32_t &te, uint32_t &end) {
  uint64_t arr = 0;  // bit k: the container at depth k is an array
  uint32_t d = 0;    // open containers
  uint64_t act = 0;  // live slots per container level 0..3
  uint64_t fin = 0;  // slots whose captured value is the container at that level
  uint64_t cnt = 0;  // element index per array level 0..3
  uint32_t hit = P.fused_mask;  // slots for which the next value matters
  uint32_t hit_fin = 0;         // ... as the path's final value
  enum { VALUE, MEMBER, AFTER } state = VALUE;
  i = skip_ws(rd, i);
  uint32_t c = rd.byte(i);  // the byte at i, not whitespace, in VALUE and MEMBER
  for (;;) {
    if (state == MEMBER) {
      // ---- "key": value, at i (c is its first byte) ----
      if (c != '"') return ST_INVALID;
      const uint32_t ks = i + 1;
      bool esc = false;
      if (!scan_string(rd, i, esc)) return ST_INVALID;
      const uint32_t L = d - 1;
      const uint32_t A = L < 4 ? lvl_get(act, L) : 0u;
      hit = 0;
      hit_fin = 0;
      // top-level keys of no wanted length skip the table lookups
      const uint32_t kl = i - 1 - ks;
      if (A && (L != 0 || esc || ((P.key_lens0 >> (kl < 31u ? kl : 31u)) & 1u)))
        key_hits(rd, P, A, L, ks, i - 1, esc, hit, hit_fin);
      c = rd.byte(i);
      if (is_jws(c)) {
        i = skip_ws(rd, i);
        c = rd.byte(i);
      }
      if (c != ':') return ST_INVALID;
      i++;
      c = rd.byte(i);
      if (is_jws(c)) {
        i = skip_ws(rd, i);
        c = rd.byte(i);
      }
      if (c == '{' || c == '[') {
        state = VALUE;
        continue;
      }
      // scalar member value
      uint32_t kind;
      const uint32_t j0 = i;
      bool ok;
      if (c == '"') {
        kind = JK_STR;
        bool e2 = false;
        ok = scan_string(rd, i, e2);
      } else if (c == '-' || is_digit(c)) {
        kind = JK_NUM;
        ok = scan_number(rd, i);
      } else if (c == 't' || c == 'f' || c == 'n') {
        kind = c == 'n' ? (uint32_t)JK_NULL : (uint32_t)JK_BOOL;
        ok = scan_lit(rd, i, c);
      } else {
        return ST_INVALID;
      }
      if (!ok) return ST_INVALID;
      if (hit) {
        value_hits<NS>(P, hit, hit_fin, kind, j0, d, sl, te);
        close_hits<NS>(hit_fin, i, sl);
      }
      c = rd.byte(i);
      if (is_jws(c)) {
        i = skip_ws(rd, i);
        c = rd.byte(i);
      }
      if (c == ',') {
        i++;
        c = rd.byte(i);
        if (is_jws(c)) {
          i = skip_ws(rd, i);
          c = rd.byte(i);
        }
        continue;  // next member
      }
      state = AFTER;
      continue;
    }
    if (state == VALUE) {
      // ---- a value at i (c is its first byte) ----
      uint32_t kind;
      if (c == '"') kind = JK_STR;
      else if (c == '-' || is_digit(c)) kind = JK_NUM;
      else if (c == '{') kind = JK_OBJ;
      else if (c == '[') kind = JK_ARR;
      else if (c == 't' || c == 'f') kind = JK_BOOL;
      else if (c == 'n') kind = JK_NULL;
      else return ST_INVALID;
      const uint32_t next_act = hit ? value_hits<NS>(P, hit, hit_fin, kind, i, d, sl, te) : 0u;
      if (kind == JK_OBJ || kind == JK_ARR) {
        const uint32_t close = kind == JK_OBJ ? '}' : ']';
        uint32_t j = i + 1;
        c = rd.byte(j);
        if (is_jws(c)) {
          j = skip_ws(rd, j);
          c = rd.byte(j);
        }
        if (c == close) {
          i = j + 1;
          if (hit_fin) close_hits<NS>(hit_fin, i, sl);
          state = AFTER;
          continue;
        }
        if (d < WJQ_MAX_DEPTH) {
          if (kind == JK_ARR) arr |= 1ull << d;
          else arr &= ~(1ull << d);
        } else if constexpr (has_deep_stack<R>::value) {
          // the wave-wide phase keeps deeper levels in LDS
          if (d >= rd.deep_cap()) return ST_FALLBACK;
          rd.deep_set(d, kind == JK_ARR);
        } else {
          return ST_FALLBACK;
        }
        if (d < 4) {
          act = lvl_set(act, d, next_act);
          fin = lvl_set(fin, d, hit_fin);
          cnt = lvl_set(cnt, d, 0);
        }
        d++;
        i = j;
        if (kind == JK_OBJ) {
          state = MEMBER;
        } else {
          const uint32_t L = d - 1;
          const uint32_t A = L < 4 ? lvl_get(act, L) : 0u;
          hit = 0;
          hit_fin = 0;
          if (A) index_hits(P, A, L, 0, hit, hit_fin);
        }
        continue;
      }
      const uint32_t j0 = i;
      bool ok;
      if (kind == JK_STR) {
        bool esc = false;
        ok = scan_string(rd, i, esc);
      } else if (kind == JK_NUM) {
        ok = scan_number(rd, i);
      } else {
        ok = scan_lit(rd, i, c);
      }
      if (!ok) return ST_INVALID;
      if (hit_fin) close_hits<NS>(hit_fin, i, sl);
      (void)j0;
      state = AFTER;
      continue;
    }
    // ---- AFTER: a value ended at i; ',' or the container's close ----
    if (d == 0) break;
    c = rd.byte(i);
    if (is_jws(c)) {
      i = skip_ws(rd, i);
      c = rd.byte(i);
    }
    bool top_arr;
    if constexpr (has_deep_stack<R>::value) top_arr = d - 1 < WJQ_MAX_DEPTH ? ((arr >> (d - 1)) & 1) : rd.deep_get(d - 1);
    else top_arr = (arr >> (d - 1)) & 1;
    if (c == ',') {
      i++;
      c = rd.byte(i);
      if (is_jws(c)) {
        i = skip_ws(rd, i);
        c = rd.byte(i);
      }
      if (top_arr) {
        const uint32_t L = d - 1;
        const uint32_t A = L < 4 ? lvl_get(act, L) : 0u;
        hit = 0;
        hit_fin = 0;
        if (A) {
          uint32_t k = lvl_get(cnt, L);
          if (k < 0xFFFF) k++;
          cnt = lvl_set(cnt, L, k);
          index_hits(P, A, L, k, hit, hit_fin);
        }
        state = VALUE;
      } else {
        state = MEMBER;
      }
    } else if (c == (top_arr ? ']' : '}')) {
      i++;
      d--;
      if (d < 4) {
        const uint32_t F = lvl_get(fin, d);
        if (F) close_hits<NS>(F, i, sl);
        act = lvl_set(act, d, 0);
        fin = lvl_set(fin, d, 0);
      }
    } else {
      return ST_INVALID;
    }
  }
  i = skip_ws(rd, i);
  if (rd.byte(i) != WJQ_END) return ST_INVALID;
  end = i;
  return ST_OK;
}

// Skips one value of an already-validated line; returns one past it.
template <class R>
WJQ_HD uint32_t skip_valid(R &rd, uint32_t i) {
  uint32_t c = rd.byte(i);
  if (c == '"') {
    bool esc;
    scan_string(rd, i, esc);
    return i;
  }
  if (c == '{' || c == '[') {
    uint32_t depth = 0;
    for (;;) {
      c = rd.byte(i);
      if (c == '"') {
        bool esc;
        scan_string(rd, i, esc);
        continue;
      }
      if (c == '{' || c == '[') depth++;
      else if (c == '}' || c == ']') {
        depth--;
        if (depth == 0) return i + 1;
      }
      i++;
    }
  }
  // scalar: ends at a delimiter or whitespace
  for (;;) {
    c = rd.byte(i);
    if (c == ',' || c == '}' || c == ']' || is_jws(c) || c == WJQ_END) return i;
    i++;
  }
}

// Sequential lookup (json.rs:444-508) on a validated line, for paths longer
// than the walk captures. Returns false for a type error.
template <class R, class PP, class SP>
WJQ_HD bool lookup_valid(R &rd, uint32_t line, const PP &P, const SP &sp, SlotV &out) {
  uint32_t start = skip_ws(rd, line);
  uint32_t kind = kind_of(rd.byte(start));
  uint32_t endp = 0;
  for (uint32_t k = 0; k < sp.nsteps; k++) {
    if (kind == JK_NULL) {
      out.kind = JK_MISSING;
      return true;
    }
    const auto &st = P.steps[sp.step_off + k];
    bool found = false;
    uint32_t vs = 0, ve = 0;
    if (st.kind == WARPJQ_STEP_KEY) {
      if (kind != JK_OBJ) return false;
      uint32_t j = skip_ws(rd, start + 1);
      if (rd.byte(j) != '}') {
        for (;;) {
          j = skip_ws(rd, j);
          uint32_t ks = j + 1;
          bool esc = false;
          scan_string(rd, j, esc);
          uint32_t ke = j - 1;
          j = skip_ws(rd, j) + 1;  // ':'
          uint32_t s = skip_ws(rd, j);
          uint32_t e = skip_valid(rd, s);
          uint32_t dlen;
          uint32_t h = decoded_hash(rd, ks, ke, esc, &dlen);
          if (dlen == st.key_len && h == st.key_hash &&
              decoded_cmp(rd, ks, ke, esc, P.blob + st.key_off, st.key_len) == 0) {
            found = true;
            vs = s;
            ve = e;
          }
          j = skip_ws(rd, e);
          if (rd.byte(j) == ',') {
            j++;
            continue;
          }
          break;
        }
      }
    } else {
      if (kind != JK_ARR) return false;
      uint32_t j = skip_ws(rd, start + 1);
      if (rd.byte(j) != ']') {
        uint32_t n = 0;
        for (;;) {
          uint32_t s = skip_ws(rd, j);
          uint32_t e = skip_valid(rd, s);
          if (n == st.index) {
            found = true;
            vs = s;
            ve = e;
            break;
          }
          n++;
          j = skip_ws(rd, e);
          if (rd.byte(j) == ',') {
            j++;
            continue;
          }
          break;
        }
      }
    }
    if (!found) {
      out.kind = JK_MISSING;
      return true;
    }
    start = vs;
    endp = ve;
    kind = kind_of(rd.byte(vs));
  }
  out.kind = kind;
  out.off = start;
  out.len = endp - start;
  return true;
}

// ---------------------------------------------------------------------------
// predicate (cpu.rs:443-453, json.rs:662-693, ir.rs:66-76)
// ---------------------------------------------------------------------------

template <class R>
WJQ_HD uint32_t slot_rank(R &rd, const SlotV &s) {
  switch (s.kind) {
    case JK_MISSING:
    case JK_NULL: return 0;
    case JK_BOOL: return rd.byte(s.off) == 't' ? 2 : 1;
    default: return s.kind;  // NUM 3, STR 4, ARR 5, OBJ 6
  }
}

// Returns -1/0/1, or 2 when a number could not be decided.
template <class R, class PP, class CC>
WJQ_HD int compare_lit(R &rd, const PP &P, const SlotV &s, const CC &c) {
  uint32_t lrank = c.lit_kind;  // NULL 0, FALSE 1, TRUE 2, NUM 3, STR 4
  uint32_t srank = slot_rank(rd, s);
  if (srank != lrank) return srank < lrank ? -1 : 1;
  if (lrank == WARPJQ_LIT_NUM) {
    NumVal nv;
    if (!parse_number(rd, s.off, s.off + s.len, nv)) return 2;
    if (nv.v < c.lit_num) return -1;
    if (nv.v > c.lit_num) return 1;
    return 0;
  }
  if (lrank == WARPJQ_LIT_STR) {
    bool esc = body_has_escape(rd, s.off + 1, s.off + s.len - 1);
    return decoded_cmp(rd, s.off + 1, s.off + s.len - 1, esc, P.blob + c.lit_off, c.lit_len);
  }
  return 0;
}

WJQ_HD bool op_accepts(uint32_t op, int ord) {
  switch (op) {
    case WARPJQ_OP_EQ: return ord == 0;
    case WARPJQ_OP_NE: return ord != 0;
    case WARPJQ_OP_LT: return ord < 0;
    case WARPJQ_OP_LE: return ord <= 0;
    case WARPJQ_OP_GT: return ord > 0;
    default: return ord >= 0;
  }
}

template <class R>
WJQ_HD bool truthy(R &rd, const SlotV &s) {
  if (s.kind == JK_MISSING || s.kind == JK_NULL) return false;
  return !(s.kind == JK_BOOL && rd.byte(s.off) == 'f');
}

// Evaluates the RPN filter. Returns 0/1, or 2 when undecided. The stack is
// one 64-bit word: build_dprog admits only programs whose RPN never holds
// more than 64 values (reordering deep ones), so `sp` stays in 0..64.
template <int NS, class R, class PP>
WJQ_HD int eval_cond(R &rd, const PP &P, const SlotV (&sl)[NS]) {
  uint64_t stack = 0;
  uint32_t sp = 0;
  for (uint32_t k = 0; k < P.n_cond; k++) {
    const auto &o = P.cond[k];
    uint32_t v;
    if (o.op == WARPJQ_COND_CMP) {
      const auto &c = P.cmps[o.arg];
      uint32_t s = P.slot_of_path[c.path];
      SlotV x = sl[0];
#pragma unroll
      for (int q = 0; q < NS; q++)
        if ((uint32_t)q == s) x = sl[q];
      int ord = compare_lit(rd, P, x, c);
      if (ord == 2) return 2;
      v = op_accepts(c.op, ord) ? 1u : 0u;
    } else if (o.op == WARPJQ_COND_TRUTHY) {
      uint32_t s = P.slot_of_path[o.arg];
      SlotV x = sl[0];
#pragma unroll
      for (int q = 0; q < NS; q++)
        if ((uint32_t)q == s) x = sl[q];
      v = truthy(rd, x) ? 1u : 0u;
    } else if (o.op == WARPJQ_COND_NOT) {
      stack ^= 1ull << (sp - 1);
      continue;
    } else {
      uint32_t b = (uint32_t)(stack >> (sp - 1)) & 1, a = (uint32_t)(stack >> (sp - 2)) & 1;
      sp -= 2;
      v = o.op == WARPJQ_COND_AND ? (a & b) : (a | b);
    }
    stack = (stack & ~(1ull << sp)) | ((uint64_t)v << sp);
    sp++;
  }
  return (int)(stack & 1);
}

// ---------------------------------------------------------------------------
// fast path: one minified flat object (bit-parallel)
//
// Most NDJSON lines are one object whose members are strings and unsigned
// integers, written without whitespace. For such a line the grammar reduces
// to relations between neighbouring bytes, checked 64 bytes at a time on
// bit masks instead of by a byte walk:
//   * each byte is classified by two 16-entry nibble tables (v_perm_b32 on
//     the device) into , : " digit '0' backslash control;
//   * the quote mask's prefix XOR gives the bytes inside strings, so the
//     other masks are taken outside strings only;
//   * every relation is a shift by one byte: after '{' comes '"' or '}',
//     after ',' a key, after ':' a string or a digit run, after a closing
//     quote ':' ',' or '}', after a digit a digit ',' or '}', and so on;
//   * keys and values alternate: adding the key starts to the string mask
//     carries each key's run up to its closing quote ((S + KS) & ~S), and
//     those closing quotes must be exactly the bytes before the colons.
// Any byte outside that shape (whitespace, escapes, signs, fractions,
// literals, nesting, control bytes, a line running past the first
// (FLAT_WORDS - 1) x 64 + 32 bytes from its first byte rounded down to the
// reader's flat_align) makes it return false and the caller walks the line byte by byte,
// which decides everything; when it returns true the line is valid and the
// captures of the fused paths are those walk() would take. Each lane runs it
// on its own line, so there is no carry between lanes.
// ---------------------------------------------------------------------------

#define FC_COMMA 0x01u
#define FC_COLON 0x02u
#define FC_QUOTE 0x04u
#define FC_DIGIT 0x08u
#define FC_ZERO 0x10u
#define FC_BSL 0x20u
#define FC_CTRL 0x40u
#define FLAT_WORDS 5

// Class bits of the four bytes of x (byte k in bits 8k..8k+7): the AND of a
// low-nibble and a high-nibble table entry. Bytes >= 0x80 classify as 0.
WJQ_HD uint32_t fclass4(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  // low nibble: entries 0-7 and 8-15 looked up with its low three bits,
  // chosen per byte by bit 3 (selector 12 gives 0x00, 13 gives 0xFF)
  const uint32_t lo = x & 0x0F0F0F0Fu;
  const uint32_t l7 = lo & 0x07070707u;
  const uint32_t a = __builtin_amdgcn_perm(0x48484848u, 0x484C4858u, l7);
  const uint32_t b = __builtin_amdgcn_perm(0x40404061u, 0x40424848u, l7);
  const uint32_t m = __builtin_amdgcn_perm(0u, 0u, ((lo >> 3) & 0x01010101u) | 0x0C0C0C0Cu);
  const uint32_t lc = (b & m) | (a & ~m);
  // high nibble: 12-15 folded to 8-11, whose selectors give the sign of
  // table bytes 1, 3, 5, 7 (all clear): 0 for every byte >= 0x80
  const uint32_t hi = (x >> 4) & 0x0F0F0F0Fu;
  const uint32_t hs = hi & ~((hi >> 1) & 0x04040404u);
  const uint32_t hc = __builtin_amdgcn_perm(0x00002000u, 0x1A054040u, hs);
  return lc & hc;
#else
  static const uint8_t LO[16] = {0x58, 0x48, 0x4C, 0x48, 0x48, 0x48, 0x48, 0x48,
                                 0x48, 0x48, 0x42, 0x40, 0x61, 0x40, 0x40, 0x40};
  static const uint8_t HI[16] = {0x40, 0x40, 0x05, 0x1A, 0x00, 0x20, 0x00, 0x00,
                                 0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t r = 0;
  for (int k = 0; k < 4; k++) {
    const uint32_t c = (x >> (8 * k)) & 0xFFu;
    r |= (uint32_t)(LO[c & 15] & HI[c >> 4]) << (8 * k);
  }
  return r;
#endif
}

// Bit p of the result: class bit K set in byte p of the first NB x 16 bytes
// of c[0..15] (bits past them zero).
template <uint32_t K, int NB = 4>
WJQ_HD uint64_t fmask(const uint32_t (&c)[16]) {
#if defined(__HIP_DEVICE_COMPILE__)
  constexpr uint32_t k = K == 1u ? 0 : K == 2u ? 1 : K == 4u ? 2 : K == 8u ? 3 : K == 16u ? 4 : K == 32u ? 5 : 6;
  constexpr uint32_t B4 = K * 0x01010101u;
  uint32_t h[4] = {0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < NB; q++) {
    // 16 bytes: two chained dot products per 8 bytes give (pattern << k)
    const uint32_t v0 = __builtin_amdgcn_udot4(c[4 * q + 1] & B4, 0x80402010u,
                                               __builtin_amdgcn_udot4(c[4 * q] & B4, 0x08040201u, 0u, false), false);
    const uint32_t v1 = __builtin_amdgcn_udot4(c[4 * q + 3] & B4, 0x80402010u,
                                               __builtin_amdgcn_udot4(c[4 * q + 2] & B4, 0x08040201u, 0u, false), false);
    h[q] = ((v1 << 8) | v0) >> k;
  }
  return (uint64_t)(h[0] | (h[1] << 16)) | ((uint64_t)(h[2] | (h[3] << 16)) << 32);
#else
  uint64_t m = 0;
  for (int q = 0; q < 4 * NB; q++)
    for (int b = 0; b < 4; b++)
      if ((c[q] >> (8 * b)) & K) m |= 1ull << (4 * q + b);
  return m;
#endif
}

// Bit p: parity of the bits of x at or below p.
WJQ_HD uint64_t prefix_xor64(uint64_t x) {
  x ^= x << 1;
  x ^= x << 2;
  x ^= x << 4;
  x ^= x << 8;
  x ^= x << 16;
  x ^= x << 32;
  return x;
}

// The first set bit at or after bit `from` of the FLAT_WORDS-word mask m
// (position relative to word 0), or 64 * FLAT_WORDS when there is none.
WJQ_HD uint32_t flat_next(const uint64_t (&m)[FLAT_WORDS], uint32_t from) {
  uint32_t r = 64u * FLAT_WORDS;
#pragma unroll
  for (int w = FLAT_WORDS - 1; w >= 0; w--) {
    const uint32_t base = 64u * (uint32_t)w;
    uint64_t x = m[w];
    if (from >= base + 64u) x = 0;
    else if (from > base) x &= ~0ull << (from - base);
    if (x) r = base + (uint32_t)ctz64(x);
  }
  return r;
}

// Readers may declare `flat_align` (4 or 16) and `blk(pos)`, 16 bytes at
// any multiple of it; the flat walk then starts its words at the line's
// first byte rounded down to it, so fewer lines reach a fifth word. Others
// are read in aligned 16-byte blocks.
template <class R, class = void>
struct flat_align : std::integral_constant<uint32_t, 16u> {};
template <class R>
struct flat_align<R, std::void_t<decltype(R::flat_align)>> : std::integral_constant<uint32_t, R::flat_align> {};

template <class R>
WJQ_HD Blk16 flat_blk(R &rd, uint32_t pos) {
  if constexpr (flat_align<R>::value == 16u) return rd.block(pos >> 4);
  else return rd.blk(pos);
}

template <class R, class = void>
struct has_block_reads : std::false_type {};
template <class R>
struct has_block_reads<R, std::void_t<decltype(&R::block)>> : std::true_type {};

// `i` is the line's '{'. True: the line is a valid flat object; the fused
// slots, `te` and `end` are set as walk() sets them. False: nothing is
// written; walk the line.
// One 64-byte word (the last of a line's words holds 32 bytes: NB blocks)
// of flat_walk: classes, masks and the neighbour relations, carrying the
// string parity, the key-run addition and the top bits of the shifted masks
// into the next word.
struct FlatCarry {
  uint64_t bad = 0;
  uint32_t n_other = 0, str_c = 0, add_c = 0, bsl = 0;
  uint32_t pQc = 0, pCM = 0, pCL = 0, pDG = 0, pOP = 0, pKE = 0, pZR = 0;
  uint32_t e_pos = 0;
  bool ended = false;
};

template <int NB, class R>
WJQ_HD void flat_word(R &rd, uint32_t B, uint32_t o0, uint32_t w, FlatCarry &F, uint64_t &KSo, uint64_t &KEo,
                      uint64_t &EDo) {
  uint32_t c[16];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (j < NB) {
      const Blk16 v = flat_blk(rd, B + 64u * w + 16u * (uint32_t)j);
      c[4 * j] = fclass4(v.x);
      c[4 * j + 1] = fclass4(v.y);
      c[4 * j + 2] = fclass4(v.z);
      c[4 * j + 3] = fclass4(v.w);
    } else {
      c[4 * j] = c[4 * j + 1] = c[4 * j + 2] = c[4 * j + 3] = 0;
    }
  }
  uint32_t orc = 0;
#pragma unroll
  for (int j = 0; j < 4 * NB; j++) orc |= c[j];
  F.bsl |= orc & (FC_BSL * 0x01010101u);  // a backslash anywhere near: walk it
  const uint64_t lo = w == 0 ? (~0ull << o0) : ~0ull;
  uint64_t in = lo;
  if (orc & (FC_CTRL * 0x01010101u)) {  // words with no control byte skip this mask
    const uint64_t ct = fmask<FC_CTRL, NB>(c) & lo;
    if (ct) {
      const uint32_t e = (uint32_t)ctz64(ct);
      in &= e ? (~0ull >> (64u - e)) : 0ull;
      F.e_pos = B + 64u * w + e;
      F.ended = true;
    }
  }
  if (NB < 4) in &= (1ull << (16 * NB)) - 1;
  const uint64_t q = fmask<FC_QUOTE, NB>(c) & in;
  const uint64_t S = (prefix_xor64(q) ^ (F.str_c ? ~0ull : 0ull)) & in;
  F.str_c ^= (uint32_t)__builtin_popcountll(q) & 1u;
  const uint64_t Qo = q & S, Qc = q & ~S;
  const uint64_t out = in & ~S & ~q;
  const uint64_t cm = fmask<FC_COMMA, NB>(c) & out, cl = fmask<FC_COLON, NB>(c) & out;
  const uint64_t dg = fmask<FC_DIGIT, NB>(c) & out, zr = fmask<FC_ZERO, NB>(c) & out;
  const uint64_t OP = w == 0 ? (1ull << o0) : 0ull;
  const uint64_t oth = out & ~(cm | cl | dg | OP);  // must be the closing '}' alone
  F.n_other += (uint32_t)__builtin_popcountll(oth);
  const uint64_t PQc = (Qc << 1) | F.pQc, PCM = (cm << 1) | F.pCM, PCL = (cl << 1) | F.pCL;
  const uint64_t PDG = (dg << 1) | F.pDG, POP = (OP << 1) | F.pOP;
  const uint64_t KS = Qo & (PCM | POP);
  uint64_t bad = 0;
  // each key string's run, carried to its closing quote
  const uint64_t t = S + KS;
  const uint32_t c1 = t < S ? 1u : 0u;
  const uint64_t sum = t + F.add_c;
  const uint32_t c2 = sum < t ? 1u : 0u;
  F.add_c = c1 | c2;
  const uint64_t KE = sum & ~S & in;
  const uint64_t PKE = (KE << 1) | F.pKE;
  bad |= (PQc & in) & ~(cl | cm | oth);
  bad |= (PCL & in) & ~(Qo | dg);
  bad |= (PCM & in) & ~Qo;
  bad |= (POP & in) & ~(Qo | oth);
  bad |= (PDG & in) & ~(dg | cm | oth);
  const uint64_t ZRS = zr & ~PDG;  // zeros that start a digit run
  bad |= (((ZRS << 1) | F.pZR) & dg);  // a leading zero followed by a digit
  bad |= (PKE & in) ^ cl;              // colons are exactly the bytes after key ends
  F.bad |= bad;
  KSo = KS;
  KEo = KE;
  EDo = Qc | cm | oth;
  F.pQc = (uint32_t)(Qc >> 63);
  F.pCM = (uint32_t)(cm >> 63);
  F.pCL = (uint32_t)(cl >> 63);
  F.pDG = (uint32_t)(dg >> 63);
  F.pOP = (uint32_t)(OP >> 63);
  F.pKE = (uint32_t)(KE >> 63);
  F.pZR = (uint32_t)(ZRS >> 63);
}

// Bits of the FLAT_WORDS-word mask m shifted down by k (0 < k < 64), word w.
WJQ_HD uint64_t flat_shr(const uint64_t (&m)[FLAT_WORDS], int w, uint32_t k) {
  const uint64_t hi = w + 1 < FLAT_WORDS ? m[w + 1] : 0ull;
  return (m[w] >> k) | (hi << (64u - k));
}

// Records the value after key `kq` (its opening quote, from B; `kl` bytes
// long) for the fused slots in `hit`.
template <int NS, class R, class PP>
WJQ_HD void flat_take(R &rd, const PP &P, uint32_t B, uint32_t ke, uint32_t hit, const uint64_t (&EDw)[FLAT_WORDS],
                      SlotV (&sl)[NS], uint32_t &t_err) {
  const uint32_t vs = ke + 2u;  // after '":'
  const uint32_t ve0 = flat_next(EDw, vs + 1u);
  const bool str = rd.byte(B + vs) == '"';
  const uint32_t ve = str ? ve0 + 1u : ve0;
#pragma unroll
  for (int q = 0; q < NS; q++) {
    if (!((hit >> q) & 1)) continue;
    if (P.spath[q].nsteps == 1) {
      sl[q].off = B + vs;
      sl[q].len = ve - vs;
      sl[q].kind = str ? (uint32_t)JK_STR : (uint32_t)JK_NUM;
      t_err &= ~(1u << q);
    } else {
      sl[q].kind = JK_MISSING;  // a string or number indexed further
      t_err |= 1u << q;
    }
  }
}

// Fused slots among `want` whose first step is the key of `kl` bytes at `ka`.
template <int NS, class R, class PP>
WJQ_HD uint32_t flat_key_hits(R &rd, const PP &P, uint32_t want, uint32_t ka, uint32_t kl) {
  uint32_t hit = 0;
  uint32_t A = want;
  while (A) {
    const uint32_t q = (uint32_t)ctz32(A);
    A &= A - 1;
    const auto &st = P.steps[P.spath[q].step_off];
    if (st.key_len != kl) continue;
    bool same = true;
    for (uint32_t k = 0; k < kl && same; k++) same = rd.byte(ka + k) == P.blob[st.key_off + k];
    if (same) hit |= 1u << q;
  }
  return hit;
}

// Slots among `cand` (plain keys of L <= 16 bytes, SlotPath::kw) equal to
// the L bytes at x, which are read as aligned words.
template <int NS, class R, class PP>
WJQ_HD uint32_t flat_key_eq(R &rd, const PP &P, uint32_t cand, uint32_t x, uint32_t L) {
  const uint32_t nw = (L + 3u) >> 2;
  const uint32_t k0 = x >> 2, sh = x & 3u;
  uint32_t d[5];
#pragma unroll
  for (uint32_t i = 0; i < 5; i++) d[i] = i <= nw ? rd.dword(k0 + i) : 0u;
  uint32_t lw[4];
#pragma unroll
  for (uint32_t i = 0; i < 4; i++) {
    lw[i] = i < nw ? funnel_bytes(d[i + 1], d[i], sh) : 0u;
    if (i + 1u == nw && (L & 3u)) lw[i] &= (1u << (8u * (L & 3u))) - 1u;
  }
  uint32_t hit = 0;
  while (cand) {
    const uint32_t q = (uint32_t)ctz32(cand);
    cand &= cand - 1;
    const auto &sp = P.spath[q];
    const uint32_t diff = (lw[0] ^ sp.kw[0]) | (lw[1] ^ sp.kw[1]) | (lw[2] ^ sp.kw[2]) | (lw[3] ^ sp.kw[3]);
    if (!diff) hit |= 1u << q;
  }
  return hit;
}

// `i` is the line's '{'. True: the line is a valid flat object; the fused
// slots, `te` and `end` are set as walk() sets them. False: nothing is
// written; walk the line.
template <int NS, class R, class PP>
WJQ_HD bool flat_walk(R &rd, uint32_t i, const PP &P, SlotV (&sl)[NS], uint32_t &te, uint32_t &end) {
  const uint32_t B = i & ~(flat_align<R>::value - 1u);  // words start here
  const uint32_t o0 = i - B;
  uint64_t KSw[FLAT_WORDS], EDw[FLAT_WORDS], KEw[FLAT_WORDS];
  FlatCarry F;
#pragma unroll
  for (int w = 0; w < FLAT_WORDS; w++) {
    KSw[w] = EDw[w] = KEw[w] = 0;
    if (F.ended || F.bad || F.bsl || F.n_other > 1) continue;  // decided: valid end, or walk it
    if (w + 1 < FLAT_WORDS) flat_word<4>(rd, B, o0, (uint32_t)w, F, KSw[w], KEw[w], EDw[w]);
    else flat_word<2>(rd, B, o0, (uint32_t)w, F, KSw[w], KEw[w], EDw[w]);
  }
  if (!F.ended || F.bad || F.bsl || F.str_c || F.n_other != 1) return false;
  const uint32_t e_pos = F.e_pos;
  if (rd.byte(e_pos) != '\n' || rd.byte(e_pos - 1) != '}') return false;
  // valid: the root object's fused slots (walk's value_hits at the root)
  uint32_t t_err = 0, want = 0;  // want: fused slots whose first step is a key
#pragma unroll
  for (int q = 0; q < NS; q++) {
    sl[q].off = 0;
    sl[q].len = 0;
    sl[q].kind = JK_MISSING;
    if ((P.fused_mask >> q) & 1) {
      if (P.steps[P.spath[q].step_off].kind == WARPJQ_STEP_KEY) want |= 1u << q;
      else t_err |= 1u << q;
    }
  }
  if (want) {
    if (!(P.key_lens0 >> 31) && !(want & ~P.plain_keys)) {
      // every wanted key is plain and at most 16 bytes: per key length L,
      // key starts whose key end lies L + 1 bytes on, compared as words. A
      // plain key holds no '"', so equal bytes mean that end is the key's
      // own closing quote. Line order within a length (a duplicate key's
      // later value wins); one slot has one length.
      uint32_t lens = P.key_lens0;
      while (lens) {
        const uint32_t L = (uint32_t)ctz32(lens);
        lens &= lens - 1;
        uint32_t wl = 0;
        for (uint32_t A = want; A; A &= A - 1) {
          const uint32_t q = (uint32_t)ctz32(A);
          if (P.spath[q].klen == L) wl |= 1u << q;
        }
#pragma unroll
        for (int w = 0; w < FLAT_WORDS; w++) {
          uint64_t cb = KSw[w] & flat_shr(KEw, w, L + 1u);
          while (cb) {
            const uint32_t kq = 64u * (uint32_t)w + (uint32_t)ctz64(cb);
            cb &= cb - 1;
            const uint32_t hit = flat_key_eq<NS>(rd, P, wl, B + kq + 1u, L);
            if (hit) flat_take<NS>(rd, P, B, kq + L + 1u, hit, EDw, sl, t_err);
          }
        }
      }
    } else if (!(P.key_lens0 >> 31)) {
      // every wanted key is shorter than 31 bytes: candidates by length,
      // key starts whose key end lies len + 1 bytes on (KS & KE >> (len + 1)),
      // in line order (a duplicate key's later value wins)
      uint64_t cand[FLAT_WORDS];
#pragma unroll
      for (int w = 0; w < FLAT_WORDS; w++) cand[w] = 0;
      uint32_t lens = P.key_lens0;
      while (lens) {
        const uint32_t L = (uint32_t)ctz32(lens);
        lens &= lens - 1;
#pragma unroll
        for (int w = 0; w < FLAT_WORDS; w++) cand[w] |= KSw[w] & flat_shr(KEw, w, L + 1u);
      }
#pragma unroll
      for (int w = 0; w < FLAT_WORDS; w++) {
        uint64_t cb = cand[w];
        while (cb) {
          const uint32_t kq = 64u * (uint32_t)w + (uint32_t)ctz64(cb);
          cb &= cb - 1;
          const uint32_t ke = flat_next(KEw, kq + 1u);
          const uint32_t hit = flat_key_hits<NS>(rd, P, want, B + kq + 1u, ke - kq - 1u);
          if (hit) flat_take<NS>(rd, P, B, ke, hit, EDw, sl, t_err);
        }
      }
    } else {
#pragma unroll
      for (int w = 0; w < FLAT_WORDS; w++) {
        uint64_t ks = KSw[w];
        while (ks) {
          const uint32_t kq = 64u * (uint32_t)w + (uint32_t)ctz64(ks);  // opening quote, from B
          ks &= ks - 1;
          const uint32_t ke = flat_next(KEw, kq + 1u);  // closing quote
          const uint32_t kl = ke - kq - 1u;
          if (!((P.key_lens0 >> (kl < 31u ? kl : 31u)) & 1u)) continue;
          const uint32_t hit = flat_key_hits<NS>(rd, P, want, B + kq + 1u, kl);
          if (hit) flat_take<NS>(rd, P, B, ke, hit, EDw, sl, t_err);
        }
      }
    }
  }
  te = t_err;
  end = e_pos;
  return true;
}

// ---------------------------------------------------------------------------
// one line, in the CPU engine's order (cpu.rs:308-338)
// ---------------------------------------------------------------------------

struct LineResult {
  uint32_t status;
  uint32_t selected;  // passed the filter and every post path resolved
  uint32_t end;       // position of the terminator (valid lines)
  uint32_t flat;      // the bit-parallel fast path decided the line
};

template <int NS, class R, class PP>
WJQ_HD LineResult eval_line(R &rd, uint32_t s, const PP &P, SlotV (&sl)[NS], bool try_flat = true) {
  LineResult r{ST_OK, 0, 0, 0};
  uint32_t i = s;
  bool ff = false;
  for (;;) {
    uint32_t c = rd.byte(i);
    if (is_jws(c)) {
      i++;
    } else if (c == 0x0C) {
      ff = true;
      i++;
    } else {
      break;
    }
  }
  if (rd.byte(i) == WJQ_END) {
    r.status = ST_BLANK;
    return r;
  }
  if (ff) {
    r.status = ST_INVALID;
    return r;
  }
#pragma unroll
  for (int q = 0; q < NS; q++) {
    sl[q].off = 0;
    sl[q].len = 0;
    sl[q].kind = JK_MISSING;
  }
  uint32_t te = 0;
  int st;
  if constexpr (has_block_reads<R>::value && !wide_reader<R>::value) {
    r.flat = try_flat && rd.byte(i) == '{' && flat_walk<NS>(rd, i, P, sl, te, r.end);
    st = r.flat ? (int)ST_OK : walk<NS>(rd, i, P, sl, te, r.end);
  } else {
    st = walk<NS>(rd, i, P, sl, te, r.end);
  }
  if (st != ST_OK) {
    r.status = (uint32_t)st;
    return r;
  }
  uint32_t generic = ((1u << P.n_slots) - 1) & ~P.fused_mask;
  if (generic) {
#pragma unroll
    for (int q = 0; q < NS; q++) {
      if (!((generic >> q) & 1)) continue;
      if (!lookup_valid(rd, i, P, P.spath[q], sl[q])) te |= 1u << q;
    }
  }
  if (te & P.filter_mask) {
    r.status = ST_TYPE_ERROR;
    return r;
  }
  if (P.has_filter) {
    int k = eval_cond<NS>(rd, P, sl);
    if (k == 2) {
      r.status = ST_FALLBACK;
      return r;
    }
    if (!k) return r;
  }
  if (te & P.post_mask) {
    r.status = ST_TYPE_ERROR;
    return r;
  }
  r.selected = 1;
  return r;
}

// ---------------------------------------------------------------------------
// rows (output.rs:204-264): the length of the row for a selected line, and
// its bytes, through a byte sink `W` with put(uint8_t).
// ---------------------------------------------------------------------------

WJQ_HD bool csv_special(uint32_t c) { return c == ',' || c == '"' || c == '\n' || c == '\r'; }

// CSV cell for one value (output.rs:65-81). Writes when `w` is non-null;
// returns the length either way.
template <class R, class W>
WJQ_HD uint64_t csv_cell(R &rd, const SlotV &s, W *w) {
  if (s.kind == JK_MISSING || s.kind == JK_NULL) return 0;
  uint32_t a = s.off, b = s.off + s.len;
  bool esc = false;
  if (s.kind == JK_STR) {
    a++;
    b--;
    esc = body_has_escape(rd, a, b);
  }
  // pass 1: length and whether quoting is needed
  uint64_t n = 0, quotes = 0;
  bool q = false;
  {
    DecIter<R> it(rd, a, esc ? b : a);
    if (esc) {
      for (int c; (c = it.next()) >= 0; n++) {
        if (csv_special((uint32_t)c)) q = true;
        if (c == '"') quotes++;
      }
    } else {
      for (uint32_t p = a; p < b; p++, n++) {
        uint32_t c = rd.byte(p);
        if (csv_special(c)) q = true;
        if (c == '"') quotes++;
      }
    }
  }
  uint64_t len = q ? n + quotes + 2 : n;
  if (!w) return len;
  if (q) w->put('"');
  if (esc) {
    DecIter<R> it(rd, a, b);
    for (int c; (c = it.next()) >= 0;) {
      if (q && c == '"') w->put('"');
      w->put((uint8_t)c);
    }
  } else {
    for (uint32_t p = a; p < b; p++) {
      uint32_t c = rd.byte(p);
      if (q && c == '"') w->put('"');
      w->put((uint8_t)c);
    }
  }
  if (q) w->put('"');
  return len;
}

template <int NS, class R, class W, class PP>
WJQ_HD uint64_t put_value(R &rd, const PP &P, const SlotV &s, W *w) {
  if (P.csv_mode) return csv_cell(rd, s, w);
  if (s.kind == JK_MISSING) {
    if (w) {
      w->put('n');
      w->put('u');
      w->put('l');
      w->put('l');
    }
    return 4;
  }
  if (w) w->span(rd, s.off, s.len);
  return s.len;
}

template <int NS>
WJQ_HD SlotV pick(const SlotV (&sl)[NS], uint32_t s) {
  SlotV x = sl[0];
#pragma unroll
  for (int q = 0; q < NS; q++)
    if ((uint32_t)q == s) x = sl[q];
  return x;
}

// Row for a selected line starting at `line` with terminator at `end`.
template <int NS, class R, class W, class PP>
WJQ_HD uint64_t row(R &rd, const PP &P, const SlotV (&sl)[NS], uint32_t line, uint32_t end, W *w) {
  uint64_t n = 0;
  if (P.output_kind == WARPJQ_OUT_PASSTHROUGH) {
    uint32_t e = end;
    if (e > line && rd.byte(e - 1) == '\r') e--;
    if (P.csv_mode) {
      SlotV whole{line, e - line, JK_NUM};  // raw text as one cell
      n = csv_cell(rd, whole, w);
    } else {
      if (w) w->span(rd, line, e - line);
      n = e - line;
    }
  } else if (P.output_kind == WARPJQ_OUT_PATH) {
    n = put_value<NS>(rd, P, pick<NS>(sl, P.output_slot), w);
  } else {
    for (uint32_t f = 0; f < P.n_fields; f++) {
      uint32_t pl = P.prefix_len[f];
      if (w)
        for (uint32_t k = 0; k < pl; k++) w->put(P.blob[P.prefix_off[f] + k]);
      n += pl + put_value<NS>(rd, P, pick<NS>(sl, P.field_slot[f]), w);
    }
    if (w)
      for (uint32_t k = 0; k < P.suffix_len; k++) w->put(P.blob[P.suffix_off + k]);
    n += P.suffix_len;
  }
  if (w) w->put('\n');
  return n + 1;
}

// Group key of a slot (agg.rs:121-144 and gpu/mod.rs:1198-1221): Missing and
// null are one group; strings compare by decoded body; everything else by
// raw bytes. `off/len` are the bytes that identify it.
struct GKey {
  uint32_t off, len, kind, esc;
};

template <class R>
WJQ_HD GKey group_key(R &rd, const SlotV &s) {
  GKey k;
  if (s.kind == JK_MISSING || s.kind == JK_NULL) {
    k.off = 0;
    k.len = 0;
    k.kind = JK_NULL;
    k.esc = 0;
  } else if (s.kind == JK_STR) {
    k.off = s.off + 1;
    k.len = s.len - 2;
    k.kind = JK_STR;
    k.esc = body_has_escape(rd, k.off, k.off + k.len) ? 1 : 0;
  } else {
    // bools compare by length (true 4, false 5); numbers and composites by
    // raw bytes, arrays and objects forming one composite class
    k.off = s.off;
    k.len = s.len;
    k.kind = s.kind == JK_OBJ ? (uint32_t)JK_ARR : s.kind;
    k.esc = 0;
  }
  return k;
}

template <class R>
WJQ_HD uint32_t group_hash(R &rd, const GKey &k) {
  uint32_t h;
  if (k.kind == JK_STR) h = decoded_hash(rd, k.off, k.off + k.len, k.esc != 0, nullptr);
  else if (k.kind == JK_NULL || k.kind == JK_BOOL) h = 0x811c9dc5u ^ k.len;
  else h = decoded_hash(rd, k.off, k.off + k.len, false, nullptr);
  h ^= k.kind * 0x9E3779B9u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return h | 1u;  // 0 is reserved for "not yet written"
}

// Equality of two group keys from the same chunk; `ra` reads key `a` and `rb`
// reads key `b` (the kernels read the current line from LDS and a stored key
// from HBM).
template <class RA, class RB>
WJQ_HD bool group_eq(RA &ra, const GKey &a, RB &rb, const GKey &b) {
  if (a.kind != b.kind) return false;
  if (a.kind == JK_NULL) return true;
  if (a.kind == JK_BOOL) return a.len == b.len;
  if (a.kind == JK_STR && (a.esc || b.esc)) {
    DecIter<RA> x(ra, a.off, a.off + a.len);
    DecIter<RB> y(rb, b.off, b.off + b.len);
    for (;;) {
      int c = x.next(), e = y.next();
      if (c != e) return false;
      if (c < 0) return true;
    }
  }
  if (a.len != b.len) return false;
  for (uint32_t q = 0; q < a.len; q++)
    if (ra.byte(a.off + q) != rb.byte(b.off + q)) return false;
  return true;
}

}  // namespace wjqd
