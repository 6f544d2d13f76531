// ORACLE — test infrastructure only.
//
// A CPU restatement of warpjq's CPU backend, the reference's own correctness
// oracle for the GPU path (crates/warpjq-core/src/exec/cpu.rs:1-13). It is
// the checker for the HIP kernels in warpjq_amd/ and nothing else: only the
// tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
// The product never links or calls it.
//
// Restated from (file:line in the reference):
//   scanner / lookup / compare   crates/warpjq-core/src/json.rs:122-697
//   aggregates / group keys      crates/warpjq-core/src/agg.rs:15-278
//   writer and escaping          crates/warpjq-core/src/output.rs:23-352
//   chunker and Lines            crates/warpjq-core/src/chunk.rs:164-455
//   CPU engine                   crates/warpjq-core/src/exec/cpu.rs:87-453
//   run_bytes / finish_totals    crates/warpjq-core/src/exec/mod.rs:272-339
// The query front end (text -> Program) is warpjq_amd/csrc/query.cpp, pinned
// on its own by the reference's lexer/parser/lowering unit tests.
//
// Pinning: tests/test_oracle_*.py check every function here against the
// golden vectors of the reference's tests (conformance lists, json.rs,
// agg.rs, output.rs, chunk.rs and cpu.rs unit tests), against brute force on
// tiny inputs and against metamorphic invariants. The one part that nothing
// in the reference pins beyond agg.rs:401-410 is ryu's decimal layout for
// non-integral aggregates (format_number); it is restated from ryu 1.0.23's
// documented layout and is marked "parity unpinned" beyond those cases.
#include <algorithm>
#include <atomic>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <locale.h>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../warpjq_amd/csrc/query.h"

using wjq::AggKind;
using wjq::Cond;
using wjq::Literal;
using wjq::Output;
using wjq::PathId;
using wjq::Program;
using wjq::Step;

namespace orc {

// ---------------------------------------------------------------------------
// json.rs
// ---------------------------------------------------------------------------

enum Kind : uint8_t { K_MISSING = 0, K_NULL = 1, K_BOOL = 2, K_NUM = 3, K_STR = 4, K_ARR = 5, K_OBJ = 6 };

struct Slot {
  Kind kind = K_MISSING;
  const uint8_t *raw = (const uint8_t *)"null";  // MISSING renders as null
  size_t len = 4;
};

struct JErr {
  size_t at = 0;
  std::string detail;
};

// json.rs:30-45
static int kind_rank(Kind k, const uint8_t *raw, size_t len) {
  switch (k) {
    case K_MISSING:
    case K_NULL: return 0;
    case K_BOOL: return (len == 4 && memcmp(raw, "true", 4) == 0) ? 2 : 1;
    case K_NUM: return 3;
    case K_STR: return 4;
    case K_ARR: return 5;
    case K_OBJ: return 6;
  }
  return 0;
}

static size_t skip_ws(const uint8_t *b, size_t n, size_t i) {  // json.rs:122-128
  while (i < n && (b[i] == ' ' || b[i] == '\t' || b[i] == '\n' || b[i] == '\r')) i++;
  return i;
}

static bool is_hex(uint8_t c) {
  return (c >= '0' && c <= '9') || ((c | 0x20) >= 'a' && (c | 0x20) <= 'f');
}

// json.rs:336-368
static bool skip_string(const uint8_t *b, size_t n, size_t i, size_t &out, JErr &e) {
  size_t j = i + 1;
  while (j < n) {
    uint8_t c = b[j];
    if (c == '"') {
      out = j + 1;
      return true;
    }
    if (c == '\\') {
      if (j + 1 >= n) {
        e = {j, "unterminated escape"};
        return false;
      }
      uint8_t x = b[j + 1];
      switch (x) {
        case '"': case '\\': case '/': case 'b': case 'f': case 'n': case 'r': case 't':
          j += 2;
          break;
        case 'u':
          if (j + 6 > n || !is_hex(b[j + 2]) || !is_hex(b[j + 3]) || !is_hex(b[j + 4]) || !is_hex(b[j + 5])) {
            e = {j, "malformed `\\u` escape"};
            return false;
          }
          j += 6;
          break;
        default:
          e = {j, std::string("invalid escape `\\") + (char)x + "`"};
          return false;
      }
      continue;
    }
    if (c < 0x20) {
      e = {j, "unescaped control character in string"};
      return false;
    }
    j++;
  }
  e = {i, "unterminated string"};
  return false;
}

// json.rs:370-410
static bool skip_number(const uint8_t *b, size_t n, size_t i, size_t &out, JErr &e) {
  size_t j = i;
  if (b[j] == '-') j++;
  size_t int_start = j;
  while (j < n && b[j] >= '0' && b[j] <= '9') j++;
  if (j == int_start) {
    e = {i, "number with no digits"};
    return false;
  }
  if (b[int_start] == '0' && j - int_start > 1) {
    e = {i, "number has a leading zero"};
    return false;
  }
  if (j < n && b[j] == '.') {
    j++;
    size_t fs = j;
    while (j < n && b[j] >= '0' && b[j] <= '9') j++;
    if (j == fs) {
      e = {j, "number has no digits after `.`"};
      return false;
    }
  }
  if (j < n && (b[j] == 'e' || b[j] == 'E')) {
    j++;
    if (j < n && (b[j] == '+' || b[j] == '-')) j++;
    size_t es = j;
    while (j < n && b[j] >= '0' && b[j] <= '9') j++;
    if (j == es) {
      e = {j, "number has no digits in its exponent"};
      return false;
    }
  }
  out = j;
  return true;
}

// json.rs:325-334
static bool expect_lit(const uint8_t *b, size_t n, size_t i, const char *lit, size_t &out, JErr &e) {
  size_t l = strlen(lit);
  if (n >= i + l && memcmp(b + i, lit, l) == 0) {
    out = i + l;
    return true;
  }
  e = {i, std::string("expected `") + lit + "`"};
  return false;
}

// json.rs:213-323 — iterative, no depth limit.
static bool skip_value(const uint8_t *b, size_t n, size_t start, size_t &out, JErr &e) {
  enum { VALUE, KEY, AFTER } state = VALUE;
  size_t i = start;
  std::vector<uint8_t> depth;  // 1 = array
  for (;;) {
    if (state == VALUE) {
      i = skip_ws(b, n, i);
      if (i >= n) {
        e = {i, "unexpected end of input"};
        return false;
      }
      uint8_t c = b[i];
      switch (c) {
        case '{':
          i = skip_ws(b, n, i + 1);
          if (i < n && b[i] == '}') {
            i++;
            state = AFTER;
          } else {
            depth.push_back(0);
            state = KEY;
          }
          break;
        case '[':
          i = skip_ws(b, n, i + 1);
          if (i < n && b[i] == ']') {
            i++;
            state = AFTER;
          } else {
            depth.push_back(1);
            state = VALUE;
          }
          break;
        case '"':
          if (!skip_string(b, n, i, i, e)) return false;
          state = AFTER;
          break;
        case 't':
          if (!expect_lit(b, n, i, "true", i, e)) return false;
          state = AFTER;
          break;
        case 'f':
          if (!expect_lit(b, n, i, "false", i, e)) return false;
          state = AFTER;
          break;
        case 'n':
          if (!expect_lit(b, n, i, "null", i, e)) return false;
          state = AFTER;
          break;
        default:
          if (c == '-' || (c >= '0' && c <= '9')) {
            if (!skip_number(b, n, i, i, e)) return false;
            state = AFTER;
          } else {
            e = {i, std::string("unexpected character `") + (char)c + "`"};
            return false;
          }
      }
    } else if (state == KEY) {
      i = skip_ws(b, n, i);
      if (i >= n || b[i] != '"') {
        e = {i, "expected a string key"};
        return false;
      }
      if (!skip_string(b, n, i, i, e)) return false;
      i = skip_ws(b, n, i);
      if (i >= n || b[i] != ':') {
        e = {i, "expected `:` after object key"};
        return false;
      }
      i++;
      state = VALUE;
    } else {
      if (depth.empty()) {
        out = i;
        return true;
      }
      i = skip_ws(b, n, i);
      if (i >= n) {
        e = {i, "unexpected end of input"};
        return false;
      }
      uint8_t c = b[i];
      bool arr = depth.back() != 0;
      if (c == ',' && arr) {
        i++;
        state = VALUE;
      } else if (c == ',' && !arr) {
        i++;
        state = KEY;
      } else if ((c == '}' && !arr) || (c == ']' && arr)) {
        i++;
        depth.pop_back();
      } else {
        e = {i, arr ? "expected `,` or `]`" : "expected `,` or `}`"};
        return false;
      }
    }
  }
}

// json.rs:414-421
static bool validate(const uint8_t *b, size_t n, JErr &e) {
  size_t end;
  if (!skip_value(b, n, 0, end, e)) return false;
  size_t rest = skip_ws(b, n, end);
  if (rest != n) {
    e = {rest, "trailing content after JSON value"};
    return false;
  }
  return true;
}

static Kind kind_of(const uint8_t *b, size_t i) {  // json.rs:423-432
  switch (b[i]) {
    case '{': return K_OBJ;
    case '[': return K_ARR;
    case '"': return K_STR;
    case 't':
    case 'f': return K_BOOL;
    case 'n': return K_NULL;
    default: return K_NUM;
  }
}

static void push_utf8(std::string &s, uint32_t cp) {
  if (cp < 0x80) {
    s.push_back((char)cp);
  } else if (cp < 0x800) {
    s.push_back((char)(0xC0 | (cp >> 6)));
    s.push_back((char)(0x80 | (cp & 0x3F)));
  } else if (cp < 0x10000) {
    s.push_back((char)(0xE0 | (cp >> 12)));
    s.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
    s.push_back((char)(0x80 | (cp & 0x3F)));
  } else {
    s.push_back((char)(0xF0 | (cp >> 18)));
    s.push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
    s.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
    s.push_back((char)(0x80 | (cp & 0x3F)));
  }
}

// u32::from_str_radix(.., 16) over exactly the given bytes.
static bool radix16(const uint8_t *p, size_t n, uint32_t &v) {
  size_t i = 0;
  if (n > 0 && p[0] == '+') i = 1;
  if (i >= n) return false;
  uint32_t acc = 0;
  for (; i < n; i++) {
    if (!is_hex(p[i])) return false;
    uint8_t c = p[i];
    acc = acc * 16 + ((c >= '0' && c <= '9') ? c - '0' : (c | 0x20) - 'a' + 10);
  }
  v = acc;
  return true;
}

// json.rs:594-654. A lone surrogate becomes U+FFFD.
static bool unescape_into(const uint8_t *raw, size_t n, std::string &out) {
  size_t i = 0;
  while (i < n) {
    if (raw[i] != '\\') {
      out.push_back((char)raw[i]);
      i++;
      continue;
    }
    if (i + 1 >= n) return false;
    uint8_t e = raw[i + 1];
    switch (e) {
      case '"': out.push_back('"'); break;
      case '\\': out.push_back('\\'); break;
      case '/': out.push_back('/'); break;
      case 'b': out.push_back('\x08'); break;
      case 'f': out.push_back('\x0c'); break;
      case 'n': out.push_back('\n'); break;
      case 'r': out.push_back('\r'); break;
      case 't': out.push_back('\t'); break;
      case 'u': {
        if (i + 6 > n) return false;
        uint32_t cp;
        if (!radix16(raw + i + 2, 4, cp)) return false;
        i += 6;
        if (cp >= 0xD800 && cp < 0xDC00 && i + 1 < n && raw[i] == '\\' && raw[i + 1] == 'u' && i + 6 <= n) {
          uint32_t lo;
          if (radix16(raw + i + 2, 4, lo) && lo >= 0xDC00 && lo < 0xE000) {
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
            i += 6;
          }
        }
        if ((cp >= 0xD800 && cp < 0xE000) || cp > 0x10FFFF) cp = 0xFFFD;
        push_utf8(out, cp);
        continue;
      }
      default: return false;
    }
    i += 2;
  }
  return true;
}

// json.rs:582-591
static bool key_matches(const uint8_t *raw, size_t n, const std::string &key) {
  if (!memchr(raw, '\\', n)) return n == key.size() && memcmp(raw, key.data(), n) == 0;
  std::string buf;
  if (!unescape_into(raw, n, buf)) return false;
  return buf == key;
}

// json.rs:523-556 — last duplicate wins.
static int object_get(const uint8_t *b, size_t n, size_t i, const std::string &key, size_t &vs, size_t &ve, JErr &e) {
  size_t j = skip_ws(b, n, i + 1);
  if (j < n && b[j] == '}') return 0;
  bool found = false;
  for (;;) {
    j = skip_ws(b, n, j);
    if (j >= n || b[j] != '"') {
      e = {j, "expected a string key"};
      return -1;
    }
    size_t ks = j;
    if (!skip_string(b, n, j, j, e)) return -1;
    const uint8_t *kraw = b + ks + 1;
    size_t klen = j - 1 - (ks + 1);
    j = skip_ws(b, n, j);
    if (j >= n || b[j] != ':') {
      e = {j, "expected `:` after object key"};
      return -1;
    }
    size_t s = skip_ws(b, n, j + 1);
    size_t end;
    if (!skip_value(b, n, s, end, e)) return -1;
    if (key_matches(kraw, klen, key)) {
      found = true;
      vs = s;
      ve = end;
    }
    j = skip_ws(b, n, end);
    if (j < n && b[j] == ',') {
      j++;
    } else if (j < n && b[j] == '}') {
      return found ? 1 : 0;
    } else {
      e = {j, "expected `,` or `}`"};
      return -1;
    }
  }
}

// json.rs:558-578
static int array_get(const uint8_t *b, size_t n, size_t i, uint32_t idx, size_t &vs, size_t &ve, JErr &e) {
  size_t j = skip_ws(b, n, i + 1);
  if (j < n && b[j] == ']') return 0;
  uint32_t k = 0;
  for (;;) {
    size_t s = skip_ws(b, n, j);
    size_t end;
    if (!skip_value(b, n, s, end, e)) return -1;
    if (k == idx) {
      vs = s;
      ve = end;
      return 1;
    }
    k++;
    j = skip_ws(b, n, end);
    if (j < n && b[j] == ',') {
      j++;
    } else if (j < n && b[j] == ']') {
      return 0;
    } else {
      e = {j, "expected `,` or `]`"};
      return -1;
    }
  }
}

enum LookupRc { L_FOUND = 0, L_TYPE_ERROR = 1, L_INVALID = 2 };

static const char *type_name(Kind k) {
  switch (k) {
    case K_MISSING:
    case K_NULL: return "null";
    case K_BOOL: return "boolean";
    case K_NUM: return "number";
    case K_STR: return "string";
    case K_ARR: return "array";
    case K_OBJ: return "object";
  }
  return "?";
}

// json.rs:444-508
static LookupRc lookup(const uint8_t *b, size_t n, const std::vector<Step> &steps, Slot &out, std::string &msg) {
  size_t start = skip_ws(b, n, 0);
  if (start >= n) {
    msg = "empty value at byte 0";
    return L_INVALID;
  }
  Kind kind = kind_of(b, start);
  size_t end = n;
  JErr e;
  if (steps.empty()) {
    if (!skip_value(b, n, start, end, e)) {
      msg = e.detail;
      return L_INVALID;
    }
  }
  for (const Step &st : steps) {
    if (kind == K_MISSING || kind == K_NULL) {
      out = Slot{};
      return L_FOUND;
    }
    size_t s, en;
    int rc;
    if (!st.is_index) {
      if (kind != K_OBJ) {
        msg = std::string("cannot index ") + type_name(kind) + " with \"" + st.key + "\"";
        return L_TYPE_ERROR;
      }
      rc = object_get(b, n, start, st.key, s, en, e);
    } else {
      if (kind != K_ARR) {
        msg = std::string("cannot index ") + type_name(kind) + " with number";
        return L_TYPE_ERROR;
      }
      rc = array_get(b, n, start, st.index, s, en, e);
    }
    if (rc < 0) {
      msg = e.detail;
      return L_INVALID;
    }
    if (rc == 0) {
      out = Slot{};
      return L_FOUND;
    }
    start = s;
    end = en;
    kind = kind_of(b, s);
  }
  out.kind = kind;
  out.raw = b + start;
  out.len = end - start;
  return L_FOUND;
}

// Slot::as_f64 (json.rs:73-78): Rust's correctly rounded f64::from_str on
// the raw token. glibc strtod in the "C" locale has the same contract.
static locale_t c_locale() {
  static locale_t l = newlocale(LC_NUMERIC_MASK, "C", (locale_t)0);
  return l;
}
static double parse_num(const uint8_t *p, size_t n) {
  char buf[64];
  if (n < sizeof buf) {
    memcpy(buf, p, n);
    buf[n] = 0;
    return strtod_l(buf, nullptr, c_locale());
  }
  std::string s((const char *)p, n);
  return strtod_l(s.c_str(), nullptr, c_locale());
}

static bool is_truthy(const Slot &s) {  // json.rs:63-65
  if (s.kind == K_MISSING || s.kind == K_NULL) return false;
  return !(s.len == 5 && memcmp(s.raw, "false", 5) == 0);
}

static int bytes_cmp(const uint8_t *a, size_t na, const uint8_t *b, size_t nb) {
  size_t m = na < nb ? na : nb;
  int c = m ? memcmp(a, b, m) : 0;
  if (c != 0) return c < 0 ? -1 : 1;
  return na < nb ? -1 : (na > nb ? 1 : 0);
}

// json.rs:662-693
static int compare(const Slot &s, const Literal &lit) {
  int lrank = lit.rank();
  int srank = kind_rank(s.kind, s.raw, s.len);
  if (srank != lrank) return srank < lrank ? -1 : 1;
  switch (lit.kind) {
    case Literal::NUL:
    case Literal::BOOL: return 0;
    case Literal::NUM: {
      if (s.kind != K_NUM) return 0;
      double v = parse_num(s.raw, s.len);
      if (v < lit.num) return -1;
      if (v > lit.num) return 1;
      return 0;  // equal, or NaN's partial_cmp None -> Equal
    }
    case Literal::STR: {
      const uint8_t *inner = s.raw + 1;
      size_t n = s.len - 2;
      if (!memchr(inner, '\\', n))
        return bytes_cmp(inner, n, (const uint8_t *)lit.str.data(), lit.str.size());
      std::string buf;
      if (!unescape_into(inner, n, buf)) return 0;
      return bytes_cmp((const uint8_t *)buf.data(), buf.size(), (const uint8_t *)lit.str.data(), lit.str.size());
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// agg.rs
// ---------------------------------------------------------------------------

struct AggState {  // agg.rs:15-98
  uint64_t count = 0, numeric = 0;
  double sum = 0.0, min = INFINITY, max = -INFINITY;
  bool saw_non_numeric = false;

  void push_count() { count++; }
  void push_number(double v) {
    numeric++;
    sum += v;
    if (v < min) min = v;
    if (v > max) max = v;
  }
  void push_value(const Slot &s) {
    count++;
    switch (s.kind) {
      case K_MISSING:
      case K_NULL: break;
      case K_NUM: push_number(parse_num(s.raw, s.len)); break;
      default: saw_non_numeric = true;
    }
  }
  void merge(const AggState &o) {
    count += o.count;
    numeric += o.numeric;
    sum += o.sum;
    if (o.min < min) min = o.min;
    if (o.max > max) max = o.max;
    saw_non_numeric |= o.saw_non_numeric;
  }
  bool finish(AggKind k, double &v) const {
    switch (k) {
      case AggKind::Count: v = (double)count; return true;
      case AggKind::Sum: v = sum; return true;
      case AggKind::Min: v = min; return numeric > 0;
      case AggKind::Max: v = max; return numeric > 0;
      case AggKind::Avg: v = sum / (double)numeric; return numeric > 0;
    }
    return false;
  }
};

static uint64_t order_bits(double v) {  // agg.rs:205-212
  uint64_t b;
  memcpy(&b, &v, 8);
  return (b >> 63) ? ~b : (b ^ (1ull << 63));
}

// agg.rs:106-195. rank: 0 null, 1 false, 2 true, 3 number, 4 string,
// 5 composite. Equality is structural (rank + bytes); the jq order compares
// numbers by value bits, so `1` and `1.0` are distinct groups that sort as
// equal — the map below breaks that tie by raw bytes, which the reference
// leaves to HashMap iteration order (parity unpinned for such ties).
struct GroupKey {
  int rank = 0;
  std::string bytes;  // raw number text / decoded string / composite raw
  uint64_t bits = 0;
};

struct GroupKeyLess {
  bool operator()(const GroupKey &a, const GroupKey &b) const {
    if (a.rank != b.rank) return a.rank < b.rank;
    if (a.rank == 3 && a.bits != b.bits) return a.bits < b.bits;
    return a.bytes < b.bytes;  // std::string compares as unsigned bytes
  }
};

static GroupKey group_key_from_slot(const Slot &s) {
  GroupKey k;
  switch (s.kind) {
    case K_MISSING:
    case K_NULL: k.rank = 0; break;
    case K_BOOL: k.rank = (s.len == 4 && memcmp(s.raw, "true", 4) == 0) ? 2 : 1; break;
    case K_NUM:
      k.rank = 3;
      k.bytes.assign((const char *)s.raw, s.len);
      k.bits = order_bits(parse_num(s.raw, s.len));
      break;
    case K_STR: {
      k.rank = 4;
      const uint8_t *inner = s.raw + 1;
      size_t n = s.len - 2;
      if (memchr(inner, '\\', n)) {
        std::string buf;
        if (unescape_into(inner, n, buf)) {
          k.bytes = buf;
          break;
        }
      }
      k.bytes.assign((const char *)inner, n);
      break;
    }
    case K_ARR:
    case K_OBJ:
      k.rank = 5;
      k.bytes.assign((const char *)s.raw, s.len);
      break;
  }
  return k;
}

using GroupAccumulator = std::map<GroupKey, AggState, GroupKeyLess>;

// ryu 1.0.23 layout (ryu/src/pretty/mod.rs format64) applied to the
// shortest round-trip digits, which std::to_chars produces.
static std::string ryu_format(double v) {
  if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific);
  std::string s(buf, r.ptr);
  bool neg = false;
  size_t p = 0;
  if (s[0] == '-') {
    neg = true;
    p = 1;
  }
  size_t epos = s.find('e');
  std::string mant = s.substr(p, epos - p);
  int exp10 = atoi(s.c_str() + epos + 1);
  std::string digits;
  for (char c : mant)
    if (c != '.') digits.push_back(c);
  int length = (int)digits.size();
  int k = exp10 - (length - 1);
  int kk = length + k;
  std::string out = neg ? "-" : "";
  if (k >= 0 && kk <= 16) {
    out += digits + std::string(kk - length, '0') + ".0";
  } else if (kk > 0 && kk <= 16) {
    out += digits.substr(0, kk) + "." + digits.substr(kk);
  } else if (kk > -5 && kk <= 0) {
    out += "0." + std::string(-kk, '0') + digits;
  } else if (length == 1) {
    out += digits + "e" + std::to_string(kk - 1);
  } else {
    out += digits.substr(0, 1) + "." + digits.substr(1) + "e" + std::to_string(kk - 1);
  }
  return out;
}

// agg.rs:253-278
static std::string format_number(double v) {
  if (std::isnan(v)) return "null";
  if (std::isinf(v)) return v > 0 ? "1.7976931348623157e+308" : "-1.7976931348623157e+308";
  if (std::trunc(v) == v && std::fabs(v) < 9007199254740992.0) {
    return std::to_string((long long)v);
  }
  std::string s = ryu_format(v);
  size_t e = s.find('e');
  if (e != std::string::npos && s[e + 1] != '-') s.insert(e + 1, "+");
  return s;
}

// ---------------------------------------------------------------------------
// output.rs
// ---------------------------------------------------------------------------

enum Format { F_NDJSON = 0, F_CSV = 1, F_COUNT = 2 };

static void write_csv_cell(std::string &out, const uint8_t *p, size_t n) {  // output.rs:45-61
  bool q = false;
  for (size_t i = 0; i < n; i++)
    if (p[i] == ',' || p[i] == '"' || p[i] == '\n' || p[i] == '\r') {
      q = true;
      break;
    }
  if (!q) {
    out.append((const char *)p, n);
    return;
  }
  out.push_back('"');
  for (size_t i = 0; i < n; i++) {
    if (p[i] == '"') out.push_back('"');
    out.push_back((char)p[i]);
  }
  out.push_back('"');
}

static void csv_value(std::string &out, const Slot &s) {  // output.rs:65-81
  switch (s.kind) {
    case K_MISSING:
    case K_NULL: return;
    case K_STR: {
      const uint8_t *inner = s.raw + 1;
      size_t n = s.len - 2;
      if (memchr(inner, '\\', n)) {
        std::string buf;
        if (unescape_into(inner, n, buf)) {
          write_csv_cell(out, (const uint8_t *)buf.data(), buf.size());
          return;
        }
      }
      write_csv_cell(out, inner, n);
      return;
    }
    default: write_csv_cell(out, s.raw, s.len);
  }
}

struct Writer {  // output.rs:87-334 (buffering is irrelevant to the bytes)
  std::string buf;
  Format format = F_NDJSON;
  bool header_written = false;
  uint64_t rows = 0;

  void write_raw(const std::string &b, uint64_t r) {
    rows += r;
    if (format == F_COUNT) return;
    buf += b;
  }
  void write_header(const Program &p) {
    if (format != F_CSV || header_written) return;
    header_written = true;
    auto cols = p.csv_headers();
    if (!cols) return;
    for (size_t i = 0; i < cols->size(); i++) {
      if (i) buf.push_back(',');
      write_csv_cell(buf, (const uint8_t *)(*cols)[i].data(), (*cols)[i].size());
    }
    buf.push_back('\n');
  }
  void passthrough(const uint8_t *line, size_t n) {
    rows++;
    if (format == F_COUNT) return;
    if (format == F_CSV) write_csv_cell(buf, line, n);
    else buf.append((const char *)line, n);
    buf.push_back('\n');
  }
  void value(const Slot &s) {
    rows++;
    if (format == F_COUNT) return;
    if (format == F_CSV) csv_value(buf, s);
    else buf.append((const char *)s.raw, s.len);
    buf.push_back('\n');
  }
  void projection(const std::vector<std::string> &keys, const std::vector<Slot> &slots) {
    rows++;
    if (format == F_COUNT) return;
    if (format == F_CSV) {
      for (size_t i = 0; i < slots.size(); i++) {
        if (i) buf.push_back(',');
        csv_value(buf, slots[i]);
      }
    } else {
      buf.push_back('{');
      for (size_t i = 0; i < keys.size(); i++) {
        if (i) buf.push_back(',');
        wjq::write_json_string(buf, (const uint8_t *)keys[i].data(), keys[i].size());
        buf.push_back(':');
        buf.append((const char *)slots[i].raw, slots[i].len);
      }
      buf.push_back('}');
    }
    buf.push_back('\n');
  }
  void aggregate(AggKind kind, const AggState &st) {
    rows++;
    double v;
    std::string value = st.finish(kind, v) ? format_number(v) : "null";
    if (format == F_COUNT) {
      rows = 0;
      buf = value + "\n";
      format = F_NDJSON;
      return;
    }
    buf += value;
    buf.push_back('\n');
  }
  void groups(const std::string &key_name, AggKind kind, const GroupAccumulator &acc) {
    for (auto &kv : acc) {
      rows++;
      double v;
      std::string value = kv.second.finish(kind, v) ? format_number(v) : "null";
      const GroupKey &k = kv.first;
      if (format == F_CSV) {
        std::string plain = k.rank == 4 ? k.bytes : group_key_json(k);
        write_csv_cell(buf, (const uint8_t *)plain.data(), plain.size());
        buf.push_back(',');
        buf += value;
      } else {
        buf.push_back('{');
        wjq::write_json_string(buf, (const uint8_t *)key_name.data(), key_name.size());
        buf.push_back(':');
        buf += group_key_json(k);
        buf.push_back(',');
        const char *kn = wjq::agg_name(kind);
        wjq::write_json_string(buf, (const uint8_t *)kn, strlen(kn));
        buf.push_back(':');
        buf += value;
        buf.push_back('}');
      }
      buf.push_back('\n');
    }
  }
  // agg.rs:158-171
  static std::string group_key_json(const GroupKey &k) {
    switch (k.rank) {
      case 0: return "null";
      case 1: return "false";
      case 2: return "true";
      case 3: return k.bytes;
      case 4: {
        std::string s;
        wjq::write_json_string(s, (const uint8_t *)k.bytes.data(), k.bytes.size());
        return s;
      }
      default: return k.bytes;
    }
  }
  std::string finish() {
    if (format == F_COUNT) buf = std::to_string(rows) + "\n";
    return buf;
  }
};

static std::string group_key_name(const Program &p) {  // output.rs:338-347
  if (!p.group_by) return "key";
  const auto &steps = p.path(*p.group_by).steps;
  if (steps.empty()) return "key";
  const Step &s = steps.back();
  if (s.is_index) return "[" + std::to_string(s.index) + "]";
  return s.key;
}

// ---------------------------------------------------------------------------
// chunk.rs
// ---------------------------------------------------------------------------

struct Chunk {
  const uint8_t *data;
  size_t len;
  uint64_t first_line;
};

static bool is_ascii_ws(uint8_t c) {
  return c == ' ' || c == '\t' || c == '\n' || c == 0x0C || c == '\r';
}

// Lines::next (chunk.rs:415-444): split on '\n', trim one '\r', skip lines
// made only of ASCII whitespace (which includes 0x0C) but keep numbering.
template <class F>
static void for_each_line(const Chunk &c, F &&f) {
  size_t pos = 0;
  uint64_t no = c.first_line;
  while (pos < c.len) {
    const uint8_t *nl = (const uint8_t *)memchr(c.data + pos, '\n', c.len - pos);
    size_t end = nl ? (size_t)(nl - c.data) : c.len;
    size_t start = pos;
    pos = end + 1;
    uint64_t number = no++;
    size_t len = end - start;
    if (len > 0 && c.data[start + len - 1] == '\r') len--;
    bool blank = true;
    for (size_t i = 0; i < len; i++)
      if (!is_ascii_ws(c.data[start + i])) {
        blank = false;
        break;
      }
    if (blank) continue;
    if (!f(c.data + start, len, number)) return;
  }
}

struct ChunkError {
  std::string msg;
};

// Chunker::next_chunk (chunk.rs:234-278) over a mapped buffer.
template <class F>
static bool mapped_chunks(const uint8_t *data, size_t n, size_t target, size_t max_line, F &&f, ChunkError &err) {
  if (n >= 3 && data[0] == 0xEF && data[1] == 0xBB && data[2] == 0xBF) {
    data += 3;
    n -= 3;
  }
  if (target < 1) target = 1;
  size_t pos = 0;
  uint64_t line_no = 1;
  while (pos < n) {
    size_t start = pos;
    size_t nominal_end = start + target < n ? start + target : n;
    size_t end;
    if (nominal_end == n) {
      end = n;
    } else {
      size_t rel = SIZE_MAX;
      for (size_t i = nominal_end; i > start; i--)
        if (data[i - 1] == '\n') {
          rel = i - 1;
          break;
        }
      if (rel != SIZE_MAX) {
        end = rel + 1;
      } else {
        const uint8_t *nl = (const uint8_t *)memchr(data + nominal_end, '\n', n - nominal_end);
        end = nl ? (size_t)(nl - data) + 1 : n;
      }
    }
    if (end - start > max_line) {
      size_t lim = (end > start + 1) ? end - 1 : start;
      bool has_nl = lim > start && memchr(data + start, '\n', lim - start) != nullptr;
      if (!has_nl) {
        err.msg = "line " + std::to_string(line_no) + " is at least " + std::to_string(end - start) +
                  " bytes, over the " + std::to_string(max_line) + " byte limit (raise it with --max-line-bytes)";
        return false;
      }
    }
    pos = end;
    uint64_t lines = f(Chunk{data + start, end - start, line_no});
    line_no += lines;
  }
  return true;
}

// stream_chunks (chunk.rs:292-387) over an in-memory reader.
template <class F>
static bool stream_chunks(const uint8_t *src, size_t n, size_t chunk_bytes, size_t max_line, F &&f, ChunkError &err) {
  size_t rd = 0;
  std::string buf, carry;
  uint64_t line_no = 1;
  bool first = true;
  for (;;) {
    buf = carry;
    carry.clear();
    size_t base = buf.size();
    size_t want = base >= chunk_bytes ? chunk_bytes : chunk_bytes - base;
    size_t take = std::min(want, n - rd);
    buf.append((const char *)src + rd, take);
    rd += take;
    size_t filled = base + take;
    if (first) {
      if (buf.size() >= 3 && (uint8_t)buf[0] == 0xEF && (uint8_t)buf[1] == 0xBB && (uint8_t)buf[2] == 0xBF)
        buf.erase(0, 3);
      first = false;
    }
    if (buf.empty()) return true;
    bool eof = filled < base + want;
    size_t end;
    if (eof) {
      end = buf.size();
    } else {
      size_t p = buf.rfind('\n');
      if (p == std::string::npos) {
        if (buf.size() > max_line) {
          err.msg = "line " + std::to_string(line_no) + " exceeds the " + std::to_string(max_line) +
                    " byte limit (raise it with --max-line-bytes)";
          return false;
        }
        carry = buf;
        continue;
      }
      end = p + 1;
    }
    carry = buf.substr(end);
    uint64_t first_line = line_no;
    line_no += f(Chunk{(const uint8_t *)buf.data(), end, first_line});
    if (eof) return true;
  }
}

// ---------------------------------------------------------------------------
// exec/cpu.rs
// ---------------------------------------------------------------------------

enum OnInvalid { WARN = 0, SKIP = 1, ABORT = 2 };

struct Plan {  // cpu.rs:87-139
  const Program *program;
  std::vector<PathId> filter_paths, post_paths, proj_paths;
  std::vector<std::string> proj_keys;
  explicit Plan(const Program &p) : program(&p) {
    if (p.filter)
      p.filter->visit_paths([&](PathId id) {
        if (std::find(filter_paths.begin(), filter_paths.end(), id) == filter_paths.end())
          filter_paths.push_back(id);
      });
    for (PathId id : p.required_paths())
      if (std::find(filter_paths.begin(), filter_paths.end(), id) == filter_paths.end()) post_paths.push_back(id);
    if (p.output.kind == Output::PROJECT)
      for (auto &f : p.output.fields) {
        proj_keys.push_back(f.first);
        proj_paths.push_back(f.second);
      }
  }
};

struct Totals {  // exec/mod.rs:111-138
  int kind = 0;  // 0 none, 1 scalar, 2 grouped
  AggState scalar;
  GroupAccumulator groups;
  static Totals for_program(const Program &p) {
    Totals t;
    if (!p.is_aggregate()) t.kind = 0;
    else if (p.group_by) t.kind = 2;
    else t.kind = 1;
    return t;
  }
  void merge(const Totals &o) {
    if (kind == 1 && o.kind == 1) scalar.merge(o.scalar);
    if (kind == 2 && o.kind == 2)
      for (auto &kv : o.groups) groups[kv.first].merge(kv.second);
  }
};

enum Outcome { EMITTED, NOROW, MALFORMED, TYPEERR };

static void push_agg(AggState &st, const Slot *v) {
  if (v) st.push_value(*v);
  else st.push_count();
}

// cpu.rs:443-453
static bool eval_cond(const Cond &c, const std::vector<Slot> &slots) {
  switch (c.type) {
    case Cond::CMP: return wjq::cmp_accepts(c.op, compare(slots[c.path], c.lit));
    case Cond::TRUTHY: return is_truthy(slots[c.path]);
    case Cond::AND: return eval_cond(*c.a, slots) && eval_cond(*c.b, slots);
    case Cond::OR: return eval_cond(*c.a, slots) || eval_cond(*c.b, slots);
    case Cond::NOT: return !eval_cond(*c.a, slots);
  }
  return false;
}

// cpu.rs:299-371
static Outcome eval_line(const uint8_t *line, size_t n, const Plan &plan, std::vector<Slot> &slots, Writer &out,
                         Totals &totals, std::string &msg) {
  const Program &p = *plan.program;
  JErr e;
  if (!validate(line, n, e)) {
    msg = e.detail + " at byte " + std::to_string(e.at);
    return MALFORMED;
  }
  for (auto &s : slots) s = Slot{};
  for (PathId id : plan.filter_paths) {
    Slot s;
    LookupRc rc = lookup(line, n, p.path(id).steps, s, msg);
    if (rc == L_TYPE_ERROR) return TYPEERR;
    if (rc == L_INVALID) return MALFORMED;
    slots[id] = s;
  }
  if (p.filter && !eval_cond(*p.filter, slots)) return NOROW;
  for (PathId id : plan.post_paths) {
    Slot s;
    LookupRc rc = lookup(line, n, p.path(id).steps, s, msg);
    if (rc == L_TYPE_ERROR) return TYPEERR;
    if (rc == L_INVALID) return MALFORMED;
    slots[id] = s;
  }
  switch (p.output.kind) {
    case Output::PASSTHROUGH: out.passthrough(line, n); return EMITTED;
    case Output::PATH: out.value(slots[p.output.path]); return EMITTED;
    case Output::PROJECT: {
      std::vector<Slot> vals;
      for (PathId id : plan.proj_paths) vals.push_back(slots[id]);
      out.projection(plan.proj_keys, vals);
      return EMITTED;
    }
    case Output::AGG: {
      const Slot *v = p.output.arg ? &slots[*p.output.arg] : nullptr;
      if (totals.kind == 2 && p.group_by) {
        GroupKey k = group_key_from_slot(slots[*p.group_by]);
        push_agg(totals.groups[k], v);
      } else if (totals.kind == 1) {
        push_agg(totals.scalar, v);
      }
      return NOROW;
    }
  }
  return NOROW;
}

struct Partial {  // cpu.rs:141-153
  std::string bytes;
  uint64_t rows = 0, physical_lines = 0, lines_in = 0, malformed = 0, type_errors = 0;
  Totals totals;
  bool aborted = false;
  uint64_t abort_line = 0;
  std::string abort_msg;
};

static uint64_t count_physical_lines(const uint8_t *p, size_t n) {  // cpu.rs:433-441
  uint64_t c = 0;
  for (size_t i = 0; i < n; i++) c += p[i] == '\n';
  if (n > 0 && p[n - 1] != '\n') c++;
  return c;
}

// cpu.rs:373-424
static Partial process_slice(const uint8_t *s, size_t n, const Plan &plan, Format fmt, OnInvalid oi) {
  Partial part;
  Writer out;
  out.format = fmt;
  part.totals = Totals::for_program(*plan.program);
  std::vector<Slot> slots(plan.program->paths.size());
  std::string msg;
  for_each_line(Chunk{s, n, 0}, [&](const uint8_t *line, size_t len, uint64_t number) {
    part.lines_in++;
    Outcome o = eval_line(line, len, plan, slots, out, part.totals, msg);
    if (o == MALFORMED || o == TYPEERR) {
      if (oi == ABORT) {
        part.aborted = true;
        part.abort_line = number;
        part.abort_msg = msg;
        return false;
      }
      if (o == MALFORMED) part.malformed++;
      else part.type_errors++;
    }
    return true;
  });
  part.physical_lines = count_physical_lines(s, n);
  part.bytes = out.buf;
  part.rows = out.rows;
  return part;
}

// cpu.rs:256-282
static std::vector<std::pair<size_t, size_t>> split_chunk(const uint8_t *d, size_t n, size_t threads) {
  const size_t MIN_SLICE = 64 << 10;
  size_t k = std::max<size_t>(threads, 1);
  k = std::min(k, std::max<size_t>(n / MIN_SLICE, 1));
  if (k == 1) return {{0, n}};
  std::vector<size_t> b{0};
  for (size_t i = 1; i < k; i++) {
    size_t want = n * i / k;
    size_t prev = b.back();
    const uint8_t *nl = (const uint8_t *)memchr(d + want, '\n', n - want);
    size_t cut = nl ? (size_t)(nl - d) + 1 : n;
    if (cut > prev && cut < n) b.push_back(cut);
  }
  b.push_back(n);
  b.erase(std::unique(b.begin(), b.end()), b.end());
  std::vector<std::pair<size_t, size_t>> out;
  for (size_t i = 0; i + 1 < b.size(); i++) out.push_back({b[i], b[i + 1] - b[i]});
  return out;
}

struct RunStats {
  uint64_t bytes_in = 0, lines_in = 0, lines_out = 0, malformed = 0, type_errors = 0;
};

struct RunResult {
  std::string out;
  RunStats stats;
  int err_kind = 0;  // 0 ok, 1 query, 2 malformed line (abort), 3 other
  uint64_t err_line = 0;
  std::string err;
};

// CpuBackend::run (cpu.rs:160-239) + finish_totals (exec/mod.rs:321-339).
static void cpu_run(const Program &p, const uint8_t *data, size_t n, Format fmt, OnInvalid oi, size_t chunk_bytes,
                    size_t max_line, int threads, bool streamed, RunResult &res) {
  if (fmt == F_CSV && p.output.kind == Output::PASSTHROUGH) {  // exec/mod.rs:256-266
    res.err_kind = 3;
    res.err = "--csv needs a query with named columns, but this one emits whole lines";
    return;
  }
  Plan plan(p);
  Totals totals = Totals::for_program(p);
  Writer w;
  w.format = fmt;
  size_t nthreads = threads > 0 ? (size_t)threads : std::max(1u, std::thread::hardware_concurrency());
  w.write_header(p);
  bool aborted = false;
  auto on_chunk = [&](const Chunk &c) -> uint64_t {
    if (aborted) return 0;
    res.stats.bytes_in += c.len;
    auto slices = split_chunk(c.data, c.len, nthreads);
    std::vector<Partial> parts(slices.size());
    if (slices.size() == 1) {
      parts[0] = process_slice(c.data, c.len, plan, fmt, oi);
    } else {
      std::vector<std::thread> ts;
      for (size_t i = 0; i < slices.size(); i++)
        ts.emplace_back([&, i] { parts[i] = process_slice(c.data + slices[i].first, slices[i].second, plan, fmt, oi); });
      for (auto &t : ts) t.join();
    }
    uint64_t line_base = c.first_line, chunk_lines = 0;
    for (auto &pp : parts) chunk_lines += pp.physical_lines;
    for (auto &pp : parts) {
      if (pp.aborted && !aborted) {
        aborted = true;
        res.err_kind = 2;
        res.err_line = line_base + pp.abort_line;
        res.err = pp.abort_msg;
      }
      res.stats.lines_in += pp.lines_in;
      res.stats.malformed += pp.malformed;
      res.stats.type_errors += pp.type_errors;
      w.write_raw(pp.bytes, pp.rows);
      totals.merge(pp.totals);
      line_base += pp.physical_lines;
    }
    return chunk_lines;
  };
  ChunkError ce;
  bool ok = (!streamed && n > 0) ? mapped_chunks(data, n, chunk_bytes, max_line, on_chunk, ce)
                                 : stream_chunks(data, n, chunk_bytes, max_line, on_chunk, ce);
  if (!ok) {
    res.err_kind = 3;
    res.err = ce.msg;
    return;
  }
  if (aborted) return;
  if (p.output.kind == Output::AGG) {
    if (totals.kind == 1) w.aggregate(p.output.agg, totals.scalar);
    else if (totals.kind == 2) w.groups(group_key_name(p), p.output.agg, totals.groups);
  }
  res.stats.lines_out = w.rows;
  res.out = w.finish();
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C API for the tests (ctypes)
// ---------------------------------------------------------------------------

extern "C" {

struct oracle_result {
  uint8_t *out;
  uint64_t out_len;
  uint64_t bytes_in, lines_in, lines_out, malformed, type_errors;
  int32_t err_kind;
  int32_t _pad;
  uint64_t err_line;
  char err[1024];
};

void oracle_free(void *p) { free(p); }

static void copy_err(char *dst, size_t cap, const std::string &s) {
  size_t n = std::min(cap - 1, s.size());
  memcpy(dst, s.data(), n);
  dst[n] = 0;
}

int oracle_run(const char *query, const uint8_t *data, uint64_t n, int format, int on_invalid, uint64_t chunk_bytes,
               uint64_t max_line, int threads, int streamed, oracle_result *r) {
  memset(r, 0, sizeof *r);
  wjq::Program p;
  wjq::QueryError qe;
  if (!wjq::parse(query, p, qe)) {
    r->err_kind = 1;
    copy_err(r->err, sizeof r->err, qe.render(query));
    return 1;
  }
  orc::RunResult res;
  orc::cpu_run(p, data, n, (orc::Format)format, (orc::OnInvalid)on_invalid, chunk_bytes, max_line, threads,
               streamed != 0, res);
  r->err_kind = res.err_kind;
  r->err_line = res.err_line;
  copy_err(r->err, sizeof r->err, res.err);
  r->bytes_in = res.stats.bytes_in;
  r->lines_in = res.stats.lines_in;
  r->lines_out = res.stats.lines_out;
  r->malformed = res.stats.malformed;
  r->type_errors = res.stats.type_errors;
  r->out_len = res.out.size();
  r->out = (uint8_t *)malloc(res.out.size() + 1);
  memcpy(r->out, res.out.data(), res.out.size());
  return res.err_kind;
}

// 1 when `line` is exactly one valid JSON value (json.rs:414-421). On
// failure `err_at` receives the offset the scanner reported.
int oracle_validate(const uint8_t *line, uint64_t n, uint64_t *err_at) {
  orc::JErr e;
  bool ok = orc::validate(line, n, e);
  if (!ok && err_at) *err_at = e.at;
  return ok ? 1 : 0;
}

// Resolves the path of `path_query` (a query consisting of one path) against
// `line`. Returns 0 found / 1 type error / 2 invalid / -1 bad query; kind,
// offset and length of the value are written out (offset 0, len 4 and kind
// 0 for Missing, whose raw text is `null`).
int oracle_lookup(const uint8_t *line, uint64_t n, const char *path_query, int *kind, uint64_t *off,
                  uint64_t *len) {
  wjq::Program p;
  wjq::QueryError qe;
  if (!wjq::parse(path_query, p, qe)) return -1;
  std::vector<Step> steps;
  if (p.output.kind == Output::PATH) steps = p.path(p.output.path).steps;
  else if (p.output.kind != Output::PASSTHROUGH) return -1;
  orc::Slot s;
  std::string msg;
  orc::LookupRc rc = orc::lookup(line, n, steps, s, msg);
  if (rc == orc::L_FOUND) {
    *kind = s.kind;
    *off = s.kind == orc::K_MISSING ? 0 : (uint64_t)(s.raw - line);
    *len = s.len;
  }
  return rc;
}

int oracle_format_number(double v, char *buf, uint64_t cap) {
  std::string s = orc::format_number(v);
  copy_err(buf, cap, s);
  return (int)s.size();
}

// Decoded bytes of a JSON string body; -1 if an escape is malformed.
int64_t oracle_unescape(const uint8_t *raw, uint64_t n, uint8_t *out, uint64_t cap) {
  std::string s;
  if (!orc::unescape_into(raw, n, s)) return -1;
  if (s.size() > cap) return -2;
  memcpy(out, s.data(), s.size());
  return (int64_t)s.size();
}

int64_t oracle_json_string(const uint8_t *p, uint64_t n, uint8_t *out, uint64_t cap) {
  std::string s;
  wjq::write_json_string(s, p, n);
  if (s.size() > cap) return -2;
  memcpy(out, s.data(), s.size());
  return (int64_t)s.size();
}

int64_t oracle_csv_cell(const uint8_t *p, uint64_t n, uint8_t *out, uint64_t cap) {
  std::string s;
  orc::write_csv_cell(s, p, n);
  if (s.size() > cap) return -2;
  memcpy(out, s.data(), s.size());
  return (int64_t)s.size();
}

// Three-way compare of `line`'s value at the path of `cmp_query`'s filter
// (which must be `select(.path <op> lit)`) against its literal: -1/0/1, or
// -9 on a bad query, -8 when the path does not resolve to a value.
int oracle_compare(const uint8_t *line, uint64_t n, const char *cmp_query) {
  wjq::Program p;
  wjq::QueryError qe;
  if (!wjq::parse(cmp_query, p, qe) || !p.filter || p.filter->type != Cond::CMP) return -9;
  orc::Slot s;
  std::string msg;
  if (orc::lookup(line, n, p.path(p.filter->path).steps, s, msg) != orc::L_FOUND) return -8;
  return orc::compare(s, p.filter->lit);
}

// Chunk boundaries of the mapped (streamed=0) or streamed (1) chunker, as
// (offset-into-data, length, first_line) triples; returns the chunk count or
// -1 on a too-long-line error. `advance` = count newlines per chunk, which
// is what a consumer reports back.
int64_t oracle_chunks(const uint8_t *data, uint64_t n, uint64_t target, uint64_t max_line, int streamed,
                      uint64_t *triples, uint64_t cap) {
  int64_t k = 0;
  orc::ChunkError ce;
  auto f = [&](const orc::Chunk &c) -> uint64_t {
    if ((uint64_t)k < cap) {
      triples[3 * k] = 0;
      triples[3 * k + 1] = c.len;
      triples[3 * k + 2] = c.first_line;
    }
    k++;
    uint64_t nl = 0;
    for (size_t i = 0; i < c.len; i++) nl += c.data[i] == '\n';
    return nl;
  };
  bool ok = streamed ? orc::stream_chunks(data, n, target, max_line, f, ce)
                     : orc::mapped_chunks(data, n, target, max_line, f, ce);
  return ok ? k : -1;
}

// Logical lines of one chunk (chunk.rs Lines) as (offset, len, number).
int64_t oracle_lines(const uint8_t *data, uint64_t n, uint64_t first_line, uint64_t *triples, uint64_t cap) {
  int64_t k = 0;
  orc::for_each_line(orc::Chunk{data, n, first_line}, [&](const uint8_t *line, size_t len, uint64_t no) {
    if ((uint64_t)k < cap) {
      triples[3 * k] = (uint64_t)(line - data);
      triples[3 * k + 1] = len;
      triples[3 * k + 2] = no;
    }
    k++;
    return true;
  });
  return k;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// CSV input (no reference implementation: this build's own specification,
// restated here independently of the device transcoder so the GPU tests
// compare two implementations of it; parity unpinned)
//
// The first line is the header (after a UTF-8 byte order mark, which is
// dropped, as the NDJSON reader drops one per input); its fields name the
// columns. Every later line
// is one record, rendered as the NDJSON object {"name":value,...} in column
// order, and the header itself as a blank line, so line numbers stay the
// file's. A trailing CR is dropped from every line. A field is quoted when it
// starts with '"'; inside, "" is one quote; the closing quote must end the
// line or be followed by a comma. An unquoted field may not contain a quote.
// Values: an empty unquoted field is null, an unquoted field that is a JSON
// number is that number, anything else (every quoted field) is a string.
// A line with bad quoting or a field count other than the header's becomes
// the invalid line "{" (counted malformed); an empty line stays blank.
// ---------------------------------------------------------------------------

namespace csvo {

// Fields of one line (without its newline or CR); false on bad quoting.
static bool split(const uint8_t *b, size_t n, std::vector<std::pair<std::string, bool>> &f) {
  f.clear();
  size_t i = 0;
  for (;;) {
    std::string v;
    bool quoted = false;
    if (i < n && b[i] == '"') {
      quoted = true;
      size_t j = i + 1;
      for (;;) {
        const uint8_t *q = (const uint8_t *)memchr(b + j, '"', n - j);
        if (!q) return false;
        const size_t k = (size_t)(q - b);
        v.append((const char *)b + j, k - j);
        if (k + 1 < n && b[k + 1] == '"') {
          v.push_back('"');
          j = k + 2;
          continue;
        }
        i = k + 1;
        break;
      }
      if (i < n && b[i] != ',') return false;
    } else {
      size_t k = i;
      while (k < n && b[k] != ',') {
        if (b[k] == '"') return false;
        k++;
      }
      v.assign((const char *)b + i, k - i);
      i = k;
    }
    f.emplace_back(std::move(v), quoted);
    if (i >= n) return true;
    i++;  // the comma; a line ending in one has an empty last field
  }
}

static bool json_number(const std::string &s) {
  size_t i = 0, n = s.size();
  if (i < n && s[i] == '-') i++;
  if (i >= n) return false;
  if (s[i] == '0') {
    i++;
  } else if (s[i] >= '1' && s[i] <= '9') {
    while (i < n && s[i] >= '0' && s[i] <= '9') i++;
  } else {
    return false;
  }
  if (i < n && s[i] == '.') {
    i++;
    size_t d = i;
    while (i < n && s[i] >= '0' && s[i] <= '9') i++;
    if (i == d) return false;
  }
  if (i < n && (s[i] == 'e' || s[i] == 'E')) {
    i++;
    if (i < n && (s[i] == '+' || s[i] == '-')) i++;
    size_t d = i;
    while (i < n && s[i] >= '0' && s[i] <= '9') i++;
    if (i == d) return false;
  }
  return i == n;
}

static void json_string(std::string &o, const std::string &s) {
  static const char hex[] = "0123456789abcdef";
  o.push_back('"');
  for (unsigned char c : s) {
    if (c == '"' || c == '\\') {
      o.push_back('\\');
      o.push_back((char)c);
    } else if (c < 0x20) {
      o += "\\u00";
      o.push_back(hex[c >> 4]);
      o.push_back(hex[c & 15]);
    } else {
      o.push_back((char)c);
    }
  }
  o.push_back('"');
}

}  // namespace csvo

extern "C" {

// The NDJSON the CSV input reads as (see above); -1 when the header is bad.
int oracle_csv_to_ndjson(const uint8_t *data, uint64_t n, uint8_t **out, uint64_t *out_len) {
  *out = nullptr;
  *out_len = 0;
  std::string o;
  std::vector<std::string> keys;
  std::vector<std::pair<std::string, bool>> f;
  uint64_t pos = (n >= 3 && data[0] == 0xEF && data[1] == 0xBB && data[2] == 0xBF) ? 3 : 0;
  if (pos >= n) return -1;  // no header line
  bool header = true;
  while (pos < n) {
    const uint8_t *nl = (const uint8_t *)memchr(data + pos, '\n', n - pos);
    const uint64_t end = nl ? (uint64_t)(nl - data) : n;
    uint64_t len = end - pos;
    if (len && data[pos + len - 1] == '\r') len--;
    if (header) {
      if (!csvo::split(data + pos, len, f)) return -1;
      for (auto &x : f) {
        std::string k;
        csvo::json_string(k, x.first);
        keys.push_back(k);
      }
      o.push_back('\n');
      header = false;
    } else if (len == 0) {
      o.push_back('\n');
    } else if (!csvo::split(data + pos, len, f) || f.size() != keys.size()) {
      o += "{\n";
    } else {
      o.push_back('{');
      for (size_t c = 0; c < f.size(); c++) {
        if (c) o.push_back(',');
        o += keys[c];
        o.push_back(':');
        const std::string &v = f[c].first;
        if (!f[c].second && v.empty()) o += "null";
        else if (!f[c].second && csvo::json_number(v)) o += v;
        else csvo::json_string(o, v);
      }
      o += "}\n";
    }
    pos = nl ? end + 1 : n;
  }
  if (header) return -1;  // no header line
  *out = (uint8_t *)malloc(o.size() ? o.size() : 1);
  memcpy(*out, o.data(), o.size());
  *out_len = o.size();
  return 0;
}

}  // extern "C"
