// Query front end: lexer, recursive-descent parser, RPN flattening and the
// lowering to warpjq_program tables. See query.h for the reference files each
// part follows.
#include "query.h"

#include <cerrno>
#include <clocale>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace wjq {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

static bool is_ascii_ws(uint8_t c) {  // Rust u8::is_ascii_whitespace
  return c == ' ' || c == '\t' || c == '\n' || c == 0x0C || c == '\r';
}
static bool is_alpha(uint8_t c) { return (c | 0x20) >= 'a' && (c | 0x20) <= 'z'; }
static bool is_digit(uint8_t c) { return c >= '0' && c <= '9'; }
static bool is_alnum(uint8_t c) { return is_alpha(c) || is_digit(c); }
static bool is_hex(uint8_t c) {
  return is_digit(c) || ((c | 0x20) >= 'a' && (c | 0x20) <= 'f');
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

// char::from_u32(cp).unwrap_or(U+FFFD): surrogates and > 0x10FFFF are not
// scalar values.
static void push_scalar_or_replacement(std::string &s, uint32_t cp) {
  if ((cp >= 0xD800 && cp < 0xE000) || cp > 0x10FFFF) cp = 0xFFFD;
  push_utf8(s, cp);
}

// u32::from_str_radix(text, 16): an optional leading '+', then one or more
// hex digits, nothing else.
static bool from_str_radix16(const std::string &t, uint32_t &v) {
  size_t i = 0;
  if (!t.empty() && t[0] == '+') i = 1;
  if (i >= t.size()) return false;
  uint64_t acc = 0;
  for (; i < t.size(); i++) {
    uint8_t c = (uint8_t)t[i];
    if (!is_hex(c)) return false;
    acc = acc * 16 + (is_digit(c) ? c - '0' : (c | 0x20) - 'a' + 10);
    if (acc > 0xFFFFFFFFull) return false;
  }
  v = (uint32_t)acc;
  return true;
}

// Length of the UTF-8 sequence led by `c` (the query text is valid UTF-8).
static size_t utf8_len(uint8_t c) {
  if (c < 0x80) return 1;
  if ((c >> 5) == 0x6) return 2;
  if ((c >> 4) == 0xE) return 3;
  if ((c >> 3) == 0x1E) return 4;
  return 1;
}

// Correctly rounded decimal -> double in the "C" locale (glibc strtod rounds
// to nearest-even, the same contract as Rust's f64::from_str).
static double parse_f64(const std::string &text) {
  locale_t c = newlocale(LC_NUMERIC_MASK, "C", (locale_t)0);
  double v = strtod_l(text.c_str(), nullptr, c);
  freelocale(c);
  return v;
}

void write_json_string(std::string &out, const uint8_t *p, size_t n) {
  out.push_back('"');
  for (size_t i = 0; i < n; i++) {
    uint8_t c = p[i];
    switch (c) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
      case 0x08: out += "\\b"; break;
      case 0x0C: out += "\\f"; break;
      default:
        if (c < 0x20) {
          char buf[8];
          snprintf(buf, sizeof buf, "\\u%04x", c);
          out += buf;
        } else {
          out.push_back((char)c);
        }
    }
  }
  out.push_back('"');
}

const char *cmp_str(CmpOp op) {
  switch (op) {
    case OP_EQ: return "==";
    case OP_NE: return "!=";
    case OP_LT: return "<";
    case OP_LE: return "<=";
    case OP_GT: return ">";
    case OP_GE: return ">=";
  }
  return "?";
}

const char *agg_name(AggKind k) {
  switch (k) {
    case AggKind::Count: return "count";
    case AggKind::Sum: return "sum";
    case AggKind::Min: return "min";
    case AggKind::Max: return "max";
    case AggKind::Avg: return "avg";
  }
  return "?";
}

// ir.rs:48-52
static bool is_bare_key(const std::string &k) {
  if (k.empty() || is_digit((uint8_t)k[0])) return false;
  for (unsigned char c : k)
    if (!(is_alnum(c) || c == '_')) return false;
  return true;
}

// Rust's `{:?}` for a string: quoted, with \" \\ \t \r \n \0 escaped and other
// control characters as \u{hex}. Printable non-ASCII passes through.
static std::string debug_quote(const std::string &k) {
  std::string s = "\"";
  for (unsigned char c : k) {
    switch (c) {
      case '"': s += "\\\""; break;
      case '\\': s += "\\\\"; break;
      case '\t': s += "\\t"; break;
      case '\r': s += "\\r"; break;
      case '\n': s += "\\n"; break;
      case 0: s += "\\0"; break;
      default:
        if (c < 0x20 || c == 0x7F) {
          char buf[16];
          snprintf(buf, sizeof buf, "\\u{%x}", c);
          s += buf;
        } else {
          s.push_back((char)c);
        }
    }
  }
  s += "\"";
  return s;
}

std::string Path::display() const {
  if (steps.empty()) return ".";
  std::string s;
  for (const Step &st : steps) {
    if (st.is_index) {
      s += "[" + std::to_string(st.index) + "]";
    } else if (is_bare_key(st.key)) {
      s += "." + st.key;
    } else {
      s += "." + debug_quote(st.key);
    }
  }
  return s;
}

std::vector<PathId> Program::required_paths() const {
  std::vector<PathId> ids;
  auto push = [&](PathId id) {
    for (PathId x : ids)
      if (x == id) return;
    ids.push_back(id);
  };
  if (filter) filter->visit_paths(push);
  if (group_by) push(*group_by);
  switch (output.kind) {
    case Output::PASSTHROUGH: break;
    case Output::PATH: push(output.path); break;
    case Output::PROJECT:
      for (auto &f : output.fields) push(f.second);
      break;
    case Output::AGG:
      if (output.arg) push(*output.arg);
      break;
  }
  return ids;
}

std::optional<std::vector<std::string>> Program::csv_headers() const {
  switch (output.kind) {
    case Output::PROJECT: {
      std::vector<std::string> v;
      for (auto &f : output.fields) v.push_back(f.first);
      return v;
    }
    case Output::PATH: return std::vector<std::string>{path(output.path).display()};
    case Output::AGG: {
      std::vector<std::string> v;
      if (group_by) v.push_back(path(*group_by).display());
      v.push_back(agg_name(output.agg));
      return v;
    }
    case Output::PASSTHROUGH: return std::nullopt;
  }
  return std::nullopt;
}

std::string QueryError::render(const std::string &src) const {
  std::string s = "invalid query: " + message + "\n  " + src + "\n  ";
  size_t col = 0;
  for (size_t i = 0; i < offset && i < src.size(); i++)
    if (((uint8_t)src[i] & 0xC0) != 0x80) col++;
  s += std::string(col, ' ') + "^\n";
  if (!hint.empty()) s += "  help: " + hint;
  return s;
}

// ---------------------------------------------------------------------------
// lexer (lex.rs:56-317)
// ---------------------------------------------------------------------------

namespace {

enum class T { Dot, Pipe, LParen, RParen, LBracket, RBracket, LBrace, RBrace, Colon, Comma, Op, Ident, Str, Num, Int, Eof };

struct Tok {
  T t = T::Eof;
  CmpOp op = OP_EQ;
  std::string s;  // Ident / Str
  double num = 0;
  int64_t i = 0;
  size_t at = 0;
  bool operator==(const Tok &o) const {
    if (t != o.t) return false;
    switch (t) {
      case T::Op: return op == o.op;
      case T::Ident:
      case T::Str: return s == o.s;
      case T::Num: return num == o.num;
      case T::Int: return i == o.i;
      default: return true;
    }
  }
  std::string describe() const {
    switch (t) {
      case T::Dot: return "`.`";
      case T::Pipe: return "`|`";
      case T::LParen: return "`(`";
      case T::RParen: return "`)`";
      case T::LBracket: return "`[`";
      case T::RBracket: return "`]`";
      case T::LBrace: return "`{`";
      case T::RBrace: return "`}`";
      case T::Colon: return "`:`";
      case T::Comma: return "`,`";
      case T::Op: return std::string("`") + cmp_str(op) + "`";
      case T::Ident: return "`" + s + "`";
      case T::Str: return "string " + debug_quote(s);
      case T::Num: {
        char b[64];
        snprintf(b, sizeof b, "%.17g", num);
        return std::string("number `") + b + "`";
      }
      case T::Int: return "number `" + std::to_string(i) + "`";
      case T::Eof: return "end of query";
    }
    return "?";
  }
};

struct LexFail {
  QueryError e;
};

[[noreturn]] static void fail(size_t at, const std::string &src, const std::string &msg,
                              const std::string &hint = "") {
  QueryError e;
  e.message = msg;
  e.offset = at < src.size() ? at : src.size();
  e.hint = hint;
  throw LexFail{e};
}

static size_t lex_string(const std::string &src, size_t start, std::string &s) {
  const size_t n = src.size();
  size_t i = start + 1;
  while (i < n) {
    uint8_t c = (uint8_t)src[i];
    if (c == '"') return i + 1;
    if (c == '\\') {
      i += 1;
      if (i >= n) fail(start, src, "unterminated string");
      uint8_t e = (uint8_t)src[i];
      switch (e) {
        case '"': s.push_back('"'); break;
        case '\\': s.push_back('\\'); break;
        case '/': s.push_back('/'); break;
        case 'n': s.push_back('\n'); break;
        case 't': s.push_back('\t'); break;
        case 'r': s.push_back('\r'); break;
        case 'b': s.push_back('\x08'); break;
        case 'f': s.push_back('\x0c'); break;
        case 'u': {
          if (i + 5 > n) fail(i, src, "truncated `\\u` escape");
          uint32_t cp;
          if (!from_str_radix16(src.substr(i + 1, 4), cp)) fail(i, src, "bad `\\u` escape");
          i += 4;
          if (cp >= 0xD800 && cp < 0xDC00) {
            bool ok = i + 2 < n && src[i + 1] == '\\' && src[i + 2] == 'u';
            if (ok) {
              if (i + 7 > n) fail(i, src, "truncated surrogate pair");
              uint32_t lo;
              if (!from_str_radix16(src.substr(i + 3, 4), lo)) fail(i, src, "bad surrogate pair");
              if (lo >= 0xDC00 && lo < 0xE000) {
                i += 6;
                push_scalar_or_replacement(s, 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00));
              } else {
                push_utf8(s, 0xFFFD);
              }
            } else {
              push_utf8(s, 0xFFFD);
            }
          } else {
            push_scalar_or_replacement(s, cp);
          }
          break;
        }
        default: fail(i, src, std::string("unknown escape `\\") + (char)e + "`");
      }
      i += 1;
    } else {
      size_t l = utf8_len(c);
      if (i + l > n) l = n - i;
      s.append(src, i, l);
      i += l;
    }
  }
  fail(start, src, "unterminated string");
}

static size_t lex_number(const std::string &src, size_t start, Tok &tok) {
  const size_t n = src.size();
  size_t i = start;
  if (src[i] == '-') i++;
  size_t int_start = i;
  while (i < n && is_digit((uint8_t)src[i])) i++;
  if (i == int_start) fail(start, src, "expected a number");
  bool is_float = false;
  if (i < n && src[i] == '.') {
    is_float = true;
    i++;
    while (i < n && is_digit((uint8_t)src[i])) i++;
  }
  if (i < n && (src[i] == 'e' || src[i] == 'E')) {
    is_float = true;
    i++;
    if (i < n && (src[i] == '+' || src[i] == '-')) i++;
    size_t exp_start = i;
    while (i < n && is_digit((uint8_t)src[i])) i++;
    if (i == exp_start) fail(start, src, "number has no digits in its exponent");
  }
  std::string text = src.substr(start, i - start);
  if (!is_float) {
    errno = 0;
    char *end = nullptr;
    long long v = strtoll(text.c_str(), &end, 10);
    if (errno == 0 && end && *end == 0) {
      tok.t = T::Int;
      tok.i = v;
      return i;
    }
  }
  tok.t = T::Num;
  tok.num = parse_f64(text);
  return i;
}

static std::vector<Tok> lex(const std::string &src) {
  std::vector<Tok> out;
  const size_t n = src.size();
  size_t i = 0;
  while (i < n) {
    uint8_t c = (uint8_t)src[i];
    if (is_ascii_ws(c)) {
      i++;
      continue;
    }
    if (c == '#') {
      while (i < n && src[i] != '\n') i++;
      continue;
    }
    Tok tok;
    tok.at = i;
    switch (c) {
      case '.': tok.t = T::Dot; i++; break;
      case '|': tok.t = T::Pipe; i++; break;
      case '(': tok.t = T::LParen; i++; break;
      case ')': tok.t = T::RParen; i++; break;
      case '[': tok.t = T::LBracket; i++; break;
      case ']': tok.t = T::RBracket; i++; break;
      case '{': tok.t = T::LBrace; i++; break;
      case '}': tok.t = T::RBrace; i++; break;
      case ':': tok.t = T::Colon; i++; break;
      case ',': tok.t = T::Comma; i++; break;
      case '=':
        if (i + 1 < n && src[i + 1] == '=') {
          tok.t = T::Op; tok.op = OP_EQ; i += 2;
        } else {
          fail(i, src, "`=` is assignment in jq and is not supported; did you mean `==`?");
        }
        break;
      case '!':
        if (i + 1 < n && src[i + 1] == '=') {
          tok.t = T::Op; tok.op = OP_NE; i += 2;
        } else {
          fail(i, src, "expected `!=`");
        }
        break;
      case '<':
        tok.t = T::Op;
        if (i + 1 < n && src[i + 1] == '=') { tok.op = OP_LE; i += 2; } else { tok.op = OP_LT; i += 1; }
        break;
      case '>':
        tok.t = T::Op;
        if (i + 1 < n && src[i + 1] == '=') { tok.op = OP_GE; i += 2; } else { tok.op = OP_GT; i += 1; }
        break;
      case '"':
        tok.t = T::Str;
        i = lex_string(src, i, tok.s);
        break;
      default:
        if (c == '-' || is_digit(c)) {
          i = lex_number(src, i, tok);
        } else if (is_alpha(c) || c == '_') {
          size_t st = i;
          while (i < n && (is_alnum((uint8_t)src[i]) || src[i] == '_')) i++;
          tok.t = T::Ident;
          tok.s = src.substr(st, i - st);
        } else {
          std::string ch;
          push_utf8(ch, c);  // `c as char` maps the byte to U+00XX
          fail(i, src, "unexpected character `" + ch + "`");
        }
    }
    out.push_back(std::move(tok));
  }
  Tok eof;
  eof.t = T::Eof;
  eof.at = n;
  out.push_back(eof);
  return out;
}

// ---------------------------------------------------------------------------
// parser (parse.rs:33-631)
// ---------------------------------------------------------------------------

struct Unsupported {
  const char *name, *why;
};
static const Unsupported KNOWN_UNSUPPORTED[] = {
    {"reduce", "`reduce` is not in the v0.1 subset"},
    {"foreach", "`foreach` is not in the v0.1 subset"},
    {"def", "user-defined functions are not in the v0.1 subset"},
    {"if", "`if/then/else` is not in the v0.1 subset; use `select(...)`"},
    {"try", "`try/catch` is not in the v0.1 subset"},
    {"test", "regex (`test`) is on the roadmap, not in v0.1"},
    {"match", "regex (`match`) is on the roadmap, not in v0.1"},
    {"map", "`map` is not in the v0.1 subset"},
    {"keys", "`keys` is not in the v0.1 subset"},
    {"length", "`length` is not in the v0.1 subset; `count` counts lines"},
    {"tostring", "type coercions are not in the v0.1 subset"},
    {"tonumber", "type coercions are not in the v0.1 subset"},
    {"sort_by", "`sort_by` is not in the v0.1 subset"},
    {"unique", "`unique` is not in the v0.1 subset"},
    {"add", "`add` is not in the v0.1 subset; use `sum(.field)`"},
};
static const char *UNSUPPORTED_HINT =
    "see the Limitations section of the README for the full v0.1 subset";

static const char *find_unsupported(const std::string &w) {
  for (const auto &u : KNOWN_UNSUPPORTED)
    if (w == u.name) return u.why;
  return nullptr;
}

// Rust char::is_whitespace for the ASCII range plus the common Unicode
// spaces that str::trim_start removes.
static size_t trim_start_len(const std::string &s) {
  size_t i = 0;
  while (i < s.size()) {
    uint8_t c = (uint8_t)s[i];
    if (c == ' ' || (c >= 0x09 && c <= 0x0D)) { i++; continue; }
    if (c == 0xC2 && i + 1 < s.size() && ((uint8_t)s[i + 1] == 0x85 || (uint8_t)s[i + 1] == 0xA0)) { i += 2; continue; }
    break;
  }
  return i;
}

// parse.rs:103-126 — split on top-level `|`, outside strings and brackets.
static std::vector<std::pair<size_t, std::string>> stage_starts(const std::string &src) {
  std::vector<std::pair<size_t, std::string>> out;
  size_t start = 0;
  int depth = 0;
  bool in_str = false;
  for (size_t i = 0; i < src.size(); i++) {
    char c = src[i];
    if (c == '\\' && in_str) {
      i++;
      continue;
    }
    if (c == '"') in_str = !in_str;
    else if (!in_str && (c == '(' || c == '[' || c == '{')) depth++;
    else if (!in_str && (c == ')' || c == ']' || c == '}')) depth--;
    else if (c == '|' && !in_str && depth == 0) {
      out.push_back({start, src.substr(start, i - start)});
      start = i + 1;
    }
  }
  out.push_back({start, src.substr(start)});
  return out;
}

static bool unsupported_leading_keyword(const std::string &src, QueryError &err) {
  for (auto &st : stage_starts(src)) {
    size_t t = trim_start_len(st.second);
    std::string word;
    for (size_t i = t; i < st.second.size(); i++) {
      uint8_t c = (uint8_t)st.second[i];
      if (is_alnum(c) || c == '_') word.push_back((char)c);
      else break;
    }
    if (word.empty()) continue;
    if (const char *why = find_unsupported(word)) {
      err.message = why;
      err.offset = st.first + t;
      err.hint = UNSUPPORTED_HINT;
      return true;
    }
  }
  return false;
}

static const uint32_t MAX_COND_DEPTH = 64;

struct Parser {
  const std::string &src;
  std::vector<Tok> toks;
  size_t pos = 0;
  std::vector<Path> paths;
  uint32_t depth = 0;

  const Tok &peek() const { return toks[pos]; }
  const Tok &peek_at(size_t n) const {
    size_t k = pos + n;
    if (k > toks.size() - 1) k = toks.size() - 1;
    return toks[k];
  }
  size_t at() const { return toks[pos].at; }
  Tok bump() {
    Tok t = toks[pos];
    if (pos + 1 < toks.size()) pos++;
    return t;
  }
  bool is(T t) const { return peek().t == t; }
  bool eat(T t) {
    if (peek().t == t) {
      bump();
      return true;
    }
    return false;
  }
  [[noreturn]] void err_at(size_t a, const std::string &msg, const std::string &hint = "") {
    fail(a, src, msg, hint);
  }
  void expect(T t, const std::string &ctx) {
    if (eat(t)) return;
    Tok want;
    want.t = t;
    err_at(at(), "expected " + want.describe() + " " + ctx + ", found " + peek().describe());
  }
  bool peek_keyword(const char *kw) const { return peek().t == T::Ident && peek().s == kw; }

  PathId intern(Path p) {
    for (size_t i = 0; i < paths.size(); i++)
      if (paths[i] == p) return (PathId)i;
    paths.push_back(std::move(p));
    return (PathId)(paths.size() - 1);
  }

  enum class StageKind { Filter, GroupBy, Out };
  struct Stage {
    StageKind k;
    std::unique_ptr<Cond> cond;
    PathId gb = 0;
    Output out;
  };

  static const char *describe_output(const Output &o) {
    switch (o.kind) {
      case Output::PASSTHROUGH: return "`.`";
      case Output::PATH: return "a field path";
      case Output::PROJECT: return "an object projection";
      case Output::AGG: return "an aggregate";
    }
    return "?";
  }

  void program(Program &prog) {
    std::unique_ptr<Cond> filter;
    std::optional<PathId> group_by;
    std::optional<Output> output;
    size_t output_at = 0;
    for (;;) {
      size_t stage_at = at();
      Stage s = stage();
      if (s.k == StageKind::Filter) {
        if (!filter) {
          filter = std::move(s.cond);
        } else {
          auto c = std::make_unique<Cond>();
          c->type = Cond::AND;
          c->a = std::move(filter);
          c->b = std::move(s.cond);
          filter = std::move(c);
        }
      } else if (s.k == StageKind::GroupBy) {
        if (group_by) err_at(stage_at, "only one `group_by` per query is supported in v0.1");
        if (output)
          err_at(stage_at, "`group_by` must come before the output stage",
                 "write `group_by(.host) | count`, not `count | group_by(.host)`");
        group_by = s.gb;
      } else {
        if (output)
          err_at(stage_at, std::string("a second output stage follows ") + describe_output(*output) +
                               "; v0.1 supports one output stage per query");
        output = std::move(s.out);
        output_at = stage_at;
      }
      if (!eat(T::Pipe)) break;
    }
    if (!is(T::Eof)) err_at(at(), "unexpected " + peek().describe());
    Output out = output ? std::move(*output) : Output{};
    if (group_by && out.kind != Output::AGG)
      err_at(output_at > 1 ? output_at : 1, "`group_by` must be followed by an aggregate",
             "try `group_by(.host) | count` or `group_by(.host) | sum(.bytes)`");
    prog.paths = std::move(paths);
    prog.filter = std::move(filter);
    prog.group_by = group_by;
    prog.output = std::move(out);
    prog.source = src;
  }

  Stage stage() {
    size_t a = at();
    Stage st;
    if (peek().t == T::Ident) {
      std::string name = peek().s;
      if (const char *why = find_unsupported(name)) err_at(a, why, UNSUPPORTED_HINT);
      if (name == "select") {
        bump();
        expect(T::LParen, "after `select`");
        st.cond = cond();
        expect(T::RParen, "to close `select(`");
        st.k = StageKind::Filter;
        return st;
      }
      if (name == "group_by") {
        bump();
        expect(T::LParen, "after `group_by`");
        Path p = path();
        if (p.is_identity())
          err_at(a, "`group_by(.)` would group by the whole line",
                 "group by a field, e.g. `group_by(.status)`");
        PathId id = intern(std::move(p));
        expect(T::RParen, "to close `group_by(`");
        st.k = StageKind::GroupBy;
        st.gb = id;
        return st;
      }
      if (name == "count") {
        bump();
        if (is(T::LParen))
          err_at(at(), "`count` takes no argument",
                 "write `count`, or `sum(.field)` to total a field");
        st.k = StageKind::Out;
        st.out.kind = Output::AGG;
        st.out.agg = AggKind::Count;
        return st;
      }
      if (name == "sum" || name == "min" || name == "max" || name == "avg") {
        bump();
        AggKind kind = name == "sum" ? AggKind::Sum
                       : name == "min" ? AggKind::Min
                       : name == "max" ? AggKind::Max
                                       : AggKind::Avg;
        expect(T::LParen, "after `" + name + "`");
        Path p = path();
        if (p.is_identity())
          err_at(a, "`" + name + "(.)` has no field to aggregate", "try `" + name + "(.bytes)`");
        PathId id = intern(std::move(p));
        expect(T::RParen, "to close `" + name + "(`");
        st.k = StageKind::Out;
        st.out.kind = Output::AGG;
        st.out.agg = kind;
        st.out.arg = id;
        return st;
      }
      err_at(a, "unknown function `" + name + "`",
             "v0.1 supports: select, group_by, count, sum, min, max, avg");
    }
    if (is(T::LBrace)) {
      st.k = StageKind::Out;
      st.out = object();
      return st;
    }
    Path p = path();
    st.k = StageKind::Out;
    if (p.is_identity()) {
      st.out.kind = Output::PASSTHROUGH;
      return st;
    }
    st.out.kind = Output::PATH;
    st.out.path = intern(std::move(p));
    return st;
  }

  Output object() {
    expect(T::LBrace, "to start an object");
    std::vector<std::pair<std::string, PathId>> fields;
    for (;;) {
      if (is(T::RBrace)) break;
      size_t key_at = at();
      Tok kt = bump();
      std::string key;
      if (kt.t == T::Ident || kt.t == T::Str) key = kt.s;
      else err_at(key_at, "expected an object key, found " + kt.describe());
      PathId pid;
      if (eat(T::Colon)) {
        Path p = path();
        if (p.is_identity())
          err_at(key_at, "an object value of `.` would embed the whole line",
                 "name a field, e.g. `{msg: .message}`");
        pid = intern(std::move(p));
      } else {
        Path p;
        Step s;
        s.key = key;
        p.steps.push_back(s);
        pid = intern(std::move(p));
      }
      for (auto &f : fields)
        if (f.first == key) err_at(key_at, "duplicate key `" + key + "` in projection");
      fields.push_back({key, pid});
      if (!eat(T::Comma)) break;
    }
    expect(T::RBrace, "to close the object");
    if (fields.empty()) err_at(at(), "empty projection `{}` produces no output");
    Output o;
    o.kind = Output::PROJECT;
    o.fields = std::move(fields);
    return o;
  }

  std::unique_ptr<Cond> cond() {
    auto lhs = and_expr();
    while (peek_keyword("or")) {
      bump();
      auto rhs = and_expr();
      auto c = std::make_unique<Cond>();
      c->type = Cond::OR;
      c->a = std::move(lhs);
      c->b = std::move(rhs);
      lhs = std::move(c);
    }
    return lhs;
  }

  std::unique_ptr<Cond> and_expr() {
    auto lhs = term();
    while (peek_keyword("and")) {
      bump();
      auto rhs = term();
      auto c = std::make_unique<Cond>();
      c->type = Cond::AND;
      c->a = std::move(lhs);
      c->b = std::move(rhs);
      lhs = std::move(c);
    }
    return lhs;
  }

  std::unique_ptr<Cond> term() {
    size_t a = at();
    std::unique_ptr<Cond> base;
    if (eat(T::LParen)) {
      depth++;
      if (depth > MAX_COND_DEPTH)
        err_at(a, "conditions may not nest more than " + std::to_string(MAX_COND_DEPTH) + " deep");
      base = cond();
      expect(T::RParen, "to close the group");
      depth--;
    } else {
      if (peek().t == T::Ident) {
        std::string name = peek().s;
        err_at(a, "expected a field path, found `" + name + "`", "did you mean `." + name + "`?");
      }
      Path p = path();
      if (p.is_identity())
        err_at(a, "`.` cannot be used as a condition in v0.1",
               "compare a field, e.g. `select(.status == 500)`");
      PathId id = intern(std::move(p));
      base = std::make_unique<Cond>();
      if (peek().t == T::Op) {
        CmpOp op = peek().op;
        bump();
        base->type = Cond::CMP;
        base->path = id;
        base->op = op;
        base->lit = literal();
      } else {
        base->type = Cond::TRUTHY;
        base->path = id;
      }
    }
    while (is(T::Pipe) && peek_at(1).t == T::Ident && peek_at(1).s == "not") {
      bump();
      bump();
      auto c = std::make_unique<Cond>();
      c->type = Cond::NOT;
      c->a = std::move(base);
      base = std::move(c);
    }
    return base;
  }

  Literal literal() {
    size_t a = at();
    Tok t = bump();
    Literal l;
    switch (t.t) {
      case T::Num: l.kind = Literal::NUM; l.num = t.num; return l;
      case T::Int: l.kind = Literal::NUM; l.num = (double)t.i; return l;
      case T::Str: l.kind = Literal::STR; l.str = t.s; return l;
      case T::Ident:
        if (t.s == "true") { l.kind = Literal::BOOL; l.b = true; return l; }
        if (t.s == "false") { l.kind = Literal::BOOL; l.b = false; return l; }
        if (t.s == "null") { l.kind = Literal::NUL; return l; }
        err_at(a, "expected a literal value, found `" + t.s + "`",
               "the right-hand side of a comparison must be a number, string, `true`, "
               "`false`, or `null`. Field-to-field comparison is not in v0.1");
      case T::Dot:
        err_at(a, "field-to-field comparison is not supported in v0.1",
               "compare against a constant, e.g. `select(.status == 500)`");
      default: err_at(a, "expected a literal value, found " + t.describe());
    }
  }

  Path path() {
    size_t a = at();
    if (!is(T::Dot))
      err_at(a, "expected a field path, found " + peek().describe(),
             "paths start with `.`, e.g. `.status` or `.req.path`");
    bump();
    Path p;
    for (;;) {
      const Tok &t = peek();
      if (t.t == T::Ident || t.t == T::Str) {
        Step s;
        s.key = t.s;
        bump();
        p.steps.push_back(std::move(s));
      } else if (t.t == T::LBracket) {
        bump();
        size_t idx_at = at();
        Tok it = bump();
        if (it.t == T::Int && it.i >= 0) {
          Step s;
          s.is_index = true;
          s.index = (uint32_t)(uint64_t)it.i;  // `i as u32` truncates
          p.steps.push_back(s);
        } else if (it.t == T::Int) {
          err_at(idx_at, "negative array indices are not supported in v0.1");
        } else {
          err_at(idx_at, "expected an array index, found " + it.describe(),
                 "slices (`.[1:3]`) and iteration (`.[]`) are not in v0.1");
        }
        expect(T::RBracket, "to close the index");
      } else {
        break;
      }
      if (is(T::Dot)) {
        T n = peek_at(1).t;
        if (n == T::Ident || n == T::Str || n == T::LBracket) bump();
        else break;
      } else if (!is(T::LBracket)) {
        break;
      }
    }
    return p;
  }
};

}  // namespace

bool parse(const std::string &src, Program &out, QueryError &err) {
  if (unsupported_leading_keyword(src, err)) return false;
  try {
    Parser p{src, lex(src)};
    p.program(out);
    return true;
  } catch (const LexFail &f) {
    err = f.e;
    return false;
  }
}

// ---------------------------------------------------------------------------
// flat tables (query/mod.rs:95-198) and lowering (gpu/lower.rs:38-127)
// ---------------------------------------------------------------------------

static std::pair<uint32_t, uint32_t> push_blob(std::vector<uint8_t> &blob, const std::string &s) {
  uint32_t off = (uint32_t)blob.size();
  blob.insert(blob.end(), s.begin(), s.end());
  return {off, (uint32_t)s.size()};
}

static uint32_t flatten_cond(const Cond &c, Lowered &f) {
  switch (c.type) {
    case Cond::CMP: {
      warpjq_cmp cm{};
      cm.path = c.path;
      cm.op = c.op;
      switch (c.lit.kind) {
        case Literal::NUL: cm.lit_kind = WARPJQ_LIT_NULL; break;
        case Literal::BOOL: cm.lit_kind = c.lit.b ? WARPJQ_LIT_TRUE : WARPJQ_LIT_FALSE; break;
        case Literal::NUM: cm.lit_kind = WARPJQ_LIT_NUM; cm.lit_num = c.lit.num; break;
        case Literal::STR: {
          auto ol = push_blob(f.blob, c.lit.str);
          cm.lit_kind = WARPJQ_LIT_STR;
          cm.lit_off = ol.first;
          cm.lit_len = ol.second;
          break;
        }
      }
      uint32_t idx = (uint32_t)f.cmps.size();
      f.cmps.push_back(cm);
      f.cond_rpn.push_back({WARPJQ_COND_CMP, idx});
      return 1;
    }
    case Cond::TRUTHY:
      f.cond_rpn.push_back({WARPJQ_COND_TRUTHY, c.path});
      return 1;
    case Cond::NOT: {
      uint32_t d = flatten_cond(*c.a, f);
      f.cond_rpn.push_back({WARPJQ_COND_NOT, 0});
      return d;
    }
    case Cond::AND:
    case Cond::OR: {
      uint32_t da = flatten_cond(*c.a, f);
      uint32_t db = flatten_cond(*c.b, f);
      f.cond_rpn.push_back({c.type == Cond::AND ? (uint32_t)WARPJQ_COND_AND : (uint32_t)WARPJQ_COND_OR, 0});
      return da > db + 1 ? da : db + 1;
    }
  }
  return 0;
}

void lower(const Program &p, bool csv, Lowered &f) {
  f = Lowered{};
  for (const Path &path : p.paths) {
    uint32_t step_off = (uint32_t)f.steps.size();
    for (const Step &s : path.steps) {
      warpjq_step gs{};
      if (s.is_index) {
        gs.kind = WARPJQ_STEP_INDEX;
        gs.index = s.index;
      } else {
        auto ol = push_blob(f.blob, s.key);
        gs.kind = WARPJQ_STEP_KEY;
        gs.key_off = ol.first;
        gs.key_len = ol.second;
        gs.key_hash = key_hash((const uint8_t *)s.key.data(), s.key.size());
      }
      f.steps.push_back(gs);
    }
    f.paths.push_back({step_off, (uint32_t)path.steps.size()});
  }
  if (p.filter) f.cond_stack_depth = flatten_cond(*p.filter, f);

  std::vector<PathId> needed = p.required_paths();
  f.slot_of_path.assign(p.paths.size(), 0xFFFFFFFFu);
  for (size_t i = 0; i < needed.size(); i++) {
    f.slot_of_path[needed[i]] = (uint32_t)i;
    f.needed_paths.push_back(needed[i]);
  }

  switch (p.output.kind) {
    case Output::PASSTHROUGH: f.output_kind = WARPJQ_OUT_PASSTHROUGH; break;
    case Output::PATH: f.output_kind = WARPJQ_OUT_PATH; break;
    case Output::AGG: f.output_kind = WARPJQ_OUT_AGG; break;
    case Output::PROJECT: {
      f.output_kind = WARPJQ_OUT_PROJECT;
      for (size_t i = 0; i < p.output.fields.size(); i++) {
        f.field_paths.push_back(p.output.fields[i].second);
        std::string prefix;
        if (csv) {
          if (i > 0) prefix.push_back(',');
        } else {
          prefix.push_back(i == 0 ? '{' : ',');
          const std::string &k = p.output.fields[i].first;
          write_json_string(prefix, (const uint8_t *)k.data(), k.size());
          prefix.push_back(':');
        }
        f.prefix_off.push_back((uint32_t)f.blob.size());
        f.prefix_len.push_back((uint32_t)prefix.size());
        f.blob.insert(f.blob.end(), prefix.begin(), prefix.end());
      }
      if (!csv) {
        f.suffix_off = (uint32_t)f.blob.size();
        f.suffix_len = 1;
        f.blob.push_back('}');
      }
      break;
    }
  }
  f.output_path = p.output.kind == Output::PATH ? p.output.path : 0xFFFFFFFFu;
  if (p.output.kind == Output::AGG) {
    f.agg_kind = (uint32_t)p.output.agg;
    f.agg_path = p.output.arg ? (uint32_t)*p.output.arg : 0xFFFFFFFFu;
  } else {
    f.agg_kind = WARPJQ_AGG_COUNT;
    f.agg_path = 0xFFFFFFFFu;
  }
  f.group_path = p.group_by ? (uint32_t)*p.group_by : 0xFFFFFFFFu;
  f.has_filter = p.filter ? 1u : 0u;
  f.csv_mode = csv ? 1u : 0u;
}

warpjq_program Lowered::view() const {
  warpjq_program v{};
  v.steps = steps.data();
  v.n_steps = (uint32_t)steps.size();
  v.paths = paths.data();
  v.n_paths = (uint32_t)paths.size();
  v.cmps = cmps.data();
  v.n_cmps = (uint32_t)cmps.size();
  v.cond_rpn = cond_rpn.data();
  v.n_cond = (uint32_t)cond_rpn.size();
  v.blob = blob.data();
  v.n_blob = (uint32_t)blob.size();
  v.needed_paths = needed_paths.data();
  v.n_needed = (uint32_t)needed_paths.size();
  v.slot_of_path = slot_of_path.data();
  v.output_kind = output_kind;
  v.output_path = output_path;
  v.agg_kind = agg_kind;
  v.agg_path = agg_path;
  v.group_path = group_path;
  v.has_filter = has_filter;
  v.n_fields = (uint32_t)field_paths.size();
  v.field_paths = field_paths.data();
  v.prefix_off = prefix_off.data();
  v.prefix_len = prefix_len.data();
  v.suffix_off = suffix_off;
  v.suffix_len = suffix_len;
  v.csv_mode = csv_mode;
  return v;
}

}  // namespace wjq
