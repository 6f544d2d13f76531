// Test-only host build of the device line evaluator (warpjq_amd/csrc/jscan.h).
// Lets the CPU suite pin the exact source the gfx950 kernels compile against
// the oracle, line by line, without a GPU. Never linked into the product.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../warpjq_amd/csrc/dprog.h"
#include "../../warpjq_amd/csrc/query.h"

using namespace wjqd;

struct HostReader {
  const uint8_t *d;
  uint32_t n;
  uint32_t byte(uint32_t i) const { return i < n ? d[i] : WJQ_END; }
  uint32_t dword(uint32_t k) const {
    uint32_t v = 0;
    for (uint32_t b = 0; b < 4; b++) v |= byte(4 * k + b) << (8 * b);
    return v;
  }
  Blk16 block(uint32_t b) const { return Blk16{dword(4 * b), dword(4 * b + 1), dword(4 * b + 2), dword(4 * b + 3)}; }
  // as the kernels' LDS reader: the flat walk starts at the line's dword
  static constexpr uint32_t flat_align = 4u;
  Blk16 blk(uint32_t pos) const { return Blk16{dword(pos / 4), dword(pos / 4 + 1), dword(pos / 4 + 2), dword(pos / 4 + 3)}; }
};

// The same bytes read only in aligned 16-byte blocks, as the kernels' HBM
// reader does (the flat walk then starts at the line's 16-byte block).
struct HostBlockReader {
  HostReader r;
  uint32_t byte(uint32_t i) const { return r.byte(i); }
  uint32_t dword(uint32_t k) const { return r.dword(k); }
  Blk16 block(uint32_t b) const { return r.block(b); }
};

struct StrSink {
  std::string s;
  void put(uint8_t c) { s.push_back((char)c); }
  void span(HostReader &rd, uint32_t off, uint32_t len) {
    for (uint32_t k = 0; k < len; k++) s.push_back((char)rd.byte(off + k));
  }
};

struct Handle {
  wjq::Program prog;
  wjq::Lowered low;
  DProgHost dp;
};

extern "C" {

void *hs_new(const char *query, int csv, char *err, size_t err_len) {
  auto *h = new Handle();
  wjq::QueryError qe;
  if (!wjq::parse(query, h->prog, qe)) {
    snprintf(err, err_len, "%s", qe.message.c_str());
    delete h;
    return nullptr;
  }
  wjq::lower(h->prog, csv != 0, h->low);
  warpjq_program v = h->low.view();
  std::string e;
  // keep the view's arrays alive inside h->low; DProg points into them
  if (!build_dprog(v, h->dp, e)) {
    snprintf(err, err_len, "%s", e.c_str());
    delete h;
    return nullptr;
  }
  return h;
}

void hs_free(void *p) { delete (Handle *)p; }

struct hs_line {
  uint32_t status, selected, end;
  uint32_t has_num, is_int, agg_kind_seen;  // agg: value kind of the aggregate slot
  double num;
  int64_t ival;
  uint32_t gkind, gesc;
  uint64_t row_len;
};

}  // extern "C"

template <int NS>
static void eval_t(Handle *h, HostReader &rd, hs_line *o, char *row, uint64_t row_cap, char *gkey,
                   uint64_t gkey_cap, uint64_t *gkey_len) {
  const DProg &P = h->dp.d;
  SlotV sl[NS];
  LineResult r = eval_line<NS>(rd, 0, P, sl);
  o->status = r.status;
  o->selected = r.selected;
  o->end = r.end;
  if (r.status != ST_OK || !r.selected) return;
  if (P.output_kind != WARPJQ_OUT_AGG) {
    StrSink w;
    uint64_t n = wjqd::row<NS>(rd, P, sl, 0, r.end, &w);
    uint64_t n2 = wjqd::row<NS, HostReader, StrSink>(rd, P, sl, 0, r.end, nullptr);
    o->row_len = n;
    if (n2 != n || w.s.size() != n) o->row_len = ~0ull;  // length pass disagrees
    memcpy(row, w.s.data(), std::min<uint64_t>(row_cap, w.s.size()));
    return;
  }
  if (P.agg_slot != 0xFFFFFFFFu) {
    SlotV v = pick<NS>(sl, P.agg_slot);
    o->agg_kind_seen = v.kind;
    if (v.kind == JK_NUM) {
      NumVal nv;
      if (!parse_number(rd, v.off, v.off + v.len, nv)) {
        o->status = ST_FALLBACK;
        return;
      }
      o->has_num = 1;
      o->num = nv.v;
      o->is_int = nv.is_int;
      o->ival = nv.iv;
    }
  }
  if (P.group_slot != 0xFFFFFFFFu) {
    SlotV g = pick<NS>(sl, P.group_slot);
    GKey k = group_key(rd, g);
    o->gkind = k.kind;
    o->gesc = k.esc;
    std::string s;
    if (k.kind == JK_STR && k.esc) {
      DecIter<HostReader> it(rd, k.off, k.off + k.len);
      for (int c; (c = it.next()) >= 0;) s.push_back((char)c);
    } else {
      for (uint32_t q = 0; q < k.len; q++) s.push_back((char)rd.byte(k.off + q));
    }
    *gkey_len = s.size();
    memcpy(gkey, s.data(), std::min<uint64_t>(gkey_cap, s.size()));
  }
}

extern "C" {

// Evaluates one line (bytes without the terminator) as the kernels do.
int hs_eval(void *hp, const uint8_t *line, uint64_t n, hs_line *o, char *row, uint64_t row_cap, char *gkey,
            uint64_t gkey_cap, uint64_t *gkey_len) {
  Handle *h = (Handle *)hp;
  memset(o, 0, sizeof *o);
  *gkey_len = 0;
  std::vector<uint8_t> buf(line, line + n);
  buf.push_back('\n');
  HostReader rd{buf.data(), (uint32_t)buf.size()};
  switch (h->dp.ns) {
    case 1: eval_t<1>(h, rd, o, row, row_cap, gkey, gkey_cap, gkey_len); break;
    case 2: eval_t<2>(h, rd, o, row, row_cap, gkey, gkey_cap, gkey_len); break;
    case 4: eval_t<4>(h, rd, o, row, row_cap, gkey, gkey_cap, gkey_len); break;
    case 8: eval_t<8>(h, rd, o, row, row_cap, gkey, gkey_cap, gkey_len); break;
    default: eval_t<16>(h, rd, o, row, row_cap, gkey, gkey_cap, gkey_len); break;
  }
  return 0;
}

// The device number parser on one token; 1 when decided.
int hs_parse_number(const uint8_t *tok, uint64_t n, double *v, int *is_int, int64_t *iv) {
  std::vector<uint8_t> buf(tok, tok + n);
  buf.push_back('\n');
  HostReader rd{buf.data(), (uint32_t)buf.size()};
  NumVal nv;
  if (!parse_number(rd, 0, (uint32_t)n, nv)) return 0;
  *v = nv.v;
  *is_int = nv.is_int;
  *iv = nv.iv;
  return 1;
}

// The same with a digit buffer, as the wave-wide phase reads lines: every
// number is decided.
struct ExactReader : HostReader {
  uint8_t *buf;
  uint8_t *dec_buf() { return buf; }
};

int hs_parse_number_exact(const uint8_t *tok, uint64_t n, double *v) {
  std::vector<uint8_t> b(tok, tok + n);
  b.push_back('\n');
  std::vector<uint8_t> digits(WJQ_DEC_BUF);
  ExactReader rd;
  rd.d = b.data();
  rd.n = (uint32_t)b.size();
  rd.buf = digits.data();
  NumVal nv;
  if (!parse_number(rd, 0, (uint32_t)n, nv)) return 0;
  *v = nv.v;
  return 1;
}

// The exact path alone (no fast path in front of it).
double hs_slow_decimal(const uint8_t *tok, uint64_t n) {
  std::vector<uint8_t> b(tok, tok + n);
  b.push_back('\n');
  std::vector<uint8_t> digits(WJQ_DEC_BUF);
  HostReader rd{b.data(), (uint32_t)b.size()};
  uint64_t bits = slow_decimal(rd, 0, (uint32_t)n, digits.data());
  double d;
  memcpy(&d, &bits, 8);
  return d;
}

// A reader with a deep container stack and a digit buffer, as the
// wave-wide phase reads a line (levels past 64 in its LDS).
struct DeepReader : ExactReader {
  std::vector<uint8_t> *bits;
  uint32_t cap;
  uint32_t deep_cap() const { return cap; }
  void deep_set(uint32_t d, bool a) { (*bits)[d] = a; }
  bool deep_get(uint32_t d) const { return (*bits)[d] != 0; }
};

// walk() through the deep reader: status as hs_validate.
int hs_validate_deep(const uint8_t *line, uint64_t n, uint32_t cap) {
  std::vector<uint8_t> buf(line, line + n);
  buf.push_back('\n');
  std::vector<uint8_t> digits(WJQ_DEC_BUF), bits(cap + 1);
  DeepReader rd;
  rd.d = buf.data();
  rd.n = (uint32_t)buf.size();
  rd.buf = digits.data();
  rd.bits = &bits;
  rd.cap = cap;
  DProg P{};
  SlotV sl[1];
  sl[0] = SlotV{0, 0, 0};
  uint32_t te = 0, end = 0;
  uint32_t i = skip_ws(rd, 0);
  return walk<1>(rd, i, P, sl, te, end);
}

// The flat fast path against the byte walk on one line: 0 the fast path
// declined it, 1 it accepted it and the walk agrees (valid, same captures,
// type errors and end), 2 it accepted a line the walk decides otherwise.
}  // extern "C"

template <int NS, class RD>
static int flat_vs_walk_r(Handle *h, RD &rd, uint32_t i) {
  const DProg &P = h->dp.d;
  SlotV a[NS], b[NS];
  for (int q = 0; q < NS; q++) a[q] = b[q] = SlotV{0, 0, JK_MISSING};
  uint32_t ta = 0, tb = 0, ea = 0, eb = 0;
  if (!flat_walk<NS>(rd, i, P, a, ta, ea)) return 0;
  int st = walk<NS>(rd, i, P, b, tb, eb);
  if (st != ST_OK || ta != tb || ea != eb) return 2;
  for (int q = 0; q < NS; q++) {
    if (!((P.fused_mask >> q) & 1)) continue;
    if (a[q].kind != b[q].kind) return 2;
    if (a[q].kind != JK_MISSING && (a[q].off != b[q].off || a[q].len != b[q].len)) return 2;
  }
  return 1;
}

// Both word alignments of the flat walk: 2 when either disagrees with the
// walk, else the dword-aligned (LDS reader) verdict.
template <int NS>
static int flat_vs_walk_t(Handle *h, HostReader &rd, uint32_t i) {
  HostBlockReader rb{rd};
  const int a = flat_vs_walk_r<NS>(h, rd, i), b = flat_vs_walk_r<NS>(h, rb, i);
  return a == 2 || b == 2 ? 2 : a;
}

extern "C" {

int hs_flat_vs_walk(void *hp, const uint8_t *line, uint64_t n, uint32_t lead) {
  Handle *h = (Handle *)hp;
  // `lead` bytes of junk before the line (the staged neighbour's tail)
  std::vector<uint8_t> buf(lead, 'x');
  buf.insert(buf.end(), line, line + n);
  buf.push_back('\n');
  for (int k = 0; k < 80; k++) buf.push_back(k & 1 ? '"' : ',');  // bytes past the end
  HostReader rd{buf.data(), (uint32_t)buf.size()};
  uint32_t i = lead;
  if (rd.byte(i) != '{') return 0;
  switch (h->dp.ns) {
    case 1: return flat_vs_walk_t<1>(h, rd, i);
    case 2: return flat_vs_walk_t<2>(h, rd, i);
    case 4: return flat_vs_walk_t<4>(h, rd, i);
    case 8: return flat_vs_walk_t<8>(h, rd, i);
    default: return flat_vs_walk_t<16>(h, rd, i);
  }
}

// The device's string scanner + structural validation on one line.
int hs_validate(const uint8_t *line, uint64_t n) {
  std::vector<uint8_t> buf(line, line + n);
  buf.push_back('\n');
  HostReader rd{buf.data(), (uint32_t)buf.size()};
  DProg P{};
  SlotV sl[1];
  sl[0] = SlotV{0, 0, 0};
  uint32_t te = 0, end = 0;
  uint32_t i = skip_ws(rd, 0);
  return walk<1>(rd, i, P, sl, te, end);
}

}  // extern "C"
