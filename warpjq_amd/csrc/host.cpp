// Host driver over the device ABI: the MI355X counterpart of the reference's
// GPU backend (crates/warpjq-core/src/gpu/mod.rs:244-1259) for in-memory
// input.
//
//   chunking        newline-aligned chunks of at most the slot capacity,
//                   trimmed to the device line budget (gpu/mod.rs:697-810)
//   pipeline        two slots: chunk k fills and submits while chunk k-1
//                   drains (gpu/mod.rs:487-565)
//   drain / merge   counters -> stats, strict handling, aggregate and group
//                   merge, rows written in input order (gpu/mod.rs:959-1191)
//   totals          jq number rendering and sorted groups
//                   (agg.rs:182-278, output.rs:267-347, exec/mod.rs:321-339)
//
// Differences, all in the direction of "the GPU does the work":
//   * a chunk the device cannot represent (chunk_overflow / group_overflow)
//     is split in two at a newline and resubmitted instead of being redone
//     on the CPU;
//   * sums merge exactly (128-bit integer part + compensated FP64) and round
//     once at the end;
//   * there is no CPU evaluator: a line the kernels decline is an error
//     reported with its line number.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <cstdio>
#include <cstring>
#include <map>
#include <cerrno>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <string>
#include <vector>

#include "../../include/warpjq_amd.h"
#include "../../include/warpjq_kernels.h"
#include "driver.h"
#include "query.h"

namespace wjqd {


// ---- aggregates (agg.rs:15-98) with exact sum parts ----------------------

struct Acc {
  uint64_t count = 0, numeric = 0;
  unsigned __int128 isum = 0;  // two's complement
  double fsum = 0.0, fcomp = 0.0;
  double min = INFINITY, max = -INFINITY;
  bool saw_non_numeric = false;

  void add_parts(const warpjq_amd_agg_ext &e) {
    isum += ((unsigned __int128)e.isum_hi << 64) | e.isum_lo;
    double t = fsum + e.fsum;
    double bp = t - fsum;
    fcomp += (fsum - (t - bp)) + (e.fsum - bp) + e.fcomp;
    fsum = t;
  }
  void merge_minmax(uint64_t num, double mn, double mx) {
    if (num) {
      if (mn < min) min = mn;
      if (mx > max) max = mx;
    }
  }
  double sum() const {
    warpjq_amd_agg_ext e;
    e.isum_lo = (uint64_t)isum;
    e.isum_hi = (uint64_t)(isum >> 64);
    e.fsum = fsum;
    e.fcomp = fcomp;
    return warpjq_amd_ext_value(&e);
  }
  bool finish(wjq::AggKind k, double &v) const {
    switch (k) {
      case wjq::AggKind::Count: v = (double)count; return true;
      case wjq::AggKind::Sum: v = sum(); return true;
      case wjq::AggKind::Min: v = min; return numeric > 0;
      case wjq::AggKind::Max: v = max; return numeric > 0;
      case wjq::AggKind::Avg: v = sum() / (double)numeric; return numeric > 0;
    }
    return false;
  }
};

// GroupKey (agg.rs:106-212): rank 0 null, 1 false, 2 true, 3 number,
// 4 string, 5 composite; numbers order by value and are identified by their
// raw text; strings by decoded bytes.
struct GKey {
  int rank = 0;
  std::string bytes;
  uint64_t bits = 0;
};
struct GKeyLess {
  bool operator()(const GKey &a, const GKey &b) const {
    if (a.rank != b.rank) return a.rank < b.rank;
    if (a.rank == 3 && a.bits != b.bits) return a.bits < b.bits;
    return a.bytes < b.bytes;
  }
};

static locale_t c_locale() {
  static locale_t l = newlocale(LC_NUMERIC_MASK, "C", (locale_t)0);
  return l;
}

static uint64_t order_bits(double v) {
  uint64_t b;
  memcpy(&b, &v, 8);
  return (b >> 63) ? ~b : (b ^ (1ull << 63));
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

static uint32_t hex4(const uint8_t *p) {
  uint32_t v = 0;
  for (int k = 0; k < 4; k++) {
    uint8_t c = p[k];
    v = v * 16 + (c <= '9' ? c - '0' : (c | 0x20) - 'a' + 10);
  }
  return v;
}

// Decodes a validated JSON string body (json.rs:594-654 semantics).
static std::string decode_body(const uint8_t *raw, size_t n) {
  std::string out;
  size_t i = 0;
  while (i < n) {
    if (raw[i] != '\\') {
      out.push_back((char)raw[i++]);
      continue;
    }
    uint8_t e = raw[i + 1];
    switch (e) {
      case 'b': out.push_back('\x08'); break;
      case 'f': out.push_back('\x0c'); break;
      case 'n': out.push_back('\n'); break;
      case 'r': out.push_back('\r'); break;
      case 't': out.push_back('\t'); break;
      case 'u': {
        uint32_t cp = hex4(raw + i + 2);
        i += 6;
        if (cp >= 0xD800 && cp < 0xDC00 && i + 6 <= n && raw[i] == '\\' && raw[i + 1] == 'u') {
          uint32_t lo = hex4(raw + i + 2);
          if (lo >= 0xDC00 && lo < 0xE000) {
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
            i += 6;
          }
        }
        if (cp >= 0xD800 && cp < 0xE000) cp = 0xFFFD;
        push_utf8(out, cp);
        continue;
      }
      default: out.push_back((char)e);
    }
    i += 2;
  }
  return out;
}

// device_group_key (gpu/mod.rs:1198-1221).
static GKey key_from_device(const uint8_t *data, const warpjq_group &g) {
  GKey k;
  const uint8_t *raw = data + g.key_off;
  switch (g.key_kind) {
    case 0:
    case 1: k.rank = 0; break;
    case 2: k.rank = (g.key_len == 4 && memcmp(raw, "true", 4) == 0) ? 2 : 1; break;
    case 3: {
      k.rank = 3;
      k.bytes.assign((const char *)raw, g.key_len);
      k.bits = order_bits(strtod_l(k.bytes.c_str(), nullptr, c_locale()));
      break;
    }
    case 4:
      k.rank = 4;
      k.bytes = memchr(raw, '\\', g.key_len) ? decode_body(raw, g.key_len)
                                             : std::string((const char *)raw, g.key_len);
      break;
    default:
      k.rank = 5;
      k.bytes.assign((const char *)raw, g.key_len);
  }
  return k;
}

// ---- number rendering (agg.rs:253-278; ryu 1.0 layout) ---------------------

static std::string ryu_layout(double v) {
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific);
  std::string s(buf, r.ptr);
  bool neg = s[0] == '-';
  size_t p = neg ? 1 : 0;
  size_t epos = s.find('e');
  std::string digits;
  for (size_t i = p; i < epos; i++)
    if (s[i] != '.') digits.push_back(s[i]);
  int exp10 = atoi(s.c_str() + epos + 1);
  int length = (int)digits.size();
  int k = exp10 - (length - 1);
  int kk = length + k;
  std::string out = neg ? "-" : "";
  if (k >= 0 && kk <= 16) out += digits + std::string(kk - length, '0') + ".0";
  else if (kk > 0 && kk <= 16) out += digits.substr(0, kk) + "." + digits.substr(kk);
  else if (kk > -5 && kk <= 0) out += "0." + std::string(-kk, '0') + digits;
  else if (length == 1) out += digits + "e" + std::to_string(kk - 1);
  else out += digits.substr(0, 1) + "." + digits.substr(1) + "e" + std::to_string(kk - 1);
  return out;
}

static std::string format_number(double v) {
  if (std::isnan(v)) return "null";
  if (std::isinf(v)) return v > 0 ? "1.7976931348623157e+308" : "-1.7976931348623157e+308";
  if (std::trunc(v) == v && std::fabs(v) < 9007199254740992.0) return std::to_string((long long)v);
  std::string s = ryu_layout(v);
  size_t e = s.find('e');
  if (e != std::string::npos && s[e + 1] != '-') s.insert(e + 1, "+");
  return s;
}

// ---- writer (output.rs:87-334) --------------------------------------------

static void csv_cell(std::string &out, const std::string &b) {
  bool q = b.find_first_of(",\"\n\r") != std::string::npos;
  if (!q) {
    out += b;
    return;
  }
  out.push_back('"');
  for (char c : b) {
    if (c == '"') out.push_back('"');
    out.push_back(c);
  }
  out.push_back('"');
}

static std::string group_key_json(const GKey &k) {
  switch (k.rank) {
    case 0: return "null";
    case 1: return "false";
    case 2: return "true";
    case 4: {
      std::string s;
      wjq::write_json_string(s, (const uint8_t *)k.bytes.data(), k.bytes.size());
      return s;
    }
    default: return k.bytes;
  }
}

static std::string group_key_name(const wjq::Program &p) {
  if (!p.group_by) return "key";
  const auto &steps = p.path(*p.group_by).steps;
  if (steps.empty()) return "key";
  if (steps.back().is_index) return "[" + std::to_string(steps.back().index) + "]";
  return steps.back().key;
}



static uint64_t batch_bytes(const Batch &b) {
  uint64_t n = 0;
  for (const Piece &x : b) n += x.len;
  return n;
}

struct Run {
  const wjq::Program &prog;
  Format fmt;
  OnInvalid oi;
  warpjq_ctx *ctx = nullptr;
  uint64_t cap = 0, max_lines = 0;
  std::string out;
  uint64_t rows = 0;
  warpjq_amd_stats st{};
  uint64_t lines_seen = 0;
  Acc scalar;
  std::map<GKey, Acc, GKeyLess> groups;
  std::string err;
  int rc = 0;
  uint64_t err_line = 0;
  bool csv_header_done = false;
};

static int drain(Run &r, uint32_t slot, const Batch &b, std::vector<Batch> &redo) {
  warpjq_chunk_result res;
  char e[512] = {0};
  warpjq_status s = warpjq_wait(r.ctx, slot, &res, e, sizeof e);
  if (s != WARPJQ_OK) {
    r.err = std::string("waiting for a chunk: ") + e;
    return 3;
  }
  float h2d = 0, kms = 0;
  warpjq_amd_slot_timing(r.ctx, slot, &h2d, &kms, nullptr);
  r.st.h2d_ms += h2d;
  r.st.kernel_ms += kms;
  r.st.chunks++;
  if (res.chunk_overflow || res.group_overflow) {
    r.st.gpu_redone_chunks++;
    if (b.size() > 1) {  // halve the piece list
      const size_t h = b.size() / 2;
      redo.push_back(Batch(b.begin(), b.begin() + h));
      redo.push_back(Batch(b.begin() + h, b.end()));
      return 0;
    }
    // Split at the newline nearest the middle and run both halves again.
    const uint8_t *d = b[0].p;
    const uint64_t len = b[0].len;
    uint64_t mid = len / 2;
    const uint8_t *nl = (const uint8_t *)memchr(d + mid, '\n', len - mid);
    uint64_t cut = nl ? (uint64_t)(nl - d) + 1 : len;
    if (cut >= len) {
      nl = (const uint8_t *)memrchr(d, '\n', mid);
      cut = nl ? (uint64_t)(nl - d) + 1 : 0;
    }
    if (cut == 0 || cut >= len) {
      r.err = "a single line exceeds the device's per-chunk capacity";
      return 3;
    }
    const uint8_t *dd = b[0].d;
    redo.push_back(Batch{{d, cut, dd}});
    redo.push_back(Batch{{d + cut, len - cut, dd ? dd + cut : nullptr}});
    return 0;
  }
  // warpjq_submit: key offsets are relative to the chunk; a batched submit:
  // they index the slot's key copy
  const uint8_t *data = (b.size() == 1 && !b[0].d) ? b[0].p : warpjq_amd_group_keys(r.ctx, slot);
  uint64_t first_line = r.lines_seen + 1;
  r.lines_seen += res.n_lines;
  r.st.lines_in += res.n_lines - res.n_blank;
  r.st.malformed += res.n_invalid;
  r.st.type_errors += res.n_type_error;
  r.st.gpu_fallback_lines += res.n_fallback;
  if ((res.n_invalid || res.n_type_error) && r.oi == OI_ABORT) {
    // --strict stops at the first malformed or type-error line, whichever
    // comes first (exec/cpu.rs:401-405)
    const uint32_t fi = res.n_invalid ? res.first_invalid : 0xFFFFFFFFu;
    const uint32_t ft = res.n_type_error ? warpjq_amd_first_type_error(r.ctx, slot) : 0xFFFFFFFFu;
    const uint32_t f = std::min(fi, ft);
    r.err_line = f == 0xFFFFFFFFu ? first_line : first_line + f;
    r.err = ft < fi ? "a path indexes a value of the wrong kind" : "line is not valid JSON";
    return 2;
  }
  if (res.n_fallback) {
    uint32_t lo = 0xFFFFFFFFu;
    for (uint64_t k = 0; k < res.n_fallback; k++) lo = std::min(lo, res.fallback_idx[k]);
    r.err_line = first_line + lo;
    r.err = std::to_string(res.n_fallback) + " line(s) could not be decided exactly on the device";
    return 4;
  }
  if (r.prog.is_aggregate()) {
    if (r.prog.group_by) {
      const warpjq_amd_agg_ext *gx = nullptr;
      warpjq_amd_wait_ext(r.ctx, slot, nullptr, &gx);
      for (uint32_t g = 0; g < res.n_groups; g++) {
        const warpjq_group &G = res.groups[g];
        Acc &a = r.groups[key_from_device(data, G)];
        a.count += G.count;
        a.numeric += G.numeric;
        a.add_parts(gx[g]);
        a.merge_minmax(G.numeric, G.min, G.max);
      }
    } else {
      warpjq_amd_agg_ext x;
      warpjq_amd_wait_ext(r.ctx, slot, &x, nullptr);
      r.scalar.count += res.agg.count;
      r.scalar.numeric += res.agg.numeric;
      r.scalar.add_parts(x);
      r.scalar.merge_minmax(res.agg.numeric, res.agg.min, res.agg.max);
      r.scalar.saw_non_numeric |= res.agg.saw_non_numeric != 0;
    }
    return 0;
  }
  r.rows += res.n_selected;
  if (r.fmt != F_COUNT && res.n_selected) r.out.append((const char *)res.out_bytes, res.out_len);
  return 0;
}

// Newline-aligned chunks of at most `cap` bytes holding at most `max_lines`
// lines (gpu/mod.rs:697-810). When the line at `pos` alone is longer than
// `cap`, returns false with *too_long set.
// Newlines in p[0..n), found with memchr (vectorised in the C library; the
// per-byte loop it replaces ran at a few GB/s and, over a batch of many small
// files, left the GPU idle between chunks).
static uint64_t count_nl(const uint8_t *p, uint64_t n) {
  uint64_t c = 0;
  const uint8_t *e = p + n;
  while (p < e) {
    const uint8_t *q = (const uint8_t *)memchr(p, '\n', (size_t)(e - p));
    if (!q) break;
    c++;
    p = q + 1;
  }
  return c;
}

static bool next_chunk(const uint8_t *d, uint64_t n, uint64_t &pos, uint64_t cap, uint64_t max_lines,
                       uint64_t max_line, Chunk &c, std::string &err, bool *too_long = nullptr) {
  if (pos >= n) return false;
  uint64_t want = std::min<uint64_t>(cap, n - pos);
  uint64_t end;
  if (pos + want >= n) {
    end = n;
  } else {
    const uint8_t *nl = (const uint8_t *)memrchr(d + pos, '\n', want);
    if (!nl) {
      err = "a line is longer than the device chunk (" + std::to_string(cap) + " bytes)";
      if (too_long) *too_long = true;
      return false;
    }
    end = (uint64_t)(nl - d) + 1;
  }
  // line budget: count only when the head of the chunk is dense
  const uint64_t sample = std::min<uint64_t>(1 << 16, end - pos);
  const uint64_t seen = count_nl(d + pos, sample);
  uint64_t projected = sample ? seen * (end - pos) / sample : 0;
  if (projected * 2 >= max_lines) {
    uint64_t lines = 0;
    const uint8_t *q = d + pos, *e = d + end;
    while (q < e) {
      const uint8_t *nl = (const uint8_t *)memchr(q, '\n', (size_t)(e - q));
      if (!nl) break;
      if (++lines == max_lines) {
        if (nl + 1 < e) end = (uint64_t)(nl - d) + 1;
        break;
      }
      q = nl + 1;
    }
  }
  if (max_line < cap) {
    uint64_t s = pos;
    const uint8_t *q = d + pos, *e = d + end;
    while (q < e) {
      const uint8_t *nl = (const uint8_t *)memchr(q, '\n', (size_t)(e - q));
      if (!nl) break;
      const uint64_t i = (uint64_t)(nl - d);
      if (i - s > max_line) {
        err = "a line exceeds the " + std::to_string(max_line) + " byte limit";
        return false;
      }
      s = i + 1;
      q = nl + 1;
    }
    if (end - s > max_line && (end == n || d[end - 1] != '\n')) {
      err = "a line exceeds the " + std::to_string(max_line) + " byte limit";
      return false;
    }
  }
  c = {pos, end - pos};
  pos = end;
  return true;
}

// Copies into the pinned staging slots on several host threads: one thread
// copies ~10 GB/s, below the ~56 GB/s the PCIe link takes, so a chunk's copy
// into slot k+1 would otherwise outlast the H2D of slot k it overlaps.
class CopyPool {
 public:
  // Never destroyed: its threads wait on its condition variable until the
  // process ends, and destroying a condition variable that threads wait on
  // at exit can block the exit.
  static CopyPool &get() {
    static CopyPool *p = new CopyPool();
    return *p;
  }
  void copy(uint8_t *dst, const uint8_t *src, uint64_t n) {
    const uint64_t parts = std::min<uint64_t>(nthreads_ + 1, n >> 20);
    if (parts <= 1) {
      memcpy(dst, src, n);
      return;
    }
    std::unique_lock<std::mutex> lk(call_mu_);  // one parallel copy at a time
    const uint64_t step = ((n / parts) + 4095) & ~(uint64_t)4095;
    {
      std::lock_guard<std::mutex> g(mu_);
      dst_ = dst;
      src_ = src;
      n_ = n;
      step_ = step;
      next_ = 1;  // piece 0 is the caller's
      left_ = (uint32_t)((n + step - 1) / step) - 1;
    }
    cv_.notify_all();
    memcpy(dst, src, std::min(step, n));
    for (;;) {  // help until every piece is taken, then wait for the rest
      uint64_t k;
      {
        std::lock_guard<std::mutex> g(mu_);
        if (next_ * step_ >= n_) break;
        k = next_++;
      }
      piece(k);
    }
    std::unique_lock<std::mutex> w(mu_);
    done_cv_.wait(w, [&] { return left_ == 0; });
  }

 private:
  CopyPool() {
    unsigned hw = std::thread::hardware_concurrency();
    nthreads_ = std::max(1u, std::min(8u, hw ? hw - 1 : 1u));
    for (unsigned i = 0; i < nthreads_; i++) std::thread([this] { run(); }).detach();
  }
  void piece(uint64_t k) {
    const uint64_t a = k * step_, e = std::min(n_, a + step_);
    memcpy(dst_ + a, src_ + a, e - a);
    std::lock_guard<std::mutex> g(mu_);
    if (--left_ == 0) done_cv_.notify_all();
  }
  void run() {
    for (;;) {
      uint64_t k;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return next_ * step_ < n_; });
        k = next_++;
      }
      piece(k);
    }
  }
  std::mutex mu_, call_mu_;
  std::condition_variable cv_, done_cv_;
  uint8_t *dst_ = nullptr;
  const uint8_t *src_ = nullptr;
  uint64_t n_ = 0, step_ = 1, next_ = 0;
  uint32_t left_ = 0;
  unsigned nthreads_ = 1;
};

static int submit(Run &r, uint32_t slot, const Batch &b) {
  uint8_t *buf = warpjq_slot_buffer(r.ctx, slot);
  char e[512] = {0};
  warpjq_status st;
  uint64_t n = 0;
  if (b[0].d) {  // resident: no copy, one launch over the spans
    std::vector<warpjq_amd_span> sp(b.size());
    for (size_t i = 0; i < b.size(); i++) {
      sp[i] = {b[i].d, b[i].len};
      n += b[i].len;
    }
    st = warpjq_amd_submit_spans(r.ctx, slot, sp.data(), (uint32_t)sp.size(), e, sizeof e);
  } else if (b.size() == 1) {  // the reference ABI's own submit
    CopyPool::get().copy(buf, b[0].p, b[0].len);
    n = b[0].len;
    st = warpjq_submit(r.ctx, slot, n, e, sizeof e);
  } else {  // packed back to back; one copy and one launch for all pieces
    std::vector<warpjq_amd_piece> pcs(b.size());
    for (size_t i = 0; i < b.size(); i++) {
      memcpy(buf + n, b[i].p, b[i].len);
      pcs[i] = {n, b[i].len};
      n += b[i].len;
    }
    st = warpjq_amd_submit_pieces(r.ctx, slot, n, pcs.data(), (uint32_t)pcs.size(), e, sizeof e);
  }
  if (st != WARPJQ_OK) {
    r.err = std::string("submitting a chunk: ") + e;
    return 3;
  }
  r.st.bytes_in += n;
  return 0;
}

bool Batcher::next(Batch &b, std::string &err) {
  b.clear();
  uint64_t used = 0, lines = 0;
  while (fi < n_files) {
    const uint8_t *d = files[fi];
    const uint64_t n = lens[fi];
    if (!started) {
      started = true;
      pos = (n >= 3 && d[0] == 0xEF && d[1] == 0xBB && d[2] == 0xBF) ? 3 : 0;
    }
    if (pos >= n) {
      fi++;
      started = false;
      continue;
    }
    if (b.size() >= max_pieces) break;
    const uint64_t room = cap - used;
    // a large rest starts a batch of its own rather than being cut small
    if (!b.empty() && n - pos > room && room < cap / 4) break;
    uint64_t p = pos;
    Chunk c;
    std::string e;
    bool too_long = false;
    if (!next_chunk(d, n, p, room, max_lines, max_line, c, e, &too_long)) {
      if (!b.empty()) break;  // no newline within the room left: next batch
      if (too_long) {
        // one line longer than the chunk: alone, if the limit allows it
        // (gpu/mod.rs:411-458; a line past the limit is an error,
        // chunk.rs:292-387)
        const uint8_t *nl = (const uint8_t *)memchr(d + pos, '\n', n - pos);
        const uint64_t end = nl ? (uint64_t)(nl - d) + 1 : n;
        const uint64_t body = end - pos - (nl ? 1 : 0);
        if (body > max_line) {
          err = "a line exceeds the " + std::to_string(max_line) + " byte limit";
          return false;
        }
        Piece pc{d + pos, end - pos, dfiles ? dfiles[fi] + pos : nullptr};
        pc.big = true;
        b.push_back(pc);
        pos = end;
        return true;
      }
      err = e.empty() ? "a line is longer than the device chunk" : e;
      return false;
    }
    // line budget for the whole batch (the kernels report an overflow
    // otherwise, and the batch is split)
    const uint64_t sample = std::min<uint64_t>(1 << 16, c.len);
    const uint64_t seen = count_nl(d + c.off, sample);
    const uint64_t est = sample ? seen * c.len / sample + 1 : 1;
    if (!b.empty() && (lines + est) * 2 >= max_lines) break;
    lines += est;
    b.push_back({d + c.off, c.len, dfiles ? dfiles[fi] + c.off : nullptr});
    used += c.len;
    pos = p;
    if (pos < n || used >= cap) break;  // an input that was cut ends its batch
  }
  return !b.empty();
}


// ---------------------------------------------------------------------------
// Inputs named by path (chunk.rs:60-205, :292-387; gpu/mod.rs:662-914)
// ---------------------------------------------------------------------------

struct PathSource::Part {
  // regular files, mapped read-only and batched in memory
  std::vector<const uint8_t *> ptrs;
  std::vector<uint64_t> lens;
  std::vector<std::pair<void *, size_t>> maps;
  Batcher *bat = nullptr;
  // a stream (stdin, a pipe): three rotating buffers, so the two batches a
  // run may still hold stay valid while the next buffer is read
  bool stream = false;
  int fd = -1;
  bool own_fd = false;
  std::vector<uint8_t> buf[3];
  int cur = -1;
  uint64_t pos = 0, end = 0;
  std::vector<uint8_t> carry;
  bool eof = false, first = true;

  ~Part() {
    delete bat;
    for (auto &m : maps) munmap(m.first, m.second);
    if (own_fd && fd >= 0) close(fd);
  }

  // Fills buf[cur] with carried bytes plus a read of up to `cap` more, cut
  // after its last newline (stream_chunks, chunk.rs:292-387): the rest is
  // carried. A buffer with no newline keeps growing until one arrives or the
  // input ends, within max_line.
  bool fill(uint64_t cap, uint64_t max_line, std::string &err) {
    cur = (cur + 1) % 3;
    std::vector<uint8_t> &B = buf[cur];
    B.assign(carry.begin(), carry.end());
    carry.clear();
    pos = end = 0;
    for (;;) {
      const uint64_t base = B.size();
      const uint64_t want = base >= cap ? cap : cap - base;
      B.resize(base + want);
      uint64_t filled = base;
      while (filled < B.size()) {
        const ssize_t k = read(fd, B.data() + filled, B.size() - filled);
        if (k < 0) {
          if (errno == EINTR) continue;
          err = std::string("reading the input: ") + strerror(errno);
          return false;
        }
        if (k == 0) {
          eof = true;
          break;
        }
        filled += (uint64_t)k;
      }
      B.resize(filled);
      if (first && B.size() >= 3) {  // strip_bom, chunk.rs:210-212
        first = false;
        if (B[0] == 0xEF && B[1] == 0xBB && B[2] == 0xBF) B.erase(B.begin(), B.begin() + 3);
      } else if (first && eof) {
        first = false;
      }
      if (eof) {
        end = B.size();
        return true;
      }
      const uint8_t *nl = (const uint8_t *)memrchr(B.data(), '\n', B.size());
      if (!nl) {
        if (B.size() > max_line) {
          err = "a line exceeds the " + std::to_string(max_line) + " byte limit (raise it with --max-line-bytes)";
          return false;
        }
        continue;  // a line longer than what was read: read more
      }
      end = (uint64_t)(nl - B.data()) + 1;
      carry.assign(B.begin() + (ptrdiff_t)end, B.end());
      B.resize(end);
      return true;
    }
  }

  bool next_stream(Batch &b, std::string &err, uint64_t cap, uint64_t max_lines, uint64_t max_line) {
    for (;;) {
      if (cur >= 0 && pos < end) {
        const uint8_t *d = buf[cur].data();
        uint64_t p = pos;
        Chunk c;
        std::string e;
        bool too_long = false;
        if (!next_chunk(d, end, p, cap, max_lines, max_line, c, e, &too_long)) {
          if (!too_long) {
            err = e;
            return false;
          }
          // one line longer than the chunk: alone (gpu/mod.rs:411-458)
          const uint8_t *nl = (const uint8_t *)memchr(d + pos, '\n', end - pos);
          const uint64_t le = nl ? (uint64_t)(nl - d) + 1 : end;
          if (le - pos - (nl ? 1 : 0) > max_line) {
            err = "a line exceeds the " + std::to_string(max_line) + " byte limit (raise it with --max-line-bytes)";
            return false;
          }
          Piece pc{d + pos, le - pos};
          pc.big = true;
          b.push_back(pc);
          pos = le;
          return true;
        }
        b.push_back({d + c.off, c.len});
        pos = p;
        return true;
      }
      if (eof) return false;
      if (!fill(cap, max_line, err)) return false;
    }
  }
};

PathSource::PathSource(uint64_t max_line) : max_line_(max_line) {}

PathSource::~PathSource() {
  for (Part *p : parts_) delete p;
}

bool PathSource::open(const char *const *paths, uint32_t n, std::string &err) {
  auto stream_part = [&](int fd, bool own) {
    Part *p = new Part;
    p->stream = true;
    p->fd = fd;
    p->own_fd = own;
    parts_.push_back(p);
  };
  if (n == 0) {  // no files: stdin (run.rs:67-76)
    stream_part(0, false);
    return true;
  }
  for (uint32_t i = 0; i < n; i++) {
    const char *path = paths[i];
    if (!path || !strcmp(path, "-")) {
      stream_part(0, false);
      continue;
    }
    const int fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (fd < 0) {
      err = std::string("could not open `") + path + "`: " + strerror(errno);
      return false;
    }
    struct stat st;
    if (fstat(fd, &st) != 0) {
      err = std::string("could not open `") + path + "`: " + strerror(errno);
      close(fd);
      return false;
    }
    // only regular files are mapped (chunk.rs:61-80); an empty one adds nothing
    if (S_ISREG(st.st_mode)) {
      if (st.st_size == 0) {
        close(fd);
        continue;
      }
      void *m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
      const int e = errno;
      close(fd);
      if (m == MAP_FAILED) {
        err = std::string("could not map `") + path + "`: " + strerror(e);
        return false;
      }
      madvise(m, (size_t)st.st_size, MADV_SEQUENTIAL);
      madvise(m, (size_t)st.st_size, MADV_WILLNEED);
      if (parts_.empty() || parts_.back()->stream) parts_.push_back(new Part);
      Part *p = parts_.back();
      p->maps.push_back({m, (size_t)st.st_size});
      p->ptrs.push_back((const uint8_t *)m);
      p->lens.push_back((uint64_t)st.st_size);
      continue;
    }
    stream_part(fd, true);
  }
  return true;
}

bool PathSource::next(Batch &b, std::string &err) {
  b.clear();
  while (cur_ < parts_.size()) {
    Part *p = parts_[cur_];
    bool got;
    if (p->stream) {
      got = p->next_stream(b, err, cap, max_lines, max_line_);
    } else {
      if (!p->bat) p->bat = new Batcher(p->ptrs.data(), p->lens.data(), (uint32_t)p->ptrs.size(), max_line_, 4096);
      p->bat->cap = cap;
      p->bat->max_lines = max_lines;
      got = p->bat->next(b, err);
    }
    if (got) return true;
    if (!err.empty()) return false;
    cur_++;
  }
  return false;
}

}  // namespace wjqd

struct warpjq_amd_program {
  wjq::Program prog;
  wjq::Lowered low;
  warpjq_program view;
};

extern "C" {

void warpjq_amd_free(void *p) { free(p); }

warpjq_amd_program *warpjq_amd_compile(const char *query, int csv, char *err, size_t err_len) {
  auto *h = new warpjq_amd_program();
  wjq::QueryError qe;
  if (!wjq::parse(query, h->prog, qe)) {
    if (err && err_len) snprintf(err, err_len, "%s", qe.render(query).c_str());
    delete h;
    return nullptr;
  }
  wjq::lower(h->prog, csv != 0, h->low);
  h->view = h->low.view();
  return h;
}

const warpjq_program *warpjq_amd_program_view(const warpjq_amd_program *p) { return p ? &p->view : nullptr; }

void warpjq_amd_program_free(warpjq_amd_program *p) { delete p; }

}  // extern "C"

namespace wjqd {

int engine_open(Engine &g, const char *query, int format, uint64_t chunk_bytes, char *err, size_t err_len) {
  wjq::QueryError qe;
  if (!wjq::parse(query, g.prog, qe)) {
    snprintf(err, err_len, "%s", qe.render(query).c_str());
    return 1;
  }
  if (format == F_CSV && g.prog.output.kind == wjq::Output::PASSTHROUGH) {  // exec/mod.rs:256-266
    snprintf(err, err_len, "--csv needs a query with named columns, but this one emits whole lines");
    return 3;
  }
  g.csv = format == F_CSV;
  wjq::lower(g.prog, g.csv, g.low);
  g.view = g.low.view();
  g.cap = std::max<uint64_t>(1, std::min<uint64_t>(chunk_bytes, g.cap_limit));
  char e[512] = {0};
  if (warpjq_ctx_create(&g.view, g.cap, 2, &g.ctx, e, sizeof e) != WARPJQ_OK) {
    snprintf(err, err_len, "creating the device context: %s", e);
    g.ctx = nullptr;
    return 3;
  }
  g.max_lines = warpjq_max_lines(g.ctx);
  return 0;
}

int engine_reprogram(Engine &g, const char *query, int format, char *err, size_t err_len) {
  wjq::Program prog;
  wjq::QueryError qe;
  if (!wjq::parse(query, prog, qe)) {
    snprintf(err, err_len, "%s", qe.render(query).c_str());
    return 1;
  }
  if (format == F_CSV && prog.output.kind == wjq::Output::PASSTHROUGH) {  // exec/mod.rs:256-266
    snprintf(err, err_len, "--csv needs a query with named columns, but this one emits whole lines");
    return 3;
  }
  if (prog.is_aggregate() != g.prog.is_aggregate() || bool(prog.group_by) != bool(g.prog.group_by)) return 2;
  const bool csv = format == F_CSV;
  wjq::Lowered low;
  wjq::lower(prog, csv, low);
  warpjq_program view = low.view();
  char e[512] = {0};
  if (warpjq_amd_ctx_reprogram(g.ctx, &view, e, sizeof e) != WARPJQ_OK) return 2;
  g.prog = std::move(prog);
  g.low = std::move(low);
  g.view = g.low.view();
  g.csv = csv;
  return 0;
}

void engine_close(Engine &g) {
  if (g.ctx) warpjq_ctx_destroy(g.ctx);
  g.ctx = nullptr;
}

int engine_run(Engine &g, Source &src, int format, int on_invalid, uint8_t **out, uint64_t *out_len,
               warpjq_amd_stats *stats, char *err, size_t err_len) {
  *out = nullptr;
  *out_len = 0;
  memset(stats, 0, sizeof *stats);
  const wjq::Program &prog = g.prog;
  Run r{prog, (Format)format, (OnInvalid)on_invalid};
  r.ctx = g.ctx;
  r.cap = g.cap;
  r.max_lines = g.max_lines;
  src.cap = r.cap;
  src.max_lines = r.max_lines;
  if (src.dfiles && prog.is_aggregate()) {
    // resident input, aggregate: one launch takes up to 3.5 GiB of spans
    // (positions stay under 4 GiB with every span rounded up to a claim)
    src.cap = 7ull << 29;
    src.max_lines = ~0ull >> 2;
  }

  if (r.fmt == F_CSV) {  // write_header
    if (auto cols = prog.csv_headers()) {
      for (size_t i = 0; i < cols->size(); i++) {
        if (i) r.out.push_back(',');
        csv_cell(r.out, (*cols)[i]);
      }
      r.out.push_back('\n');
    }
  }

  bool inflight[2] = {false, false};
  auto put = [&](uint32_t sl, const Batch &b) {
    int rc2 = submit(r, sl, b);
    if (!rc2) inflight[sl] = true;
    return rc2;
  };
  auto take = [&](uint32_t sl, const Batch &b, std::vector<Batch> &redo) {
    inflight[sl] = false;
    return drain(r, sl, b, redo);
  };
  std::vector<Batch> pending;  // batches to resubmit after an overflow, in order
  bool have_prev = false;
  Batch prev;
  uint32_t slot = 0;
  int rc = 0;
  std::string cerr;
  // a line longer than the chunk runs alone, in order, on a context sized
  // for it (kept for the rest of the run)
  warpjq_ctx *big = nullptr;
  uint64_t big_cap = 0;
  auto run_big = [&](const Batch &b) -> int {
    const uint64_t need = b[0].len;
    if (!big || big_cap < need) {
      if (big) warpjq_ctx_destroy(big);
      big = nullptr;
      // room for the line and for rows up to 1.5x the capacity
      uint64_t c = std::max<uint64_t>(need + (need >> 1), 1ull << 20);
      c = std::min<uint64_t>((c + (1u << 20) - 1) & ~(uint64_t)((1u << 20) - 1), (4ull << 30) - (1ull << 20));
      if (c < need) {
        r.err = "a line is longer than the device can take in one chunk";
        return 3;
      }
      char e[512] = {0};
      if (warpjq_ctx_create(&g.view, c, 1, &big, e, sizeof e) != WARPJQ_OK) {
        r.err = std::string("creating a context for a long line: ") + e;
        big = nullptr;
        return 3;
      }
      big_cap = c;
    }
    warpjq_ctx *keep = r.ctx;
    r.ctx = big;
    int rc2 = submit(r, 0, b);
    if (!rc2) {
      std::vector<Batch> redo;
      rc2 = drain(r, 0, b, redo);
      if (!rc2 && !redo.empty()) {
        r.err = "a single line exceeds the device's per-chunk capacity";
        rc2 = 3;
      }
    }
    r.ctx = keep;
    return rc2;
  };
  for (;;) {
    Batch c;
    bool got;
    if (!pending.empty()) {
      c = pending.front();
      pending.erase(pending.begin());
      got = true;
    } else {
      got = src.next(c, cerr);
      if (!got && !cerr.empty()) {
        rc = 3;
        r.err = cerr;
        break;
      }
    }
    if (!got) break;
    if (c.size() == 1 && c[0].big) {
      // everything before it first, in order
      if (have_prev) {
        std::vector<Batch> redo;
        if ((rc = take(slot ^ 1, prev, redo))) break;
        have_prev = false;
        if (!redo.empty()) {
          redo.push_back(c);
          pending.insert(pending.begin(), redo.begin(), redo.end());
          continue;
        }
      }
      if ((rc = run_big(c))) break;
      continue;
    }
    if ((rc = put(slot, c))) break;
    if (have_prev) {
      std::vector<Batch> redo;
      if ((rc = take(slot ^ 1, prev, redo))) break;
      if (!redo.empty()) {
        // the previous batch must be redone before this one: drain this one
        // too, discard it, and queue redo + this batch again
        warpjq_chunk_result tmp;
        char e2[512];
        warpjq_wait(r.ctx, slot, &tmp, e2, sizeof e2);
        inflight[slot] = false;
        r.st.bytes_in -= batch_bytes(prev) + batch_bytes(c);
        redo.push_back(c);
        pending.insert(pending.begin(), redo.begin(), redo.end());
        have_prev = false;
        continue;
      }
    }
    prev = c;
    have_prev = true;
    slot ^= 1;
  }
  if (!rc && have_prev) {
    std::vector<Batch> redo;
    rc = take(slot ^ 1, prev, redo);
    if (!rc && !redo.empty()) {
      r.st.bytes_in -= batch_bytes(prev);
      std::vector<Batch> todo = redo;
      for (size_t i = 0; i < todo.size() && !rc; i++) {
        std::vector<Batch> more;
        if ((rc = put(0, todo[i]))) break;
        rc = take(0, todo[i], more);
        if (!rc && !more.empty()) {
          r.st.bytes_in -= batch_bytes(todo[i]);
          todo.insert(todo.begin() + i + 1, more.begin(), more.end());
        }
      }
    }
  }
  // leave no slot in flight: the context is reused by the next run
  for (uint32_t sl = 0; sl < 2; sl++)
    if (inflight[sl]) {
      warpjq_chunk_result tmp;
      char e2[512];
      warpjq_wait(r.ctx, sl, &tmp, e2, sizeof e2);
    }
  if (big) warpjq_ctx_destroy(big);
  stats->bytes_in = r.st.bytes_in;
  stats->lines_in = r.st.lines_in;
  stats->malformed = r.st.malformed;
  stats->type_errors = r.st.type_errors;
  stats->gpu_fallback_lines = r.st.gpu_fallback_lines;
  stats->gpu_redone_chunks = r.st.gpu_redone_chunks;
  stats->h2d_ms = r.st.h2d_ms;
  stats->kernel_ms = r.st.kernel_ms;
  stats->chunks = r.st.chunks;
  stats->err_line = r.err_line;
  if (rc) {
    snprintf(err, err_len, "%s", r.err.c_str());
    return rc;
  }
  // finish_totals (exec/mod.rs:321-339) and Writer::finish
  if (prog.is_aggregate()) {
    wjq::AggKind kind = prog.output.agg;
    if (!prog.group_by) {
      double v;
      std::string value = r.scalar.finish(kind, v) ? format_number(v) : "null";
      if (r.fmt == F_COUNT) {
        r.out = value + "\n";
        r.rows = 0;
        r.fmt = F_NDJSON;
      } else {
        r.out += value + "\n";
        r.rows++;
      }
    } else {
      std::string name = group_key_name(prog);
      for (auto &kv : r.groups) {
        r.rows++;
        double v;
        std::string value = kv.second.finish(kind, v) ? format_number(v) : "null";
        if (r.fmt == F_CSV) {
          csv_cell(r.out, kv.first.rank == 4 ? kv.first.bytes : group_key_json(kv.first));
          r.out += "," + value;
        } else {
          r.out.push_back('{');
          wjq::write_json_string(r.out, (const uint8_t *)name.data(), name.size());
          r.out += ":" + group_key_json(kv.first) + ",";
          const char *kn = wjq::agg_name(kind);
          wjq::write_json_string(r.out, (const uint8_t *)kn, strlen(kn));
          r.out += ":" + value + "}";
        }
        r.out.push_back('\n');
      }
    }
  }
  if (r.fmt == F_COUNT) r.out = std::to_string(r.rows) + "\n";
  stats->lines_out = r.rows;
  *out = (uint8_t *)malloc(r.out.size() + 1);
  if (!*out) {
    snprintf(err, err_len, "out of host memory for the output");
    return 3;
  }
  memcpy(*out, r.out.data(), r.out.size());
  *out_len = r.out.size();
  return 0;
}

}  // namespace wjqd

extern "C" {

static int run_once(const char *query, wjqd::Source &src, int format, int on_invalid, uint64_t chunk_bytes,
                    uint8_t **out, uint64_t *out_len, warpjq_amd_stats *stats, char *err, size_t err_len) {
  *out = nullptr;
  *out_len = 0;
  memset(stats, 0, sizeof *stats);
  wjqd::Engine g;
  int rc = wjqd::engine_open(g, query, format, chunk_bytes, err, err_len);
  if (rc) return rc;
  rc = wjqd::engine_run(g, src, format, on_invalid, out, out_len, stats, err, err_len);
  wjqd::engine_close(g);
  return rc;
}

int warpjq_amd_run_buffer(const char *query, const uint8_t *data, uint64_t n, int format, int on_invalid,
                          uint64_t chunk_bytes, uint64_t max_line_bytes, uint8_t **out, uint64_t *out_len,
                          warpjq_amd_stats *stats, char *err, size_t err_len) {
  wjqd::Batcher src(&data, &n, 1, max_line_bytes, 1);
  return run_once(query, src, format, on_invalid, chunk_bytes, out, out_len, stats, err, err_len);
}

int warpjq_amd_run_csv(const char *query, const uint8_t *data, uint64_t n, int format, int on_invalid,
                       uint64_t chunk_bytes, uint64_t max_line_bytes, uint8_t **out, uint64_t *out_len,
                       warpjq_amd_stats *stats, char *err, size_t err_len) {
  wjqd::HostBytes nd;
  uint64_t records = 0;
  float kms = 0;
  std::string e;
  uint8_t *dev = nullptr;
  *out = nullptr;
  *out_len = 0;
  memset(stats, 0, sizeof *stats);
  if (!data && n) {
    snprintf(err, err_len, "run_csv: null input");
    return 3;
  }
  if (!wjqd::csv_transcode(data, n, nd, &dev, &records, &kms, e)) {
    wjqd::csv_free(dev);
    snprintf(err, err_len, "%s", e.c_str());
    return 3;
  }
  // the run reads the transcoded bytes where they already are, in device
  // memory; the host copy serves overflow splits and row cuts
  const uint8_t *hp = nd.p;
  uint64_t hn = nd.n;
  const uint8_t *dp = dev;
  wjqd::Batcher src(&hp, &hn, 1, max_line_bytes, 1);
  src.dfiles = &dp;
  const int rc = run_once(query, src, format, on_invalid, chunk_bytes, out, out_len, stats, err, err_len);
  wjqd::csv_free(dev);
  stats->bytes_in = n;
  stats->kernel_ms += kms;
  return rc;
}

int warpjq_amd_run_files(const char *query, const uint8_t *const *files, const uint64_t *lens, uint32_t n_files,
                         int format, int on_invalid, uint64_t chunk_bytes, uint64_t max_line_bytes, uint8_t **out,
                         uint64_t *out_len, warpjq_amd_stats *stats, char *err, size_t err_len) {
  if (n_files && (!files || !lens)) {
    *out = nullptr;
    *out_len = 0;
    snprintf(err, err_len, "run_files: null file list");
    return 3;
  }
  wjqd::Batcher src(files, lens, n_files, max_line_bytes, 4096);
  return run_once(query, src, format, on_invalid, chunk_bytes, out, out_len, stats, err, err_len);
}

int warpjq_amd_run_paths(const char *query, const char *const *paths, uint32_t n_paths, int format, int on_invalid,
                         uint64_t chunk_bytes, uint64_t max_line_bytes, uint8_t **out, uint64_t *out_len,
                         warpjq_amd_stats *stats, char *err, size_t err_len) {
  *out = nullptr;
  *out_len = 0;
  memset(stats, 0, sizeof *stats);
  if (n_paths && !paths) {
    snprintf(err, err_len, "run_paths: null path list");
    return 3;
  }
  wjqd::PathSource src(max_line_bytes);
  std::string e;
  if (!src.open(paths, n_paths, e)) {
    snprintf(err, err_len, "%s", e.c_str());
    return 3;
  }
  return run_once(query, src, format, on_invalid, chunk_bytes, out, out_len, stats, err, err_len);
}

}  // extern "C"
