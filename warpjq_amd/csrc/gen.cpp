// Seeded synthetic NDJSON for benchmarks and large parity checks. Each line
// depends only on (seed, line index) through a counter-based generator, so
// any range can be regenerated independently and threads split the work.
//
// nginx schema (one flat object per line, ~220 bytes):
//   ts      1700000000 + line index
//   host    one of 8, uniform
//   method  one of 5, uniform
//   path    one of 32, Zipf s = 1.0
//   status  200 80%, 404 10%, 500 5%, 301/304/403/502/503 1% each
//   bytes   floor(lognormal(mu = 8, sigma = 2)), capped at 2^40
//   ua      40-120 printable ASCII characters, no '"' or '\'
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/warpjq_amd.h"

namespace {

struct Rng {
  uint64_t s;
  uint64_t next() {  // splitmix64
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

const char *HOSTS[8] = {"web-01.prod.us-east-1.example.com", "web-02.prod.us-east-1.example.com",
                        "api-1.prod.eu-west-2.internal",     "api-2.prod.eu-west-2.internal",
                        "cdn-edge.ap-south-1.example.net",   "static-assets.us-west-2.example.net",
                        "auth.prod.us-east-1.example.com",   "edge-7.sa-east-1.example.org"};
const char *METHODS[5] = {"GET", "POST", "PUT", "DELETE", "HEAD"};
const char *PATHS[32] = {"/",          "/index.html",    "/api/v1/users",  "/api/v1/orders", "/api/v1/items",
                         "/login",     "/logout",        "/search",        "/cart",          "/checkout",
                         "/health",    "/metrics",       "/static/app.js", "/static/app.css", "/images/logo.png",
                         "/favicon.ico", "/api/v2/feed", "/api/v2/comments", "/api/v2/likes", "/profile",
                         "/settings",  "/billing",       "/invoices",      "/reports",       "/admin",
                         "/upload",    "/download",      "/export",        "/import",        "/notifications",
                         "/messages",  "/robots.txt"};
const int STATUS[8] = {200, 404, 500, 301, 304, 403, 502, 503};

struct Tables {
  double zipf_cdf[32];
  char ua_chars[96];
  int n_ua;
  Tables() {
    double z = 0;
    for (int k = 0; k < 32; k++) z += 1.0 / (k + 1);
    double acc = 0;
    for (int k = 0; k < 32; k++) {
      acc += 1.0 / (k + 1) / z;
      zipf_cdf[k] = acc;
    }
    zipf_cdf[31] = 1.0;
    n_ua = 0;
    for (int c = 0x20; c < 0x7F; c++)
      if (c != '"' && c != '\\') ua_chars[n_ua++] = (char)c;
  }
};
const Tables T;

int status_index(double u) {
  if (u < 0.80) return 0;
  if (u < 0.90) return 1;
  if (u < 0.95) return 2;
  int k = 3 + (int)((u - 0.95) / 0.01);
  return k > 7 ? 7 : k;
}

// Appends line `i` (with its newline) to `out`; returns its status index.
int nginx_line(uint64_t seed, uint64_t i, std::string &out) {
  Rng r{seed ^ (i * 0xD1B54A32D192ED03ull)};
  r.next();
  const char *host = HOSTS[r.next() % 8];
  const char *method = METHODS[r.next() % 5];
  double u = r.uniform();
  int pk = 0;
  while (pk < 31 && u > T.zipf_cdf[pk]) pk++;
  int si = status_index(r.uniform());
  double u1 = r.uniform(), u2 = r.uniform();
  if (u1 < 1e-300) u1 = 1e-300;
  double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  double b = std::floor(std::exp(8.0 + 2.0 * z));
  uint64_t bytes = b >= 1099511627776.0 ? 1099511627776ull : (uint64_t)b;
  int ua_len = 40 + (int)(r.next() % 81);
  char buf[512];
  int n = snprintf(buf, sizeof buf,
                   "{\"ts\":%llu,\"host\":\"%s\",\"method\":\"%s\",\"path\":\"%s\",\"status\":%d,\"bytes\":%llu,"
                   "\"ua\":\"",
                   (unsigned long long)(1700000000ull + i), host, method, PATHS[pk], STATUS[si],
                   (unsigned long long)bytes);
  out.append(buf, n);
  uint64_t bits = 0;
  int have = 0;
  for (int k = 0; k < ua_len; k++) {
    if (have < 7) {
      bits = r.next();
      have = 64;
    }
    out.push_back(T.ua_chars[(bits & 0x7F) % T.n_ua]);
    bits >>= 7;
    have -= 7;
  }
  out.append("\"}\n");
  return si;
}

// nested schema (~330 bytes, objects six deep):
//   {"ts":T,"user":U,"bytes":B,"resp":{"latency":L,"code":C,"meta":{"region":R,
//    "tags":[..],"d":{"e":{"g":{"f":K}}}}},"x\u002did":"..","ok":true|false}
//   user    "u" + 11 base-36 chars, 1M distinct values, Zipf s = 1.1; the key
//           is written "us\u0065r" on 2% of lines
//   bytes   an integer floor(lognormal(10, 3)) capped at 2^53 on 70% of
//           lines, %.17g of U[0, 1e6) on 28%, null on 2%
//   latency %.3f of lognormal(3, 1)
//   0.1% of lines are malformed: a leading zero, NaN, or the line cut short
constexpr uint32_t N_USERS = 1u << 20;

const std::vector<double> &user_cdf() {
  static const std::vector<double> cdf = [] {
    std::vector<double> c(N_USERS);
    double acc = 0;
    for (uint32_t k = 0; k < N_USERS; k++) {
      acc += std::pow((double)(k + 1), -1.1);
      c[k] = acc;
    }
    for (auto &x : c) x /= acc;
    c.back() = 1.0;
    return c;
  }();
  return cdf;
}

const char *REGIONS[6] = {"us-east-1", "us-west-2", "eu-west-2", "ap-south-1", "sa-east-1", "eu-central-1"};
const int CODES[6] = {200, 201, 204, 400, 404, 500};

// Appends line `i`; returns 1 if it is malformed.
int nested_line(uint64_t seed, uint64_t i, std::string &out) {
  Rng r{seed ^ (i * 0xA24BAED4963EE407ull) ^ 0x5DEECE66Dull};
  r.next();
  const std::vector<double> &cdf = user_cdf();
  const double u = r.uniform();
  uint32_t lo = 0, hi = N_USERS - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) / 2;
    if (cdf[mid] < u) lo = mid + 1;
    else hi = mid;
  }
  uint64_t x = ((uint64_t)lo * 0x9E3779B97F4A7C15ull) % 131621703842267136ull;  // 36^11
  char user[13];
  user[0] = 'u';
  for (int k = 11; k >= 1; k--) {
    const int dgt = (int)(x % 36);
    user[k] = (char)(dgt < 10 ? '0' + dgt : 'a' + dgt - 10);
    x /= 36;
  }
  user[12] = 0;
  const bool esc_user = r.uniform() < 0.02;
  char bytes[40];
  const double bu = r.uniform();
  if (bu < 0.70) {
    double u1 = r.uniform(), u2 = r.uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    double b = std::floor(std::exp(10.0 + 3.0 * z));
    if (b > 9007199254740992.0) b = 9007199254740992.0;
    snprintf(bytes, sizeof bytes, "%llu", (unsigned long long)b);
  } else if (bu < 0.98) {
    snprintf(bytes, sizeof bytes, "%.17g", r.uniform() * 1e6);
  } else {
    snprintf(bytes, sizeof bytes, "null");
  }
  double u1 = r.uniform(), u2 = r.uniform();
  if (u1 < 1e-300) u1 = 1e-300;
  const double lat = std::exp(3.0 + std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
  const int code = CODES[r.next() % 6];
  const char *region = REGIONS[r.next() % 6];
  const int ntags = (int)(r.next() % 4);
  std::string tags;
  for (int k = 0; k < ntags; k++) {
    if (k) tags.push_back(',');
    tags += "\"t" + std::to_string(r.next() % 50) + "\"";
  }
  const uint64_t xid = r.next();
  const bool ok = (r.next() & 1) != 0;
  const double bad = r.uniform();
  int malformed = 0;
  if (bad < 0.0003) {  // leading zero
    snprintf(bytes, sizeof bytes, "0%llu", (unsigned long long)(1 + r.next() % 99999));
    malformed = 1;
  } else if (bad < 0.0006) {
    snprintf(bytes, sizeof bytes, "NaN");
    malformed = 1;
  }
  char buf[768];
  int n = snprintf(buf, sizeof buf,
                   "{\"ts\":%llu,\"%s\":\"%s\",\"bytes\":%s,\"resp\":{\"latency\":%.3f,\"code\":%d,\"meta\":{"
                   "\"region\":\"%s\",\"tags\":[%s],\"d\":{\"e\":{\"g\":{\"f\":%d}}}}},"
                   "\"x\\u002did\":\"%016llx\",\"ok\":%s}",
                   (unsigned long long)(1700000000ull + i), esc_user ? "us\\u0065r" : "user", user, bytes, lat, code,
                   region, tags.c_str(), (int)(r.next() % 1000), (unsigned long long)xid, ok ? "true" : "false");
  if (bad >= 0.0006 && bad < 0.001) {  // cut short
    n = 10 + (int)(r.next() % (uint64_t)(n - 20));
    malformed = 1;
  }
  out.append(buf, n);
  out.push_back('\n');
  return malformed;
}

}  // namespace

extern "C" {

// Nested-schema lines, whole, until adding one more would pass target_bytes
// (same contract as warpjq_amd_gen_nginx); *n_malformed counts the lines
// made malformed on purpose.
int warpjq_amd_gen_nested(uint64_t seed, uint64_t target_bytes, uint32_t threads, uint64_t pad, uint8_t **out,
                          uint64_t *n_bytes, uint64_t *n_lines, uint64_t *n_malformed) {
  if (threads == 0) threads = 1;
  (void)user_cdf();
  const uint64_t est_lines = target_bytes / 150 + 64;
  const uint64_t per = (est_lines + threads - 1) / threads;
  std::vector<std::string> parts(threads);
  std::vector<std::vector<uint64_t>> ends(threads);
  std::vector<std::vector<uint8_t>> bad(threads);
  std::vector<std::thread> ts;
  for (uint32_t t = 0; t < threads; t++) {
    ts.emplace_back([&, t] {
      uint64_t a = t * per, b = std::min<uint64_t>(est_lines, a + per);
      parts[t].reserve((b - a) * 340);
      for (uint64_t i = a; i < b; i++) {
        bad[t].push_back((uint8_t)nested_line(seed, i, parts[t]));
        ends[t].push_back(parts[t].size());
        if (parts[t].size() > target_bytes) break;
      }
    });
  }
  for (auto &th : ts) th.join();
  uint64_t total = 0, lines = 0, nbad = 0;
  std::vector<uint64_t> take(threads, 0);
  for (uint32_t t = 0; t < threads && total < target_bytes; t++) {
    const auto &e = ends[t];
    uint64_t lo = 0, hi = e.size();
    while (lo < hi) {
      uint64_t mid = (lo + hi) / 2;
      if (total + e[mid] <= target_bytes) lo = mid + 1;
      else hi = mid;
    }
    take[t] = lo ? e[lo - 1] : 0;
    total += take[t];
    lines += lo;
    for (uint64_t k = 0; k < lo; k++) nbad += bad[t][k];
    if (lo < e.size()) break;
  }
  uint8_t *buf = (uint8_t *)malloc(total + pad);
  if (!buf) return 3;
  uint64_t off = 0;
  for (uint32_t t = 0; t < threads; t++) {
    if (!take[t]) continue;
    memcpy(buf + off, parts[t].data(), take[t]);
    off += take[t];
  }
  memset(buf + total, 0, pad);
  *out = buf;
  *n_bytes = total;
  *n_lines = lines;
  *n_malformed = nbad;
  return 0;
}


// Generates whole nginx-schema lines until adding one more would pass
// target_bytes. *out is malloc'd (free with warpjq_amd_free) and padded with
// `pad` zero bytes past *n_bytes. status_hist[8] counts lines per status in
// the order 200, 404, 500, 301, 304, 403, 502, 503.
int warpjq_amd_gen_nginx(uint64_t seed, uint64_t target_bytes, uint32_t threads, uint64_t pad, uint8_t **out,
                         uint64_t *n_bytes, uint64_t *n_lines, uint64_t *status_hist) {
  if (threads == 0) threads = 1;
  // Lines average ~220 B; give each thread a contiguous range of line
  // indices sized to overshoot, then trim to the byte target.
  const uint64_t est_lines = target_bytes / 180 + 64;
  const uint64_t per = (est_lines + threads - 1) / threads;
  std::vector<std::string> parts(threads);
  std::vector<std::vector<uint64_t>> ends(threads);  // cumulative line ends per part
  std::vector<std::vector<uint8_t>> stat(threads);
  std::vector<std::thread> ts;
  for (uint32_t t = 0; t < threads; t++) {
    ts.emplace_back([&, t] {
      uint64_t a = t * per, b = std::min<uint64_t>(est_lines, a + per);
      parts[t].reserve((b - a) * 232);
      for (uint64_t i = a; i < b; i++) {
        stat[t].push_back((uint8_t)nginx_line(seed, i, parts[t]));
        ends[t].push_back(parts[t].size());
        if (parts[t].size() > target_bytes) break;
      }
    });
  }
  for (auto &th : ts) th.join();
  uint64_t total = 0, lines = 0;
  for (int k = 0; k < 8; k++) status_hist[k] = 0;
  std::vector<uint64_t> take(threads, 0), take_lines(threads, 0);
  for (uint32_t t = 0; t < threads && total < target_bytes; t++) {
    const auto &e = ends[t];
    // whole lines of this part that still fit
    uint64_t lo = 0, hi = e.size();
    while (lo < hi) {
      uint64_t mid = (lo + hi) / 2;
      if (total + e[mid] <= target_bytes) lo = mid + 1;
      else hi = mid;
    }
    take_lines[t] = lo;
    take[t] = lo ? e[lo - 1] : 0;
    total += take[t];
    lines += lo;
    for (uint64_t k = 0; k < lo; k++) status_hist[stat[t][k]]++;
    if (lo < e.size()) break;
  }
  uint8_t *buf = (uint8_t *)malloc(total + pad);
  if (!buf) return 3;
  uint64_t off = 0;
  for (uint32_t t = 0; t < threads; t++) {
    if (!take[t]) continue;
    memcpy(buf + off, parts[t].data(), take[t]);
    off += take[t];
  }
  memset(buf + total, 0, pad);
  *out = buf;
  *n_bytes = total;
  *n_lines = lines;
  return 0;
}

}  // extern "C"
