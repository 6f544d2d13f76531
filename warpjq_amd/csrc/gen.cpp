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

}  // namespace

extern "C" {

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
