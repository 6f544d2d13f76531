// CSV input, transcoded on the device into the NDJSON the main kernels read.
//
// The reference reads NDJSON only; CSV input is this build's own format, so
// its results are pinned only against the oracle's independent restatement
// of the same specification (oracle/oracle.cpp, "CSV input"):
//   * the first line is the header (after a UTF-8 byte order mark, which is
//     dropped, as the NDJSON reader drops one per input) and names the
//     columns; it becomes a blank line, so line numbers stay the file's;
//   * every later line is one record, the object {"name":value,...} in
//     column order; a trailing CR is dropped;
//   * a field is quoted when it starts with '"'; inside, "" is one quote; the
//     closing quote ends the line or is followed by a comma; an unquoted field
//     holds no quote;
//   * an empty unquoted field is null, an unquoted JSON number stays a number,
//     anything else (every quoted field) is a string (quote, backslash and
//     control bytes escaped);
//   * a line with bad quoting or another field count than the header's
//     becomes the invalid line "{" (counted malformed); an empty line stays
//     blank.
// Device passes: newlines per 4 KiB tile, an exclusive scan, every newline's
// position; then per record its output length, a scan, and the output bytes.
// One lane walks one record (records are short and independent), from LDS
// where its group of 64 records fits there (k_csv_pass).
#include <sys/mman.h>

#include <hipcub/hipcub.hpp>
#include <mutex>
#include <system_error>
#include <thread>

#include "driver.h"
#include "kernels.h"

namespace {

constexpr uint32_t CSV_TILE = 4096;  // bytes per wave in the newline passes
constexpr uint64_t CSV_DEV_PAD = 256;  // readable bytes the kernels may touch past a span

struct CsvHdr {
  uint32_t ncols;
  const uint8_t *keys;      // the JSON-encoded column names, back to back
  const uint32_t *key_off;  // [ncols]
  const uint32_t *key_len;  // [ncols]
};

// Newlines in each 4 KiB tile of d[0..n) (bytes past n are zero).
__global__ __launch_bounds__(256) void k_csv_nl_count(const uint8_t *d, uint64_t ntiles, uint64_t *cnt) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < ntiles; t += nw) {
    const uint8_t *p = d + t * CSV_TILE + 64u * lane;
    uint32_t c = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) c += __popc(nl_mask16(*(const uint4 *)(p + 16u * j)));
    c = wave_sum(c);
    if (lane == 0) cnt[t] = c;
  }
}

// The position of every newline, in order: pos[pre[t] + k] for the k-th
// newline of tile t.
__global__ __launch_bounds__(256) void k_csv_nl_pos(const uint8_t *d, uint64_t ntiles, const uint64_t *pre,
                                                     uint64_t *pos) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < ntiles; t += nw) {
    const uint64_t b0 = t * CSV_TILE + 64u * lane;
    uint64_t m = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) m |= (uint64_t)nl_mask16(*(const uint4 *)(d + b0 + 16u * j)) << (16u * j);
    uint32_t tot;
    const uint32_t ex = wave_excl_scan((uint32_t)__popcll(m), tot);
    uint64_t k = pre[t] + ex;
    while (m) {
      pos[k++] = b0 + (uint64_t)__builtin_ctzll(m);
      m &= m - 1;
    }
  }
}

__device__ __forceinline__ bool csv_digit(uint32_t c) { return c - '0' < 10u; }

// [a, b) of the record is a JSON number.
__device__ __forceinline__ bool csv_number(const uint8_t *r, uint32_t a, uint32_t b) {
  uint32_t i = a;
  if (i < b && r[i] == '-') i++;
  if (i >= b) return false;
  if (r[i] == '0') {
    i++;
  } else if (r[i] >= '1' && r[i] <= '9') {
    while (i < b && csv_digit(r[i])) i++;
  } else {
    return false;
  }
  if (i < b && r[i] == '.') {
    const uint32_t d = ++i;
    while (i < b && csv_digit(r[i])) i++;
    if (i == d) return false;
  }
  if (i < b && (r[i] == 'e' || r[i] == 'E')) {
    i++;
    if (i < b && (r[i] == '+' || r[i] == '-')) i++;
    const uint32_t d = i;
    while (i < b && csv_digit(r[i])) i++;
    if (i == d) return false;
  }
  return i == b;
}

// The fields of record r[0..n): false on bad quoting, else their count.
__device__ __forceinline__ bool csv_count(const uint8_t *r, uint32_t n, uint32_t &fields) {
  uint32_t i = 0;
  fields = 0;
  for (;;) {
    if (i < n && r[i] == '"') {
      i++;
      for (;;) {
        while (i < n && r[i] != '"') i++;
        if (i >= n) return false;
        if (i + 1 < n && r[i + 1] == '"') {
          i += 2;
          continue;
        }
        i++;
        break;
      }
      if (i < n && r[i] != ',') return false;
    } else {
      while (i < n && r[i] != ',') {
        if (r[i] == '"') return false;
        i++;
      }
    }
    fields++;
    if (i >= n) return true;
    i++;
  }
}

struct CountSink {
  uint64_t n = 0;
  __device__ void put(uint8_t) { n++; }
  __device__ void put_n(const uint8_t *, uint32_t len) { n += len; }
};

struct ByteSink {
  uint8_t *o;
  __device__ void put(uint8_t c) { *o++ = c; }
  __device__ void put_n(const uint8_t *s, uint32_t len) {
    for (uint32_t k = 0; k < len; k++) o[k] = s[k];
    o += len;
  }
};

template <class S>
__device__ __forceinline__ void csv_escaped(S &s, uint32_t c) {
  if (c == '"' || c == '\\') {
    s.put('\\');
    s.put((uint8_t)c);
  } else if (c < 0x20) {
    const char *hex = "0123456789abcdef";
    s.put('\\');
    s.put('u');
    s.put('0');
    s.put('0');
    s.put((uint8_t)hex[c >> 4]);
    s.put((uint8_t)hex[c & 15]);
  } else {
    s.put((uint8_t)c);
  }
}

// The NDJSON line (with its newline) of record r[0..n), already counted and
// quoted correctly (csv_count).
template <class S>
__device__ __forceinline__ void csv_object(const uint8_t *r, uint32_t n, const CsvHdr &h, S &s) {
  s.put('{');
  uint32_t i = 0;
  for (uint32_t c = 0;; c++) {
    if (c) s.put(',');
    s.put_n(h.keys + h.key_off[c], h.key_len[c]);
    s.put(':');
    if (i < n && r[i] == '"') {
      s.put('"');
      i++;
      for (;;) {
        const uint32_t ch = r[i];
        if (ch == '"') {
          if (i + 1 < n && r[i + 1] == '"') {
            csv_escaped(s, '"');
            i += 2;
            continue;
          }
          i++;
          break;
        }
        csv_escaped(s, ch);
        i++;
      }
      s.put('"');
    } else {
      uint32_t k = i;
      while (k < n && r[k] != ',') k++;
      if (k == i) {
        s.put_n((const uint8_t *)"null", 4);
      } else if (csv_number(r, i, k)) {
        s.put_n(r + i, k - i);
      } else {
        s.put('"');
        for (uint32_t j = i; j < k; j++) csv_escaped(s, r[j]);
        s.put('"');
      }
      i = k;
    }
    if (i >= n) break;
    i++;
  }
  s.put('}');
  s.put('\n');
}

// Record k: its bytes and length without the newline and a trailing CR;
// false (and n = 0) for a record of 4 GiB or more.
__device__ __forceinline__ const uint8_t *csv_record(const uint8_t *d, const uint64_t *pos, uint64_t k, uint32_t &n,
                                                     bool &ok) {
  const uint64_t a = k ? pos[k - 1] + 1 : 0, b = pos[k];
  uint64_t len = b - a;
  if (len && d[b - 1] == '\r') len--;
  ok = len <= 0xFFFFFFFFull;
  n = ok ? (uint32_t)len : 0u;
  return d + a;
}

template <class S>
__device__ __forceinline__ void csv_line(const uint8_t *r, uint32_t n, const CsvHdr &h, S &s) {
  uint32_t f;
  if (n == 0) {
    s.put('\n');
  } else if (!csv_count(r, n, f) || f != h.ncols) {
    s.put('{');
    s.put('\n');
  } else {
    csv_object(r, n, h, s);
  }
}

// Records go in groups of 64, one wave per group. A group whose bytes fit
// in LDS is staged there with dword loads; each lane reads its record from
// LDS and (EMIT) writes its object into LDS, and the wave stores the
// group's output with dword stores, the ragged first and last bytes singly
// (the neighbouring groups own the rest of those dwords). A larger group is
// walked by each lane from global memory.
constexpr uint32_t CSV_IN_LDS = 16384, CSV_OUT_LDS = 24576;

// EMIT false: len[k] = the length of record k's line (too_long set for a
// record of 4 GiB or more). EMIT true: the line of record k written at
// out[shift + off[k]].
template <bool EMIT>
__global__ __launch_bounds__(64) void k_csv_pass(const uint8_t *d, const uint64_t *pos, uint64_t nrec, CsvHdr h,
                                                 uint64_t *len, uint32_t *too_long, const uint64_t *off, uint8_t *out,
                                                 uint64_t shift) {
  __shared__ __align__(16) uint8_t lin[CSV_IN_LDS];
  __shared__ __align__(16) uint8_t lout[EMIT ? CSV_OUT_LDS : 4];
  const uint32_t lane = threadIdx.x;
  const uint64_t ngroups = (nrec + 63) / 64;
  for (uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
    const uint64_t k0 = g * 64, k1 = min(nrec, k0 + 64), k = k0 + lane;
    const uint64_t a = k0 ? pos[k0 - 1] + 1 : 0, b = pos[k1 - 1];
    const uint64_t a4 = a & ~3ull, in_n = b + 1 - a4;
    uint64_t o0 = 0, o1 = 0, o4 = 0;
    if (EMIT) {
      o0 = shift + off[k0];
      o1 = shift + off[k1];
      o4 = o0 & ~3ull;
    }
    if (in_n <= CSV_IN_LDS && (!EMIT || o1 - o4 <= CSV_OUT_LDS)) {
      // dwords from a4 through the one holding b: at most 3 bytes past b,
      // inside the input's zero pad
      const uint32_t nw = (uint32_t)((in_n + 3) >> 2);
      for (uint32_t i = lane; i < nw; i += 64) ((uint32_t *)lin)[i] = *(const uint32_t *)(d + a4 + 4ull * i);
      __syncthreads();
      if (k < k1) {
        const uint64_t ra = k ? pos[k - 1] + 1 : 0, rb = pos[k];
        uint32_t n = (uint32_t)(rb - ra);
        if (n && lin[rb - 1 - a4] == '\r') n--;
        const uint8_t *r = lin + (ra - a4);
        if (EMIT) {
          ByteSink s{lout + (shift + off[k] - o4)};
          csv_line(r, n, h, s);
        } else {
          CountSink s;
          csv_line(r, n, h, s);
          len[k] = s.n;
        }
      }
      __syncthreads();
      if (EMIT) {
        const uint64_t m0 = (o0 + 3) & ~3ull, m1 = o1 & ~3ull;
        if (m0 >= m1) {
          for (uint64_t i = o0 + lane; i < o1; i += 64) out[i] = lout[i - o4];
        } else {
          if (lane < m0 - o0) out[o0 + lane] = lout[o0 + lane - o4];
          if (lane < o1 - m1) out[m1 + lane] = lout[m1 + lane - o4];
          for (uint64_t i = m0 + 4ull * lane; i < m1; i += 256)
            *(uint32_t *)(out + i) = *(const uint32_t *)(lout + (i - o4));
        }
        __syncthreads();  // the next group overwrites lout
      }
    } else if (k < k1) {
      uint32_t n;
      bool ok;
      const uint8_t *r = csv_record(d, pos, k, n, ok);
      if (EMIT) {
        ByteSink s{out + shift + off[k]};
        csv_line(r, n, h, s);
      } else {
        if (!ok) *too_long = 1u;
        CountSink s;
        csv_line(r, n, h, s);
        len[k] = s.n;
      }
    }
  }
}

// Host: the fields of the header line, as the device reads a record.
bool header_fields(const uint8_t *b, size_t n, std::vector<std::string> &f) {
  f.clear();
  size_t i = 0;
  for (;;) {
    std::string v;
    if (i < n && b[i] == '"') {
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
    f.push_back(std::move(v));
    if (i >= n) return true;
    i++;
  }
}

std::string json_key(const std::string &s) {
  static const char hex[] = "0123456789abcdef";
  std::string o = "\"";
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
  return o;
}

#define CSV_TRY(x, what)                                               \
  do {                                                                 \
    hipError_t e_ = (x);                                               \
    if (e_ != hipSuccess) {                                            \
      err = std::string(what) + ": " + hipGetErrorString(e_);          \
      return false;                                                    \
    }                                                                  \
  } while (0)

// Device buffers of one transcoding, freed together.
struct CsvBufs {
  std::vector<void *> p;
  ~CsvBufs() {
    for (void *x : p) (void)hipFree(x);
  }
  template <class T>
  hipError_t get(T **x, size_t bytes) {
    *x = nullptr;
    hipError_t e = hipMalloc((void **)x, std::max<size_t>(bytes, 16));
    if (e == hipSuccess) p.push_back(*x);
    return e;
  }
};

template <class T>
bool excl_scan(T *in_out, uint64_t n, hipStream_t st, CsvBufs &B, std::string &err) {
  size_t tb = 0;
  CSV_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in_out, in_out, (int)n, st), "scan (size)");
  void *tmp = nullptr;
  CSV_TRY(B.get(&tmp, tb), "hipMalloc(scan)");
  CSV_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tb, in_out, in_out, (int)n, st), "scan");
  return true;
}

// Pinned staging for the copies in and out: two 32 MiB buffers, taken on
// first use and kept for the life of the process. One transcoding at a time
// holds them (the mutex).
constexpr uint64_t STAGE = 32ull << 20;

struct Staging {
  std::mutex mu;
  uint8_t *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
};

Staging &staging() {
  static Staging s;
  return s;
}

bool staging_ready(Staging &S, std::string &err) {
  for (int b = 0; b < 2; b++) {
    if (!S.buf[b]) CSV_TRY(hipHostMalloc((void **)&S.buf[b], STAGE, hipHostMallocDefault), "hipHostMalloc(stage)");
    if (!S.ev[b]) CSV_TRY(hipEventCreateWithFlags(&S.ev[b], hipEventDisableTiming), "hipEventCreate(stage)");
  }
  return true;
}

// memcpy split over up to 8 threads (the caller's included), 4 KiB-aligned
// parts, so that first-touch page faults of a fresh destination run in
// parallel too.
void par_copy(uint8_t *dst, const uint8_t *src, uint64_t n) {
  const uint64_t want = std::min<uint64_t>(8, std::max<uint64_t>(1, n >> 22));
  const unsigned T = (unsigned)std::min<uint64_t>(want, std::max(1u, std::thread::hardware_concurrency()));
  if (T <= 1) {
    memcpy(dst, src, n);
    return;
  }
  const uint64_t per = ((n + T - 1) / T + 4095) & ~4095ull;
  std::vector<std::thread> th;
  for (unsigned t = 1; t < T; t++) {
    const uint64_t a = std::min(n, per * t), b = std::min(n, per * (t + 1));
    if (a >= b) continue;
    try {
      th.emplace_back([=] { memcpy(dst + a, src + a, b - a); });
    } catch (const std::system_error &) {  // no thread to be had: this part here
      memcpy(dst + a, src + a, b - a);
    }
  }
  memcpy(dst, src, std::min(n, per));
  for (std::thread &x : th) x.join();
}

// host h[0..n) -> device d: each 32 MiB stage copied into one pinned buffer
// while the DMA reads the other
bool h2d_staged(Staging &S, uint8_t *d, const uint8_t *h, uint64_t n, hipStream_t st, std::string &err) {
  for (uint64_t off = 0, k = 0; off < n; off += STAGE, k++) {
    const int b = (int)(k & 1);
    const uint64_t len = std::min(STAGE, n - off);
    CSV_TRY(hipEventSynchronize(S.ev[b]), "sync(stage)");  // the DMA that last read buf[b] is done
    par_copy(S.buf[b], h + off, len);
    CSV_TRY(hipMemcpyAsync(d + off, S.buf[b], len, hipMemcpyHostToDevice, st), "H2D(csv)");
    CSV_TRY(hipEventRecord(S.ev[b], st), "hipEventRecord(stage)");
  }
  return true;
}

// device d[0..n) -> host h: stage k+1's DMA runs while stage k is copied out
bool d2h_staged(Staging &S, uint8_t *h, const uint8_t *d, uint64_t n, hipStream_t st, std::string &err) {
  const uint64_t nst = (n + STAGE - 1) / STAGE;
  auto issue = [&](uint64_t k) -> hipError_t {
    const uint64_t off = k * STAGE;
    hipError_t e = hipMemcpyAsync(S.buf[k & 1], d + off, std::min(STAGE, n - off), hipMemcpyDeviceToHost, st);
    return e == hipSuccess ? hipEventRecord(S.ev[k & 1], st) : e;
  };
  if (nst) CSV_TRY(issue(0), "D2H(ndjson)");
  for (uint64_t k = 0; k < nst; k++) {
    // buf[(k + 1) & 1] was last copied out in step k - 1, which has returned
    if (k + 1 < nst) CSV_TRY(issue(k + 1), "D2H(ndjson)");
    CSV_TRY(hipEventSynchronize(S.ev[k & 1]), "sync(stage)");
    const uint64_t off = k * STAGE;
    par_copy(h + off, S.buf[k & 1], std::min(STAGE, n - off));
  }
  return true;
}

// Uninitialised host memory on 2 MiB boundaries, advised for huge pages
// (fewer first-touch faults); freed with free().
uint8_t *host_alloc(uint64_t n) {
  const uint64_t al = 2ull << 20, sz = std::max<uint64_t>(al, (n + al - 1) & ~(al - 1));
  void *p = nullptr;
  if (posix_memalign(&p, al, sz)) return nullptr;
  (void)madvise(p, sz, MADV_HUGEPAGE);
  return (uint8_t *)p;
}

}  // namespace

namespace wjqd {

bool csv_transcode(const uint8_t *data, uint64_t n, HostBytes &out, uint8_t **dev_out, uint64_t *records,
                   float *kernel_ms, std::string &err) {
  free(out.p);
  out.p = nullptr;
  out.n = 0;
  *dev_out = nullptr;
  *records = 0;
  if (kernel_ms) *kernel_ms = 0;
  if (n >= 3 && data[0] == 0xEF && data[1] == 0xBB && data[2] == 0xBF) {  // a byte order mark
    data += 3;
    n -= 3;
  }
  const uint8_t *nl = (const uint8_t *)memchr(data, '\n', n);
  const uint64_t hend = nl ? (uint64_t)(nl - data) : n;
  uint64_t hlen = hend;
  if (hlen && data[hlen - 1] == '\r') hlen--;
  std::vector<std::string> cols;
  if (n == 0 || !header_fields(data, hlen, cols)) {
    err = "CSV input: the header line is missing or badly quoted";
    return false;
  }
  // the key blob
  std::string keys;
  std::vector<uint32_t> koff, klen;
  for (const std::string &c : cols) {
    const std::string k = json_key(c);
    koff.push_back((uint32_t)keys.size());
    klen.push_back((uint32_t)k.size());
    keys += k;
  }
  // the body, ending in a newline, zero-padded to whole tiles
  const uint64_t b0 = nl ? hend + 1 : n;
  const uint64_t body = n - b0;
  const bool add_nl = body && data[n - 1] != '\n';
  const uint64_t blen = body + (add_nl ? 1 : 0);
  const uint64_t ntiles = (blen + CSV_TILE - 1) / CSV_TILE;
  if (ntiles > 0xFFFFFFFFull || blen > (1ull << 40)) {
    err = "CSV input: too large";
    return false;
  }
  if (blen == 0) {  // header only: one blank line
    out.p = host_alloc(1);
    uint8_t *o = nullptr;
    if (!out.p) {
      err = "CSV input: out of host memory";
      return false;
    }
    out.p[0] = '\n';
    out.n = 1;
    CSV_TRY(hipMalloc((void **)&o, 1 + CSV_DEV_PAD), "hipMalloc(ndjson)");
    *dev_out = o;  // the caller's from here on
    CSV_TRY(hipMemcpy(o, out.p, 1, hipMemcpyHostToDevice), "H2D(ndjson)");
    return true;
  }
  Staging &S = staging();
  std::lock_guard<std::mutex> hold(S.mu);
  if (!staging_ready(S, err)) return false;
  CsvBufs B;
  hipStream_t st = nullptr;
  CSV_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);  // no staged copy outlives the call
      (void)hipStreamDestroy(s);
    }
  } sg{st};
  CSV_TRY(hipEventRecord(S.ev[0], st), "hipEventRecord(stage)");
  CSV_TRY(hipEventRecord(S.ev[1], st), "hipEventRecord(stage)");
  uint8_t *d = nullptr;
  CSV_TRY(B.get(&d, ntiles * CSV_TILE + 64), "hipMalloc(csv)");
  CSV_TRY(hipMemsetAsync(d + blen, 0, ntiles * CSV_TILE + 64 - blen, st), "memset(csv)");
  if (!h2d_staged(S, d, data + b0, body, st, err)) return false;
  if (add_nl) CSV_TRY(hipMemsetAsync(d + body, '\n', 1, st), "memset(csv)");
  uint8_t *dk = nullptr;
  uint32_t *dko = nullptr, *dkl = nullptr;
  CSV_TRY(B.get(&dk, keys.size()), "hipMalloc(keys)");
  CSV_TRY(B.get(&dko, koff.size() * 4), "hipMalloc(keys)");
  CSV_TRY(B.get(&dkl, klen.size() * 4), "hipMalloc(keys)");
  CSV_TRY(hipMemcpyAsync(dk, keys.data(), keys.size(), hipMemcpyHostToDevice, st), "H2D(keys)");
  CSV_TRY(hipMemcpyAsync(dko, koff.data(), koff.size() * 4, hipMemcpyHostToDevice, st), "H2D(keys)");
  CSV_TRY(hipMemcpyAsync(dkl, klen.data(), klen.size() * 4, hipMemcpyHostToDevice, st), "H2D(keys)");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CSV_TRY(hipEventCreate(&e0), "hipEventCreate");
  CSV_TRY(hipEventCreate(&e1), "hipEventCreate");
  struct EventGuard {
    hipEvent_t a, b;
    ~EventGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } eg{e0, e1};
  CSV_TRY(hipEventRecord(e0, st), "hipEventRecord");
  // newline positions
  uint64_t *cnt = nullptr;
  CSV_TRY(B.get(&cnt, ntiles * 8), "hipMalloc(counts)");
  const uint32_t g1 = (uint32_t)std::min<uint64_t>(4096, (ntiles + 3) / 4);
  hipLaunchKernelGGL(k_csv_nl_count, dim3(g1), dim3(256), 0, st, d, ntiles, cnt);
  CSV_TRY(hipGetLastError(), "k_csv_nl_count");
  uint64_t last_cnt = 0, nrec = 0;
  CSV_TRY(hipMemcpyAsync(&last_cnt, cnt + ntiles - 1, 8, hipMemcpyDeviceToHost, st), "D2H(count)");
  if (!excl_scan(cnt, ntiles, st, B, err)) return false;
  CSV_TRY(hipMemcpyAsync(&nrec, cnt + ntiles - 1, 8, hipMemcpyDeviceToHost, st), "D2H(count)");
  CSV_TRY(hipStreamSynchronize(st), "sync");
  nrec += last_cnt;
  if (nrec >= 0x7FFFFFFFull) {  // the scans count items in an int
    err = "CSV input: more than 2^31 - 2 records";
    return false;
  }
  uint64_t *pos = nullptr, *len = nullptr;
  uint32_t *too_long = nullptr;
  CSV_TRY(B.get(&too_long, 4), "hipMalloc(flag)");
  CSV_TRY(hipMemsetAsync(too_long, 0, 4, st), "memset(flag)");
  CSV_TRY(B.get(&pos, nrec * 8), "hipMalloc(positions)");
  CSV_TRY(B.get(&len, (nrec + 1) * 8), "hipMalloc(lengths)");
  hipLaunchKernelGGL(k_csv_nl_pos, dim3(g1), dim3(256), 0, st, d, ntiles, cnt, pos);
  CSV_TRY(hipGetLastError(), "k_csv_nl_pos");
  const CsvHdr h{(uint32_t)cols.size(), dk, dko, dkl};
  const uint32_t g2 = (uint32_t)std::min<uint64_t>(8192, (nrec + 63) / 64);
  hipLaunchKernelGGL(k_csv_pass<false>, dim3(g2), dim3(64), 0, st, d, pos, nrec, h, len, too_long, nullptr, nullptr,
                     0ull);
  CSV_TRY(hipGetLastError(), "k_csv_pass<len>");
  CSV_TRY(hipMemsetAsync(len + nrec, 0, 8, st), "memset(lengths)");
  if (!excl_scan(len, nrec + 1, st, B, err)) return false;
  uint64_t total = 0;
  uint32_t long_rec = 0;
  CSV_TRY(hipMemcpyAsync(&total, len + nrec, 8, hipMemcpyDeviceToHost, st), "D2H(total)");
  CSV_TRY(hipMemcpyAsync(&long_rec, too_long, 4, hipMemcpyDeviceToHost, st), "D2H(flag)");
  CSV_TRY(hipStreamSynchronize(st), "sync");
  if (long_rec) {
    err = "CSV input: a record of 4 GiB or more";
    return false;
  }
  // o[0] is the header's blank line, the records follow; the device copy
  // keeps the pad the kernels may read past a span
  uint8_t *o = nullptr;
  CSV_TRY(hipMalloc((void **)&o, 1 + total + CSV_DEV_PAD), "hipMalloc(ndjson)");
  *dev_out = o;  // the caller's from here on, also on an error below
  CSV_TRY(hipMemsetAsync(o, '\n', 1, st), "memset(ndjson)");
  hipLaunchKernelGGL(k_csv_pass<true>, dim3(g2), dim3(64), 0, st, d, pos, nrec, h, nullptr, nullptr, len, o, 1ull);
  CSV_TRY(hipGetLastError(), "k_csv_pass<emit>");
  CSV_TRY(hipEventRecord(e1, st), "hipEventRecord");
  out.p = host_alloc(1 + total);
  if (!out.p) {
    err = "CSV input: out of host memory";
    return false;
  }
  out.n = 1 + total;
  if (!d2h_staged(S, out.p, o, 1 + total, st, err)) return false;
  CSV_TRY(hipStreamSynchronize(st), "sync");
  if (kernel_ms) (void)hipEventElapsedTime(kernel_ms, e0, e1);
  *records = nrec;
  return true;
}

void csv_free(uint8_t *dev) {
  if (dev) (void)hipFree(dev);
}

}  // namespace wjqd
