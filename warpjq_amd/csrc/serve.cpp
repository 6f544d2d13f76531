// The served form: datasets kept resident in HBM under a byte budget, and a
// request queue drained by worker threads, each running one request at a
// time through its own device contexts (its own HIP streams), so that
// requests overlap on the GPU and a finished request's worker takes the next
// admissible one at once (continuous batching). The reference has no serving
// loop; each request's result is defined as the reference's result for the
// same query over its datasets read as one stream (Input::Chained,
// crates/warpjq-core/src/chunk.rs:46-57, :198-203).
//
// Admission: a request runs only when all of its datasets are resident.
// Loading one reserves its bytes against the budget, evicting the least
// recently used datasets that no running request holds; a request that
// cannot be placed waits, and requests behind it whose datasets are already
// resident go ahead. A request whose datasets alone exceed the budget fails.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/warpjq_amd.h"
#include "driver.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint64_t DEV_PAD = 256;  // readable bytes the kernels may touch past a span
constexpr size_t ENGINES_PER_WORKER = 8;
constexpr uint64_t MAX_LINE_BYTES = 64ull << 20;  // the reference's default limit (exec/mod.rs:43-63)

struct Dataset {
  const uint8_t *host = nullptr;
  uint64_t n = 0;
  uint64_t longest = 0;  // longest line, found once when the dataset is added
  uint8_t *dev = nullptr;
  bool loading = false;
  uint32_t pins = 0;
  uint64_t last_use = 0;
  // the bytes, the kernels' pad, and the newline index (one byte per 64-byte
  // block) that row requests count each claim's lines from
  uint64_t idx_at() const { return (n + DEV_PAD + 255) & ~255ull; }
  uint64_t charge() const { return idx_at() + wjqd::nlindex_bytes(n); }
};

struct Request {
  int64_t id = 0;
  std::string query;
  std::vector<uint32_t> ds;
  int fmt = 0, oi = 1;
  Clock::time_point arrive, done;
  bool finished = false;
  int rc = 0;
  uint8_t *out = nullptr;
  uint64_t out_len = 0;
  warpjq_amd_stats st{};
  std::string err;
};

}  // namespace

struct warpjq_amd_server {
  uint64_t budget = 0, reserved = 0, chunk_bytes = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Dataset> ds;  // elements keep their addresses; the deque itself is read under `mu`
  std::vector<uint8_t *> to_free;  // evicted device buffers, freed by a worker outside `mu`
  std::deque<Request *> waiting;
  std::map<int64_t, std::unique_ptr<Request>> reqs;
  int64_t next_id = 1;
  uint64_t tick = 0, loads = 0, evictions = 0, completed = 0, opened = 0, reprogrammed = 0;
  std::vector<double> lat_ms;
  bool stop = false;
  std::vector<std::thread> workers;
};

namespace {

// Bytes that loading `r`'s missing datasets would add to the reservation.
uint64_t missing_bytes(const warpjq_amd_server &s, const Request &r) {
  uint64_t need = 0;
  for (size_t k = 0; k < r.ds.size(); k++) {
    const Dataset &d = s.ds[r.ds[k]];
    bool dup = std::find(r.ds.begin(), r.ds.begin() + k, r.ds[k]) != r.ds.begin() + k;
    if (!d.dev && !d.loading && !dup) need += d.charge();
  }
  return need;
}

// Frees unpinned datasets, least recently used first, until `need` more
// bytes fit the budget. False if they cannot.
bool make_room(warpjq_amd_server &s, uint64_t need, const Request &r) {
  while (s.reserved + need > s.budget) {
    Dataset *victim = nullptr;
    for (size_t i = 0; i < s.ds.size(); i++) {
      Dataset &d = s.ds[i];
      if (!d.dev || d.pins || d.loading) continue;
      if (std::find(r.ds.begin(), r.ds.end(), (uint32_t)i) != r.ds.end()) continue;
      if (!victim || d.last_use < victim->last_use) victim = &d;
    }
    if (!victim) return false;
    // freed after the lock is released: hipFree waits for the device
    wjqd::nlindex_unregister(victim->dev);
    s.to_free.push_back(victim->dev);
    victim->dev = nullptr;
    s.reserved -= victim->charge();
    s.evictions++;
  }
  return true;
}

// Picks the next request to run (called with the lock held): the oldest
// that fits, after the head has had its chance to evict.
Request *admit(warpjq_amd_server &s) {
  for (auto it = s.waiting.begin(); it != s.waiting.end(); ++it) {
    Request *r = *it;
    const uint64_t need = missing_bytes(s, *r);
    bool ok = need == 0 || (it == s.waiting.begin() && make_room(s, need, *r)) ||
              (need && s.reserved + need <= s.budget);
    // a dataset another worker is still loading: wait for it
    for (uint32_t i : r->ds)
      if (s.ds[i].loading) ok = false;
    if (!ok) continue;
    s.waiting.erase(it);
    for (uint32_t i : r->ds) {
      Dataset &d = s.ds[i];
      d.pins++;
      d.last_use = ++s.tick;
      if (!d.dev && !d.loading) {
        d.loading = true;
        s.reserved += d.charge();
      }
    }
    return r;
  }
  return nullptr;
}

void finish(warpjq_amd_server &s, Request *r) {
  std::lock_guard<std::mutex> g(s.mu);
  for (uint32_t i : r->ds) s.ds[i].pins--;
  r->done = Clock::now();
  r->finished = true;
  s.completed++;
  s.lat_ms.push_back(std::chrono::duration<double, std::milli>(r->done - r->arrive).count());
  s.cv_done.notify_all();
  s.cv_work.notify_all();  // unpinned datasets may let a waiting request in
}

struct EngineSlot {
  std::string key;
  wjqd::Engine eng;
  uint64_t last = 0;
};

void worker(warpjq_amd_server *s) {
  std::list<EngineSlot> engines;
  uint64_t uses = 0;
  for (;;) {
    Request *r = nullptr;
    std::vector<uint8_t *> evicted;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop) break;
        r = admit(*s);
        if (r) break;
        s->cv_work.wait(lk);
      }
      evicted.swap(s->to_free);
      if (!r) {
        lk.unlock();
        for (uint8_t *p : evicted) (void)hipFree(p);
        break;
      }
    }
    for (uint8_t *p : evicted) (void)hipFree(p);
    // load what this request needs and another request has not; the
    // entries' addresses are taken under the lock (add_dataset may grow the
    // deque meanwhile) and their host bytes and sizes never change
    std::vector<Dataset *> mine;
    {
      std::lock_guard<std::mutex> g(s->mu);
      for (uint32_t i : r->ds) {
        Dataset *d = &s->ds[i];
        if (d->loading && !d->dev && std::find(mine.begin(), mine.end(), d) == mine.end()) mine.push_back(d);
      }
    }
    std::string lerr;
    for (Dataset *dptr : mine) {
      Dataset &d = *dptr;
      uint8_t *p = nullptr;
      hipError_t e = hipMalloc((void **)&p, d.charge());
      if (e == hipSuccess && d.n) e = hipMemcpy(p, d.host, d.n, hipMemcpyHostToDevice);
      std::string ierr;
      const bool indexed = e == hipSuccess && wjqd::nlindex_build(p, d.n, p + d.idx_at(), ierr);
      if (indexed) wjqd::nlindex_register(p, d.n, p + d.idx_at());
      std::lock_guard<std::mutex> g(s->mu);
      d.loading = false;
      if (!indexed) {
        if (p) (void)hipFree(p);
        s->reserved -= d.charge();
        lerr = e != hipSuccess ? std::string("loading a dataset: ") + hipGetErrorString(e) : ierr;
      } else {
        d.dev = p;
        s->loads++;
      }
      s->cv_work.notify_all();
    }
    if (!lerr.empty()) {
      r->rc = 3;
      r->err = lerr;
      finish(*s, r);
      continue;
    }
    // the engine for this query text and format (compiled once per worker)
    const std::string key = std::to_string(r->fmt) + ":" + r->query;
    auto it = std::find_if(engines.begin(), engines.end(), [&](const EngineSlot &x) { return x.key == key; });
    char err[1024] = {0};
    if (it == engines.end()) {
      // a miss: point the least recently used engine whose buffers fit this
      // query at it (a table upload) rather than opening a context
      std::vector<std::list<EngineSlot>::iterator> order;
      for (auto e = engines.begin(); e != engines.end(); ++e) order.push_back(e);
      std::sort(order.begin(), order.end(), [](auto x, auto y) { return x->last < y->last; });
      int rc = 2;
      for (auto e : order) {
        rc = wjqd::engine_reprogram(e->eng, r->query.c_str(), r->fmt, err, sizeof err);
        if (rc == 0) {
          e->key = key;
          it = e;
          std::lock_guard<std::mutex> g(s->mu);
          s->reprogrammed++;
          break;
        }
        if (rc != 2) break;
      }
      if (rc != 0 && rc != 2) {
        r->rc = rc;
        r->err = err;
        finish(*s, r);
        continue;
      }
    }
    if (it == engines.end()) {
      if (engines.size() >= ENGINES_PER_WORKER) {
        auto old = std::min_element(engines.begin(), engines.end(),
                                    [](const EngineSlot &a, const EngineSlot &b) { return a.last < b.last; });
        wjqd::engine_close(old->eng);
        engines.erase(old);
      }
      engines.emplace_back();
      engines.back().key = key;
      engines.back().eng.cap_limit = s->chunk_bytes;
      int rc = wjqd::engine_open(engines.back().eng, r->query.c_str(), r->fmt, s->chunk_bytes, err, sizeof err);
      if (rc) {
        engines.pop_back();
        r->rc = rc;
        r->err = err;
        finish(*s, r);
        continue;
      }
      it = std::prev(engines.end());
      std::lock_guard<std::mutex> g(s->mu);
      s->opened++;
    }
    it->last = ++uses;
    std::vector<const uint8_t *> hp, dp;
    std::vector<uint64_t> lens;
    uint64_t longest = 0;
    {
      std::lock_guard<std::mutex> g(s->mu);
      for (uint32_t i : r->ds) {
        hp.push_back(s->ds[i].host);
        dp.push_back(s->ds[i].dev);
        lens.push_back(s->ds[i].n);
        longest = std::max(longest, s->ds[i].longest);
      }
    }
    if (longest > MAX_LINE_BYTES) {  // chunk.rs:364-367, checked once per dataset
      r->rc = 3;
      r->err = "a line exceeds the " + std::to_string(MAX_LINE_BYTES) + " byte limit";
      finish(*s, r);
      continue;
    }
    // no per-request scan of the bytes: line lengths were checked above
    wjqd::Batcher src(hp.data(), lens.data(), (uint32_t)hp.size(), ~0ull, 65536);
    src.dfiles = dp.data();
    r->rc = wjqd::engine_run(it->eng, src, r->fmt, r->oi, &r->out, &r->out_len, &r->st, err, sizeof err);
    if (r->rc) r->err = err;
    finish(*s, r);
  }
  for (auto &e : engines) wjqd::engine_close(e.eng);
}

}  // namespace

extern "C" {

warpjq_amd_server *warpjq_amd_server_create(uint64_t hbm_budget, uint32_t workers, uint64_t chunk_bytes, char *err,
                                            size_t err_len) {
  if (workers == 0 || workers > 16 || hbm_budget == 0 || chunk_bytes > (1ull << 30)) {
    if (err && err_len) snprintf(err, err_len, "server_create: 1-16 workers, a non-zero budget, chunks up to 1 GiB");
    return nullptr;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    if (err && err_len) snprintf(err, err_len, "server_create: no HIP device");
    return nullptr;
  }
  auto *s = new warpjq_amd_server();
  s->budget = hbm_budget;
  s->chunk_bytes = chunk_bytes ? chunk_bytes : (64ull << 20);
  for (uint32_t i = 0; i < workers; i++) s->workers.emplace_back(worker, s);
  return s;
}

void warpjq_amd_server_destroy(warpjq_amd_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  for (auto &t : s->workers) t.join();
  for (uint8_t *p : s->to_free) (void)hipFree(p);
  for (auto &kv : s->reqs) free(kv.second->out);
  for (auto &d : s->ds)
    if (d.dev) {
      wjqd::nlindex_unregister(d.dev);
      (void)hipFree(d.dev);
    }
  delete s;
}

int64_t warpjq_amd_server_add_dataset(warpjq_amd_server *s, const uint8_t *bytes, uint64_t n) {
  if (!s || (!bytes && n)) return -1;
  std::lock_guard<std::mutex> g(s->mu);
  Dataset d;
  d.host = bytes;
  d.n = n;
  for (uint64_t a = 0; a < n;) {  // the dataset's index: its longest line
    const uint8_t *nl = (const uint8_t *)memchr(bytes + a, '\n', n - a);
    const uint64_t e = nl ? (uint64_t)(nl - bytes) : n;
    d.longest = std::max(d.longest, e - a);
    a = e + 1;
  }
  s->ds.push_back(d);
  return (int64_t)s->ds.size() - 1;
}

int64_t warpjq_amd_server_submit(warpjq_amd_server *s, const char *query, const uint32_t *datasets, uint32_t n_datasets,
                                 int format, int on_invalid) {
  if (!s || !query || (!datasets && n_datasets)) return -1;
  auto r = std::make_unique<Request>();
  r->query = query;
  r->fmt = format;
  r->oi = on_invalid;
  r->arrive = Clock::now();
  std::lock_guard<std::mutex> g(s->mu);
  uint64_t total = 0;
  for (uint32_t k = 0; k < n_datasets; k++) {
    if (datasets[k] >= s->ds.size()) return -1;
    r->ds.push_back(datasets[k]);
    if (std::find(datasets, datasets + k, datasets[k]) == datasets + k) total += s->ds[datasets[k]].charge();
  }
  r->id = s->next_id++;
  Request *p = r.get();
  s->reqs[p->id] = std::move(r);
  if (total > s->budget) {  // can never be admitted
    p->rc = 3;
    p->err = "the request's datasets exceed the HBM budget";
    p->finished = true;
    p->done = p->arrive;
    s->cv_done.notify_all();
    return p->id;
  }
  s->waiting.push_back(p);
  s->cv_work.notify_one();
  return p->id;
}

int warpjq_amd_server_wait(warpjq_amd_server *s, int64_t id, uint8_t **out, uint64_t *out_len,
                           warpjq_amd_stats *stats, double *latency_ms, char *err, size_t err_len) {
  if (!s || !out || !out_len) return 3;
  std::unique_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) {
      if (err && err_len) snprintf(err, err_len, "server_wait: unknown request %lld", (long long)id);
      return 3;
    }
    Request *p = it->second.get();
    s->cv_done.wait(lk, [&] { return p->finished; });
    r = std::move(it->second);
    s->reqs.erase(it);
  }
  *out = r->out;
  *out_len = r->out_len;
  if (stats) *stats = r->st;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->done - r->arrive).count();
  if (err && err_len) snprintf(err, err_len, "%s", r->err.c_str());
  return r->rc;
}

void warpjq_amd_server_get_stats(warpjq_amd_server *s, warpjq_amd_server_stats *o) {
  if (!s || !o) return;
  std::lock_guard<std::mutex> g(s->mu);
  memset(o, 0, sizeof *o);
  o->budget = s->budget;
  o->reserved = s->reserved;
  o->loads = s->loads;
  o->evictions = s->evictions;
  o->completed = s->completed;
  o->engines_opened = s->opened;
  o->engines_reprogrammed = s->reprogrammed;
  for (const Dataset &d : s->ds)
    if (d.dev) o->resident++;
  if (!s->lat_ms.empty()) {
    std::vector<double> v = s->lat_ms;
    std::sort(v.begin(), v.end());
    auto q = [&](double f) { return v[std::min(v.size() - 1, (size_t)(f * (v.size() - 1) + 0.5))]; };
    o->p50_ms = q(0.50);
    o->p99_ms = q(0.99);
    o->max_ms = v.back();
  }
}

}  // extern "C"
