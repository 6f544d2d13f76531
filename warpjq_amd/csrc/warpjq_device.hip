// warpjq's device ABI (include/warpjq_kernels.h) implemented for MI355X
// (gfx950): the non-template kernels, and the host side of the ABI. The
// per-line kernels live in kernels.h and are instantiated per slot-table
// width in kern_<N>.hip.
#include <atomic>
#include <map>
#include <mutex>

#include "driver.h"
#include "kernels.h"

namespace wjqd {
// Sets each span's newline index and dataset offset when it lies inside a
// registered resident dataset (defined with the registry below).
void nlindex_attach(Span *sp, uint32_t n);
}  // namespace wjqd

// ---------------------------------------------------------------------------
// k_reset
// ---------------------------------------------------------------------------

__global__ void k_reset(DevCtr *ctr, unsigned long long *tile_state, uint32_t ntiles) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    ctr->n_lines = ctr->n_blank = ctr->n_invalid = ctr->n_type_error = ctr->n_fallback = 0;
    ctr->n_sel = ctr->out_total = 0;
    ctr->first_invalid = 0xFFFFFFFFu;
    ctr->first_bad = ~0ull;
    ctr->overflow = ctr->saw_non_numeric = ctr->group_overflow = 0;
    ctr->fb_count = ctr->n_groups = 0;

    ctr->blob_used = 0;
    ctr->long_count = ctr->long_next = 0;
    ctr->first_te_bad = ~0ull;
    ctr->first_te = 0xFFFFFFFFu;
    ctr->deep_count = 0;
  }
  if (i < NREG) ctr->rticket[i * RSTRIDE] = 0;
  for (uint32_t t = i; t < ntiles; t += gridDim.x * blockDim.x) tile_state[t] = 0;
}

// Clears the whole group table once, at context creation.
__global__ void k_group_init(GroupTab g) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < g.cap; i += gridDim.x * blockDim.x) {
    g.keys[i] = EMPTY_KEY;
    g.hash[i] = 0;
    g.count[i] = g.numeric[i] = g.isum_lo[i] = g.isum_hi[i] = 0;
    g.fsum[i] = g.fcomp[i] = 0.0;
    g.mnb[i] = ~0ull;
    g.mxb[i] = 0;
  }
}

// ---------------------------------------------------------------------------
// aggregates: finalisation and compaction
// ---------------------------------------------------------------------------

struct AggOut {
  warpjq_agg_partial agg;
  warpjq_amd_agg_ext ext;
};

// 256 threads: each folds a strided subset in order, then a fixed tree, so
// the result is deterministic for a given grid.
__global__ __launch_bounds__(256) void k_agg_final(const AggPart *parts, uint32_t nparts, const DevCtr *ctr,
                                                    AggOut *out) {
  __shared__ unsigned long long s_cnt[256], s_num[256], s_lo[256], s_hi[256];
  __shared__ double s_s[256], s_c[256], s_mn[256], s_mx[256];
  const uint32_t t = threadIdx.x;
  unsigned long long cnt = 0, num = 0, lo = 0, hi = 0;
  double s = 0.0, c = 0.0, mn = INFINITY, mx = -INFINITY;
  for (uint32_t i = t; i < nparts; i += 256) {
    const AggPart &p = parts[i];
    cnt += p.count;
    num += p.numeric;
    unsigned long long nlo = lo + p.isum_lo;
    hi += p.isum_hi + (nlo < lo ? 1 : 0);
    lo = nlo;
    double x = s + p.fsum;
    double bp = x - s;
    c += (s - (x - bp)) + (p.fsum - bp) + p.fcomp;
    s = x;
    if (p.mn < mn) mn = p.mn;
    if (p.mx > mx) mx = p.mx;
  }
  s_cnt[t] = cnt;
  s_num[t] = num;
  s_lo[t] = lo;
  s_hi[t] = hi;
  s_s[t] = s;
  s_c[t] = c;
  s_mn[t] = mn;
  s_mx[t] = mx;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (t < w) {
      uint32_t o = t + w;
      s_cnt[t] += s_cnt[o];
      s_num[t] += s_num[o];
      unsigned long long nlo = s_lo[t] + s_lo[o];
      s_hi[t] += s_hi[o] + (nlo < s_lo[t] ? 1 : 0);
      s_lo[t] = nlo;
      double x = s_s[t] + s_s[o];
      double bp = x - s_s[t];
      s_c[t] += (s_s[t] - (x - bp)) + (s_s[o] - bp) + s_c[o];
      s_s[t] = x;
      if (s_mn[o] < s_mn[t]) s_mn[t] = s_mn[o];
      if (s_mx[o] > s_mx[t]) s_mx[t] = s_mx[o];
    }
    __syncthreads();
  }
  if (t == 0) {
    out->ext.isum_lo = s_lo[0];
    out->ext.isum_hi = s_hi[0];
    out->ext.fsum = s_s[0];
    out->ext.fcomp = s_c[0];
    out->agg.count = s_cnt[0];
    out->agg.numeric = s_num[0];
    out->agg.sum = 0.0;  // filled on the host from the exact parts
    out->agg.min = s_mn[0];
    out->agg.max = s_mx[0];
    out->agg.saw_non_numeric = ctr->saw_non_numeric;
    out->agg._pad = 0;
  }
}

// Key offsets: relative to the chunk's first byte (the ABI's meaning), or,
// for a batched submit, keys are copied into `blob` and key_off points there
// (there is no single buffer the caller could index).
__global__ void k_group_compact(GroupTab g, DevCtr *ctr, warpjq_group *out, warpjq_amd_agg_ext *ext, const Span *spans,
                                uint32_t n_spans, uint8_t *blob, uint32_t blob_cap) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < g.cap; i += gridDim.x * blockDim.x) {
    unsigned long long k = g.keys[i];
    if (k == EMPTY_KEY) continue;
    GKey gk = unpack_key(k);
    uint32_t o = atomicAdd(&ctr->n_groups, 1u);
    warpjq_group r;
    if (gk.kind == JK_NULL) {
      r.key_off = 0;
    } else if (!blob) {
      r.key_off = gk.off - spans[0].start;
    } else {
      const uint32_t at = atomicAdd(&ctr->blob_used, gk.len);
      if ((uint64_t)at + gk.len <= blob_cap) {
        DevReader rd = span_reader(spans, n_spans, gk.off);
        for (uint32_t q = 0; q < gk.len; q++) blob[at + q] = (uint8_t)rd.byte(gk.off + q);
      } else {
        atomicOr(&ctr->group_overflow, 1u);
      }
      r.key_off = at;
    }
    r.key_len = gk.len;
    r.key_kind = gk.kind;
    r._pad = 0;
    r.count = g.count[i];
    r.numeric = g.numeric[i];
    r.sum = 0.0;
    r.min = g.numeric[i] ? from_ordered(g.mnb[i]) : INFINITY;
    r.max = g.numeric[i] ? from_ordered(g.mxb[i]) : -INFINITY;
    out[o] = r;
    warpjq_amd_agg_ext e;
    e.isum_lo = g.isum_lo[i];
    e.isum_hi = g.isum_hi[i];
    e.fsum = g.fsum[i];
    e.fcomp = g.fcomp[i];
    ext[o] = e;
    g.keys[i] = EMPTY_KEY;
    g.hash[i] = 0;
    g.count[i] = g.numeric[i] = g.isum_lo[i] = g.isum_hi[i] = 0;
    g.fsum[i] = g.fcomp[i] = 0.0;
    g.mnb[i] = ~0ull;
    g.mxb[i] = 0;
  }
}

// ---------------------------------------------------------------------------
// row outputs: reduce-then-scan over per-line row lengths
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t nlines_capped(const DevCtr *ctr, uint64_t max_lines) {
  uint64_t n = ctr->n_lines;
  return n < max_lines ? n : max_lines;
}

// Block b sums lines [b*1024, (b+1)*1024): selected count and bytes.
__global__ __launch_bounds__(256) void k_row_reduce(const uint32_t *row_len, const DevCtr *ctr, uint64_t max_lines,
                                                     uint64_t *blk_cnt, uint64_t *blk_bytes) {
  uint64_t n = nlines_capped(ctr, max_lines);
  uint64_t base = (uint64_t)blockIdx.x * ROWSCAN_BLOCK;
  if (base >= n) return;
  uint64_t c = 0, b = 0;
  for (uint32_t k = threadIdx.x; k < ROWSCAN_BLOCK; k += 256) {
    uint64_t i = base + k;
    if (i < n) {
      uint32_t r = row_len[i];
      c += r ? 1 : 0;
      b += r;
    }
  }
  __shared__ uint64_t sc[4], sbt[4];
  c = wave_sum(c);
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0) {
    sc[threadIdx.x >> 6] = c;
    sbt[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    blk_cnt[blockIdx.x] = sc[0] + sc[1] + sc[2] + sc[3];
    blk_bytes[blockIdx.x] = sbt[0] + sbt[1] + sbt[2] + sbt[3];
  }
}

// One block: exclusive scan of the block sums; totals into the counters.
__global__ __launch_bounds__(1024) void k_row_blockscan(uint64_t *blk_cnt, uint64_t *blk_bytes, DevCtr *ctr,
                                                        uint64_t max_lines, uint64_t out_cap) {
  uint64_t n = nlines_capped(ctr, max_lines);
  uint32_t nb = (uint32_t)((n + ROWSCAN_BLOCK - 1) / ROWSCAN_BLOCK);
  __shared__ uint64_t wc[16], wb[16];
  __shared__ uint64_t run_c, run_b;
  if (threadIdx.x == 0) {
    run_c = 0;
    run_b = 0;
  }
  __syncthreads();
  for (uint32_t base = 0; base < nb; base += 1024) {
    uint32_t i = base + threadIdx.x;
    uint64_t c = i < nb ? blk_cnt[i] : 0, b = i < nb ? blk_bytes[i] : 0;
    // inclusive wave scan
    uint64_t xc = c, xb = b;
    uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint64_t yc = __shfl_up(xc, o, 64), yb = __shfl_up(xb, o, 64);
      if (lane >= (uint32_t)o) {
        xc += yc;
        xb += yb;
      }
    }
    if (lane == 63) {
      wc[threadIdx.x >> 6] = xc;
      wb[threadIdx.x >> 6] = xb;
    }
    __syncthreads();
    uint64_t pc = run_c, pb = run_b;
    for (uint32_t w = 0; w < (threadIdx.x >> 6); w++) {
      pc += wc[w];
      pb += wb[w];
    }
    if (i < nb) {
      blk_cnt[i] = pc + xc - c;
      blk_bytes[i] = pb + xb - b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (uint32_t w = 0; w < 16; w++) {
        run_c += wc[w];
        run_b += wb[w];
      }
    }
    __syncthreads();
    if (threadIdx.x < 16) {
      wc[threadIdx.x] = 0;
      wb[threadIdx.x] = 0;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ctr->n_sel = run_c;
    ctr->out_total = run_b;
    if (run_b > out_cap || ctr->n_lines > max_lines) ctr->overflow = 1;
  }
}

// Block b writes the selected indices and row offsets of its lines.
__global__ __launch_bounds__(256) void k_row_scatter(const uint32_t *row_len, const DevCtr *ctr, uint64_t max_lines,
                                                      const uint64_t *blk_cnt, const uint64_t *blk_bytes,
                                                      uint32_t *sel_idx, uint64_t *row_off) {
  uint64_t n = nlines_capped(ctr, max_lines);
  uint64_t base = (uint64_t)blockIdx.x * ROWSCAN_BLOCK;
  if (base >= n) return;
  __shared__ uint64_t wc[4], wb[4];
  __shared__ uint64_t run_c, run_b;
  if (threadIdx.x == 0) {
    run_c = blk_cnt[blockIdx.x];
    run_b = blk_bytes[blockIdx.x];
  }
  __syncthreads();
  uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint32_t k0 = 0; k0 < ROWSCAN_BLOCK; k0 += 256) {
    uint64_t i = base + k0 + threadIdx.x;
    uint32_t r = i < n ? row_len[i] : 0;
    uint64_t c = r ? 1 : 0, b = r;
    uint64_t xc = c, xb = b;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint64_t yc = __shfl_up(xc, o, 64), yb = __shfl_up(xb, o, 64);
      if (lane >= (uint32_t)o) {
        xc += yc;
        xb += yb;
      }
    }
    if (lane == 63) {
      wc[w] = xc;
      wb[w] = xb;
    }
    __syncthreads();
    uint64_t pc = run_c, pb = run_b;
    for (uint32_t q = 0; q < w; q++) {
      pc += wc[q];
      pb += wb[q];
    }
    if (r) {
      uint64_t sidx = pc + xc - c;
      sel_idx[sidx] = (uint32_t)i;
      row_off[sidx] = pb + xb - b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      run_c += wc[0] + wc[1] + wc[2] + wc[3];
      run_b += wb[0] + wb[1] + wb[2] + wb[3];
    }
    __syncthreads();
  }
}

__global__ void k_row_finish(const DevCtr *ctr, uint64_t *row_off) {
  if (threadIdx.x == 0 && blockIdx.x == 0) row_off[ctr->n_sel] = ctr->out_total;
}

// ---------------------------------------------------------------------------
// aggregate queries: (claim, rank) records -> line indices in the chunk
// ---------------------------------------------------------------------------

// One block: tile_state[0..ntiles) line counts -> exclusive prefix sums.
__global__ __launch_bounds__(1024) void k_res_prefix(unsigned long long *cnt, uint32_t ntiles) {
  __shared__ unsigned long long wsum[16];
  __shared__ unsigned long long run;
  if (threadIdx.x == 0) run = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint32_t base = 0; base < ntiles; base += 1024) {
    uint32_t i = base + threadIdx.x;
    unsigned long long v = i < ntiles ? cnt[i] : 0, x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      unsigned long long y = __shfl_up(x, o, 64);
      if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    unsigned long long p = run;
    for (uint32_t q = 0; q < w; q++) p += wsum[q];
    if (i < ntiles) cnt[i] = p + x - v;
    __syncthreads();
    if (threadIdx.x == 0)
      for (uint32_t q = 0; q < 16; q++) run += wsum[q];
    __syncthreads();
  }
}

// This lane's share of the newlines at q positions [a, b) (d[q] is the byte
// at q; 16-byte blocks dealt to the wave's lanes).
__device__ __forceinline__ uint32_t nl_between(const uint8_t *d, uint32_t a, uint32_t b, uint32_t lane) {
  uint32_t m32 = 0;
  for (uint32_t k = (a >> 4) + lane; (k << 4) < b; k += 64) {
    const uint32_t p = k << 4;
    uint32_t m = nl_mask16(*(const uint4 *)(d + p));
    if (p < a) m &= 0xFFFFu << (a - p);
    if (p + 16u > b) m &= (1u << (b - p)) - 1u;
    m32 += __popc(m);
  }
  return m32;
}

// Line starts per claim for row outputs, by the rule k_main's enqueue
// applies: q in [max(t0, start), min(t0 + claim, vend)) with q == start or
// byte q - 1 == '\n'. One wave per claim, 16 bytes per lane per step. A span
// inside a resident dataset with a newline index takes whole 64-byte blocks
// from the index and reads only the partial blocks at the claim's two ends.
__global__ __launch_bounds__(256) void k_line_count(const Span *spans, uint32_t n_spans, uint32_t ntiles, uint32_t claim,
                                                     unsigned long long *cnt) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < ntiles; t += nw) {
    const uint32_t t0 = t * claim;
    const Span sp = spans[span_of(spans, n_spans, t0)];
    const uint32_t lo = max(t0, sp.start), hi = min(t0 + claim, sp.vend);
    unsigned long long n = 0;
    if (lo < hi) {
      const uint8_t *d = sp.base - sp.vbase;           // d[q]: the byte at q position q
      const uint32_t a0 = lo > sp.start ? lo - 1 : sp.start;  // newlines in [a0, a1) start a line
      const uint32_t a1 = hi - 1;
      uint32_t m32 = 0;
      if (sp.nlidx) {
        // dataset offsets of a0 and a1; whole index blocks [kA, kB) between
        const uint64_t A = (uint64_t)(a0 - sp.start) + sp.ioff, B = (uint64_t)(a1 - sp.start) + sp.ioff;
        const uint64_t kA = (A + 63u) >> 6, kB = B >> 6;
        if (kA < kB) {
          for (uint64_t k = kA + lane; k < kB; k += 64) m32 += sp.nlidx[k];
          const uint32_t qA = (uint32_t)((kA << 6) - sp.ioff) + sp.start;
          const uint32_t qB = (uint32_t)((kB << 6) - sp.ioff) + sp.start;
          m32 += nl_between(d, a0, qA, lane) + nl_between(d, qB, a1, lane);
        } else {
          m32 = nl_between(d, a0, a1, lane);
        }
      } else {
        m32 = nl_between(d, a0, a1, lane);
      }
      n = wave_sum((unsigned long long)m32) + (lo == sp.start ? 1u : 0u);
    }
    if (lane == 0) cnt[t] = n;
  }
}

// Newlines per 64-byte block of d[0..n) (the last block partial).
__global__ __launch_bounds__(256) void k_nl_index(const uint8_t *d, uint64_t n, uint8_t *idx) {
  const uint64_t nb = (n + 63) >> 6;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
      const uint64_t p = (k << 6) + 16u * j;
      if (p >= n) break;
      uint32_t m = nl_mask16(*(const uint4 *)(d + p));
      if (p + 16u > n) m &= (1u << (uint32_t)(n - p)) - 1u;
      c += __popc(m);
    }
    idx[k] = (uint8_t)c;
  }
}

// Exclusive prefix of cnt[0..n) in place, three steps: per-1024 scans with
// block totals, a scan of the totals, the totals added back.
__global__ __launch_bounds__(1024) void k_tile_scan_blocks(unsigned long long *cnt, uint32_t n,
                                                            unsigned long long *bsum) {
  __shared__ unsigned long long wsum[16];
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
  const unsigned long long v = i < n ? cnt[i] : 0;
  unsigned long long x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  unsigned long long p = 0;
  for (uint32_t q = 0; q < w; q++) p += wsum[q];
  if (i < n) cnt[i] = p + x - v;
  if (threadIdx.x == 1023) bsum[blockIdx.x] = p + x;
}

__global__ __launch_bounds__(1024) void k_tile_scan_totals(unsigned long long *bsum, uint32_t nb) {
  __shared__ unsigned long long wsum[16];
  __shared__ unsigned long long run;
  if (threadIdx.x == 0) run = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint32_t base = 0; base < nb; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const unsigned long long v = i < nb ? bsum[i] : 0;
    unsigned long long x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned long long y = __shfl_up(x, o, 64);
      if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    unsigned long long p = run;
    for (uint32_t q = 0; q < w; q++) p += wsum[q];
    if (i < nb) bsum[i] = p + x - v;
    __syncthreads();
    if (threadIdx.x == 0)
      for (uint32_t q = 0; q < 16; q++) run += wsum[q];
    __syncthreads();
  }
}

__global__ __launch_bounds__(1024) void k_tile_scan_add(unsigned long long *cnt, uint32_t n,
                                                         const unsigned long long *bsum) {
  const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
  if (i < n) cnt[i] += bsum[blockIdx.x];
}

__global__ void k_res_records(DevCtr *ctr, const unsigned long long *prefix, uint32_t *fb_idx, const uint32_t *fb_tile,
                              uint32_t fb_cap) {
  uint32_t n = min(ctr->fb_count, fb_cap);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    fb_idx[k] += (uint32_t)prefix[fb_tile[k]];
  if (blockIdx.x == 0 && threadIdx.x == 0 && ctr->first_bad != ~0ull) {
    unsigned long long b = ctr->first_bad;
    ctr->first_invalid = (uint32_t)(prefix[b >> 32] + (b & 0xFFFFFFFFull));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && ctr->first_te_bad != ~0ull) {
    unsigned long long b = ctr->first_te_bad;
    ctr->first_te = (uint32_t)(prefix[b >> 32] + (b & 0xFFFFFFFFull));
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

namespace {

// A kernel chain captured for one span table and program generation.
struct GraphEntry {
  std::vector<Span> spans;
  uint64_t gen;
  hipGraphExec_t exec;
  uint64_t last;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;  // before H2D, after H2D, after kernels
  hipEvent_t evm0 = nullptr, evm1 = nullptr;              // around the main kernel
  uint8_t *h_pinned = nullptr;
  uint8_t *d_data = nullptr;
  const uint8_t *d_src = nullptr;  // what the last submit read (staging or a resident buffer)
  uint64_t n_bytes = 0;
  bool resident = false;

  DevCtr *d_ctr = nullptr;
  DevCtr *h_ctr = nullptr;
  unsigned long long *d_tile_state = nullptr;
  AggPart *d_parts = nullptr;
  AggOut *d_aggout = nullptr, *h_aggout = nullptr;

  GroupTab gt{};
  warpjq_group *d_groups = nullptr, *h_groups = nullptr;
  warpjq_amd_agg_ext *d_gext = nullptr, *h_gext = nullptr;

  uint32_t *d_row_len = nullptr, *d_sel_start = nullptr, *d_sel_idx = nullptr;
  uint64_t *d_row_off = nullptr, *d_blk_cnt = nullptr, *d_blk_bytes = nullptr;
  uint8_t *d_out = nullptr;
  uint32_t *h_sel_idx = nullptr;
  uint64_t *h_row_off = nullptr;
  uint8_t *h_out = nullptr;

  uint32_t *d_fb_idx = nullptr, *d_fb_off = nullptr, *d_fb_len = nullptr, *d_fb_tile = nullptr;
  uint32_t ntiles = 0;
  uint32_t *h_fb_idx = nullptr, *h_fb_off = nullptr, *h_fb_len = nullptr, *h_fb_span = nullptr;

  // the submit's spans (device copy, pinned staging for it, host view)
  Span *d_spans = nullptr, *h_spans = nullptr;
  warpjq_amd_span *h_pieces = nullptr;
  uint32_t *d_long = nullptr;  // lines for the wave-wide phase, 3 words each, zero between launches
  uint32_t *d_deep_rec = nullptr, *d_deep_end = nullptr;  // row outputs: captures of lines past 64 levels
  SlotV *d_deep_sl = nullptr;
  unsigned long long *d_tsum = nullptr;  // block totals of the line-count scan
  std::vector<GraphEntry> graphs;  // captured chains of resident submits
  uint64_t graph_tick = 0;
  bool keys_blob = false;  // group keys copied out (batched submits)
  uint32_t n_spans = 0;
  uint8_t *d_keyblob = nullptr, *h_keyblob = nullptr;

  bool submitted = false;
  uint32_t first_te = 0xFFFFFFFFu;  // lowest type-error line index of the last wait
};

}  // namespace

static uint32_t g_num_cus = 0;
#define MAX_WG_PER_CU 32u

struct warpjq_ctx {
  DProgHost host;
  // device copies of the program tables
  void *d_tables = nullptr;
  uint32_t tab_bytes = 0;
  uint32_t tab_cap = 0;  // bytes allocated at d_tables
  DProg dprog{};
  Slot *slots = nullptr;
  uint32_t n_slots = 0;
  uint64_t chunk_cap = 0;
  uint64_t max_lines = 0;
  uint32_t fb_cap = 0;
  uint64_t out_cap = 0;
  uint32_t group_cap = 0;     // slots allocated per slot's table
  uint32_t group_active = 0;  // slots in use (a power of two <= group_cap; the autotuner may lower it)
  uint32_t grid = 0;
  uint32_t max_tiles = 0;
  uint32_t claim = CLAIM_DEFAULT;
  uint32_t arena = ARENA_DEFAULT;
  uint32_t graph_cap = 0;  // chains cached per slot (0: no graphs)
  uint64_t prog_gen = 0;   // bumped when anything a captured chain bakes in changes
  bool grouped = false, rows = false;
  int device = 0;
};

static void set_err(char *err, size_t len, const char *op, const char *detail) {
  if (err && len) snprintf(err, len, "%s: %s", op, detail);
}

#define HIP_TRY(expr, op)                                      \
  do {                                                         \
    hipError_t _e = (expr);                                    \
    if (_e != hipSuccess) {                                    \
      set_err(err, err_len, op, hipGetErrorString(_e));        \
      return _e == hipErrorOutOfMemory ? WARPJQ_ERR_OOM : WARPJQ_ERR_CUDA; \
    }                                                          \
  } while (0)

template <class T>
static hipError_t dalloc(T **p, size_t n) {
  if (n == 0) n = 1;
  return hipMalloc((void **)p, n * sizeof(T));
}
template <class T>
static hipError_t halloc(T **p, size_t n) {
  if (n == 0) n = 1;
  return hipHostMalloc((void **)p, n * sizeof(T), hipHostMallocDefault);
}

static void free_slot(Slot &s) {
  if (s.stream) (void)hipStreamSynchronize(s.stream);
  (void)hipFree(s.d_data);
  (void)hipHostFree(s.h_pinned);
  (void)hipFree(s.d_ctr);
  (void)hipHostFree(s.h_ctr);
  (void)hipFree(s.d_tile_state);
  (void)hipFree(s.d_parts);
  (void)hipFree(s.d_aggout);
  (void)hipHostFree(s.h_aggout);
  (void)hipFree(s.gt.keys);
  (void)hipFree(s.gt.hash);
  (void)hipFree(s.gt.count);
  (void)hipFree(s.gt.numeric);
  (void)hipFree(s.gt.isum_lo);
  (void)hipFree(s.gt.isum_hi);
  (void)hipFree(s.gt.mnb);
  (void)hipFree(s.gt.mxb);
  (void)hipFree(s.gt.fsum);
  (void)hipFree(s.gt.fcomp);
  (void)hipFree(s.d_groups);
  (void)hipHostFree(s.h_groups);
  (void)hipFree(s.d_gext);
  (void)hipHostFree(s.h_gext);
  (void)hipFree(s.d_row_len);
  (void)hipFree(s.d_sel_start);
  (void)hipFree(s.d_sel_idx);
  (void)hipFree(s.d_row_off);
  (void)hipFree(s.d_blk_cnt);
  (void)hipFree(s.d_blk_bytes);
  (void)hipFree(s.d_out);
  (void)hipHostFree(s.h_sel_idx);
  (void)hipHostFree(s.h_row_off);
  (void)hipHostFree(s.h_out);
  (void)hipFree(s.d_fb_idx);
  (void)hipFree(s.d_fb_off);
  (void)hipFree(s.d_fb_len);
  (void)hipFree(s.d_fb_tile);
  (void)hipFree(s.d_spans);
  (void)hipFree(s.d_long);
  (void)hipFree(s.d_deep_rec);
  (void)hipFree(s.d_deep_end);
  (void)hipFree(s.d_deep_sl);
  (void)hipFree(s.d_tsum);
  for (GraphEntry &g : s.graphs) (void)hipGraphExecDestroy(g.exec);
  s.graphs.clear();
  (void)hipHostFree(s.h_spans);
  delete[] s.h_pieces;
  (void)hipHostFree(s.h_fb_span);
  (void)hipFree(s.d_keyblob);
  (void)hipHostFree(s.h_keyblob);
  (void)hipHostFree(s.h_fb_idx);
  (void)hipHostFree(s.h_fb_off);
  (void)hipHostFree(s.h_fb_len);
  if (s.ev0) (void)hipEventDestroy(s.ev0);
  if (s.ev1) (void)hipEventDestroy(s.ev1);
  if (s.ev2) (void)hipEventDestroy(s.ev2);
  if (s.evm0) (void)hipEventDestroy(s.evm0);
  if (s.evm1) (void)hipEventDestroy(s.evm1);
  if (s.stream) (void)hipStreamDestroy(s.stream);
  s = Slot{};
}

static warpjq_status alloc_slot(warpjq_ctx *ctx, Slot &s, char *err, size_t err_len) {
  HIP_TRY(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
  HIP_TRY(hipEventCreate(&s.ev0), "hipEventCreate");
  HIP_TRY(hipEventCreate(&s.ev1), "hipEventCreate");
  HIP_TRY(hipEventCreate(&s.ev2), "hipEventCreate");
  HIP_TRY(hipEventCreate(&s.evm0), "hipEventCreate");
  HIP_TRY(hipEventCreate(&s.evm1), "hipEventCreate");
  HIP_TRY(halloc(&s.h_pinned, ctx->chunk_cap), "hipHostMalloc(staging)");
  HIP_TRY(dalloc(&s.d_data, ctx->chunk_cap + DATA_PAD), "hipMalloc(data)");
  HIP_TRY(hipMemsetAsync(s.d_data, 0, ctx->chunk_cap + DATA_PAD, s.stream), "hipMemset(data)");
  HIP_TRY(dalloc(&s.d_ctr, 1), "hipMalloc(ctr)");
  HIP_TRY(halloc(&s.h_ctr, 1), "hipHostMalloc(ctr)");
  HIP_TRY(dalloc(&s.d_tile_state, ctx->max_tiles), "hipMalloc(tile_state)");
  HIP_TRY(dalloc(&s.d_parts, (size_t)g_num_cus * MAX_WG_PER_CU), "hipMalloc(parts)");
  HIP_TRY(dalloc(&s.d_aggout, 1), "hipMalloc(agg)");
  HIP_TRY(halloc(&s.h_aggout, 1), "hipHostMalloc(agg)");
  const uint32_t fc = ctx->fb_cap;
  HIP_TRY(dalloc(&s.d_fb_idx, fc), "hipMalloc(fb)");
  HIP_TRY(dalloc(&s.d_fb_off, fc), "hipMalloc(fb)");
  HIP_TRY(dalloc(&s.d_fb_len, fc), "hipMalloc(fb)");
  HIP_TRY(dalloc(&s.d_fb_tile, fc), "hipMalloc(fb)");
  HIP_TRY(halloc(&s.h_fb_span, fc), "hipHostMalloc(fb)");
  HIP_TRY(dalloc(&s.d_spans, MAX_SPANS), "hipMalloc(spans)");
  HIP_TRY(dalloc(&s.d_long, (size_t)ctx->max_tiles * 3), "hipMalloc(long lines)");
  HIP_TRY(dalloc(&s.d_tsum, (size_t)ctx->max_tiles / 1024 + 2), "hipMalloc(tile sums)");
  HIP_TRY(hipMemset(s.d_long, 0, (size_t)ctx->max_tiles * 12), "hipMemset(long lines)");
  HIP_TRY(halloc(&s.h_spans, MAX_SPANS), "hipHostMalloc(spans)");
  s.h_pieces = new (std::nothrow) warpjq_amd_span[MAX_SPANS];
  if (!s.h_pieces) return WARPJQ_ERR_OOM;
  HIP_TRY(halloc(&s.h_fb_idx, fc), "hipHostMalloc(fb)");
  HIP_TRY(halloc(&s.h_fb_off, fc), "hipHostMalloc(fb)");
  HIP_TRY(halloc(&s.h_fb_len, fc), "hipHostMalloc(fb)");
  if (ctx->grouped) {
    const uint32_t G = ctx->group_cap;
    s.gt.cap = G;
    HIP_TRY(dalloc(&s.gt.keys, G), "hipMalloc(group keys)");
    HIP_TRY(dalloc(&s.gt.hash, G), "hipMalloc(group hash)");
    HIP_TRY(dalloc(&s.gt.count, G), "hipMalloc(group)");
    HIP_TRY(dalloc(&s.gt.numeric, G), "hipMalloc(group)");
    HIP_TRY(dalloc(&s.gt.isum_lo, G), "hipMalloc(group)");
    HIP_TRY(dalloc(&s.gt.isum_hi, G), "hipMalloc(group)");
    HIP_TRY(dalloc(&s.gt.mnb, G), "hipMalloc(group)");
    HIP_TRY(dalloc(&s.gt.mxb, G), "hipMalloc(group)");
    HIP_TRY(dalloc(&s.gt.fsum, G), "hipMalloc(group)");
    HIP_TRY(dalloc(&s.gt.fcomp, G), "hipMalloc(group)");
    HIP_TRY(dalloc(&s.d_groups, G), "hipMalloc(groups)");
    HIP_TRY(halloc(&s.h_groups, G), "hipHostMalloc(groups)");
    HIP_TRY(dalloc(&s.d_gext, G), "hipMalloc(groups)");
    HIP_TRY(halloc(&s.h_gext, G), "hipHostMalloc(groups)");
    HIP_TRY(dalloc(&s.d_keyblob, KEYBLOB_BYTES), "hipMalloc(keys)");
    HIP_TRY(halloc(&s.h_keyblob, KEYBLOB_BYTES), "hipHostMalloc(keys)");
    hipLaunchKernelGGL(k_group_init, dim3(1024), dim3(256), 0, s.stream, s.gt);
    HIP_TRY(hipGetLastError(), "k_group_init");
  }
  if (ctx->rows) {
    HIP_TRY(dalloc(&s.d_deep_rec, DEEP_CAP), "hipMalloc(deep)");
    HIP_TRY(dalloc(&s.d_deep_end, DEEP_CAP), "hipMalloc(deep)");
    HIP_TRY(dalloc(&s.d_deep_sl, (size_t)DEEP_CAP * 16), "hipMalloc(deep)");
    const uint64_t ml = ctx->max_lines;
    const uint64_t nb = (ml + ROWSCAN_BLOCK - 1) / ROWSCAN_BLOCK;
    HIP_TRY(dalloc(&s.d_row_len, ml), "hipMalloc(row_len)");
    HIP_TRY(dalloc(&s.d_sel_start, ml), "hipMalloc(sel_start)");
    HIP_TRY(dalloc(&s.d_sel_idx, ml), "hipMalloc(sel_idx)");
    HIP_TRY(dalloc(&s.d_row_off, ml + 1), "hipMalloc(row_off)");
    HIP_TRY(dalloc(&s.d_blk_cnt, nb), "hipMalloc(blk)");
    HIP_TRY(dalloc(&s.d_blk_bytes, nb), "hipMalloc(blk)");
    HIP_TRY(dalloc(&s.d_out, ctx->out_cap), "hipMalloc(out)");
    HIP_TRY(halloc(&s.h_sel_idx, ml), "hipHostMalloc(sel_idx)");
    HIP_TRY(halloc(&s.h_row_off, ml + 1), "hipHostMalloc(row_off)");
    HIP_TRY(halloc(&s.h_out, ctx->out_cap), "hipHostMalloc(out)");
  }
  HIP_TRY(hipStreamSynchronize(s.stream), "init sync");
  return WARPJQ_OK;
}

template <class T>
static size_t put_table(std::vector<uint8_t> &buf, const T *src, size_t n) {
  size_t off = (buf.size() + 15) & ~(size_t)15;
  buf.resize(off + n * sizeof(T) + 16);
  if (n) memcpy(buf.data() + off, src, n * sizeof(T));
  return off;
}


// Persistent grid of k_main: every workgroup resident at once.
static uint32_t main_grid(const warpjq_ctx *ctx) {
  const bool lb = ctx->rows;
  const uint32_t lds = main_lds_bytes(ctx->arena, ctx->tab_bytes);
  int occ;
  switch (ctx->host.ns) {
    case 1: occ = occupancy_main_1(ctx->grouped, lb, lds); break;
    case 2: occ = occupancy_main_2(ctx->grouped, lb, lds); break;
    case 4: occ = occupancy_main_4(ctx->grouped, lb, lds); break;
    case 8: occ = occupancy_main_8(ctx->grouped, lb, lds); break;
    default: occ = occupancy_main_16(ctx->grouped, lb, lds); break;
  }
  if (occ < 1) occ = 1;
  if (occ > (int)MAX_WG_PER_CU) occ = MAX_WG_PER_CU;
  return g_num_cus * (uint32_t)occ;
}

extern "C" {

warpjq_status warpjq_probe(char *err, size_t err_len) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    set_err(err, err_len, "no usable HIP device", e != hipSuccess ? hipGetErrorString(e) : "device count is zero");
    return WARPJQ_ERR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, 0);
  if (e != hipSuccess) {
    set_err(err, err_len, "hipGetDeviceProperties", hipGetErrorString(e));
    return WARPJQ_ERR_NO_DEVICE;
  }
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_err(err, err_len, "unsupported GPU", "this build targets gfx950 (MI355X) only");
    return WARPJQ_ERR_NO_DEVICE;
  }
  return WARPJQ_OK;
}

warpjq_status warpjq_device_name(char *buf, size_t len) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return WARPJQ_ERR_NO_DEVICE;
  snprintf(buf, len, "%s (%s, %d CUs, %zu MiB)", prop.name, prop.gcnArchName, prop.multiProcessorCount,
           (size_t)(prop.totalGlobalMem >> 20));
  return WARPJQ_OK;
}

warpjq_status warpjq_abi_check(const uint64_t *sizes, size_t n) {
  const uint64_t mine[] = {sizeof(warpjq_step),        sizeof(warpjq_path),  sizeof(warpjq_cmp),
                           sizeof(warpjq_cond_op),     sizeof(warpjq_agg_partial), sizeof(warpjq_group)};
  if (!sizes || n != sizeof(mine) / sizeof(mine[0])) return WARPJQ_ERR_ABI;
  for (size_t i = 0; i < n; i++)
    if (sizes[i] != mine[i]) return WARPJQ_ERR_ABI;
  return WARPJQ_OK;
}

// LDS one workgroup may take on this device, and what the grouped main
// kernel's static table adds to its dynamic LDS.
static uint32_t lds_limit() {
  static uint32_t lim = 0;
  if (!lim) {
    hipDeviceProp_t p;
    int dev = 0;
    (void)hipGetDevice(&dev);
    lim = hipGetDeviceProperties(&p, dev) == hipSuccess && p.sharedMemPerBlock ? (uint32_t)p.sharedMemPerBlock : 65536u;
  }
  return lim;
}

// Whether the main kernel's LDS for `arena` and `tab` bytes of tables fits a
// workgroup.
static bool lds_fits(const warpjq_ctx *ctx, uint32_t arena, uint32_t tab);

// Uploads the program tables (kept for the context's life, grown when a
// reprogrammed query needs more) and points ctx->dprog at them; sets the LDS
// table size and the grid that follows from it. Tables are staged in LDS per
// workgroup: up to 16 KiB beside the default arena; larger ones (a long
// string literal, many paths) are staged whole with the arena shrunk until
// both fit the workgroup's LDS, which lowers occupancy but not results.
// ctx->host must hold the program.
static warpjq_status put_program(warpjq_ctx *ctx, const warpjq_program *prog, char *err, size_t err_len) {
  std::vector<uint8_t> buf;
  const DProg &h = ctx->host.d;
  size_t o_steps = put_table(buf, prog->steps, prog->n_steps);
  size_t o_cmps = put_table(buf, prog->cmps, prog->n_cmps);
  size_t o_cond = put_table(buf, ctx->host.cond.data(), ctx->host.cond.size());
  size_t o_blob = put_table(buf, prog->blob, prog->n_blob);
  size_t o_spath = put_table(buf, ctx->host.spath.data(), ctx->host.spath.size());
  size_t o_sop = put_table(buf, prog->slot_of_path, prog->n_paths);
  size_t o_fs = put_table(buf, ctx->host.field_slot.data(), ctx->host.field_slot.size());
  size_t o_po = put_table(buf, prog->prefix_off, prog->n_fields);
  size_t o_pl = put_table(buf, prog->prefix_len, prog->n_fields);
  buf.resize((buf.size() + 15) & ~(size_t)15);
  const uint32_t tb = (uint32_t)buf.size();
  if (!lds_fits(ctx, ctx->arena, tb)) {
    // shrink the arena (and then the claim) until tables and arena fit
    uint32_t claim = ctx->claim, arena = ctx->arena;
    while (!lds_fits(ctx, arena, tb) && arena > claim + TAIL_MAX) arena -= 256;
    while (!lds_fits(ctx, arena, tb) && claim > CLAIM_MIN) {
      claim -= STEP_BYTES;
      arena = claim + TAIL_MAX;
    }
    if (!lds_fits(ctx, arena, tb)) {
      set_err(err, err_len, "query is too large for the kernels",
              "its tables and the smallest line arena do not fit one workgroup's LDS");
      return WARPJQ_ERR_INVALID_ARG;
    }
    ctx->claim = claim;
    ctx->arena = arena;
  }
  if (ctx->tab_cap < std::max<uint32_t>(tb, TAB_LDS_MAX)) {
    if (ctx->d_tables) HIP_TRY(hipFree(ctx->d_tables), "uploading the compiled query");
    ctx->d_tables = nullptr;
    ctx->tab_cap = std::max<uint32_t>(tb, TAB_LDS_MAX);
    HIP_TRY(hipMalloc(&ctx->d_tables, ctx->tab_cap), "uploading the compiled query");
  }
  if (!buf.empty())
    HIP_TRY(hipMemcpy(ctx->d_tables, buf.data(), buf.size(), hipMemcpyHostToDevice), "uploading the compiled query");
  ctx->tab_bytes = (uint32_t)buf.size();
  ctx->grid = main_grid(ctx);
  uint8_t *base = (uint8_t *)ctx->d_tables;
  DProg d = h;
  d.steps = (const warpjq_step *)(base + o_steps);
  d.cmps = (const warpjq_cmp *)(base + o_cmps);
  d.cond = (const warpjq_cond_op *)(base + o_cond);
  d.blob = base + o_blob;
  d.spath = (const SlotPath *)(base + o_spath);
  d.slot_of_path = (const uint32_t *)(base + o_sop);
  d.field_slot = (const uint32_t *)(base + o_fs);
  d.prefix_off = (const uint32_t *)(base + o_po);
  d.prefix_len = (const uint32_t *)(base + o_pl);
  ctx->dprog = d;
  return WARPJQ_OK;
}

static bool lds_fits(const warpjq_ctx *ctx, uint32_t arena, uint32_t tab) {
  const uint32_t stat = ctx->grouped ? (uint32_t)sizeof(LGroup) : 0u;
  return main_lds_bytes(arena, tab) + stat <= lds_limit();
}

warpjq_status warpjq_ctx_create(const warpjq_program *prog, uint64_t max_chunk_bytes, uint32_t n_slots,
                                warpjq_ctx **out, char *err, size_t err_len) {
  if (!prog || !out || n_slots == 0 || max_chunk_bytes == 0) return WARPJQ_ERR_INVALID_ARG;
  if (max_chunk_bytes > MAX_CHUNK) {
    set_err(err, err_len, "chunk is too large", "chunk offsets are 32-bit");
    return WARPJQ_ERR_INVALID_ARG;
  }
  warpjq_ctx *ctx = new (std::nothrow) warpjq_ctx();
  if (!ctx) return WARPJQ_ERR_OOM;
  std::string why;
  if (!build_dprog(*prog, ctx->host, why)) {
    set_err(err, err_len, "query is not supported by the kernels", why.c_str());
    delete ctx;
    return WARPJQ_ERR_INVALID_ARG;
  }
  (void)hipGetDevice(&ctx->device);
  if (!g_num_cus) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) g_num_cus = (uint32_t)prop.multiProcessorCount;
    if (!g_num_cus) g_num_cus = 256;
  }
  ctx->n_slots = n_slots;
  ctx->chunk_cap = max_chunk_bytes;
  ctx->max_lines = max_chunk_bytes / MIN_LINE_BYTES + 2;
  ctx->fb_cap = (uint32_t)(ctx->max_lines / 16 + 1024);
  ctx->out_cap = max_chunk_bytes + max_chunk_bytes / 2 + (1u << 20);
  ctx->grouped = prog->output_kind == WARPJQ_OUT_AGG && prog->group_path != 0xFFFFFFFFu;
  ctx->rows = prog->output_kind != WARPJQ_OUT_AGG;
  if (ctx->grouped) {
    ctx->claim = GROUP_CLAIM_DEFAULT;
    ctx->arena = GROUP_ARENA_DEFAULT;
  }
  uint64_t g = 65536;
  while (g < max_chunk_bytes / 128 && g < (1u << 22)) g <<= 1;
  ctx->group_cap = (uint32_t)g;
  ctx->group_active = (uint32_t)g;
  // tiles for the largest chunk a slot or a resident buffer may hold
  ctx->max_tiles = (uint32_t)((MAX_CHUNK + 16 + CLAIM_MIN - 1) / CLAIM_MIN) + 1;

  if (warpjq_status st = put_program(ctx, prog, err, err_len)) {
    warpjq_ctx_destroy(ctx);
    return st;
  }

  ctx->slots = new (std::nothrow) Slot[n_slots];
  if (!ctx->slots) {
    warpjq_ctx_destroy(ctx);
    return WARPJQ_ERR_OOM;
  }
  for (uint32_t s = 0; s < n_slots; s++) {
    warpjq_status st = alloc_slot(ctx, ctx->slots[s], err, err_len);
    if (st != WARPJQ_OK) {
      warpjq_ctx_destroy(ctx);
      return st;
    }
  }
  *out = ctx;
  return WARPJQ_OK;
}

void warpjq_ctx_destroy(warpjq_ctx *ctx) {
  if (!ctx) return;
  if (ctx->slots) {
    for (uint32_t i = 0; i < ctx->n_slots; i++) free_slot(ctx->slots[i]);
    delete[] ctx->slots;
  }
  (void)hipFree(ctx->d_tables);
  delete ctx;
}

warpjq_status warpjq_amd_ctx_reprogram(warpjq_ctx *ctx, const warpjq_program *prog, char *err, size_t err_len) {
  if (!ctx || !prog) return WARPJQ_ERR_INVALID_ARG;
  const bool grouped = prog->output_kind == WARPJQ_OUT_AGG && prog->group_path != 0xFFFFFFFFu;
  const bool rows = prog->output_kind != WARPJQ_OUT_AGG;
  if (grouped != ctx->grouped || rows != ctx->rows) {
    set_err(err, err_len, "reprogram", "the new query needs other buffers (grouped / row output differ)");
    return WARPJQ_ERR_INVALID_ARG;
  }
  DProgHost host;
  std::string why;
  if (!build_dprog(*prog, host, why)) {
    set_err(err, err_len, "query is not supported by the kernels", why.c_str());
    return WARPJQ_ERR_INVALID_ARG;
  }
  for (uint32_t i = 0; i < ctx->n_slots; i++) HIP_TRY(hipStreamSynchronize(ctx->slots[i].stream), "reprogram");
  DProgHost old = ctx->host;
  ctx->host = host;
  warpjq_status st = put_program(ctx, prog, err, err_len);
  if (st != WARPJQ_OK) ctx->host = old;  // tables too large: the old program stays
  ctx->prog_gen++;
  return st;
}

uint8_t *warpjq_slot_buffer(warpjq_ctx *ctx, uint32_t slot) {
  if (!ctx || slot >= ctx->n_slots) return nullptr;
  return ctx->slots[slot].h_pinned;
}

uint64_t warpjq_slot_capacity(const warpjq_ctx *ctx) { return ctx ? ctx->chunk_cap : 0; }

uint64_t warpjq_max_lines(const warpjq_ctx *ctx) { return ctx ? ctx->max_lines : 0; }

}  // extern "C"

// The span table's copy to the device and the fixed-size results' copy back,
// on either side of the kernels (a captured graph holds kernel nodes only:
// the copies are queued around its launch).
static warpjq_status copy_spans_in(Slot &s, uint32_t ns, char *err, size_t err_len) {
  HIP_TRY(hipMemcpyAsync(s.d_spans, s.h_spans, (size_t)ns * sizeof(Span), hipMemcpyHostToDevice, s.stream),
          "H2D(spans)");
  return WARPJQ_OK;
}
static warpjq_status copy_results_out(warpjq_ctx *ctx, Slot &s, char *err, size_t err_len) {
  HIP_TRY(hipMemcpyAsync(s.h_ctr, s.d_ctr, sizeof(DevCtr), hipMemcpyDeviceToHost, s.stream), "D2H(counters)");
  if (ctx->dprog.output_kind == WARPJQ_OUT_AGG && !ctx->grouped)
    HIP_TRY(hipMemcpyAsync(s.h_aggout, s.d_aggout, sizeof(AggOut), hipMemcpyDeviceToHost, s.stream), "D2H(agg)");
  return WARPJQ_OK;
}

// Queues the kernel chain over the span table already in s.h_spans[0..ns)
// on the slot's stream. `timed`: hipEvents around the main kernel (left out
// of a captured graph); `copies`: the span and result copies too (left out
// of a captured graph, whose caller queues them around the launch).
static warpjq_status enqueue_chain(warpjq_ctx *ctx, Slot &s, uint32_t ns, uint32_t ntiles, uint64_t total, bool timed,
                                   bool copies, char *err, size_t err_len) {
  hipStream_t st = s.stream;
  const uint32_t claim = ctx->claim;
  if (copies && total > 0)
    if (warpjq_status cs = copy_spans_in(s, ns, err, err_len)) return cs;
  hipLaunchKernelGGL(k_reset, dim3(std::max<uint32_t>(1, std::min<uint32_t>(1024, (ntiles + 255) / 256))), dim3(256),
                     0, st, s.d_ctr, s.d_tile_state, ntiles);
  if (total > 0) {
    KArgs a{};
    a.spans = s.d_spans;
    a.n_spans = ns;
    a.ntiles = ntiles;
    a.claim = claim;
    a.arena = ctx->arena;
    a.P = ctx->dprog;
    a.tab = (const uint8_t *)ctx->d_tables;
    a.tab_bytes = ctx->tab_bytes;
    a.ctr = s.d_ctr;
    a.tile_state = s.d_tile_state;
    a.parts = s.d_parts;
    a.gt = s.gt;
    a.gt.cap = ctx->group_active;
    a.row_len = s.d_row_len;
    a.sel_start = s.d_sel_start;
    a.max_lines = ctx->rows ? ctx->max_lines : 0;
    a.fb_idx = s.d_fb_idx;
    a.fb_off = s.d_fb_off;
    a.fb_len = s.d_fb_len;
    a.fb_tile = s.d_fb_tile;
    a.fb_cap = ctx->fb_cap;
    a.long_q = s.d_long;
    a.deep_rec = s.d_deep_rec;
    a.deep_end = s.d_deep_end;
    a.deep_sl = s.d_deep_sl;
    a.deep_cap = ctx->rows ? DEEP_CAP : 0;
    a.long_cap = ctx->max_tiles;
    uint32_t grid = std::min<uint32_t>(ctx->grid, ntiles);
    if (ctx->rows) {  // each claim's first line index, so claims can go in any order
      const uint32_t nb = (ntiles + 1023) / 1024;
      hipLaunchKernelGGL(k_line_count, dim3(std::min<uint32_t>(4096, (ntiles + 3) / 4)), dim3(256), 0, st, s.d_spans, ns,
                         ntiles, claim, s.d_tile_state);
      hipLaunchKernelGGL(k_tile_scan_blocks, dim3(nb), dim3(1024), 0, st, s.d_tile_state, ntiles, s.d_tsum);
      hipLaunchKernelGGL(k_tile_scan_totals, dim3(1), dim3(1024), 0, st, s.d_tsum, nb);
      hipLaunchKernelGGL(k_tile_scan_add, dim3(nb), dim3(1024), 0, st, s.d_tile_state, ntiles, s.d_tsum);
    }
    if (timed) HIP_TRY(hipEventRecord(s.evm0, st), "hipEventRecord");
    switch (ctx->host.ns) {
      case 1: launch_main_1(ctx->grouped, ctx->rows, grid, st, a); break;
      case 2: launch_main_2(ctx->grouped, ctx->rows, grid, st, a); break;
      case 4: launch_main_4(ctx->grouped, ctx->rows, grid, st, a); break;
      case 8: launch_main_8(ctx->grouped, ctx->rows, grid, st, a); break;
      default: launch_main_16(ctx->grouped, ctx->rows, grid, st, a); break;
    }
    HIP_TRY(hipGetLastError(), "k_main");
    if (timed) HIP_TRY(hipEventRecord(s.evm1, st), "hipEventRecord");
    if (ctx->dprog.output_kind == WARPJQ_OUT_AGG) {
      if (ctx->grouped) {
        GroupTab gt = s.gt;
        gt.cap = ctx->group_active;
        hipLaunchKernelGGL(k_group_compact, dim3(1024), dim3(256), 0, st, gt, s.d_ctr, s.d_groups, s.d_gext, s.d_spans,
                           ns, s.keys_blob ? s.d_keyblob : nullptr, KEYBLOB_BYTES);
      } else {
        hipLaunchKernelGGL(k_agg_final, dim3(1), dim3(256), 0, st, s.d_parts, grid, s.d_ctr, s.d_aggout);
      }
    } else {
      uint64_t lines_bound = std::min<uint64_t>(ctx->max_lines, total + ns);
      uint32_t nb = (uint32_t)((lines_bound + ROWSCAN_BLOCK - 1) / ROWSCAN_BLOCK);
      hipLaunchKernelGGL(k_row_reduce, dim3(nb), dim3(256), 0, st, s.d_row_len, s.d_ctr, ctx->max_lines, s.d_blk_cnt,
                         s.d_blk_bytes);
      hipLaunchKernelGGL(k_row_blockscan, dim3(1), dim3(1024), 0, st, s.d_blk_cnt, s.d_blk_bytes, s.d_ctr,
                         ctx->max_lines, ctx->out_cap);
      hipLaunchKernelGGL(k_row_scatter, dim3(nb), dim3(256), 0, st, s.d_row_len, s.d_ctr, ctx->max_lines,
                         s.d_blk_cnt, s.d_blk_bytes, s.d_sel_idx, s.d_row_off);
      hipLaunchKernelGGL(k_row_finish, dim3(1), dim3(64), 0, st, s.d_ctr, s.d_row_off);
      uint32_t eg = std::min<uint32_t>(ctx->grid, (uint32_t)((lines_bound + 255) / 256));
      if (eg == 0) eg = 1;
      switch (ctx->host.ns) {
        case 1: launch_emit_1(eg, st, s.d_spans, ns, ctx->dprog, (const uint8_t *)ctx->d_tables, ctx->tab_bytes, s.d_ctr, s.d_sel_idx, s.d_row_off, s.d_sel_start, s.d_out, a); break;
        case 2: launch_emit_2(eg, st, s.d_spans, ns, ctx->dprog, (const uint8_t *)ctx->d_tables, ctx->tab_bytes, s.d_ctr, s.d_sel_idx, s.d_row_off, s.d_sel_start, s.d_out, a); break;
        case 4: launch_emit_4(eg, st, s.d_spans, ns, ctx->dprog, (const uint8_t *)ctx->d_tables, ctx->tab_bytes, s.d_ctr, s.d_sel_idx, s.d_row_off, s.d_sel_start, s.d_out, a); break;
        case 8: launch_emit_8(eg, st, s.d_spans, ns, ctx->dprog, (const uint8_t *)ctx->d_tables, ctx->tab_bytes, s.d_ctr, s.d_sel_idx, s.d_row_off, s.d_sel_start, s.d_out, a); break;
        default: launch_emit_16(eg, st, s.d_spans, ns, ctx->dprog, (const uint8_t *)ctx->d_tables, ctx->tab_bytes, s.d_ctr, s.d_sel_idx, s.d_row_off, s.d_sel_start, s.d_out, a); break;
      }
    }
    HIP_TRY(hipGetLastError(), "kernel chain");
  }
  if (copies) return copy_results_out(ctx, s, err, err_len);
  return WARPJQ_OK;
}

// Queues the kernel chain over the device spans `in[0..n_in)` on `s`.
static warpjq_status launch_chain(warpjq_ctx *ctx, Slot &s, const warpjq_amd_span *in, uint32_t n_in, char *err,
                                  size_t err_len) {
  hipStream_t st = s.stream;
  const uint32_t claim = ctx->claim;
  // place the spans in the kernels' position space, each on a claim boundary
  uint64_t v = 0, total = 0;
  uint32_t ns = 0;
  for (uint32_t i = 0; i < n_in; i++) {
    if (in[i].n_bytes == 0) continue;
    const uint32_t mis = (uint32_t)((uintptr_t)in[i].d_bytes & 15u);
    Span sp;
    sp.base = in[i].d_bytes - mis;
    sp.vbase = (uint32_t)v;
    sp.start = (uint32_t)(v + mis);
    sp.vend = (uint32_t)(v + mis + in[i].n_bytes);
    sp._pad = 0;
    sp.nlidx = nullptr;
    sp.ioff = 0;
    s.h_spans[ns++] = sp;
    total += in[i].n_bytes;
    v = ((v + mis + in[i].n_bytes + claim - 1) / claim) * claim;
  }
  if (ns == 0) {  // keep one empty span so results can be read uniformly
    s.h_spans[0] = Span{in && n_in ? in[0].d_bytes : nullptr, 0, 0, 0, 0, nullptr, 0};
  }
  s.n_spans = ns ? ns : 1;
  // row outputs over spans inside resident datasets count their lines from
  // the datasets' newline indices (one registry lookup pass per submit)
  if (ctx->rows && ns) wjqd::nlindex_attach(s.h_spans, ns);
  const uint32_t ntiles = (uint32_t)(v / claim);
  s.ntiles = ntiles;
  // a resident submit seen before on this slot: replay its captured chain
  if (ctx->graph_cap && s.resident && total > 0) {
    GraphEntry *hit = nullptr;
    for (GraphEntry &g : s.graphs)
      if (g.gen == ctx->prog_gen && g.spans.size() == ns &&
          memcmp(g.spans.data(), s.h_spans, (size_t)ns * sizeof(Span)) == 0)
        hit = &g;
    if (!hit) {
      HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      warpjq_status cs = enqueue_chain(ctx, s, ns, ntiles, total, false, false, err, err_len);
      hipGraph_t graph = nullptr;
      hipError_t ce = hipStreamEndCapture(st, &graph);
      if (cs != WARPJQ_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return cs;
      }
      HIP_TRY(ce, "hipStreamEndCapture");
      hipGraphExec_t exec = nullptr;
      hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIP_TRY(ie, "hipGraphInstantiate");
      if (s.graphs.size() >= ctx->graph_cap) {  // drop the least recently used
        auto old = std::min_element(s.graphs.begin(), s.graphs.end(),
                                    [](const GraphEntry &a, const GraphEntry &b) { return a.last < b.last; });
        (void)hipGraphExecDestroy(old->exec);
        s.graphs.erase(old);
      }
      s.graphs.push_back(GraphEntry{std::vector<Span>(s.h_spans, s.h_spans + ns), ctx->prog_gen, exec, 0});
      hit = &s.graphs.back();
    }
    hit->last = ++s.graph_tick;
    if (warpjq_status cs = copy_spans_in(s, ns, err, err_len)) return cs;
    HIP_TRY(hipEventRecord(s.evm0, st), "hipEventRecord");
    HIP_TRY(hipGraphLaunch(hit->exec, st), "hipGraphLaunch");
    HIP_TRY(hipEventRecord(s.evm1, st), "hipEventRecord");
    if (warpjq_status cs = copy_results_out(ctx, s, err, err_len)) return cs;
  } else {
    warpjq_status es = enqueue_chain(ctx, s, ns, ntiles, total, true, true, err, err_len);
    if (es != WARPJQ_OK) return es;
  }
  HIP_TRY(hipEventRecord(s.ev2, st), "hipEventRecord");
  s.submitted = true;
  return WARPJQ_OK;
}

extern "C" {

warpjq_status warpjq_submit(warpjq_ctx *ctx, uint32_t slot, uint64_t n_bytes, char *err, size_t err_len) {
  if (!ctx || slot >= ctx->n_slots) return WARPJQ_ERR_INVALID_ARG;
  if (n_bytes > ctx->chunk_cap) return WARPJQ_ERR_INVALID_ARG;
  Slot &s = ctx->slots[slot];
  s.n_bytes = n_bytes;
  s.resident = false;
  s.d_src = s.d_data;
  HIP_TRY(hipEventRecord(s.ev0, s.stream), "hipEventRecord");
  if (n_bytes) HIP_TRY(hipMemcpyAsync(s.d_data, s.h_pinned, n_bytes, hipMemcpyHostToDevice, s.stream), "H2D chunk");
  HIP_TRY(hipEventRecord(s.ev1, s.stream), "hipEventRecord");
  const warpjq_amd_span one{s.d_data, n_bytes};
  s.keys_blob = false;
  return launch_chain(ctx, s, &one, 1, err, err_len);
}

warpjq_status warpjq_amd_submit_device(warpjq_ctx *ctx, uint32_t slot, const uint8_t *d_bytes, uint64_t n_bytes,
                                       char *err, size_t err_len) {
  if (!ctx || slot >= ctx->n_slots || (!d_bytes && n_bytes)) return WARPJQ_ERR_INVALID_ARG;
  if (n_bytes > MAX_CHUNK || (ctx->rows && n_bytes > ctx->chunk_cap)) return WARPJQ_ERR_INVALID_ARG;
  Slot &s = ctx->slots[slot];
  s.n_bytes = n_bytes;
  s.resident = true;
  s.d_src = d_bytes;
  HIP_TRY(hipEventRecord(s.ev0, s.stream), "hipEventRecord");
  HIP_TRY(hipEventRecord(s.ev1, s.stream), "hipEventRecord");
  const warpjq_amd_span one{d_bytes, n_bytes};
  s.keys_blob = false;
  return launch_chain(ctx, s, &one, 1, err, err_len);
}

warpjq_status warpjq_amd_submit_spans(warpjq_ctx *ctx, uint32_t slot, const warpjq_amd_span *spans, uint32_t n_spans,
                                      char *err, size_t err_len) {
  if (!ctx || slot >= ctx->n_slots || (!spans && n_spans) || n_spans > MAX_SPANS) return WARPJQ_ERR_INVALID_ARG;
  uint64_t v = 0, total = 0;
  for (uint32_t i = 0; i < n_spans; i++) {
    if (!spans[i].d_bytes && spans[i].n_bytes) return WARPJQ_ERR_INVALID_ARG;
    if (!spans[i].n_bytes) continue;
    const uint64_t mis = (uintptr_t)spans[i].d_bytes & 15u;
    v = ((v + mis + spans[i].n_bytes + ctx->claim - 1) / ctx->claim) * ctx->claim;
    total += spans[i].n_bytes;
  }
  if (v > MAX_CHUNK) {
    set_err(err, err_len, "batch is too large", "the spans of one submit must fit 4 GiB - 1 MiB of positions");
    return WARPJQ_ERR_INVALID_ARG;
  }
  if (ctx->rows && total > ctx->chunk_cap) {
    set_err(err, err_len, "batch is too large", "row outputs are bounded by the slot capacity");
    return WARPJQ_ERR_INVALID_ARG;
  }
  Slot &s = ctx->slots[slot];
  s.n_bytes = total;
  s.resident = true;
  s.d_src = n_spans ? spans[0].d_bytes : nullptr;
  HIP_TRY(hipEventRecord(s.ev0, s.stream), "hipEventRecord");
  HIP_TRY(hipEventRecord(s.ev1, s.stream), "hipEventRecord");
  s.keys_blob = true;
  return launch_chain(ctx, s, spans, n_spans, err, err_len);
}

warpjq_status warpjq_amd_submit_pieces(warpjq_ctx *ctx, uint32_t slot, uint64_t n_bytes, const warpjq_amd_piece *pieces,
                                       uint32_t n_pieces, char *err, size_t err_len) {
  if (!ctx || slot >= ctx->n_slots || n_bytes > ctx->chunk_cap || (!pieces && n_pieces) || n_pieces > MAX_SPANS)
    return WARPJQ_ERR_INVALID_ARG;
  Slot &s = ctx->slots[slot];
  warpjq_amd_span *sp = s.h_pieces;
  for (uint32_t i = 0; i < n_pieces; i++) {
    if (pieces[i].off > n_bytes || pieces[i].len > n_bytes - pieces[i].off) {
      set_err(err, err_len, "submit_pieces", "a piece lies outside the slot's bytes");
      return WARPJQ_ERR_INVALID_ARG;
    }
    sp[i].d_bytes = s.d_data + pieces[i].off;
    sp[i].n_bytes = pieces[i].len;
  }
  s.n_bytes = n_bytes;
  s.resident = false;
  s.d_src = s.d_data;
  HIP_TRY(hipEventRecord(s.ev0, s.stream), "hipEventRecord");
  if (n_bytes) HIP_TRY(hipMemcpyAsync(s.d_data, s.h_pinned, n_bytes, hipMemcpyHostToDevice, s.stream), "H2D chunk");
  HIP_TRY(hipEventRecord(s.ev1, s.stream), "hipEventRecord");
  s.keys_blob = true;
  return launch_chain(ctx, s, sp, n_pieces, err, err_len);
}

uint8_t *warpjq_amd_device_upload(const uint8_t *bytes, uint64_t n, char *err, size_t err_len) {
  if (!bytes && n) return nullptr;
  uint8_t *p = nullptr;
  hipError_t e = hipMalloc((void **)&p, n + 256);
  if (e == hipSuccess && n) e = hipMemcpy(p, bytes, n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(p + n, '\n', 256);
  if (e != hipSuccess) {
    set_err(err, err_len, "device_upload", hipGetErrorString(e));
    if (p) (void)hipFree(p);
    return nullptr;
  }
  return p;
}

void warpjq_amd_device_free(uint8_t *p) {
  if (p) (void)hipFree(p);
}

const uint32_t *warpjq_amd_fallback_spans(const warpjq_ctx *ctx, uint32_t slot) {
  if (!ctx || slot >= ctx->n_slots) return nullptr;
  return ctx->slots[slot].h_fb_span;
}

uint32_t warpjq_amd_first_type_error(const warpjq_ctx *ctx, uint32_t slot) {
  if (!ctx || slot >= ctx->n_slots) return 0xFFFFFFFFu;
  return ctx->slots[slot].first_te;
}

const uint8_t *warpjq_amd_group_keys(const warpjq_ctx *ctx, uint32_t slot) {
  if (!ctx || slot >= ctx->n_slots) return nullptr;
  return ctx->slots[slot].h_keyblob;
}

warpjq_status warpjq_wait(warpjq_ctx *ctx, uint32_t slot, warpjq_chunk_result *out, char *err, size_t err_len) {
  if (!ctx || slot >= ctx->n_slots || !out) return WARPJQ_ERR_INVALID_ARG;
  Slot &s = ctx->slots[slot];
  HIP_TRY(hipStreamSynchronize(s.stream), "stream sync");
  HIP_TRY(hipGetLastError(), "kernel");
  memset(out, 0, sizeof *out);
  out->first_invalid = 0xFFFFFFFFu;
  s.first_te = 0xFFFFFFFFu;
  if (!s.submitted || s.n_bytes == 0) return WARPJQ_OK;
  DevCtr c = *s.h_ctr;
  if (!ctx->rows && !c.overflow && (c.first_bad != ~0ull || c.first_te_bad != ~0ull || c.fb_count) &&
      c.fb_count <= ctx->fb_cap) {
    // aggregate queries recorded (claim, rank): turn them into line indices,
    // the claims' line counts scanned as the row path scans them
    const uint32_t nb = (s.ntiles + 1023) / 1024;
    hipLaunchKernelGGL(k_tile_scan_blocks, dim3(nb), dim3(1024), 0, s.stream, s.d_tile_state, s.ntiles, s.d_tsum);
    hipLaunchKernelGGL(k_tile_scan_totals, dim3(1), dim3(1024), 0, s.stream, s.d_tsum, nb);
    hipLaunchKernelGGL(k_tile_scan_add, dim3(nb), dim3(1024), 0, s.stream, s.d_tile_state, s.ntiles, s.d_tsum);
    hipLaunchKernelGGL(k_res_records, dim3(std::max<uint32_t>(1, std::min<uint32_t>(1024, (c.fb_count + 255) / 256))),
                       dim3(256), 0, s.stream, s.d_ctr, s.d_tile_state, s.d_fb_idx, s.d_fb_tile, ctx->fb_cap);
    HIP_TRY(hipGetLastError(), "k_res");
    HIP_TRY(hipMemcpyAsync(s.h_ctr, s.d_ctr, sizeof(DevCtr), hipMemcpyDeviceToHost, s.stream), "D2H(counters)");
    HIP_TRY(hipStreamSynchronize(s.stream), "stream sync");
    c = *s.h_ctr;
  }
  out->n_lines = c.n_lines;
  out->n_blank = c.n_blank;
  out->n_invalid = c.n_invalid;
  out->n_type_error = c.n_type_error;
  out->n_fallback = c.n_fallback;
  out->first_invalid = c.first_invalid;
  s.first_te = c.first_te_bad == ~0ull ? 0xFFFFFFFFu : ctx->rows ? (uint32_t)c.first_te_bad : c.first_te;
  out->chunk_overflow = c.overflow || c.fb_count > ctx->fb_cap;
  out->agg.saw_non_numeric = c.saw_non_numeric;
  if (out->chunk_overflow) {
    out->n_lines = 0;
    return WARPJQ_OK;
  }
  uint32_t nfb = c.fb_count;
  if (nfb) {
    HIP_TRY(hipMemcpyAsync(s.h_fb_idx, s.d_fb_idx, nfb * 4ull, hipMemcpyDeviceToHost, s.stream), "D2H(fallback)");
    HIP_TRY(hipMemcpyAsync(s.h_fb_off, s.d_fb_off, nfb * 4ull, hipMemcpyDeviceToHost, s.stream), "D2H(fallback)");
    HIP_TRY(hipMemcpyAsync(s.h_fb_len, s.d_fb_len, nfb * 4ull, hipMemcpyDeviceToHost, s.stream), "D2H(fallback)");
    HIP_TRY(hipStreamSynchronize(s.stream), "sync(fallback)");
    // q positions -> (span, offset from the span's first byte)
    for (uint32_t k = 0; k < nfb; k++) {
      const uint32_t q = s.h_fb_off[k];
      uint32_t lo = 0, hi = s.n_spans;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (s.h_spans[mid].vbase <= q) lo = mid;
        else hi = mid;
      }
      s.h_fb_span[k] = lo;
      s.h_fb_off[k] = q - s.h_spans[lo].start;
    }
  }
  out->fallback_idx = s.h_fb_idx;
  out->fallback_off = s.h_fb_off;
  out->fallback_len = s.h_fb_len;
  out->n_fallback = nfb;
  if (ctx->dprog.output_kind == WARPJQ_OUT_AGG) {
    if (ctx->grouped) {
      uint32_t ng = std::min<uint32_t>(c.n_groups, ctx->group_cap);
      if (s.keys_blob && c.blob_used)
        HIP_TRY(hipMemcpyAsync(s.h_keyblob, s.d_keyblob, std::min<uint32_t>(c.blob_used, KEYBLOB_BYTES),
                               hipMemcpyDeviceToHost, s.stream),
                "D2H(group keys)");
      if (ng) {
        HIP_TRY(hipMemcpyAsync(s.h_groups, s.d_groups, (size_t)ng * sizeof(warpjq_group), hipMemcpyDeviceToHost,
                               s.stream),
                "D2H(groups)");
        HIP_TRY(hipMemcpyAsync(s.h_gext, s.d_gext, (size_t)ng * sizeof(warpjq_amd_agg_ext), hipMemcpyDeviceToHost,
                               s.stream),
                "D2H(groups)");
      }
      HIP_TRY(hipStreamSynchronize(s.stream), "sync(groups)");
      for (uint32_t i = 0; i < ng; i++) s.h_groups[i].sum = warpjq_amd_ext_value(&s.h_gext[i]);
      out->groups = s.h_groups;
      out->n_groups = ng;
      out->group_overflow = c.group_overflow;
    } else {
      HIP_TRY(hipStreamSynchronize(s.stream), "sync(agg)");
      out->agg = s.h_aggout->agg;
      out->agg.sum = warpjq_amd_ext_value(&s.h_aggout->ext);
      out->agg.saw_non_numeric = c.saw_non_numeric;
    }
    return WARPJQ_OK;
  }
  uint64_t nsel = c.n_sel;
  out->n_selected = nsel;
  out->out_bytes = s.h_out;
  out->out_line_idx = s.h_sel_idx;
  out->out_row_off = s.h_row_off;
  if (nsel) {
    HIP_TRY(hipMemcpyAsync(s.h_row_off, s.d_row_off, (nsel + 1) * 8, hipMemcpyDeviceToHost, s.stream), "D2H(rows)");
    HIP_TRY(hipMemcpyAsync(s.h_sel_idx, s.d_sel_idx, nsel * 4, hipMemcpyDeviceToHost, s.stream), "D2H(rows)");
    HIP_TRY(hipMemcpyAsync(s.h_out, s.d_out, c.out_total, hipMemcpyDeviceToHost, s.stream), "D2H(rows)");
    out->out_len = c.out_total;
  }
  HIP_TRY(hipStreamSynchronize(s.stream), "sync(rows)");
  return WARPJQ_OK;
}

double warpjq_amd_ext_value(const warpjq_amd_agg_ext *e) {
  // exact 128-bit integer part, rounded once, plus the compensated float part
  __int128 i = (__int128)(((unsigned __int128)e->isum_hi << 64) | e->isum_lo);
  double hi = (double)i;
  __int128 rem = i - (__int128)hi;
  double s = hi + e->fsum;
  double bp = s - hi;
  double err = (hi - (s - bp)) + (e->fsum - bp);
  return s + (err + (double)rem + e->fcomp);
}

warpjq_status warpjq_amd_wait_ext(warpjq_ctx *ctx, uint32_t slot, warpjq_amd_agg_ext *agg_ext,
                                  const warpjq_amd_agg_ext **group_ext) {
  if (!ctx || slot >= ctx->n_slots) return WARPJQ_ERR_INVALID_ARG;
  Slot &s = ctx->slots[slot];
  if (agg_ext) {
    if (ctx->dprog.output_kind == WARPJQ_OUT_AGG && !ctx->grouped && s.n_bytes) *agg_ext = s.h_aggout->ext;
    else memset(agg_ext, 0, sizeof *agg_ext);
  }
  if (group_ext) *group_ext = ctx->grouped ? s.h_gext : nullptr;
  return WARPJQ_OK;
}

warpjq_status warpjq_amd_slot_timing(warpjq_ctx *ctx, uint32_t slot, float *h2d_ms, float *kernel_ms,
                                     float *main_ms) {
  if (!ctx || slot >= ctx->n_slots) return WARPJQ_ERR_INVALID_ARG;
  Slot &s = ctx->slots[slot];
  float a = 0, b = 0, m = 0;
  if (hipEventElapsedTime(&a, s.ev0, s.ev1) != hipSuccess) a = 0;
  if (hipEventElapsedTime(&b, s.ev1, s.ev2) != hipSuccess) b = 0;
  if (s.n_bytes == 0 || hipEventElapsedTime(&m, s.evm0, s.evm1) != hipSuccess) m = 0;
  if (h2d_ms) *h2d_ms = a;
  if (kernel_ms) *kernel_ms = b;
  if (main_ms) *main_ms = m;
  return WARPJQ_OK;
}

uint32_t warpjq_amd_kernel_width(const warpjq_ctx *ctx) { return ctx ? ctx->host.ns : 0; }

warpjq_status warpjq_amd_set_claim(warpjq_ctx *ctx, uint32_t claim_bytes) {
  if (!ctx || claim_bytes < CLAIM_MIN || claim_bytes > CLAIM_MAX || claim_bytes % STEP_BYTES) return WARPJQ_ERR_INVALID_ARG;
  if (claim_bytes + TAIL_MAX > ctx->arena) return WARPJQ_ERR_INVALID_ARG;
  ctx->claim = claim_bytes;
  ctx->prog_gen++;
  return WARPJQ_OK;
}

warpjq_status warpjq_amd_set_arena(warpjq_ctx *ctx, uint32_t arena_bytes) {
  if (!ctx || arena_bytes % 256 || arena_bytes > ARENA_MAX || arena_bytes < ctx->claim + TAIL_MAX ||
      !lds_fits(ctx, arena_bytes, ctx->tab_bytes))
    return WARPJQ_ERR_INVALID_ARG;
  ctx->arena = arena_bytes;
  ctx->grid = main_grid(ctx);
  ctx->prog_gen++;
  return WARPJQ_OK;
}

uint32_t warpjq_amd_grid(const warpjq_ctx *ctx) { return ctx ? ctx->grid : 0; }

warpjq_status warpjq_amd_autotune(warpjq_ctx *ctx, const uint8_t *d_bytes, uint64_t n_bytes, uint32_t *claim_out,
                                  uint32_t *arena_out, float *best_ms_out, char *err, size_t err_len) {
  if (!ctx || (!d_bytes && n_bytes)) return WARPJQ_ERR_INVALID_ARG;
  // claim and arena pairs around the default; the arena sets the workgroups
  // per CU, the claim the bytes per ticket and the fill granularity
  static const uint32_t cand[][2] = {{8, 18}, {8, 16}, {8, 20}, {8, 24}, {7, 16}, {6, 14}, {6, 16},
                                     {4, 12}, {4, 8},  {8, 32}, {5, 12}, {8, 12}};
  const uint64_t sample = std::min<uint64_t>(n_bytes, ctx->rows ? ctx->chunk_cap : (1ull << 30));
  const uint32_t claim0 = ctx->claim, arena0 = ctx->arena;
  float best = 1e30f;
  uint32_t bc = claim0, ba = arena0;
  for (const auto &c : cand) {
    const uint32_t cl = c[0] << 10, ar = c[1] << 10;
    if (cl + TAIL_MAX > ar || ar > ARENA_MAX || !lds_fits(ctx, ar, ctx->tab_bytes)) continue;
    ctx->arena = ar;  // both set together: the order of the two setters would matter
    ctx->claim = cl;
    ctx->grid = main_grid(ctx);
    ctx->prog_gen++;
    float t = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
      warpjq_status st = warpjq_amd_submit_device(ctx, 0, d_bytes, sample, err, err_len);
      if (st == WARPJQ_OK) {
        warpjq_chunk_result r;
        st = warpjq_wait(ctx, 0, &r, err, err_len);
      }
      if (st != WARPJQ_OK) {
        ctx->claim = claim0;
        ctx->arena = arena0;
        ctx->grid = main_grid(ctx);
        return st;
      }
      float m = 0.f;
      warpjq_amd_slot_timing(ctx, 0, nullptr, nullptr, &m);
      if (rep && m < t) t = m;  // the first run warms caches and code
    }
    if (t < best) {
      best = t;
      bc = cl;
      ba = ar;
    }
  }
  ctx->claim = bc;
  ctx->arena = ba;
  ctx->grid = main_grid(ctx);
  ctx->prog_gen++;
  if (ctx->grouped) {
    // the group table's slots in use: the smallest tables that still hold
    // twice the sample's groups keep the probes in fewer cache lines; the
    // fastest of those is kept (a chunk with more groups than it holds is
    // split by the host, so the choice never changes a result)
    uint32_t ng = 0;
    {
      warpjq_status st = warpjq_amd_submit_device(ctx, 0, d_bytes, sample, err, err_len);
      warpjq_chunk_result r;
      if (st == WARPJQ_OK) st = warpjq_wait(ctx, 0, &r, err, err_len);
      if (st != WARPJQ_OK) return st;
      ng = r.n_groups;
    }
    uint32_t bg = ctx->group_cap;
    float gbest = 1e30f;
    for (uint32_t g = 1u << 14; g <= ctx->group_cap; g <<= 2) {
      if (g < 2ull * ng) continue;
      ctx->group_active = g;
      ctx->prog_gen++;
      float t = 1e30f;
      for (int rep = 0; rep < 3; rep++) {
        warpjq_status st = warpjq_amd_submit_device(ctx, 0, d_bytes, sample, err, err_len);
        warpjq_chunk_result r;
        if (st == WARPJQ_OK) st = warpjq_wait(ctx, 0, &r, err, err_len);
        if (st != WARPJQ_OK) {
          ctx->group_active = ctx->group_cap;
          return st;
        }
        float m = 0.f;
        warpjq_amd_slot_timing(ctx, 0, nullptr, nullptr, &m);
        if (rep && m < t && !r.group_overflow) t = m;
      }
      if (t < gbest) {
        gbest = t;
        bg = g;
      }
    }
    ctx->group_active = bg;
    ctx->prog_gen++;
    if (gbest < best) best = gbest;
  }
  if (claim_out) *claim_out = bc;
  if (arena_out) *arena_out = ba;
  if (best_ms_out) *best_ms_out = best;
  return WARPJQ_OK;
}

uint32_t warpjq_amd_group_table(const warpjq_ctx *ctx) { return ctx ? ctx->group_active : 0; }

warpjq_status warpjq_amd_set_group_table(warpjq_ctx *ctx, uint32_t slots) {
  if (!ctx || !ctx->grouped || slots < 4096 || (slots & (slots - 1)) || slots > ctx->group_cap)
    return WARPJQ_ERR_INVALID_ARG;
  for (uint32_t i = 0; i < ctx->n_slots; i++)
    if (hipStreamSynchronize(ctx->slots[i].stream) != hipSuccess) return WARPJQ_ERR_CUDA;
  ctx->group_active = slots;
  ctx->prog_gen++;
  return WARPJQ_OK;
}

warpjq_status warpjq_amd_set_graphs(warpjq_ctx *ctx, uint32_t per_slot) {
  if (!ctx || per_slot > 65536) return WARPJQ_ERR_INVALID_ARG;
  ctx->graph_cap = per_slot;
  for (uint32_t i = 0; i < ctx->n_slots; i++) {
    Slot &s = ctx->slots[i];
    while (s.graphs.size() > per_slot) {
      (void)hipGraphExecDestroy(s.graphs.front().exec);
      s.graphs.erase(s.graphs.begin());
    }
  }
  return WARPJQ_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// the no-allocation entry point
//
// warpjq_amd_aggregate_device_noalloc runs an ungrouped aggregate over bytes
// already in device memory with the same kernels as warpjq_amd_submit_device
// + warpjq_wait, but every buffer those kernels need beyond the input and the
// result (counters, program tables, the span, per-wave partials, per-claim
// line counts, the wave-wide phase's queue, declined-line records) lives in
// `scratch`, a fixed array of NOALLOC_SCRATCH bytes that each GPU holds for
// the life of the process. No device memory is allocated or freed per call.
// Every call on one GPU uses that GPU's copy, and what the previous call left
// there is cleared (by kernels) before use; calls that overlap on one GPU race.
// ---------------------------------------------------------------------------

#define NOALLOC_SCRATCH (8u << 20)
#define NOALLOC_FB 4096u
#define NOALLOC_DEVICES 64

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH];

namespace {

std::mutex g_na_mu;
uint8_t *g_na_ptr[NOALLOC_DEVICES];  // the device address of `scratch`, per device

struct NaLayout {
  size_t ctr, aggout, tab, span, parts, fb_idx, fb_off, fb_len, fb_tile, tile, longq, end;
};

NaLayout na_layout(uint32_t ntiles, uint32_t nparts) {
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  NaLayout L;
  L.ctr = 0;
  L.aggout = up(L.ctr + sizeof(DevCtr));
  L.tab = up(L.aggout + sizeof(AggOut));
  L.span = up(L.tab + TAB_LDS_MAX);
  L.parts = up(L.span + sizeof(Span));
  L.fb_idx = up(L.parts + (size_t)nparts * sizeof(AggPart));
  L.fb_off = up(L.fb_idx + NOALLOC_FB * 4u);
  L.fb_len = up(L.fb_off + NOALLOC_FB * 4u);
  L.fb_tile = up(L.fb_len + NOALLOC_FB * 4u);
  L.tile = up(L.fb_tile + NOALLOC_FB * 4u);
  L.longq = up(L.tile + (size_t)ntiles * 8u);
  L.end = up(L.longq + (size_t)ntiles * 12u);
  return L;
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

// The device address of this device's `scratch`, looked up once per device
// (with that device current) and kept: host code never passes the array's
// own name to a kernel.
static warpjq_status noalloc_scratch(uint8_t **p, char *err, size_t err_len) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= NOALLOC_DEVICES) {
    set_err(err, err_len, "aggregate_device_noalloc", "device index out of range");
    return WARPJQ_ERR_INVALID_ARG;
  }
  std::lock_guard<std::mutex> g(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *q = nullptr;
    HIP_TRY(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_na_ptr[dev] = (uint8_t *)q;
  }
  *p = g_na_ptr[dev];
  return WARPJQ_OK;
}

extern "C" {

uint64_t warpjq_amd_noalloc_max_bytes(void) {
  // the largest input whose per-claim arrays fit next to the fixed part
  const NaLayout L0 = na_layout(0, 256u * MAX_WG_PER_CU);
  const uint64_t per_tile = 20;
  const uint64_t tiles = (NOALLOC_SCRATCH - L0.end - 512) / per_tile;
  return tiles * CLAIM_DEFAULT - 16;
}

warpjq_status warpjq_amd_aggregate_device_noalloc(const warpjq_program *prog, const uint8_t *d_bytes, uint64_t n_bytes,
                                                  warpjq_chunk_result *out, warpjq_amd_agg_ext *ext, char *err,
                                                  size_t err_len) {
  if (!prog || !out || (!d_bytes && n_bytes)) return WARPJQ_ERR_INVALID_ARG;
  memset(out, 0, sizeof *out);
  out->first_invalid = 0xFFFFFFFFu;
  if (prog->output_kind != WARPJQ_OUT_AGG || prog->group_path != 0xFFFFFFFFu) {
    set_err(err, err_len, "aggregate_device_noalloc", "only ungrouped aggregate queries (count, sum, min, max, avg)");
    return WARPJQ_ERR_INVALID_ARG;
  }
  if (n_bytes > warpjq_amd_noalloc_max_bytes()) {
    set_err(err, err_len, "aggregate_device_noalloc", "input larger than the fixed device scratch allows");
    return WARPJQ_ERR_INVALID_ARG;
  }
  DProgHost host;
  std::string why;
  if (!build_dprog(*prog, host, why)) {
    set_err(err, err_len, "query is not supported by the kernels", why.c_str());
    return WARPJQ_ERR_INVALID_ARG;
  }
  // the program tables, laid out as put_program lays them
  std::vector<uint8_t> buf;
  size_t o_steps = put_table(buf, prog->steps, prog->n_steps);
  size_t o_cmps = put_table(buf, prog->cmps, prog->n_cmps);
  size_t o_cond = put_table(buf, host.cond.data(), host.cond.size());
  size_t o_blob = put_table(buf, prog->blob, prog->n_blob);
  size_t o_spath = put_table(buf, host.spath.data(), host.spath.size());
  size_t o_sop = put_table(buf, prog->slot_of_path, prog->n_paths);
  size_t o_fs = put_table(buf, host.field_slot.data(), host.field_slot.size());
  size_t o_po = put_table(buf, prog->prefix_off, prog->n_fields);
  size_t o_pl = put_table(buf, prog->prefix_len, prog->n_fields);
  buf.resize((buf.size() + 15) & ~(size_t)15);
  if (buf.size() > TAB_LDS_MAX) {
    set_err(err, err_len, "query is too large for the kernels", "its tables exceed the 16 KiB staged in LDS");
    return WARPJQ_ERR_INVALID_ARG;
  }
  uint8_t *S = nullptr;
  if (warpjq_status st = noalloc_scratch(&S, err, err_len)) return st;
  if (!g_num_cus) {
    hipDeviceProp_t p;
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) g_num_cus = (uint32_t)p.multiProcessorCount;
    if (!g_num_cus) g_num_cus = 256;
  }
  const uint32_t claim = CLAIM_DEFAULT, arena = ARENA_DEFAULT;
  const uint32_t tab_bytes = (uint32_t)buf.size();
  const uint32_t lds = main_lds_bytes(arena, tab_bytes);
  int occ;
  switch (host.ns) {
    case 1: occ = occupancy_main_1(false, false, lds); break;
    case 2: occ = occupancy_main_2(false, false, lds); break;
    case 4: occ = occupancy_main_4(false, false, lds); break;
    case 8: occ = occupancy_main_8(false, false, lds); break;
    default: occ = occupancy_main_16(false, false, lds); break;
  }
  occ = std::max(1, std::min(occ, (int)MAX_WG_PER_CU));
  const uint32_t mis = (uint32_t)((uintptr_t)d_bytes & 15u);
  const uint64_t vend = n_bytes ? mis + n_bytes : 0;
  const uint32_t ntiles = (uint32_t)((vend + claim - 1) / claim);
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(g_num_cus * (uint32_t)occ, ntiles));
  const NaLayout L = na_layout(ntiles, g_num_cus * (uint32_t)occ);
  if (L.end > NOALLOC_SCRATCH) {
    set_err(err, err_len, "aggregate_device_noalloc", "input larger than the fixed device scratch allows");
    return WARPJQ_ERR_INVALID_ARG;
  }
  hipStream_t st = 0;
  Span sp{d_bytes ? d_bytes - mis : nullptr, 0, mis, (uint32_t)vend, 0, nullptr, 0};
  HIP_TRY(hipMemcpyAsync(S + L.tab, buf.data(), buf.size(), hipMemcpyHostToDevice, st), "H2D(program)");
  HIP_TRY(hipMemcpyAsync(S + L.span, &sp, sizeof sp, hipMemcpyHostToDevice, st), "H2D(span)");
  DevCtr *ctr = (DevCtr *)(S + L.ctr);
  unsigned long long *tile = (unsigned long long *)(S + L.tile);
  hipLaunchKernelGGL(k_reset, dim3(std::max<uint32_t>(1, std::min<uint32_t>(1024, (ntiles + 255) / 256))), dim3(256), 0,
                     st, ctr, tile, ntiles);
  // the wave-wide phase's queue starts zero (its readers clear what they take)
  hipLaunchKernelGGL(k_fill_u32, dim3(std::max<uint32_t>(1, std::min<uint32_t>(1024, (3u * ntiles + 255) / 256))),
                     dim3(256), 0, st, (uint32_t *)(S + L.longq), (uint64_t)ntiles * 3u, 0u);
  HIP_TRY(hipGetLastError(), "k_reset");
  if (n_bytes) {
    const uint8_t *tab = S + L.tab;
    DProg d = host.d;
    d.steps = (const warpjq_step *)(tab + o_steps);
    d.cmps = (const warpjq_cmp *)(tab + o_cmps);
    d.cond = (const warpjq_cond_op *)(tab + o_cond);
    d.blob = tab + o_blob;
    d.spath = (const SlotPath *)(tab + o_spath);
    d.slot_of_path = (const uint32_t *)(tab + o_sop);
    d.field_slot = (const uint32_t *)(tab + o_fs);
    d.prefix_off = (const uint32_t *)(tab + o_po);
    d.prefix_len = (const uint32_t *)(tab + o_pl);
    KArgs a{};
    a.spans = (const Span *)(S + L.span);
    a.n_spans = 1;
    a.ntiles = ntiles;
    a.claim = claim;
    a.arena = arena;
    a.P = d;
    a.tab = tab;
    a.tab_bytes = tab_bytes;
    a.ctr = ctr;
    a.tile_state = tile;
    a.parts = (AggPart *)(S + L.parts);
    a.max_lines = 0;
    a.fb_idx = (uint32_t *)(S + L.fb_idx);
    a.fb_off = (uint32_t *)(S + L.fb_off);
    a.fb_len = (uint32_t *)(S + L.fb_len);
    a.fb_tile = (uint32_t *)(S + L.fb_tile);
    a.fb_cap = NOALLOC_FB;
    a.long_q = (uint32_t *)(S + L.longq);
    a.long_cap = ntiles;
    a.deep_cap = 0;
    switch (host.ns) {
      case 1: launch_main_1(false, false, grid, st, a); break;
      case 2: launch_main_2(false, false, grid, st, a); break;
      case 4: launch_main_4(false, false, grid, st, a); break;
      case 8: launch_main_8(false, false, grid, st, a); break;
      default: launch_main_16(false, false, grid, st, a); break;
    }
    HIP_TRY(hipGetLastError(), "k_main");
    hipLaunchKernelGGL(k_agg_final, dim3(1), dim3(256), 0, st, a.parts, grid, ctr, (AggOut *)(S + L.aggout));
    HIP_TRY(hipGetLastError(), "k_agg_final");
  }
  DevCtr c;
  HIP_TRY(hipMemcpyAsync(&c, ctr, sizeof c, hipMemcpyDeviceToHost, st), "D2H(counters)");
  HIP_TRY(hipStreamSynchronize(st), "stream sync");
  if (!n_bytes) return WARPJQ_OK;
  if (!c.overflow && (c.first_bad != ~0ull || c.first_te_bad != ~0ull || c.fb_count) && c.fb_count <= NOALLOC_FB) {
    hipLaunchKernelGGL(k_res_prefix, dim3(1), dim3(1024), 0, st, tile, ntiles);
    hipLaunchKernelGGL(k_res_records, dim3(std::max<uint32_t>(1, std::min<uint32_t>(1024, (c.fb_count + 255) / 256))),
                       dim3(256), 0, st, ctr, tile, (uint32_t *)(S + L.fb_idx), (const uint32_t *)(S + L.fb_tile),
                       NOALLOC_FB);
    HIP_TRY(hipGetLastError(), "k_res");
    HIP_TRY(hipMemcpyAsync(&c, ctr, sizeof c, hipMemcpyDeviceToHost, st), "D2H(counters)");
    HIP_TRY(hipStreamSynchronize(st), "stream sync");
  }
  AggOut ao;
  HIP_TRY(hipMemcpyAsync(&ao, S + L.aggout, sizeof ao, hipMemcpyDeviceToHost, st), "D2H(agg)");
  HIP_TRY(hipStreamSynchronize(st), "stream sync");
  out->n_lines = c.n_lines;
  out->n_blank = c.n_blank;
  out->n_invalid = c.n_invalid;
  out->n_type_error = c.n_type_error;
  out->n_fallback = c.fb_count;  // counted; their offsets stay on the device
  out->first_invalid = c.first_invalid;
  out->chunk_overflow = c.overflow || c.fb_count > NOALLOC_FB;
  if (out->chunk_overflow) {
    out->n_lines = 0;
    return WARPJQ_OK;
  }
  out->agg = ao.agg;
  out->agg.sum = warpjq_amd_ext_value(&ao.ext);
  out->agg.saw_non_numeric = c.saw_non_numeric;
  if (ext) *ext = ao.ext;
  return WARPJQ_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// resident newline indices (the served form's datasets)
// ---------------------------------------------------------------------------

namespace wjqd {

namespace {
struct NlIndex {
  uint64_t n;
  const uint8_t *idx;
};
std::mutex g_nl_mu;
std::map<uintptr_t, NlIndex> g_nl;  // by dataset base address
std::atomic<uint32_t> g_nl_n{0};    // entries in g_nl
}  // namespace

uint64_t nlindex_bytes(uint64_t n) { return (n + 63) / 64; }

bool nlindex_build(const uint8_t *d, uint64_t n, uint8_t *idx, std::string &err) {
  if (!n) return true;
  const uint64_t nb = nlindex_bytes(n);
  const uint32_t grid = (uint32_t)std::min<uint64_t>(4096, (nb + 255) / 256);
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    err = "hipStreamCreate (index)";
    return false;
  }
  hipLaunchKernelGGL(k_nl_index, dim3(grid), dim3(256), 0, st, d, n, idx);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipStreamDestroy(st);
  if (e != hipSuccess) {
    err = std::string("building a newline index: ") + hipGetErrorString(e);
    return false;
  }
  return true;
}

void nlindex_register(const uint8_t *d, uint64_t n, const uint8_t *idx) {
  std::lock_guard<std::mutex> g(g_nl_mu);
  g_nl[(uintptr_t)d] = NlIndex{n, idx};
  g_nl_n.store((uint32_t)g_nl.size());
}

void nlindex_unregister(const uint8_t *d) {
  std::lock_guard<std::mutex> g(g_nl_mu);
  g_nl.erase((uintptr_t)d);
  g_nl_n.store((uint32_t)g_nl.size());
}

// The registered dataset holding d[0..n) (caller holds g_nl_mu).
static void nlindex_find_locked(const uint8_t *d, uint64_t n, const uint8_t **idx, uint64_t *off) {
  *idx = nullptr;
  *off = 0;
  auto it = g_nl.upper_bound((uintptr_t)d);
  if (it == g_nl.begin()) return;
  --it;
  const uintptr_t base = it->first;
  if ((uintptr_t)d + n <= base + it->second.n) {
    *idx = it->second.idx;
    *off = (uint64_t)((uintptr_t)d - base);
  }
}

void nlindex_find(const uint8_t *d, uint64_t n, const uint8_t **idx, uint64_t *off) {
  std::lock_guard<std::mutex> g(g_nl_mu);
  nlindex_find_locked(d, n, idx, off);
}

void nlindex_attach(Span *sp, uint32_t n) {
  if (g_nl_n.load() == 0) return;
  std::lock_guard<std::mutex> g(g_nl_mu);
  for (uint32_t i = 0; i < n; i++) {
    const uint8_t *d = sp[i].base + (sp[i].start - sp[i].vbase);  // the span's first byte
    nlindex_find_locked(d, sp[i].vend - sp[i].start, &sp[i].nlidx, &sp[i].ioff);
  }
}

}  // namespace wjqd
