// warpjq's device ABI (include/warpjq_kernels.h) implemented for MI355X
// (gfx950, wave64) in HIP.
//
// Per submitted chunk, on the slot's stream:
//   H2D (pinned staging -> HBM)
//   k_reset        counters, tile states, ticket
//   k_main<NS>     ONE pass over the chunk. Persistent single-wave workgroups
//                  claim 16 KiB tiles in order from a ticket. Each tile is
//                  read once with coalesced 16-byte loads (64 lanes x 16 B =
//                  1 KiB per step); per-lane newline masks and a wave prefix
//                  sum turn it into line-start offsets queued in LDS. The
//                  tile publishes its line count and takes its first line's
//                  global index by decoupled look-back. Then every lane
//                  evaluates one queued line at a time (jscan.h: validate +
//                  capture + filter), so lines from consecutive tiles fill all
//                  64 lanes. Aggregates stay in registers; group_by goes
//                  through a per-wave LDS table into an HBM hash table.
//   k_agg_final    ungrouped partials -> the chunk's warpjq_agg_partial
//   k_group_compact occupied HBM group slots -> warpjq_group[], clearing them
//   k_row_*        (row outputs) scan of per-line row lengths -> ascending
//                  selected line indices and row offsets
//   k_emit<NS>     re-evaluates each selected line and writes its row
//   D2H of the fixed-size results (counters, aggregate)
// warpjq_wait synchronises the stream and copies the variable-size results.
//
// Reference ABI semantics followed: kernels.cu:2233-2685 (probe, abi_check,
// ctx_create sizing, submit, wait), kernels.h:155-230 (result meaning).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/warpjq_amd.h"
#include "../../include/warpjq_kernels.h"
#include "dprog.h"
#include "jscan.h"

using namespace wjqd;

// ---------------------------------------------------------------------------
// tunables
// ---------------------------------------------------------------------------

#define CLAIM_MIN 1024u              // smallest claim a wave takes (one scan step)
#define CLAIM_MAX 65536u             // largest claim (the queue index is 16-bit safe)
#define STEP_BYTES 1024u             // 64 lanes x 16 B
#define STAGE_BYTES 16384u           // LDS staging of one round's line span
#define QCAP 512u                    // per-wave LDS queue of line starts
#define QMASK (QCAP - 1u)
#define LGROUP 64u                   // per-wave LDS group table entries
#define GROUP_MAX_PROBE 128u
#define MIN_LINE_BYTES 24u           // kernels.cu:1984-1988 sizing rule
#define DATA_PAD 256u
#define MAX_CHUNK (((uint64_t)1 << 32) - ((uint64_t)1 << 20))
#define ROWSCAN_BLOCK 1024u          // lines per block in the row-offset scan
#define EMPTY_KEY (~0ull)

// ---------------------------------------------------------------------------
// device-side structures
// ---------------------------------------------------------------------------

struct DevCtr {
  unsigned long long n_lines, n_blank, n_invalid, n_type_error, n_fallback, n_sel, out_total;
  unsigned int first_invalid, overflow, saw_non_numeric, group_overflow;
  unsigned int fb_count, n_groups, ticket, tiles_done;
};

// Per-wave ungrouped aggregate partial.
struct AggPart {
  unsigned long long count, numeric;
  unsigned long long isum_lo, isum_hi;  // exact sum of integer tokens (i128)
  double fsum, fcomp;                   // compensated sum of the other tokens
  double mn, mx;
};

struct GroupTab {
  unsigned long long *keys;  // packed off:40 | len:19 | esc:1 | kind:4
  unsigned int *hash;
  unsigned long long *count, *numeric, *isum_lo, *isum_hi, *mnb, *mxb;
  double *fsum, *fcomp;
  unsigned int cap;  // power of two
};

struct KArgs {
  const uint8_t *d;
  uint32_t n;
  uint32_t ntiles;
  uint32_t claim;  // bytes per claimed tile (multiple of STEP_BYTES)
  DProg P;
  DevCtr *ctr;
  unsigned long long *tile_state;
  AggPart *parts;
  GroupTab gt;
  // row outputs
  uint32_t *row_len;    // per line (0 = not selected)
  uint32_t *sel_start;  // per line, written for selected lines
  uint64_t max_lines;
  // fallback list
  uint32_t *fb_idx, *fb_off, *fb_len;
  uint32_t fb_cap;
};

// ---------------------------------------------------------------------------
// reading the chunk
// ---------------------------------------------------------------------------

// Caches one 16-byte block per lane; positions at or past n read as '\n'.
// Blocks inside the round's staged span [slo, shi) come from LDS, the rest
// from global memory (L2).
struct DevReader {
  const uint8_t *d;
  uint32_t n;
  uint32_t cb;
  uint32_t v0, v1, v2, v3;
  const uint4 *stg;
  uint32_t slo, shi;

  __device__ DevReader(const uint8_t *d_, uint32_t n_)
      : d(d_), n(n_), cb(0xFFFFFFFFu), v0(0), v1(0), v2(0), v3(0), stg(nullptr), slo(0), shi(0) {}

  __device__ __forceinline__ uint32_t dword(uint32_t k) {
    uint32_t b = k >> 2;
    if (b != cb) {
      uint32_t off = (b << 4) - slo;
      uint4 v = off < shi - slo ? stg[off >> 4] : *((const uint4 *)d + b);
      v0 = v.x;
      v1 = v.y;
      v2 = v.z;
      v3 = v.w;
      cb = b;
    }
    uint32_t s = k & 3;
    uint32_t x = s == 0 ? v0 : s == 1 ? v1 : s == 2 ? v2 : v3;
    uint32_t p = k << 2;
    if (p + 3 >= n) {
#pragma unroll
      for (uint32_t q = 0; q < 4; q++)
        if (p + q >= n) x = (x & ~(0xFFu << (8 * q))) | (WJQ_END << (8 * q));
    }
    return x;
  }
  __device__ __forceinline__ uint32_t byte(uint32_t i) {
    if (i >= n) return WJQ_END;
    return (dword(i >> 2) >> ((i & 3) * 8)) & 0xFFu;
  }
};

// Byte sink writing a row into HBM, bounded by the output capacity.
struct DevSink {
  uint8_t *out;
  uint64_t pos;
  __device__ __forceinline__ void put(uint8_t c) { out[pos++] = c; }
  __device__ __forceinline__ void span(DevReader &rd, uint32_t off, uint32_t len) {
    for (uint32_t k = 0; k < len; k++) out[pos + k] = (uint8_t)rd.byte(off + k);
    pos += len;
  }
};

// ---------------------------------------------------------------------------
// wave helpers (wave64)
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t &total) {
  uint32_t lane = lane_id();
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  total = __shfl(x, 63, 64);
  return x - v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ uint64_t ordered_bits(double v) {
  uint64_t b = f64_to_bits(v);
  return (b >> 63) ? ~b : (b ^ (1ull << 63));
}
__device__ __forceinline__ double from_ordered(uint64_t o) {
  uint64_t b = (o >> 63) ? (o ^ (1ull << 63)) : ~o;
  return bits_to_f64(b);
}

__device__ __forceinline__ unsigned long long ld_acq(unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

#define FLAG_A (1ull << 62)
#define FLAG_P (2ull << 62)
#define VAL_MASK ((1ull << 62) - 1)

// Decoupled look-back: publishes this tile's count and returns the sum of
// counts of all earlier tiles. Wave-uniform.
__device__ uint64_t tile_lookback(unsigned long long *state, uint32_t t, uint64_t agg) {
  uint32_t lane = lane_id();
  if (t == 0) {
    if (lane == 0) st_rel(&state[0], FLAG_P | agg);
    return 0;
  }
  if (lane == 0) st_rel(&state[t], FLAG_A | agg);
  uint64_t base = 0;
  int64_t j = (int64_t)t - 1;
  for (;;) {
    int64_t idx = j - (int64_t)lane;
    unsigned long long s = idx >= 0 ? ld_acq(&state[idx]) : FLAG_P;
    uint64_t flag = s >> 62;
    uint64_t pm = __ballot(flag == 2);
    uint64_t zm = __ballot(flag == 0);
    uint64_t rel = pm ? ((pm & (~pm + 1)) << 1) - 1 : ~0ull;  // lanes up to the first P
    if (zm & rel) {
      __builtin_amdgcn_s_sleep(2);
      continue;
    }
    uint64_t v = ((rel >> lane) & 1) ? (s & VAL_MASK) : 0;
    base += wave_sum(v);
    if (pm) break;
    j -= 64;
  }
  if (lane == 0) st_rel(&state[t], FLAG_P | (base + agg));
  return base;
}

// ---------------------------------------------------------------------------
// group table
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned long long pack_key(const GKey &k) {
  return ((unsigned long long)k.off << 24) | ((unsigned long long)(k.len & 0x7FFFF) << 5) |
         ((unsigned long long)(k.esc & 1) << 4) | (k.kind & 0xF);
}
__device__ __forceinline__ GKey unpack_key(unsigned long long p) {
  GKey k;
  k.off = (uint32_t)(p >> 24);
  k.len = (uint32_t)((p >> 5) & 0x7FFFF);
  k.esc = (uint32_t)((p >> 4) & 1);
  k.kind = (uint32_t)(p & 0xF);
  return k;
}

// Adds one value (or a partial) into accumulators at *lo/*hi etc. Exact for
// integers (128-bit two's complement), compensated for floats: the value the
// atomic add stored is recomputed from the returned old sum, and the
// rounding error of that addition is accumulated separately.
__device__ __forceinline__ void acc_i128(unsigned long long *lo, unsigned long long *hi, uint64_t vlo, uint64_t vhi) {
  if (vlo) {
    unsigned long long old = atomicAdd(lo, (unsigned long long)vlo);
    if (old + vlo < old) vhi += 1;
  }
  if (vhi) atomicAdd(hi, (unsigned long long)vhi);
}
__device__ __forceinline__ void acc_f64(double *s, double *c, double v, double comp) {
  if (v != 0.0) {
    double old = atomicAdd(s, v);
    double t = old + v;
    double bp = t - old;
    double e = (old - (t - bp)) + (v - bp);
    comp += e;
  }
  if (comp != 0.0) atomicAdd(c, comp);
}

struct GVal {
  uint64_t count, numeric, ilo, ihi, mnb, mxb;
  double fsum, fcomp;
};

__device__ __forceinline__ void group_apply(const GroupTab &g, uint32_t pos, const GVal &v) {
  atomicAdd(&g.count[pos], (unsigned long long)v.count);
  if (v.numeric) {
    atomicAdd(&g.numeric[pos], (unsigned long long)v.numeric);
    acc_i128(&g.isum_lo[pos], &g.isum_hi[pos], v.ilo, v.ihi);
    acc_f64(&g.fsum[pos], &g.fcomp[pos], v.fsum, v.fcomp);
    atomicMin(&g.mnb[pos], (unsigned long long)v.mnb);
    atomicMax(&g.mxb[pos], (unsigned long long)v.mxb);
  }
}

// Finds or claims the slot of `k` in the HBM table; ~0u on overflow. The key
// is published whole by one CAS, so a reader never sees a half-written key.
__device__ uint32_t group_find(DevReader &rd, const GroupTab &g, const GKey &k, uint32_t h) {
  unsigned long long pk = pack_key(k);
  uint32_t pos = h & (g.cap - 1);
  for (uint32_t probe = 0; probe < GROUP_MAX_PROBE; probe++) {
    unsigned long long cur = __hip_atomic_load(&g.keys[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == EMPTY_KEY) {
      unsigned long long old = atomicCAS(&g.keys[pos], EMPTY_KEY, pk);
      if (old == EMPTY_KEY) {
        __hip_atomic_store(&g.hash[pos], h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return pos;
      }
      cur = old;
    }
    unsigned int hv = __hip_atomic_load(&g.hash[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (hv == 0 || hv == h) {
      if (cur == pk || group_eq(rd, k, unpack_key(cur))) return pos;
    }
    pos = (pos + 1) & (g.cap - 1);
  }
  return ~0u;
}

// Per-wave LDS pre-aggregation table.
struct LGroup {
  unsigned long long key[LGROUP];
  unsigned int hash[LGROUP];
  unsigned long long count[LGROUP], numeric[LGROUP], ilo[LGROUP], ihi[LGROUP], mnb[LGROUP], mxb[LGROUP];
  double fsum[LGROUP], fcomp[LGROUP];
};

__device__ __forceinline__ void lds_acc_f64(double *s, double *c, double v) {
  double old = atomicAdd(s, v);
  double t = old + v;
  double bp = t - old;
  double e = (old - (t - bp)) + (v - bp);
  if (e != 0.0) atomicAdd(c, e);
}

// Inserts one line's value; falls through to HBM when the LDS table is full.
__device__ void group_line(DevReader &rd, const GroupTab &g, LGroup &L, const GKey &k, uint32_t h, bool has_num,
                           bool is_int, int64_t iv, double v, unsigned int *ovf) {
  unsigned long long pk = pack_key(k);
  uint32_t pos = h & (LGROUP - 1);
  int found = -1;
  for (uint32_t probe = 0; probe < 8; probe++) {
    unsigned long long cur = L.key[pos];
    if (cur == EMPTY_KEY) {
      unsigned long long old = atomicCAS(&L.key[pos], EMPTY_KEY, pk);
      if (old == EMPTY_KEY) {
        L.hash[pos] = h;
        found = (int)pos;
        break;
      }
      cur = old;
    }
    unsigned int hv = L.hash[pos];
    if ((hv == 0 || hv == h) && (cur == pk || group_eq(rd, k, unpack_key(cur)))) {
      found = (int)pos;
      break;
    }
    pos = (pos + 1) & (LGROUP - 1);
  }
  uint64_t vlo = 0, vhi = 0;
  double fv = 0.0;
  if (has_num) {
    if (is_int) {
      vlo = (uint64_t)iv;
      vhi = iv < 0 ? ~0ull : 0ull;
    } else {
      fv = v;
    }
  }
  if (found >= 0) {
    uint32_t p = (uint32_t)found;
    atomicAdd(&L.count[p], 1ull);
    if (has_num) {
      atomicAdd(&L.numeric[p], 1ull);
      if (is_int) {
        unsigned long long old = atomicAdd(&L.ilo[p], (unsigned long long)vlo);
        uint64_t hadd = vhi + ((old + vlo < old) ? 1 : 0);
        if (hadd) atomicAdd(&L.ihi[p], (unsigned long long)hadd);
      } else {
        lds_acc_f64(&L.fsum[p], &L.fcomp[p], fv);
      }
      uint64_t ob = ordered_bits(v);
      atomicMin(&L.mnb[p], (unsigned long long)ob);
      atomicMax(&L.mxb[p], (unsigned long long)ob);
    }
    return;
  }
  uint32_t gp = group_find(rd, g, k, h);
  if (gp == ~0u) {
    atomicOr(ovf, 1u);
    return;
  }
  GVal gv;
  gv.count = 1;
  gv.numeric = has_num ? 1 : 0;
  gv.ilo = vlo;
  gv.ihi = vhi;
  gv.fsum = fv;
  gv.fcomp = 0.0;
  gv.mnb = has_num ? ordered_bits(v) : ~0ull;
  gv.mxb = has_num ? ordered_bits(v) : 0ull;
  group_apply(g, gp, gv);
}

// ---------------------------------------------------------------------------
// k_main
// ---------------------------------------------------------------------------

template <bool G>
struct LGroupStore {
  __device__ static LGroup *get() {
    __shared__ LGroup lg;
    return &lg;
  }
};
template <>
struct LGroupStore<false> {
  __device__ static LGroup *get() { return nullptr; }
};

template <int NS, bool GRP>
__global__ __launch_bounds__(64) void k_main(KArgs a) {
  __shared__ uint32_t qs[QCAP];
  __shared__ uint32_t qi[QCAP];
  __shared__ uint4 stage[STAGE_BYTES / 16];
  LGroup *lgp = LGroupStore<GRP>::get();

  const uint32_t lane = lane_id();
  const DProg &P = a.P;
  const bool grouped = GRP;
  const bool rows = P.output_kind != WARPJQ_OUT_AGG;

  if (GRP) {
    LGroup &lg = *lgp;
    for (uint32_t e = lane; e < LGROUP; e += 64) {
      lg.key[e] = EMPTY_KEY;
      lg.hash[e] = 0;
      lg.count[e] = lg.numeric[e] = lg.ilo[e] = lg.ihi[e] = 0;
      lg.fsum[e] = lg.fcomp[e] = 0.0;
      lg.mnb[e] = ~0ull;
      lg.mxb[e] = 0;
    }
    __syncthreads();
  }

  // per-lane accumulators
  uint32_t c_blank = 0, c_inv = 0, c_te = 0, c_fb = 0, c_sel = 0;
  uint64_t lines_total = 0;  // wave-uniform
  uint64_t a_cnt = 0, a_num = 0, a_ilo = 0, a_ihi = 0;
  double a_fs = 0.0, a_fc = 0.0, a_mn = INFINITY, a_mx = -INFINITY;
  bool saw_nn = false;
  unsigned int ovf_local = 0;   // chunk-level capacity exceeded
  unsigned int govf_local = 0;  // group table overflow

  uint32_t qhead = 0, qlen = 0;  // wave-uniform
  // a tile whose line starts did not all fit in the queue
  bool pend = false;
  uint32_t pend_t = 0, pend_step = 0;
  uint64_t pend_idx = 0;
  bool tiles_left = true;

  DevReader rd(a.d, a.n);
  rd.stg = stage;
  uint32_t claim_end = 0;

  for (;;) {
    // ---- refill: rounds never mix claims, so a round's lines are one
    // contiguous span of the chunk ----
    while (qlen == 0 && (pend || tiles_left)) {
      qhead = 0;
      uint32_t t, step0;
      uint64_t idx0 = 0;
      bool known_base;
      if (pend) {
        t = pend_t;
        step0 = pend_step;
        idx0 = pend_idx;
        known_base = true;
        pend = false;
      } else {
        uint32_t tk = 0;
        if (lane == 0) tk = atomicAdd(&a.ctr->ticket, 1u);
        t = __shfl(tk, 0, 64);
        if (t >= a.ntiles) {
          tiles_left = false;
          break;
        }
        step0 = 0;
        known_base = false;
      }
      const uint32_t t0 = t * a.claim;
      const uint32_t tq0 = qhead + qlen;  // first queue entry of this tile
      uint32_t local = 0;                 // line starts seen in this tile from step0
      uint32_t stored = 0;
      bool full = false;
      uint32_t carry;
      {
        uint32_t p0 = t0 + step0 * STEP_BYTES;
        uint32_t c = 1;
        if (p0 > 0) c = a.d[p0 - 1] == '\n' ? 1 : 0;
        carry = c;
      }
      for (uint32_t step = step0; step < a.claim / STEP_BYTES; step++) {
        uint32_t p = t0 + step * STEP_BYTES + lane * 16;
        if (t0 + step * STEP_BYTES >= a.n) break;
        uint32_t m = 0;
        if (p < a.n) {
          uint4 v = *((const uint4 *)(a.d + p));
          uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
          for (int q = 0; q < 4; q++) {
            uint32_t x = w[q] ^ 0x0A0A0A0Au;
            uint32_t z = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);  // exact zero bytes
            uint32_t bits = ((z >> 7) & 1) | ((z >> 14) & 2) | ((z >> 21) & 4) | ((z >> 28) & 8);
            m |= bits << (4 * q);
          }
          if (p + 16 > a.n) m &= (1u << (a.n - p)) - 1u;
        }
        uint32_t up = __shfl_up(m >> 15, 1, 64) & 1u;
        uint32_t prev = lane == 0 ? carry : up;
        uint32_t st = ((m << 1) | prev) & 0xFFFFu;
        if (p >= a.n) st = 0;
        else if (p + 16 > a.n) st &= (1u << (a.n - p)) - 1u;
        carry = __shfl(m >> 15, 63, 64) & 1u;
        uint32_t cnt = __popc(st);
        uint32_t tot;
        uint32_t ex = wave_excl_scan(cnt, tot);
        if (!full && qlen + stored + tot > QCAP) {
          full = true;
          pend_step = step;
          pend_idx = local;  // local rank of the first unstored line
        }
        if (!full) {
          uint32_t r = 0;
          while (st) {
            uint32_t b = __builtin_ctz(st);
            st &= st - 1;
            uint32_t e = (tq0 + stored + ex + r) & QMASK;
            qs[e] = p + b;
            qi[e] = (uint32_t)(idx0 + local + ex + r);  // local rank, or global when the base is known
            r++;
          }
          stored += tot;
        }
        local += tot;
      }
      __syncthreads();
      if (!known_base) {
        uint64_t base = tile_lookback(a.tile_state, t, local);
        lines_total += local;
        for (uint32_t e = lane; e < stored; e += 64) qi[(tq0 + e) & QMASK] += (uint32_t)base;
        if (full) {
          pend = true;
          pend_t = t;
          pend_idx += base;
        }
      } else if (full) {
        pend = true;
        pend_t = t;
        pend_idx += idx0;
      }
      qlen += stored;
      claim_end = t0 + a.claim;
      __syncthreads();
    }
    if (qlen == 0) break;

    // ---- one round: lane k evaluates queued line k ----
    uint32_t take = qlen < 64 ? qlen : 64;
    {
      // stage the round's span [lo, hi) into LDS with coalesced 16-byte loads
      uint32_t lo = qs[qhead & QMASK] & ~15u;
      uint32_t hi = qlen > take ? qs[(qhead + take) & QMASK] : claim_end + 1024u;
      uint32_t cap_hi = (a.n + 16u) & ~15u;
      if (hi > cap_hi) hi = cap_hi;
      if (hi > lo + STAGE_BYTES) hi = lo + STAGE_BYTES;
      hi = (hi + 15u) & ~15u;
      for (uint32_t off = lane * 16u; lo + off < hi; off += 1024u) stage[off >> 4] = *((const uint4 *)(a.d + lo + off));
      rd.slo = lo;
      rd.shi = hi;
      __syncthreads();
    }
    if (lane < take) {
      uint32_t e = (qhead + lane) & QMASK;
      uint32_t s = qs[e];
      uint32_t idx = qi[e];
      SlotV sl[NS];
      LineResult r = eval_line<NS>(rd, s, P, sl);
      if (r.status == ST_OK && r.selected && P.output_kind == WARPJQ_OUT_AGG) {
        bool has_num = false, is_int = false, undecided = false;
        double v = 0.0;
        int64_t iv = 0;
        if (P.agg_slot != 0xFFFFFFFFu) {
          SlotV x = pick<NS>(sl, P.agg_slot);
          if (x.kind == JK_NUM) {
            NumVal nv;
            if (parse_number(rd, x.off, x.off + x.len, nv)) {
              has_num = true;
              is_int = nv.is_int;
              iv = nv.iv;
              v = nv.v;
            } else {
              undecided = true;
            }
          } else if (x.kind != JK_MISSING && x.kind != JK_NULL) {
            saw_nn = true;
          }
        }
        GKey gk{0, 0, 0, 0};
        uint32_t gh = 0;
        if (!undecided && grouped) {
          gk = group_key(rd, pick<NS>(sl, P.group_slot));
          if (gk.len > 0x7FFFF) undecided = true;
          else gh = group_hash(rd, gk);
        }
        if (undecided) {
          r.status = ST_FALLBACK;
        } else {
          if (grouped) {
            group_line(rd, a.gt, *lgp, gk, gh, has_num, is_int, iv, v, &govf_local);
          } else {
            a_cnt++;
            if (has_num) {
              a_num++;
              if (is_int) {
                uint64_t lo = (uint64_t)iv;
                uint64_t old = a_ilo;
                a_ilo += lo;
                a_ihi += (iv < 0 ? ~0ull : 0ull) + (a_ilo < old ? 1 : 0);
              } else {
                double t = a_fs + v;  // Neumaier
                if (fabs(a_fs) >= fabs(v)) a_fc += (a_fs - t) + v;
                else a_fc += (v - t) + a_fs;
                a_fs = t;
              }
              if (v < a_mn) a_mn = v;
              if (v > a_mx) a_mx = v;
            }
          }
        }
      }
      if (rows) {
        uint32_t rl = 0;
        if (r.status == ST_OK && r.selected) {
          uint64_t n = row<NS, DevReader, DevSink>(rd, P, sl, s, r.end, nullptr);
          rl = (uint32_t)n;
          c_sel++;
        }
        if (idx < a.max_lines) {
          a.row_len[idx] = rl;
          if (rl) a.sel_start[idx] = s;
        } else {
          ovf_local = 1;
        }
      }
      switch (r.status) {
        case ST_BLANK: c_blank++; break;
        case ST_INVALID:
          c_inv++;
          atomicMin(&a.ctr->first_invalid, idx);
          break;
        case ST_TYPE_ERROR: c_te++; break;
        case ST_FALLBACK: {
          c_fb++;
          uint32_t e = s;
          while (rd.byte(e) != WJQ_END) e++;
          uint32_t len = e - s;
          if (len > 0 && rd.byte(e - 1) == '\r') len--;
          uint32_t k = atomicAdd(&a.ctr->fb_count, 1u);
          if (k < a.fb_cap) {
            a.fb_idx[k] = idx;
            a.fb_off[k] = s;
            a.fb_len[k] = len;
          } else {
            ovf_local = 1;
          }
          if (rows && idx < a.max_lines) a.row_len[idx] = 0;
          break;
        }
        default: break;
      }
    }
    qhead = (qhead + take) & QMASK;
    qlen -= take;
    __syncthreads();
  }

  // ---- flush ----
  unsigned long long sb = wave_sum((unsigned long long)c_blank), si = wave_sum((unsigned long long)c_inv),
                     st = wave_sum((unsigned long long)c_te), sf = wave_sum((unsigned long long)c_fb),
                     ss = wave_sum((unsigned long long)c_sel);
  uint64_t nn = __ballot(saw_nn);
  uint64_t ov = __ballot(ovf_local != 0);
  uint64_t gov = __ballot(govf_local != 0);
  if (lane == 0) {
    if (lines_total) atomicAdd(&a.ctr->n_lines, (unsigned long long)lines_total);
    if (sb) atomicAdd(&a.ctr->n_blank, sb);
    if (si) atomicAdd(&a.ctr->n_invalid, si);
    if (st) atomicAdd(&a.ctr->n_type_error, st);
    if (sf) atomicAdd(&a.ctr->n_fallback, sf);
    if (ss) atomicAdd(&a.ctr->n_sel, ss);
    if (nn) atomicOr(&a.ctr->saw_non_numeric, 1u);
    if (ov) atomicOr(&a.ctr->overflow, 1u);
    if (gov) atomicOr(&a.ctr->group_overflow, 1u);
  }
  if (P.output_kind == WARPJQ_OUT_AGG && !grouped) {
    // exact integer part: 128-bit add across lanes
    unsigned long long lo = a_ilo, hi = a_ihi;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      unsigned long long olo = __shfl_xor(lo, o, 64), ohi = __shfl_xor(hi, o, 64);
      unsigned long long nlo = lo + olo;
      hi = hi + ohi + (nlo < lo ? 1 : 0);
      lo = nlo;
    }
    // compensated float part: two-sum tree
    double s = a_fs, c = a_fc;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      double os = __shfl_xor(s, o, 64), oc = __shfl_xor(c, o, 64);
      double t = s + os;
      double bp = t - s;
      double e = (s - (t - bp)) + (os - bp);
      s = t;
      c = c + oc + e;
    }
    double mn = a_mn, mx = a_mx;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mn = fmin(mn, __shfl_xor(mn, o, 64));
      mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    unsigned long long cnt = wave_sum((unsigned long long)a_cnt), num = wave_sum((unsigned long long)a_num);
    if (lane == 0) {
      AggPart &pp = a.parts[blockIdx.x];
      pp.count = cnt;
      pp.numeric = num;
      pp.isum_lo = lo;
      pp.isum_hi = hi;
      pp.fsum = s;
      pp.fcomp = c;
      pp.mn = mn;
      pp.mx = mx;
    }
  }
  if (GRP) {
    __syncthreads();
    LGroup &lg = *lgp;
    for (uint32_t e = lane; e < LGROUP; e += 64) {
      unsigned long long k = lg.key[e];
      if (k == EMPTY_KEY) continue;
      GKey gk = unpack_key(k);
      uint32_t pos = group_find(rd, a.gt, gk, lg.hash[e]);
      if (pos == ~0u) {
        atomicOr(&a.ctr->group_overflow, 1u);
        continue;
      }
      GVal v;
      v.count = lg.count[e];
      v.numeric = lg.numeric[e];
      v.ilo = lg.ilo[e];
      v.ihi = lg.ihi[e];
      v.fsum = lg.fsum[e];
      v.fcomp = lg.fcomp[e];
      v.mnb = lg.mnb[e];
      v.mxb = lg.mxb[e];
      group_apply(a.gt, pos, v);
    }
  }
}

// One lane per selected row: re-evaluate and write its bytes.
template <int NS>
__global__ __launch_bounds__(256) void k_emit(const uint8_t *d, uint32_t n, DProg P, const DevCtr *ctr,
                                               const uint32_t *sel_idx, const uint64_t *row_off,
                                               const uint32_t *sel_start, uint8_t *out) {
  if (ctr->overflow) return;
  uint64_t nsel = ctr->n_sel;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nsel;
       r += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t idx = sel_idx[r];
    uint32_t s = sel_start[idx];
    DevReader rd(d, n);
    SlotV sl[NS];
    LineResult lr = eval_line<NS>(rd, s, P, sl);
    DevSink w{out, row_off[r]};
    (void)lr;
    row<NS, DevReader, DevSink>(rd, P, sl, s, lr.end, &w);
  }
}


// ---------------------------------------------------------------------------
// per-width launchers (each width is instantiated in its own translation unit)
// ---------------------------------------------------------------------------

#define WJQ_DECLARE_LAUNCHERS(NS)                                                                  \
  void launch_main_##NS(bool grouped, uint32_t grid, hipStream_t st, const KArgs &a);             \
  void launch_emit_##NS(uint32_t grid, hipStream_t st, const uint8_t *d, uint32_t n, const DProg &P, \
                        const DevCtr *ctr, const uint32_t *sel_idx, const uint64_t *row_off,        \
                        const uint32_t *sel_start, uint8_t *out);
WJQ_DECLARE_LAUNCHERS(1)
WJQ_DECLARE_LAUNCHERS(2)
WJQ_DECLARE_LAUNCHERS(4)
WJQ_DECLARE_LAUNCHERS(8)
WJQ_DECLARE_LAUNCHERS(16)

#define WJQ_DEFINE_LAUNCHERS(NS)                                                                     \
  void launch_main_##NS(bool grouped, uint32_t grid, hipStream_t st, const KArgs &a) {              \
    if (grouped) hipLaunchKernelGGL((k_main<NS, true>), dim3(grid), dim3(64), 0, st, a);             \
    else hipLaunchKernelGGL((k_main<NS, false>), dim3(grid), dim3(64), 0, st, a);                    \
  }                                                                                                  \
  void launch_emit_##NS(uint32_t grid, hipStream_t st, const uint8_t *d, uint32_t n, const DProg &P, \
                        const DevCtr *ctr, const uint32_t *sel_idx, const uint64_t *row_off,        \
                        const uint32_t *sel_start, uint8_t *out) {                                  \
    hipLaunchKernelGGL((k_emit<NS>), dim3(grid), dim3(256), 0, st, d, n, P, ctr, sel_idx, row_off,    \
                       sel_start, out);                                                             \
  }
