// warpjq's device ABI (include/warpjq_kernels.h) implemented for MI355X
// (gfx950, wave64) in HIP.
//
// Per submitted chunk, on the slot's stream:
//   H2D (pinned staging -> HBM)
//   k_reset        counters, tile states, tickets
//   k_line_count + k_tile_scan_*  (row outputs) line starts per claim and
//                  their exclusive prefix: each claim's first line index
//   k_main<NS>     ONE pass over the chunk. Persistent single-wave workgroups
//                  claim tiles ("claims", a few KiB) from regional tickets,
//                  a few consecutive claims per atomic.
//                  A claim is copied HBM -> LDS with direct-to-LDS 16-byte
//                  loads (global_load_lds_dwordx4) into a per-wave LDS
//                  "arena", followed by the first bytes of the next claim up
//                  to the end of the claim's last line. Newline masks and a
//                  wave prefix sum over the arena give the claim's line
//                  starts, queued in LDS. Every lane then evaluates one queued
//                  line at a time (jscan.h: validate + capture + filter),
//                  reading only LDS, so a round of 64 lines mixes lines of
//                  consecutive claims and every lane stays busy; the arena is
//                  compacted before each new claim. Aggregates stay in
//                  registers; group_by goes through a per-wave LDS table into
//                  an HBM hash table. Row outputs need each line's index in
//                  the chunk: the claim reads its first index from the
//                  pre-pass. Aggregate queries skip the pre-pass and store
//                  per-claim line counts; the rare records that need an index
//                  (first invalid line, declined lines) carry (claim, rank)
//                  and are resolved after the pass.
//   k_agg_final    ungrouped partials -> the chunk's warpjq_agg_partial
//   k_group_compact occupied HBM group slots -> warpjq_group[], clearing them
//   k_row_*        (row outputs) scan of per-line row lengths -> ascending
//                  selected line indices and row offsets
//   k_emit<NS>     re-evaluates each selected line and writes its row
//   D2H of the fixed-size results (counters, aggregate)
// warpjq_wait synchronises the stream and copies the variable-size results.
//
// Positions inside the kernels are virtual "q positions". A submit is a list
// of spans (one chunk, or many files / newline-complete pieces of files for
// the batched form); span s is placed at q = vbase_s, a multiple of the claim
// size, with its 16-byte-aligned-down base there, so its first byte is at
// start_s = vbase_s + (its misalignment) and every claim lies in one span.
// All 16-byte loads are then aligned, arena offsets agree with q positions
// modulo 16, and bytes past a span's end read as '\n' (each span's last line
// ends there, as at the end of a file).
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

#define LDS_AS __attribute__((address_space(3)))

// ---------------------------------------------------------------------------
// tunables
// ---------------------------------------------------------------------------

#define CLAIM_MIN 1024u              // smallest claim a wave takes (one scan step)
#define CLAIM_MAX 8192u              // largest claim (line ranks fit 13 bits)
#define CLAIM_DEFAULT 8192u
#define STEP_BYTES 1024u             // 64 lanes x 16 B
#define TAIL_PIECE 256u              // bytes of the next claim fetched per try
#define TAIL_MAX 1024u               // a claim's last line may end this far past it
#define ARENA_DEFAULT 18432u         // per-wave LDS arena (bytes)
// grouped queries: the per-wave LDS group table (about 4 KiB) would leave
// room for 7 workgroups per CU with the default arena; 6 KiB claims in a
// 14 KiB arena keep 8, with room to prefetch the next claim (measured on the
// nginx group_by(.host): 1.30 vs 1.58 ms per GiB)
#define GROUP_CLAIM_DEFAULT 6144u
#define GROUP_ARENA_DEFAULT 14336u
#define ARENA_MAX 32768u             // offsets fit 15 bits; bit 15 marks unstaged lines
#define QCAP 128u                    // per-wave LDS queue of line starts
#define FLAT_SKIP 16u                // rounds without the flat fast path after a round with no flat line
#define QMASK (QCAP - 1u)
#define LGROUP 32u                   // per-wave LDS group table entries
#define GROUP_MAX_PROBE 128u
#define MIN_LINE_BYTES 24u           // kernels.cu:1984-1988 sizing rule
#define DATA_PAD 256u
#define MAX_CHUNK (((uint64_t)1 << 32) - ((uint64_t)1 << 20))
#define ROWSCAN_BLOCK 1024u          // lines per block in the row-offset scan
#define EMPTY_KEY (~0ull)
#define TAB_LDS_MAX 16384u           // program tables staged in LDS per workgroup
#define MAX_SPANS 65536u             // spans per submit
#define KEYBLOB_BYTES (16u << 20)    // group-key bytes returned by a multi-span submit
#define DEEP_CAP 4096u               // selected lines nested past 64 levels per submit (row outputs)

// ---------------------------------------------------------------------------
// device-side structures
// ---------------------------------------------------------------------------

// k_main takes claims from NREG regional tickets, GRAB claims per atomic
// (device-scope atomics cost microseconds under contention).
#define NREG 32
#define RSTRIDE 32
#define GRAB 4

struct DevCtr {
  unsigned long long n_lines, n_blank, n_invalid, n_type_error, n_fallback, n_sel, out_total;
  unsigned long long first_bad;  // (claim << 32 | rank) of the first invalid line, aggregate queries
  unsigned int first_invalid, overflow, saw_non_numeric, group_overflow;
  unsigned int fb_count, n_groups;
  unsigned int rticket[NREG * RSTRIDE];  // one claim ticket per region, 128 B apart
  unsigned int blob_used;   // bytes of group keys copied out (multi-span submits)
  unsigned int long_count, long_next;  // lines for the wave-wide phase: added, taken
  unsigned long long first_te_bad;  // the first type-error line: its index (row outputs) or (claim << 32 | rank)
  unsigned int first_te;            // ... as a line index after the pass
  unsigned int deep_count;          // selected lines past 64 levels whose captures k_emit reads back
};

// Per-wave ungrouped aggregate partial.
struct AggPart {
  unsigned long long count, numeric;
  unsigned long long isum_lo, isum_hi;  // exact sum of integer tokens (i128)
  double fsum, fcomp;                   // compensated sum of the other tokens
  double mn, mx;
};

struct GroupTab {
  unsigned long long *keys;  // packed off:32 | len:27 | esc:1 | kind:4
  unsigned int *hash;
  unsigned long long *count, *numeric, *isum_lo, *isum_hi, *mnb, *mxb;
  double *fsum, *fcomp;
  unsigned int cap;  // power of two
};

// One span of a submit (see above). Spans are sorted by vbase.
struct Span {
  const uint8_t *base;  // 16-byte-aligned-down address of its first byte
  uint32_t vbase;       // q position of `base`
  uint32_t start;       // q position of its first byte
  uint32_t vend;        // q position one past its last byte
  uint32_t _pad;
  // a resident dataset's newline index (newlines per 64-byte block of the
  // dataset) when the span lies in one, else null; ioff is the dataset
  // offset of the span's first byte
  const uint8_t *nlidx;
  uint64_t ioff;
};

struct KArgs {
  const Span *spans;
  uint32_t n_spans;
  uint32_t ntiles;
  uint32_t claim;    // bytes per claim (multiple of STEP_BYTES)
  uint32_t arena;    // LDS arena bytes per wave (multiple of 256)
  DProg P;
  const uint8_t *tab;  // the program tables P points into (copied to LDS per workgroup)
  uint32_t tab_bytes;
  DevCtr *ctr;
  // row outputs: each claim's first line index (line-count pre-pass);
  // aggregates: line count per claim
  unsigned long long *tile_state;
  AggPart *parts;
  GroupTab gt;
  // row outputs
  uint32_t *row_len;    // per line (0 = not selected)
  uint32_t *sel_start;  // per line, q position, written for selected lines
  uint64_t max_lines;
  // declined lines (ST_FALLBACK)
  uint32_t *fb_idx, *fb_off, *fb_len, *fb_tile;
  uint32_t fb_cap;
  // lines too long to stage with their claim, walked after the claims:
  // {q start, rec, tile + 1} each, the last word written last (0 = not yet)
  uint32_t *long_q;
  uint32_t long_cap;
  // row outputs: captures of selected lines nested past 64 levels, which the
  // per-lane re-walk in k_emit cannot redo
  uint32_t *deep_rec, *deep_end;
  SlotV *deep_sl;  // 16 per entry
  uint32_t deep_cap;
};

// ---------------------------------------------------------------------------
// readers
// ---------------------------------------------------------------------------

// Reads the chunk from HBM; caches one 16-byte block. Positions at or past
// n read as '\n'.
struct DevReader {
  const uint8_t *d;
  uint32_t n;
  uint32_t cb;
  Blk16 v;

  __device__ DevReader(const uint8_t *d_, uint32_t n_) : d(d_), n(n_), cb(0xFFFFFFFFu), v{0, 0, 0, 0} {}

  __device__ __forceinline__ Blk16 block(uint32_t b) {
    if (b != cb) {
      uint4 x = *((const uint4 *)d + b);
      v = Blk16{x.x, x.y, x.z, x.w};
      uint32_t p = b << 4;
      if (p + 15 >= n) {
        uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (uint32_t q = 0; q < 16; q++)
          if (p + q >= n) w[q >> 2] = (w[q >> 2] & ~(0xFFu << (8 * (q & 3)))) | (WJQ_END << (8 * (q & 3)));
        v = Blk16{w[0], w[1], w[2], w[3]};
      }
      cb = b;
    }
    return v;
  }
  __device__ __forceinline__ uint32_t dword(uint32_t k) {
    Blk16 x = block(k >> 2);
    uint32_t s = k & 3;
    return s == 0 ? x.x : s == 1 ? x.y : s == 2 ? x.z : x.w;
  }
  __device__ __forceinline__ uint32_t byte(uint32_t i) { return (dword(i >> 2) >> ((i & 3) * 8)) & 0xFFu; }
};

// Index of the span holding q position q (spans sorted by vbase).
__device__ __forceinline__ uint32_t span_of(const Span *sp, uint32_t n, uint32_t q) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sp[mid].vbase <= q) lo = mid;
    else hi = mid;
  }
  return lo;
}

// A reader over the span holding q position q.
__device__ __forceinline__ DevReader span_reader(const Span *sp, uint32_t n, uint32_t q) {
  const Span &s = sp[span_of(sp, n, q)];
  return DevReader(s.base - s.vbase, s.vend);
}

// Reads a line staged in the wave's LDS arena: position i lives at arena
// byte i + delta (delta is a multiple of 16, so blocks are aligned). Keeps no
// state, so nothing it holds has to follow the walker's control flow: every
// access is one ds_read. Callers only read up to the line's terminator, which
// is staged (bytes past the chunk end are staged as '\n').
typedef unsigned int wjq_v4u __attribute__((ext_vector_type(4)));
struct LdsReader {
  const LDS_AS uint8_t *b;
  uint32_t delta;

  __device__ LdsReader(const LDS_AS uint8_t *b_, uint32_t delta_) : b(b_), delta(delta_) {}

  __device__ __forceinline__ Blk16 block(uint32_t k) const {
    wjq_v4u x = *(const LDS_AS wjq_v4u *)(b + ((k << 4) + delta));
    return Blk16{x.x, x.y, x.z, x.w};
  }
  __device__ __forceinline__ uint32_t dword(uint32_t k) const {
    return *(const LDS_AS uint32_t *)(b + ((k << 2) + delta));
  }
  __device__ __forceinline__ uint32_t byte(uint32_t i) const { return b[i + delta]; }
  // 16 bytes at any multiple of 4 (LDS reads need dword alignment only)
  static constexpr uint32_t flat_align = 4u;
  __device__ __forceinline__ Blk16 blk(uint32_t pos) const {
    typedef unsigned int v4u_a4 __attribute__((ext_vector_type(4), aligned(4)));
    v4u_a4 x = *(const LDS_AS v4u_a4 *)(b + (pos + delta));
    return Blk16{x.x, x.y, x.z, x.w};
  }
};

// Reads one line through a window of the wave's LDS arena that the whole
// wave refills with direct-to-LDS copies (64 x 16 bytes per instruction, one
// wait per window), instead of one dependent HBM load per 16 bytes. Every
// lane must read the same positions in step: the long-line phase walks one
// line with all 64 lanes. Positions at or past n read as '\n'.
//
// Past the window the arena holds a digit buffer (numbers the 128-bit path
// leaves open are converted exactly there, every lane computing the same
// digits) and a bit stack for container levels past the 64 kept in
// registers, so this phase decides lines the per-lane walk declines.
struct WinReader {
  const uint8_t *d;
  uint32_t n;
  LDS_AS uint8_t *w;
  uint32_t wblk;  // window size in 16-byte blocks (a multiple of 64)
  uint32_t wb;    // first block held
  LDS_AS uint8_t *ext;  // WJQ_DEC_BUF digit bytes, then the deep stack
  uint32_t dcap;        // container levels the walk may open
  bool deep_used;       // the line went past the levels kept in registers

  __device__ WinReader(const uint8_t *d_, uint32_t n_, LDS_AS uint8_t *w_, uint32_t wbytes, uint32_t ext_bytes)
      : d(d_), n(n_), w(w_), wblk(wbytes >> 4), wb(0x80000000u), ext(w_ + wbytes),
        dcap(WJQ_MAX_DEPTH + (ext_bytes - WJQ_DEC_BUF) * 8u), deep_used(false) {}

  __device__ __forceinline__ LDS_AS uint8_t *dec_buf() { return ext; }
  __device__ __forceinline__ uint32_t deep_cap() const { return dcap; }
  __device__ __forceinline__ void deep_set(uint32_t lvl, bool is_arr) {
    const uint32_t k = lvl - WJQ_MAX_DEPTH;
    LDS_AS uint8_t *p = ext + WJQ_DEC_BUF + (k >> 3);
    const uint32_t m = 1u << (k & 7u);
    const uint32_t v = *p;
    *p = (uint8_t)(is_arr ? (v | m) : (v & ~m));
    deep_used = true;
  }
  __device__ __forceinline__ bool deep_get(uint32_t lvl) const {
    const uint32_t k = lvl - WJQ_MAX_DEPTH;
    return (ext[WJQ_DEC_BUF + (k >> 3)] >> (k & 7u)) & 1u;
  }

  // The window from block b on (every lane takes part).
  __device__ __forceinline__ void fill(uint32_t b) {
    wb = b;
    const uint32_t p0 = b << 4;
    const uint32_t lane = __lane_id();
    for (uint32_t j = 0; j < (wblk << 4); j += 1024) {
      const uint32_t x = j + lane * 16u;
      if (p0 + x < n) __builtin_amdgcn_global_load_lds((const void *)(d + p0 + x), (LDS_AS void *)(w + j), 16, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
  }
  __device__ __forceinline__ Blk16 block(uint32_t b) {
    if (b - wb >= wblk) fill(b);
    wjq_v4u x = *(const LDS_AS wjq_v4u *)(w + ((b - wb) << 4));
    Blk16 v{x.x, x.y, x.z, x.w};
    const uint32_t p = b << 4;
    if (p + 15 >= n) {
      uint32_t q4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (uint32_t q = 0; q < 16; q++)
        if (p + q >= n) q4[q >> 2] = (q4[q >> 2] & ~(0xFFu << (8 * (q & 3)))) | (WJQ_END << (8 * (q & 3)));
      v = Blk16{q4[0], q4[1], q4[2], q4[3]};
    }
    return v;
  }
  __device__ __forceinline__ uint32_t dword(uint32_t k) {
    Blk16 x = block(k >> 2);
    uint32_t s = k & 3;
    return s == 0 ? x.x : s == 1 ? x.y : s == 2 ? x.z : x.w;
  }
  __device__ __forceinline__ uint32_t byte(uint32_t i) { return (dword(i >> 2) >> ((i & 3) * 8)) & 0xFFu; }
  // Makes the window hold [a, b) (every lane takes part), so code that only
  // some lanes run can read those bytes without a refill. False when the
  // range is longer than the window.
  __device__ __forceinline__ bool ensure(uint32_t a, uint32_t b) {
    if (b <= a) return true;
    const uint32_t b0 = a >> 4, b1 = (b - 1) >> 4;
    if (b1 - b0 + 1 > wblk) return false;
    if (b0 - wb >= wblk || b1 - wb >= wblk) fill(b0);
    return true;
  }
  // First position >= j holding a byte a string scan must look at: lane k
  // tests block (j >> 4) + k, 1 KiB per step, and the lowest flagged lane
  // gives the answer to all.
  __device__ __forceinline__ uint32_t next_special_wide(uint32_t j) {
    const uint32_t lane = __lane_id();
    const uint32_t bj = j >> 4;
    uint32_t b0 = bj;
    for (;;) {
      if (b0 - wb >= wblk || b0 + 64u - wb > wblk) fill(b0);
      Blk16 v = block(b0 + lane);
      uint32_t m = movemask16(str_special(v.x), str_special(v.y), str_special(v.z), str_special(v.w));
      if (lane == 0 && b0 == bj) m &= 0xFFFFu << (j & 15u);  // bytes before j, first step only
      const uint64_t bal = __ballot(m != 0);
      if (bal) {
        const uint32_t L = (uint32_t)__builtin_ctzll(bal);
        const uint32_t mL = __shfl(m, (int)L, 64);
        return ((b0 + L) << 4) + (uint32_t)__builtin_ctz(mL);
      }
      b0 += 64;
    }
  }
};

// The program tables staged in LDS: the same fields as DProg, read with
// ds_read through address-space-3 pointers (lane-varying indices into the
// step table would otherwise be dependent HBM loads on every key).
struct LProg {
  const LDS_AS warpjq_step *steps;
  const LDS_AS warpjq_cmp *cmps;
  const LDS_AS warpjq_cond_op *cond;
  const LDS_AS uint8_t *blob;
  const LDS_AS SlotPath *spath;
  const LDS_AS uint32_t *slot_of_path;
  const LDS_AS uint32_t *field_slot;
  const LDS_AS uint32_t *prefix_off;
  const LDS_AS uint32_t *prefix_len;
  uint32_t n_slots, n_cond, fused_mask, key_lens0, plain_keys, filter_mask, post_mask, has_filter, output_kind, output_slot,
      agg_kind,
      agg_slot, group_slot, n_fields, suffix_off, suffix_len, csv_mode;
};

// Copies the table block into LDS (whole workgroup) and rebases P onto it.
__device__ __forceinline__ LProg stage_prog(const DProg &P, const uint8_t *tab, uint32_t tab_bytes,
                                            LDS_AS uint8_t *lds) {
  for (uint32_t o = threadIdx.x * 4u; o < tab_bytes; o += blockDim.x * 4u)
    *(LDS_AS uint32_t *)(lds + o) = *(const uint32_t *)(tab + o);
  __syncthreads();
  LProg L;
#define WJQ_REBASE(f, T) L.f = (const LDS_AS T *)(lds + (uint32_t)((const uint8_t *)P.f - tab))
  WJQ_REBASE(steps, warpjq_step);
  WJQ_REBASE(cmps, warpjq_cmp);
  WJQ_REBASE(cond, warpjq_cond_op);
  WJQ_REBASE(blob, uint8_t);
  WJQ_REBASE(spath, SlotPath);
  WJQ_REBASE(slot_of_path, uint32_t);
  WJQ_REBASE(field_slot, uint32_t);
  WJQ_REBASE(prefix_off, uint32_t);
  WJQ_REBASE(prefix_len, uint32_t);
#undef WJQ_REBASE
  L.n_slots = P.n_slots;
  L.n_cond = P.n_cond;
  L.fused_mask = P.fused_mask;
  L.key_lens0 = P.key_lens0;
  L.plain_keys = P.plain_keys;
  L.filter_mask = P.filter_mask;
  L.post_mask = P.post_mask;
  L.has_filter = P.has_filter;
  L.output_kind = P.output_kind;
  L.output_slot = P.output_slot;
  L.agg_kind = P.agg_kind;
  L.agg_slot = P.agg_slot;
  L.group_slot = P.group_slot;
  L.n_fields = P.n_fields;
  L.suffix_off = P.suffix_off;
  L.suffix_len = P.suffix_len;
  L.csv_mode = P.csv_mode;
  return L;
}

// 16-byte LDS accesses through address-space-3 pointers (ds_read_b128 /
// ds_write_b128).
__device__ __forceinline__ uint4 lds_ld16(const LDS_AS uint8_t *p) {
  wjq_v4u v = *(const LDS_AS wjq_v4u *)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void lds_st16(LDS_AS uint8_t *p, uint4 v) {
  wjq_v4u x = {v.x, v.y, v.z, v.w};
  *(LDS_AS wjq_v4u *)p = x;
}

// Byte sink writing a row into HBM, bounded by the output capacity.
struct DevSink {
  uint8_t *out;
  uint64_t pos;
  __device__ __forceinline__ void put(uint8_t c) { out[pos++] = c; }
  template <class R>
  __device__ __forceinline__ void span(R &rd, uint32_t off, uint32_t len) {
    for (uint32_t k = 0; k < len; k++) out[pos + k] = (uint8_t)rd.byte(off + k);
    pos += len;
  }
};

// ---------------------------------------------------------------------------
// wave helpers (wave64)
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Exclusive prefix sum over the wave (every lane active), with DPP row
// shifts and row broadcasts: no LDS round trips. Lanes a DPP step leaves out
// (row/bank mask, or a source before the row) add 0.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t &total) {
  const int x = (int)v;
  int t = x;
  t += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);  // row_shr:1
  t += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);  // row_shr:2
  t += __builtin_amdgcn_update_dpp(0, x, 0x113, 0xF, 0xF, true);  // row_shr:3
  t += __builtin_amdgcn_update_dpp(0, t, 0x114, 0xF, 0xE, true);  // row_shr:4, lanes 4-15 of a row
  t += __builtin_amdgcn_update_dpp(0, t, 0x118, 0xF, 0xC, true);  // row_shr:8, lanes 8-15 of a row
  t += __builtin_amdgcn_update_dpp(0, t, 0x142, 0xA, 0xF, true);  // row_bcast:15 into rows 1 and 3
  t += __builtin_amdgcn_update_dpp(0, t, 0x143, 0xC, 0xF, true);  // row_bcast:31 into rows 2 and 3
  total = (uint32_t)__builtin_amdgcn_readlane(t, 63);
  return (uint32_t)t - v;
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

// Bit 7 of each byte set when that byte is '\n' (exact per byte).
__device__ __forceinline__ uint32_t nl_bytes(uint32_t w) {
  const uint32_t x = w ^ 0x0A0A0A0Au;
  return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
}

// Bit k set when byte k of the 16 is '\n' (exact per byte).
__device__ __forceinline__ uint32_t nl_mask16(uint4 v) {
  return movemask16(nl_bytes(v.x), nl_bytes(v.y), nl_bytes(v.z), nl_bytes(v.w));
}

// The 16 bytes at q position p with every byte at or past n replaced by '\n'.
__device__ __forceinline__ uint4 end_patch(uint4 v, uint32_t p, uint32_t n) {
  uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (uint32_t q = 0; q < 4; q++) {
#pragma unroll
    for (uint32_t k = 0; k < 4; k++)
      if (p + 4 * q + k >= n) w[q] = (w[q] & ~(0xFFu << (8 * k))) | (WJQ_END << (8 * k));
  }
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// ---------------------------------------------------------------------------
// group table
// ---------------------------------------------------------------------------

// off:32 | len:27 | esc:1 | kind:4. EMPTY_KEY (all ones) would need off +
// len >= 2^32, which no key reaches.
#define GKEY_LEN_MAX 0x7FFFFFFu
__device__ __forceinline__ unsigned long long pack_key(const GKey &k) {
  return ((unsigned long long)k.off << 32) | ((unsigned long long)(k.len & GKEY_LEN_MAX) << 5) |
         ((unsigned long long)(k.esc & 1) << 4) | (k.kind & 0xF);
}
__device__ __forceinline__ GKey unpack_key(unsigned long long p) {
  GKey k;
  k.off = (uint32_t)(p >> 32);
  k.len = (uint32_t)((p >> 5) & GKEY_LEN_MAX);
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

// Finds or claims the slot of `k` (read through `rk`) in the HBM table; ~0u
// on overflow. Stored keys are read from HBM through `gr`. The key is
// published whole by one CAS, so a reader never sees a half-written key.
template <class RK>
__device__ __forceinline__ uint32_t group_find(RK &rk, const Span *spans, uint32_t n_spans, const GroupTab &g,
                                               const GKey &k, uint32_t h) {
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
      if (cur == pk) return pos;
      const GKey o = unpack_key(cur);
      DevReader gr = span_reader(spans, n_spans, o.off);
      if (group_eq(rk, k, gr, o)) return pos;
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
  uint4 kb[LGROUP * 3];  // the key bytes of entries with short, escape-free keys
};

#define KEY_INLINE 48u

__device__ __forceinline__ bool key_inline(const GKey &k) { return k.len <= KEY_INLINE && !k.esc; }

// The key's first 48 bytes, zero past its end (dword reads of the line).
template <class R>
__device__ __forceinline__ void key_words(R &rd, const GKey &k, uint32_t w[12]) {
  const uint32_t o = k.off, sh = (o & 3u) * 8u;
  uint32_t d[13];
#pragma unroll
  for (uint32_t j = 0; j < 13; j++) d[j] = (j * 4u < k.len + (o & 3u)) ? rd.dword((o >> 2) + j) : 0u;
#pragma unroll
  for (uint32_t q = 0; q < 12; q++) {
    uint32_t x = sh ? ((d[q] >> sh) | (d[q + 1] << (32u - sh))) : d[q];
    if (q * 4u >= k.len) x = 0;
    else if (q * 4u + 4u > k.len) x &= (1u << (8u * (k.len - q * 4u))) - 1u;
    w[q] = x;
  }
}

__device__ __forceinline__ void lds_acc_f64(double *s, double *c, double v) {
  double old = atomicAdd(s, v);
  double t = old + v;
  double bp = t - old;
  double e = (old - (t - bp)) + (v - bp);
  if (e != 0.0) atomicAdd(c, e);
}

// Inserts one line's value; falls through to HBM when the LDS table is full.
template <class R>
__device__ __forceinline__ void group_line(R &rd, const Span *spans, uint32_t n_spans, const GroupTab &g, LGroup &L,
                                           const GKey &k, uint32_t h, bool has_num, bool is_int, int64_t iv, double v,
                                           unsigned int &ovf) {
  unsigned long long pk = pack_key(k);
  uint32_t pos = h & (LGROUP - 1);
  int found = -1;
  const bool kin = key_inline(k);
  uint32_t kw[12];
  if (kin) key_words(rd, k, kw);
  for (uint32_t probe = 0; probe < 8; probe++) {
    unsigned long long cur = L.key[pos];
    if (cur == EMPTY_KEY) {
      unsigned long long old = atomicCAS(&L.key[pos], EMPTY_KEY, pk);
      if (old == EMPTY_KEY) {
        L.hash[pos] = h;
        if (kin) {
          L.kb[3 * pos] = make_uint4(kw[0], kw[1], kw[2], kw[3]);
          L.kb[3 * pos + 1] = make_uint4(kw[4], kw[5], kw[6], kw[7]);
          L.kb[3 * pos + 2] = make_uint4(kw[8], kw[9], kw[10], kw[11]);
        }
        found = (int)pos;
        break;
      }
      cur = old;
    }
    unsigned int hv = L.hash[pos];
    if (hv == 0 || hv == h) {
      bool same = cur == pk;
      if (!same) {
        const GKey o = unpack_key(cur);
        if (kin && key_inline(o)) {  // both keys' bytes at hand: compare in LDS
          const uint4 x = L.kb[3 * pos], y = L.kb[3 * pos + 1], z = L.kb[3 * pos + 2];
          same = o.kind == k.kind && o.len == k.len && x.x == kw[0] && x.y == kw[1] && x.z == kw[2] &&
                 x.w == kw[3] && y.x == kw[4] && y.y == kw[5] && y.z == kw[6] && y.w == kw[7] &&
                 z.x == kw[8] && z.y == kw[9] && z.z == kw[10] && z.w == kw[11];
        } else {
          DevReader gr = span_reader(spans, n_spans, o.off);
          same = group_eq(rd, k, gr, o);
        }
      }
      if (same) {
        found = (int)pos;
        break;
      }
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
  uint32_t gp = group_find(rd, spans, n_spans, g, k, h);
  if (gp == ~0u) {
    ovf = 1;
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
// k_main: per-line work
// ---------------------------------------------------------------------------

// Per-lane tallies, flushed once at the end of the kernel.
struct LaneAcc {
  uint32_t blank, inv, te, fb, sel;
  uint64_t te_first;  // lowest type-error record (line index, or claim << 32 | rank)
  uint64_t cnt, num, ilo, ihi;
  double fs, fc, mn, mx;
  bool saw_nn;
  unsigned int ovf, govf;
};

__device__ __forceinline__ void acc_zero(LaneAcc &L) {
  L.blank = L.inv = L.te = L.fb = L.sel = 0;
  L.te_first = ~0ull;
  L.cnt = L.num = L.ilo = L.ihi = 0;
  L.fs = L.fc = 0.0;
  L.mn = INFINITY;
  L.mx = -INFINITY;
  L.saw_nn = false;
  L.ovf = L.govf = 0;
}

// L += d (d's float part added with the same compensation as one value).
__device__ __forceinline__ void acc_merge(LaneAcc &L, const LaneAcc &d) {
  L.blank += d.blank;
  L.inv += d.inv;
  L.te += d.te;
  if (d.te_first < L.te_first) L.te_first = d.te_first;
  L.fb += d.fb;
  L.sel += d.sel;
  L.cnt += d.cnt;
  L.num += d.num;
  const uint64_t old = L.ilo;
  L.ilo += d.ilo;
  L.ihi += d.ihi + (L.ilo < old ? 1 : 0);
  const double t = L.fs + d.fs;
  if (fabs(L.fs) >= fabs(d.fs)) L.fc += (L.fs - t) + d.fs;
  else L.fc += (d.fs - t) + L.fs;
  L.fs = t;
  L.fc += d.fc;
  if (d.mn < L.mn) L.mn = d.mn;
  if (d.mx > L.mx) L.mx = d.mx;
  L.saw_nn = L.saw_nn || d.saw_nn;
  L.ovf |= d.ovf;
  L.govf |= d.govf;
}

// One line: evaluate, then fold its outcome into the lane's tallies, the
// group table or the per-line row arrays. `rec` is the line's index in the
// chunk (row outputs) or its rank in claim `tile` (aggregates).
// `act` false: evaluate and tally into L only (the long-line phase runs one
// line on every lane; lane 0 alone touches shared state).
// Returns true when a per-lane reader could not decide the line (nesting past
// 64 levels, a number the 128-bit path leaves open, a group key longer than
// a window): nothing is tallied and the caller hands the line to the
// wave-wide phase, whose reader decides it.
template <int NS, bool GRP, bool LB, class R>
__device__ __forceinline__ bool process_line(R &rd, const KArgs &a, const LProg &P, uint32_t s, uint32_t rec,
                                             uint32_t tile, LaneAcc &L, LGroup *lgp, bool act = true,
                                             bool try_flat = true, uint32_t *flat = nullptr) {
  constexpr bool WIDE = wide_reader<R>::value;
  SlotV sl[NS];
  LineResult r = eval_line<NS>(rd, s, P, sl, try_flat);
  if (flat) *flat = r.flat;
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
        L.saw_nn = true;
      }
    }
    GKey gk{0, 0, 0, 0};
    uint32_t gh = 0;
    if (!undecided && GRP) {
      gk = group_key(rd, pick<NS>(sl, P.group_slot));
      if (gk.len > GKEY_LEN_MAX) undecided = true;
      else gh = group_hash(rd, gk);
    }
    if (undecided) {
      r.status = ST_FALLBACK;
    } else if (GRP) {
      if constexpr (WIDE) {
        // lane 0 alone inserts: it reads the key from HBM, not the window
        if (act) {
          DevReader kr = span_reader(a.spans, a.n_spans, gk.off);
          group_line(kr, a.spans, a.n_spans, a.gt, *lgp, gk, gh, has_num, is_int, iv, v, L.govf);
        }
      } else {
        if (act) group_line(rd, a.spans, a.n_spans, a.gt, *lgp, gk, gh, has_num, is_int, iv, v, L.govf);
      }
    } else {
      L.cnt++;
      if (has_num) {
        L.num++;
        if (is_int) {
          uint64_t lo = (uint64_t)iv;
          uint64_t old = L.ilo;
          L.ilo += lo;
          L.ihi += (iv < 0 ? ~0ull : 0ull) + (L.ilo < old ? 1 : 0);
        } else {
          double t = L.fs + v;  // Neumaier
          if (fabs(L.fs) >= fabs(v)) L.fc += (L.fs - t) + v;
          else L.fc += (v - t) + L.fs;
          L.fs = t;
        }
        if (v < L.mn) L.mn = v;
        if (v > L.mx) L.mx = v;
      }
    }
  }
  if constexpr (!WIDE) {
    if (r.status == ST_FALLBACK) return true;
  }
  if (LB) {
    uint32_t rl = 0;
    if (r.status == ST_OK && r.selected) {
      uint64_t n = row<NS, R, DevSink>(rd, P, sl, s, r.end, nullptr);
      rl = (uint32_t)n;
      L.sel++;
      if constexpr (WIDE) {
        if (rd.deep_used && act) {
          const uint32_t k = atomicAdd(&a.ctr->deep_count, 1u);
          if (k < a.deep_cap) {
            a.deep_rec[k] = rec;
            a.deep_end[k] = r.end;
#pragma unroll
            for (int q = 0; q < NS; q++) a.deep_sl[16u * k + q] = sl[q];
          } else {
            L.ovf = 1;
          }
        }
      }
    }
    if (rec < a.max_lines) {
      if (act) {
        a.row_len[rec] = rl;
        if (rl) a.sel_start[rec] = s;
      }
    } else {
      L.ovf = 1;
    }
  }
  switch (r.status) {
    case ST_BLANK: L.blank++; break;
    case ST_INVALID:
      L.inv++;
      if (act) {
        if (LB) atomicMin(&a.ctr->first_invalid, rec);
        else atomicMin(&a.ctr->first_bad, ((unsigned long long)tile << 32) | rec);
      }
      break;
    case ST_TYPE_ERROR: {
      L.te++;
      const uint64_t tr = LB ? (uint64_t)rec : (((uint64_t)tile << 32) | rec);
      if (tr < L.te_first) L.te_first = tr;
      break;
    }
    case ST_FALLBACK: {
      // only the wave-wide phase gets here: nesting past its LDS stack, or a
      // group key past 128 MiB
      L.fb++;
      uint32_t e = s;
      while (rd.byte(e) != WJQ_END) e++;
      uint32_t len = e - s;
      if (len > 0 && rd.byte(e - 1) == '\r') len--;
      if (act) {
        uint32_t k = atomicAdd(&a.ctr->fb_count, 1u);
        if (k < a.fb_cap) {
          a.fb_idx[k] = rec;
          a.fb_tile[k] = tile;
          a.fb_off[k] = s;  // q position; the host maps it back through the spans
          a.fb_len[k] = len;
        } else {
          L.ovf = 1;
        }
        if (LB && rec < a.max_lines) a.row_len[rec] = 0;
      }
      break;
    }
    default: break;
  }
  return false;
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

extern __shared__ uint4 wjq_dyn_lds[];

#define CRING 8u  // claims with queued lines a wave tracks at once

// Dynamic LDS: [arena | queue | claim ring | program tables].
__host__ __device__ constexpr uint32_t main_lds_bytes(uint32_t arena, uint32_t tab_bytes) {
  return arena + QCAP * 4 + 128 + ((tab_bytes + 15u) & ~15u);
}

// The queue of line starts: an LDS ring of QCAP entries, each
//   bits 0-14 arena offset of the line start, bit 15 the line is not wholly
//   staged, bits 16-18 claim-ring slot, bits 19-31 rank of the line in its
//   claim,
// and a ring of CRING claims with queued lines: claim index, arena offset
// minus q position, index of the claim's first line (row outputs).
#define QE_UNSTAGED 0x8000u
struct LQueue {
  LDS_AS uint32_t *qe;
  LDS_AS uint32_t *cr_t, *cr_delta, *cr_base;
};

// The claim whose line starts are still being queued.
struct Claim {
  uint32_t t, t0, end;  // claim index, first q position, end of its line starts
  uint32_t start;       // q position of its span's first byte
  uint32_t slot;        // its claim-ring slot
  uint32_t at;          // arena offset of q position t0
  uint32_t resume;      // first line start not yet queued (q position)
  uint32_t rank;        // its rank in the claim
  uint32_t base;        // index of the claim's first line (row outputs)
  uint32_t carry0;      // 1 when a line starts at t0
  uint32_t last_unstaged;  // start of the claim's last line when it is not wholly staged, else ~0
};

// Queues line starts of claim `c` from c.resume on, at most `room` of them,
// and returns how many start at or after c.resume (the claim's line count on
// the first pass). Lane l scans the contiguous slice [t0 + l*S, t0 + (l+1)*S)
// of the staged claim (S = claim/64, 16..128 bytes), so one wave prefix sum
// per pass orders every start. Advances c.resume / c.rank; *queued gets the
// number queued and *last this lane's highest start (the caller reduces it
// over the wave when it needs the claim's). Called wave-uniformly.
__device__ __forceinline__ uint32_t enqueue_lines(const KArgs &a, const LDS_AS uint8_t *ar, LQueue &Q, uint32_t qtail,
                                                  uint32_t room, Claim &c, uint32_t *queued, uint32_t *last) {
  const uint32_t lane = lane_id();
  const uint32_t S = a.claim >> 6;
  const uint32_t nb = S >> 4;  // 16-byte blocks per lane, 1..8
  const uint32_t p0 = c.t0 + lane * S;
  const LDS_AS uint8_t *base = ar + (c.at - c.t0);  // base[q] is the staged byte at q position q
  uint32_t prev = lane == 0 ? c.carry0 : (base[p0 - 1] == '\n' ? 1u : 0u);
  // bit b of w[j]: a line starts at q position p0 + 32 j + b (a newline before it)
  uint32_t w[4];
#pragma unroll
  for (uint32_t k = 0; k < 8; k++) {
    uint32_t x = 0;
    if (k < nb) {
      const uint32_t m = nl_mask16(lds_ld16(base + p0 + 16u * k));
      x = ((m << 1) | prev) & 0xFFFFu;
      prev = m >> 15;
    }
    if (k & 1u) w[k >> 1] |= x << 16;
    else w[k >> 1] = x;
  }
  // a span's first line starts at its first byte and nothing before it does;
  // keep the starts in [max(start, resume), end)
  const uint32_t lo = max(c.start, c.resume);
  uint32_t cnt = 0, hi = 0;
#pragma unroll
  for (uint32_t j = 0; j < 4; j++) {
    const uint32_t pj = p0 + 32u * j;
    uint32_t x = w[j];
    if (c.start - p0 < S && c.start - pj < 32u) x |= 1u << (c.start - pj);
    if (lo > pj) x = lo - pj >= 32u ? 0u : x & (~0u << (lo - pj));
    if (c.end <= pj) x = 0;
    else if (c.end - pj < 32u) x &= (1u << (c.end - pj)) - 1u;
    w[j] = x;
    cnt += __popc(x);
    if (x) hi = pj + 31u - (uint32_t)__builtin_clz(x);
  }
  uint32_t tot;
  const uint32_t ex = wave_excl_scan(cnt, tot);
  uint32_t next = c.end;  // the start ranked `room`, if the queue fills
  uint32_t r = ex;
  // one pass per start of this lane (lines longer than the slice: at most
  // one), lowest first, selected without branches
  while (w[0] | w[1] | w[2] | w[3]) {
    const bool s0 = w[0] != 0, s1 = !s0 && w[1] != 0, s2 = !s0 && !s1 && w[2] != 0;
    const bool s3 = !s0 && !s1 && !s2;
    const uint32_t x = s0 ? w[0] : s1 ? w[1] : s2 ? w[2] : w[3];
    const uint32_t q = p0 + (s0 ? 0u : s1 ? 32u : s2 ? 64u : 96u) + (uint32_t)__builtin_ctz(x);
    const uint32_t xl = x & (x - 1u);
    w[0] = s0 ? xl : w[0];
    w[1] = s1 ? xl : w[1];
    w[2] = s2 ? xl : w[2];
    w[3] = s3 ? xl : w[3];
    if (r < room) {
      const uint32_t e = (qtail + r) & QMASK;
      Q.qe[e] = (q - c.t0 + c.at) | (q == c.last_unstaged ? QE_UNSTAGED : 0u) | (c.slot << 16) |
                ((c.rank + r) << 19);
    } else if (r == room) {
      next = q;
    }
    r++;
  }
  *last = hi;  // this lane's; the caller reduces it over the wave when it needs it
  if (tot > room) {
    const uint64_t own = __ballot(next != c.end);
    next = (uint32_t)__builtin_amdgcn_readlane((int)next, (int)__builtin_ctzll(own));
    c.resume = next;
    c.rank += room;
    *queued = room;
  } else {
    c.resume = c.end;
    c.rank += tot;
    *queued = tot;
  }
  return tot;
}

// Two waves per SIMD is what the LDS arena allows; the register budget is
// held to that so the allocator never spills into AGPRs and drops to one.
template <int NS, bool GRP, bool LB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_main(KArgs a) {
  LDS_AS uint8_t *ar = (LDS_AS uint8_t *)wjq_dyn_lds;
  LQueue Q;
  Q.qe = (LDS_AS uint32_t *)(ar + a.arena);
  Q.cr_t = Q.qe + QCAP;
  Q.cr_delta = Q.cr_t + CRING;
  Q.cr_base = Q.cr_delta + CRING;
  LGroup *lgp = LGroupStore<GRP>::get();
  const LProg P = stage_prog(a.P, a.tab, a.tab_bytes, ar + main_lds_bytes(a.arena, 0));

  const uint32_t lane = lane_id();

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

  LaneAcc L;
  acc_zero(L);

  uint64_t lines_total = 0;     // wave-uniform
  uint32_t qhead = 0, qlen = 0;  // wave-uniform
  uint32_t used = 0;             // arena bytes holding queued lines (from 0)
  bool tiles_left = true;
  bool pending = false;          // a claim still has unqueued line starts
  Claim c{};
  const uint32_t need = a.claim + TAIL_MAX;
  uint32_t cseq = 0;  // claims given a ring slot so far
  uint32_t hseq = 0;  // the claim of the queue's head (when the queue is not empty)

  // Claims: the chunk is split into NREG regions with a ticket each, and a
  // wave takes GRAB consecutive claims per atomic. A wave starts in region
  // blockIdx % NREG - a region of its own XCD, as workgroups are dealt to the
  // 8 XCDs round-robin, so a region's neighbouring claims share one L2 - and
  // moves on when it runs dry. Row outputs know each claim's first line index
  // from the line-count pre-pass, so they take claims in any order too.
  uint32_t region = blockIdx.x % NREG, regions_tried = 0;
  // claims left from the last grab, [grab[0], grab[1]), kept in the claim
  // ring's spare LDS words (SGPRs are all in use)
  LDS_AS uint32_t *grab = Q.cr_base + CRING;
  if (lane == 0) grab[0] = grab[1] = 0;
  __syncthreads();
  auto take_claim = [&]() -> uint32_t {
    const uint32_t cur = grab[0], end = grab[1];
    if (cur < end) {
      if (lane == 0) grab[0] = cur + 1;
      return cur;
    }
    while (regions_tried < NREG) {
      const uint32_t lo = (uint32_t)(((uint64_t)region * a.ntiles) / NREG);
      const uint32_t hi = (uint32_t)(((uint64_t)(region + 1) * a.ntiles) / NREG);
      uint32_t tk = 0;
      if (lo < hi && lane == 0) tk = atomicAdd(&a.ctr->rticket[region * RSTRIDE], (unsigned)GRAB);
      tk = (uint32_t)__builtin_amdgcn_readfirstlane((int)tk);
      if (lo + tk < hi) {
        if (lane == 0) {
          grab[0] = lo + tk + 1;
          grab[1] = min(lo + tk + GRAB, hi);
        }
        return lo + tk;
      }
      region = (region + 1) % NREG;
      regions_tried++;
    }
    return ~0u;
  };
  // Starts the direct-to-LDS copy of claim t (and the first 256 bytes past
  // it) to arena offset `at`; returns whether a line starts at its first
  // byte (read through the scalar cache; waited for on use). Row outputs
  // also start loading the claim's first line index into `pre`, so its
  // latency overlaps the copy's.
  auto issue_fetch = [&](uint32_t t, uint32_t at, unsigned long long &pre) -> uint32_t {
    const uint32_t t0 = t * a.claim;
    const Span sp = a.spans[span_of(a.spans, a.n_spans, t0)];
    const uint8_t *src = sp.base + (t0 - sp.vbase);
    for (uint32_t j = 0; j < a.claim + TAIL_PIECE; j += STEP_BYTES) {
      const uint32_t x = j + lane * 16u;
      if (x < a.claim + TAIL_PIECE && t0 + x < sp.vend)
        __builtin_amdgcn_global_load_lds((const void *)(src + x), (LDS_AS void *)(ar + at + j), 16, 0, 0);
    }
    uint32_t carry0 = 1;  // a span's first claim: its first line start is placed by c.start
    if (t0 > sp.vbase) carry0 = (*(const uint32_t *)(src - 4) >> 24) == '\n' ? 1u : 0u;
    if (LB) pre = a.tile_state[t];
    return carry0;
  };
  // Keeps the bytes of queued lines at the front of the arena.
  auto compact = [&]() {
    if (qlen == 0) {
      used = 0;
      return;
    }
    const uint32_t lo = (Q.qe[qhead & QMASK] & 0x7FFFu) & ~15u;
    if (!lo) return;
    for (uint32_t o = 0; o < used - lo; o += STEP_BYTES) {
      const uint32_t x = o + lane * 16u;
      uint4 v;
      if (x < used - lo) v = lds_ld16(ar + lo + x);
      __syncthreads();
      if (x < used - lo) lds_st16(ar + x, v);
      __syncthreads();
    }
    for (uint32_t e = lane; e < qlen; e += 64) {
      const uint32_t k = (qhead + e) & QMASK;
      Q.qe[k] -= lo;
    }
    if (lane < CRING) Q.cr_delta[lane] -= lo;
    used -= lo;
    __syncthreads();
  };
  // Waits for claim t's bytes at `at`, stages its last line's end, and queues
  // its line starts.
  auto process_claim = [&](uint32_t t, uint32_t at, uint32_t carry0, unsigned long long pre) {
    const uint32_t t0 = t * a.claim;
    const Span sp = a.spans[span_of(a.spans, a.n_spans, t0)];
    const uint32_t vend = sp.vend;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    uint32_t span = a.claim + TAIL_PIECE;  // bytes staged for this claim
    if (t0 + span > vend) {
      for (uint32_t x = lane * 16u; x < span; x += STEP_BYTES) {
        const uint32_t p = t0 + x;
        if (p + 16u > vend) lds_st16(ar + at + x, end_patch(lds_ld16(ar + at + x), p, vend));
      }
      __syncthreads();
    }
    // the claim's last line must end inside the staged bytes
    bool found = t0 + a.claim >= vend || ar[at + a.claim - 1] == '\n';
    uint32_t piece = a.claim;
    while (!found) {
      const uint32_t x = piece + lane * 16u;
      uint32_t m = 0;
      if (lane < TAIL_PIECE / 16u) m = nl_mask16(lds_ld16(ar + at + x));
      found = __ballot(m != 0) != 0;
      piece += TAIL_PIECE;
      if (found || piece >= a.claim + TAIL_MAX) break;
      // one more piece
      if (lane < TAIL_PIECE / 16u && t0 + piece + lane * 16u < vend)
        __builtin_amdgcn_global_load_lds((const void *)(sp.base + (t0 - sp.vbase) + piece + lane * 16u),
                                         (LDS_AS void *)(ar + at + piece), 16, 0, 0);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      span = piece + TAIL_PIECE;
      if (t0 + span > vend) {
        const uint32_t xx = piece + lane * 16u;
        if (lane < TAIL_PIECE / 16u && t0 + xx + 16u > vend)
          lds_st16(ar + at + xx, end_patch(lds_ld16(ar + at + xx), t0 + xx, vend));
        __syncthreads();
      }
    }
    c.t = t;
    c.t0 = t0;
    c.end = min(t0 + a.claim, vend);
    c.start = sp.start;
    c.at = at;
    c.resume = t0;
    c.rank = 0;
    c.base = 0;
    c.carry0 = carry0;
    c.last_unstaged = ~0u;
    c.slot = cseq & (CRING - 1u);
    uint32_t got, lst;
    const uint32_t total = enqueue_lines(a, ar, Q, qhead + qlen, QCAP - qlen, c, &got, &lst);
    if (!found && total) {
      // the claim's last line does not end inside the staged bytes
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) lst = max(lst, (uint32_t)__shfl_xor(lst, o, 64));
      c.last_unstaged = lst;
      if (got == total) {
        __syncthreads();
        const uint32_t e = (qhead + qlen + got - 1) & QMASK;
        if (lane == 0) Q.qe[e] |= QE_UNSTAGED;
      }
    }
    lines_total += total;
    if (LB) c.base = (uint32_t)pre;  // the pre-pass's exclusive prefix
    else if (lane == 0) a.tile_state[t] = total;
    if (total) {
      if (lane == 0) {
        Q.cr_t[c.slot] = t;
        Q.cr_delta[c.slot] = at - t0;
        Q.cr_base[c.slot] = c.base;
      }
      if (qlen == 0) hseq = cseq;
      cseq++;
    }
    qlen += got;
    pending = c.resume < c.end;
    if (total) used = at + span;
    __syncthreads();
  };

  uint32_t flat_skip = 0;  // rounds left that skip the flat fast path (wave-uniform)
  bool pf = false;  // a claim's copy was started before the last round
  uint32_t pf_t = 0, pf_at = 0, pf_carry0 = 0;
  unsigned long long pf_pre = 0;

  for (;;) {
    // ---- fill the queue to one full round when possible ----
    while (qlen < 64) {
      if (pending) {
        uint32_t got, lst;
        if (qlen == 0) hseq = cseq - 1u;  // the pending claim is the newest
        enqueue_lines(a, ar, Q, qhead + qlen, QCAP - qlen, c, &got, &lst);
        qlen += got;
        pending = c.resume < c.end;
        __syncthreads();
        if (pending) break;  // the queue is full
        continue;
      }
      if (pf) {
        pf = false;
        process_claim(pf_t, pf_at, pf_carry0, pf_pre);
        continue;
      }
      if (!tiles_left) break;
      if (qlen && cseq - hseq >= CRING) break;  // every ring slot holds queued lines
      compact();
      if (a.arena - used < need) break;  // run a (partial) round first
      const uint32_t t = take_claim();
      if (t == ~0u) {
        tiles_left = false;
        break;
      }
      unsigned long long pre = 0;
      const uint32_t carry0 = issue_fetch(t, used, pre);
      process_claim(t, used, carry0, pre);
    }
    if (qlen == 0) {
      if (!pending && !tiles_left && !pf) break;
      continue;
    }
    // start copying the next claim so it lands during the round
    if (!pf && !pending && tiles_left && cseq - hseq < CRING - 1u) {
      compact();
      if (a.arena - used >= need) {
        const uint32_t t = take_claim();
        if (t == ~0u) {
          tiles_left = false;
        } else {
          pf = true;
          pf_t = t;
          pf_at = used;
          pf_carry0 = issue_fetch(t, used, pf_pre);
        }
      }
    }

    // ---- one round: lane k evaluates queued line k ----
    const uint32_t take = qlen < 64 ? qlen : 64;
    // after a full-ish round in which no line was a flat object, the next
    // FLAT_SKIP rounds go straight to the byte walk (the same results: the
    // walk decides every line the fast path does)
    const bool try_flat = flat_skip == 0;
    uint32_t flat_hit = 0;
    if (lane < take) {
      const uint32_t e = (qhead + lane) & QMASK;
      const uint32_t qv = Q.qe[e];
      const uint32_t slot = (qv >> 16) & (CRING - 1u);
      const uint32_t off = qv & 0x7FFFu;
      const uint32_t s = off - Q.cr_delta[slot];
      const uint32_t rank = qv >> 19;
      const uint32_t rec = LB ? Q.cr_base[slot] + rank : rank;
      const uint32_t tile = Q.cr_t[slot];
      bool wide = (qv & QE_UNSTAGED) != 0;  // too long to stage
      if (!wide) {
        LdsReader rd(ar, off - s);
        wide = process_line<NS, GRP, LB>(rd, a, P, s, rec, tile, L, lgp, true, try_flat, &flat_hit);
      }
      if (wide) {
        // walked by the whole wave after the claims; a full queue makes the
        // host split the chunk
        const uint32_t k = atomicAdd(&a.ctr->long_count, 1u);
        if (k < a.long_cap) {
          a.long_q[3 * k] = s;
          a.long_q[3 * k + 1] = rec;
          __threadfence();
          __hip_atomic_store(&a.long_q[3 * k + 2], tile + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          L.ovf = 1;
        }
      }
    }
    if (try_flat) {
      if (take >= 32 && __ballot(flat_hit != 0) == 0) flat_skip = FLAT_SKIP;
    } else {
      flat_skip--;
    }
    qhead = (qhead + take) & QMASK;
    qlen -= take;
    if (qlen) {
      // the claim of the new head: the next sequence number with its slot
      const uint32_t hs = (Q.qe[qhead] >> 16) & (CRING - 1u);
      hseq += (hs - hseq) & (CRING - 1u);
    }
    __syncthreads();
  }

  // ---- long lines (and lines the per-lane walk declined): one at a time,
  // every lane walking it in step through the arena (free now) as a window,
  // the last 4 KiB (less in small arenas) kept for the digit buffer and the
  // deep container stack. Each line is taken by some wave: the wave that
  // added it looks here after adding it.
  {
    const uint32_t wbytes = a.arena >= 5120u ? ((a.arena - 4096u) & ~1023u) : 1024u;
    for (;;) {
      uint32_t k = ~0u;
      if (lane == 0) {
        for (;;) {
          const uint32_t nx = __hip_atomic_load(&a.ctr->long_next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const uint32_t ct = __hip_atomic_load(&a.ctr->long_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (nx >= ct || nx >= a.long_cap) break;
          if (atomicCAS(&a.ctr->long_next, nx, nx + 1u) == nx) {
            k = nx;
            break;
          }
        }
      }
      k = __shfl(k, 0, 64);
      if (k == ~0u) break;
      uint32_t ls = 0, lrec = 0, lt = 0;
      if (lane == 0) {
        // the adding lane stores the last word right after taking the index
        while ((lt = __hip_atomic_load(&a.long_q[3 * k + 2], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) == 0) {
        }
        ls = a.long_q[3 * k];
        lrec = a.long_q[3 * k + 1];
        a.long_q[3 * k + 2] = 0;  // clean for the next launch
      }
      ls = __shfl(ls, 0, 64);
      lrec = __shfl(lrec, 0, 64);
      lt = __shfl(lt, 0, 64);
      const Span sp = a.spans[span_of(a.spans, a.n_spans, ls)];
      WinReader rd(sp.base - sp.vbase, sp.vend, ar, wbytes, a.arena - wbytes);
      LaneAcc D;
      acc_zero(D);
      process_line<NS, GRP, LB>(rd, a, P, ls, lrec, lt - 1u, D, lgp, lane == 0);
      if (lane == 0) acc_merge(L, D);
      __syncthreads();
    }
  }

  // ---- flush ----
  unsigned long long sb = wave_sum((unsigned long long)L.blank), si = wave_sum((unsigned long long)L.inv),
                     st = wave_sum((unsigned long long)L.te), sf = wave_sum((unsigned long long)L.fb),
                     ss = wave_sum((unsigned long long)L.sel);
  uint64_t tf = L.te_first;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t x = __shfl_xor(tf, o, 64);
    if (x < tf) tf = x;
  }
  if (lane == 0 && tf != ~0ull) atomicMin(&a.ctr->first_te_bad, (unsigned long long)tf);
  uint64_t nn = __ballot(L.saw_nn);
  uint64_t ov = __ballot(L.ovf != 0);
  uint64_t gov = __ballot(L.govf != 0);
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
  if (P.output_kind == WARPJQ_OUT_AGG && !GRP) {
    // exact integer part: 128-bit add across lanes
    unsigned long long lo = L.ilo, hi = L.ihi;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      unsigned long long olo = __shfl_xor(lo, o, 64), ohi = __shfl_xor(hi, o, 64);
      unsigned long long nlo = lo + olo;
      hi = hi + ohi + (nlo < lo ? 1 : 0);
      lo = nlo;
    }
    // compensated float part: two-sum tree
    double s = L.fs, cc = L.fc;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      double os = __shfl_xor(s, o, 64), oc = __shfl_xor(cc, o, 64);
      double t = s + os;
      double bp = t - s;
      double e = (s - (t - bp)) + (os - bp);
      s = t;
      cc = cc + oc + e;
    }
    double mn = L.mn, mx = L.mx;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mn = fmin(mn, __shfl_xor(mn, o, 64));
      mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    unsigned long long cnt = wave_sum((unsigned long long)L.cnt), num = wave_sum((unsigned long long)L.num);
    if (lane == 0) {
      AggPart &pp = a.parts[blockIdx.x];
      pp.count = cnt;
      pp.numeric = num;
      pp.isum_lo = lo;
      pp.isum_hi = hi;
      pp.fsum = s;
      pp.fcomp = cc;
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
      DevReader gr = span_reader(a.spans, a.n_spans, gk.off);
      uint32_t pos = group_find(gr, a.spans, a.n_spans, a.gt, gk, lg.hash[e]);
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
__global__ __launch_bounds__(256) void k_emit(const Span *spans, uint32_t n_spans, DProg P0, const uint8_t *tab,
                                               uint32_t tab_bytes, DevCtr *ctr, const uint32_t *sel_idx,
                                               const uint64_t *row_off, const uint32_t *sel_start, uint8_t *out,
                                               const uint32_t *deep_rec, const uint32_t *deep_end,
                                               const SlotV *deep_sl, uint32_t deep_cap) {
  const LProg P = stage_prog(P0, tab, tab_bytes, (LDS_AS uint8_t *)wjq_dyn_lds);
  if (ctr->overflow) return;
  uint64_t nsel = ctr->n_sel;
  if (P.output_kind == WARPJQ_OUT_PASSTHROUGH && !P.csv_mode) {
    // a whole-line row is the line's bytes up to a final '\r', then '\n':
    // no re-walk. A wave takes 64 rows: each lane loads one row's start and
    // output range (one round of dependent loads for all 64), then the wave
    // copies the rows one after another, 64 contiguous bytes per lane step,
    // a row's loads issued before its stores.
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    uint32_t sv0 = 0, sv1 = 0;  // the span last used: q positions [sv0, sv1)
    const uint8_t *sbase = nullptr;
    for (uint64_t r0 = (((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64u; r0 < nsel; r0 += nw * 64u) {
      const uint64_t r = r0 + lane;
      uint32_t s = 0, len = 0;
      uint64_t o = 0;
      if (r < nsel) {
        s = sel_start[sel_idx[r]];
        o = row_off[r];
        len = (uint32_t)(row_off[r + 1] - o) - 1u;
      }
      const uint32_t m = (uint32_t)min<uint64_t>(64u, nsel - r0);
      for (uint32_t k = 0; k < m; k++) {
        const uint32_t sk = (uint32_t)__builtin_amdgcn_readlane((int)s, (int)k);
        const uint32_t lk = (uint32_t)__builtin_amdgcn_readlane((int)len, (int)k);
        const uint64_t ok = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(o >> 32), (int)k) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)o, (int)k);
        if (sk < sv0 || sk >= sv1) {
          const Span sp = spans[span_of(spans, n_spans, sk)];
          sv0 = sp.vbase;
          sv1 = sp.vend;
          sbase = sp.base - sp.vbase;
        }
        const uint8_t *src = sbase + sk;
        for (uint32_t j0 = 0; j0 < lk; j0 += 256u) {
          uint8_t b[4];
#pragma unroll
          for (uint32_t u = 0; u < 4; u++) {
            const uint32_t j = j0 + 64u * u + lane;
            b[u] = j < lk ? src[j] : (uint8_t)0;
          }
#pragma unroll
          for (uint32_t u = 0; u < 4; u++) {
            const uint32_t j = j0 + 64u * u + lane;
            if (j < lk) out[ok + j] = b[u];
          }
        }
        if (lane == 0) out[ok + lk] = '\n';
      }
    }
    return;
  }
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nsel;
       r += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t idx = sel_idx[r];
    uint32_t s = sel_start[idx];
    DevReader rd = span_reader(spans, n_spans, s);
    SlotV sl[NS];
    LineResult lr = eval_line<NS>(rd, s, P, sl);
    if (lr.status != ST_OK) {
      // A selected line is re-walked here by one lane. Nested past the 64
      // levels a lane keeps, the walk stops early: the wave-wide phase stored
      // its captures. A number this lane cannot settle stops only the filter
      // (end is set), and the captures are whole.
      bool found = false;
      const uint32_t nd = min(ctr->deep_count, deep_cap);
      for (uint32_t k = 0; !found && k < nd; k++)
        if (deep_rec[k] == idx) {
#pragma unroll
          for (int q = 0; q < NS; q++) sl[q] = deep_sl[16u * k + q];
          lr.end = deep_end[k];
          found = true;
        }
      if (!found && !(lr.status == ST_FALLBACK && lr.end != 0)) atomicOr(&ctr->overflow, 1u);
    }
    DevSink w{out, row_off[r]};
    row<NS, DevReader, DevSink>(rd, P, sl, s, lr.end, &w);
  }
}

// ---------------------------------------------------------------------------
// per-width launchers (each width is instantiated in its own translation unit)
// ---------------------------------------------------------------------------

#define WJQ_DECLARE_LAUNCHERS(NS)                                                                        \
  void launch_main_##NS(bool grouped, bool lb, uint32_t grid, hipStream_t st, const KArgs &a);          \
  int occupancy_main_##NS(bool grouped, bool lb, uint32_t lds);                                         \
  void launch_emit_##NS(uint32_t grid, hipStream_t st, const Span *spans, uint32_t n_spans, const DProg &P, \
                        const uint8_t *tab, uint32_t tab_bytes, DevCtr *ctr, const uint32_t *sel_idx,       \
                        const uint64_t *row_off, const uint32_t *sel_start, uint8_t *out, const KArgs &a);
WJQ_DECLARE_LAUNCHERS(1)
WJQ_DECLARE_LAUNCHERS(2)
WJQ_DECLARE_LAUNCHERS(4)
WJQ_DECLARE_LAUNCHERS(8)
WJQ_DECLARE_LAUNCHERS(16)

#define WJQ_DEFINE_LAUNCHERS(NS)                                                                          \
  void launch_main_##NS(bool grouped, bool lb, uint32_t grid, hipStream_t st, const KArgs &a) {          \
    const uint32_t lds = main_lds_bytes(a.arena, a.tab_bytes);                                            \
    if (lb) hipLaunchKernelGGL((k_main<NS, false, true>), dim3(grid), dim3(64), lds, st, a);              \
    else if (grouped) hipLaunchKernelGGL((k_main<NS, true, false>), dim3(grid), dim3(64), lds, st, a);    \
    else hipLaunchKernelGGL((k_main<NS, false, false>), dim3(grid), dim3(64), lds, st, a);                \
  }                                                                                                       \
  int occupancy_main_##NS(bool grouped, bool lb, uint32_t lds) {                                          \
    int b = 0;                                                                                            \
    hipError_t e;                                                                                         \
    if (lb) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_main<NS, false, true>, 64, lds);       \
    else if (grouped) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_main<NS, true, false>, 64, lds); \
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_main<NS, false, false>, 64, lds);         \
    return e == hipSuccess ? b : 0;                                                                       \
  }                                                                                                       \
  void launch_emit_##NS(uint32_t grid, hipStream_t st, const Span *spans, uint32_t n_spans, const DProg &P, \
                        const uint8_t *tab, uint32_t tab_bytes, DevCtr *ctr, const uint32_t *sel_idx,       \
                        const uint64_t *row_off, const uint32_t *sel_start, uint8_t *out, const KArgs &a) { \
    hipLaunchKernelGGL((k_emit<NS>), dim3(grid), dim3(256), (tab_bytes + 15u) & ~15u, st, spans, n_spans, P, tab, \
                       tab_bytes, ctr, sel_idx, row_off, sel_start, out, a.deep_rec, a.deep_end, a.deep_sl,     \
                       a.deep_cap);                                                                        \
  }
