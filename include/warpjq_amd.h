/* Additive exports of the MI355X backend. None of these changes the
 * behaviour of the ten reference entry points in warpjq_kernels.h; they
 * expose what this backend adds: exact aggregate parts, HBM-resident input,
 * per-slot device timing, and the host driver (chunked file / buffer runs
 * with the reference GPU backend's merge semantics).
 */
#ifndef WARPJQ_AMD_H
#define WARPJQ_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "warpjq_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Exact parts of a sum: a 128-bit integer total of every integer-valued
 * token (|v| <= 2^53), plus a compensated FP64 total of the other tokens
 * (fsum + fcomp). New in this backend; the reference sums naively in f64
 * (crates/warpjq-core/src/agg.rs:62-71). */
typedef struct {
  uint64_t isum_lo;
  uint64_t isum_hi;
  double fsum;
  double fcomp;
} warpjq_amd_agg_ext;

/* The parts rounded once to the nearest double. */
double warpjq_amd_ext_value(const warpjq_amd_agg_ext *e);

/* After warpjq_wait: the ungrouped exact parts, and for grouped queries a
 * pointer to n_groups parts parallel to warpjq_chunk_result.groups (valid
 * until the next submit on the slot). No counterpart in the reference. */
warpjq_status warpjq_amd_wait_ext(warpjq_ctx *ctx, uint32_t slot, warpjq_amd_agg_ext *agg_ext,
                                  const warpjq_amd_agg_ext **group_ext);

/* Like warpjq_submit, but the chunk already lives in device memory (the
 * served / resident form): no H2D copy. `d_bytes` must stay valid and
 * unmodified until warpjq_wait returns, and at least 256 bytes past n_bytes
 * must be readable. Aggregate queries accept up to 4 GiB - 1 MiB; row
 * outputs up to the slot capacity. No counterpart in the reference. */
warpjq_status warpjq_amd_submit_device(warpjq_ctx *ctx, uint32_t slot, const uint8_t *d_bytes,
                                       uint64_t n_bytes, char *err, size_t err_len);

/* Device time of the last submit on `slot` (hipEvent on the slot's stream):
 * the H2D copy, the whole kernel chain, and the one-pass main kernel alone.
 * Valid after warpjq_wait. No counterpart in the reference. */
warpjq_status warpjq_amd_slot_timing(warpjq_ctx *ctx, uint32_t slot, float *h2d_ms, float *kernel_ms,
                                     float *main_ms);

/* Slot-table width of the kernel instantiation a context uses (1..16). */
uint32_t warpjq_amd_kernel_width(const warpjq_ctx *ctx);

/* Bytes each wave claims per ticket in the one-pass kernel (1 KiB steps,
 * 1-8 KiB; default 8 KiB; at most the arena size minus 1 KiB). A tuning
 * knob; results do not depend on it. */
warpjq_status warpjq_amd_set_claim(warpjq_ctx *ctx, uint32_t claim_bytes);

/* LDS bytes each wave stages lines in (multiple of 256, at most 32 KiB, at
 * least the claim plus 1 KiB; default 18 KiB). Sets how many workgroups the
 * one-pass kernel keeps resident. A tuning knob; results do not depend on it.
 * No counterpart in the reference. */
warpjq_status warpjq_amd_set_arena(warpjq_ctx *ctx, uint32_t arena_bytes);

/* Workgroups (one wave each) the one-pass kernel launches: CUs times the
 * resident workgroups per CU for the context's arena. */
uint32_t warpjq_amd_grid(const warpjq_ctx *ctx);

/* What one driver run did (the fields of RunStats, exec/mod.rs:65-88, plus
 * device time). */
typedef struct {
  uint64_t bytes_in, lines_in, lines_out, malformed, type_errors;
  uint64_t gpu_fallback_lines, gpu_redone_chunks;
  double h2d_ms, kernel_ms; /* summed over chunks, hipEvent */
  uint64_t chunks;
  uint64_t err_line; /* 1-based line of a --strict failure or a declined line */
} warpjq_amd_stats;

/* Runs `query` over the bytes [data, data+n) through the device with the
 * reference GPU backend's chunk pipeline and merge (gpu/mod.rs:244-1191):
 * format 0 NDJSON / 1 CSV / 2 count-only; on_invalid 0 warn / 1 skip /
 * 2 abort (--strict); chunk_bytes caps the device chunk (the reference GPU
 * default is 64 MiB). On success *out holds the output bytes (free with
 * warpjq_amd_free). Returns 0 ok, 1 query error, 2 malformed line under
 * --strict (stats->err_line), 3 other error, 4 a line the device could not
 * decide (stats->err_line). */
int warpjq_amd_run_buffer(const char *query, const uint8_t *data, uint64_t n, int format, int on_invalid,
                          uint64_t chunk_bytes, uint64_t max_line_bytes, uint8_t **out, uint64_t *out_len,
                          warpjq_amd_stats *stats, char *err, size_t err_len);

void warpjq_amd_free(void *p);

/* Query text -> lowered program (the reference builds it in Rust:
 * query/parse.rs:59-75, query/mod.rs:95-131, gpu/lower.rs:38-160). The
 * returned handle owns the tables the view points into. NULL on a query
 * error, with the rendered message in `err`. */
typedef struct warpjq_amd_program warpjq_amd_program;
warpjq_amd_program *warpjq_amd_compile(const char *query, int csv, char *err, size_t err_len);
const warpjq_program *warpjq_amd_program_view(const warpjq_amd_program *p);
void warpjq_amd_program_free(warpjq_amd_program *p);

/* Seeded synthetic nginx-schema NDJSON (ts, host, method, path, status,
 * bytes, ua; ~220 B/line) for benchmarks: whole lines up to target_bytes,
 * generated on `threads` host threads; `pad` zero bytes follow the data.
 * status_hist[8] counts 200, 404, 500, 301, 304, 403, 502, 503. Free *out
 * with warpjq_amd_free. No counterpart in the reference (its generator,
 * crates/warpjq-core/src/gen.rs, has a different schema). */
int warpjq_amd_gen_nginx(uint64_t seed, uint64_t target_bytes, uint32_t threads, uint64_t pad, uint8_t **out,
                         uint64_t *n_bytes, uint64_t *n_lines, uint64_t *status_hist);

#ifdef __cplusplus
}
#endif

#endif /* WARPJQ_AMD_H */
