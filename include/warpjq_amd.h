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

/* One span of a batched submit: bytes already in device memory, newline
 * complete (a file, or a newline-aligned piece of one), with at least 256
 * readable bytes past n_bytes. */
typedef struct {
  const uint8_t *d_bytes;
  uint64_t n_bytes;
} warpjq_amd_span;

/* The batched form: every span of `spans` in ONE persistent launch whose
 * waves take claims from any span (work stealing), as one stream in span
 * order, the way the reference chains several inputs
 * (crates/warpjq-core/src/chunk.rs:46-57, :198-203): chunks never cross a
 * span, and a span's last line ends at its end. Results are those of one
 * chunk holding the spans in order (line indices continue across spans).
 * At most 65,536 spans whose sizes, each rounded up to the claim size, sum to
 * at most 4 GiB - 1 MiB; row outputs also within the slot capacity. After
 * warpjq_wait, fallback offsets are relative to their span
 * (warpjq_amd_fallback_spans gives the span of each), and groups[].key_off
 * indexes warpjq_amd_group_keys (16 MiB of key bytes per submit; more sets
 * group_overflow). No counterpart in the reference. */
warpjq_status warpjq_amd_submit_spans(warpjq_ctx *ctx, uint32_t slot, const warpjq_amd_span *spans,
                                      uint32_t n_spans, char *err, size_t err_len);
/* The batched form over the slot's pinned buffer: copies its first n_bytes
 * to the device (as warpjq_submit does) and runs every piece (off, len
 * within those bytes; newline complete) in one launch, with the semantics of
 * warpjq_amd_submit_spans. For a host that packs many small files into one
 * slot. No counterpart in the reference. */
typedef struct {
  uint64_t off;
  uint64_t len;
} warpjq_amd_piece;
warpjq_status warpjq_amd_submit_pieces(warpjq_ctx *ctx, uint32_t slot, uint64_t n_bytes,
                                       const warpjq_amd_piece *pieces, uint32_t n_pieces, char *err,
                                       size_t err_len);

/* Copies n host bytes into a new device allocation followed by 256 newline
 * bytes (so it can be submitted resident), synchronously; NULL on failure.
 * Free with warpjq_amd_device_free. No counterpart in the reference. */
uint8_t *warpjq_amd_device_upload(const uint8_t *bytes, uint64_t n, char *err, size_t err_len);
void warpjq_amd_device_free(uint8_t *p);

/* After warpjq_wait on a batched submit: the span (piece) index of each
 * fallback line, parallel to fallback_idx; and the key bytes that
 * groups[].key_off indexes. Valid until the next submit on the slot. No
 * counterpart in the reference. */
const uint32_t *warpjq_amd_fallback_spans(const warpjq_ctx *ctx, uint32_t slot);

/* After warpjq_wait: the chunk-relative index of the first line counted as a
 * type error (n_type_error), 0xFFFFFFFF when none, the way first_invalid
 * reports malformed lines. A strict run stops at the lower of the two, as
 * the reference CPU backend does (crates/warpjq-core/src/exec/cpu.rs:401-405);
 * the reference's device ABI has no such field. */
uint32_t warpjq_amd_first_type_error(const warpjq_ctx *ctx, uint32_t slot);
const uint8_t *warpjq_amd_group_keys(const warpjq_ctx *ctx, uint32_t slot);

/* Device time of the last submit on `slot` (hipEvent on the slot's stream):
 * the H2D copy, the whole kernel chain, and the one-pass main kernel alone.
 * Valid after warpjq_wait. No counterpart in the reference. */
warpjq_status warpjq_amd_slot_timing(warpjq_ctx *ctx, uint32_t slot, float *h2d_ms, float *kernel_ms,
                                     float *main_ms);

/* Points an idle context (no slot in flight) at another program whose
 * output needs the same buffers (both grouped, both ungrouped aggregates, or
 * both row outputs): uploads its tables, keeps every allocation. Lets a
 * server reuse contexts across queries that differ only in constants. On
 * error the old program stays. No counterpart in the reference. */
warpjq_status warpjq_amd_ctx_reprogram(warpjq_ctx *ctx, const warpjq_program *prog, char *err, size_t err_len);

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

/* Autotuner: times the one-pass kernel on the first 1 GiB of the resident
 * bytes (the slot capacity for row outputs) under a dozen claim / arena
 * pairs (the bytes per ticket and the lines each wave stages at once), and
 * for grouped contexts then under group-table sizes from 2^14 slots up
 * (warpjq_amd_group_table reads the choice); keeps the fastest for this
 * context and returns the claim / arena pair with its time.
 * Uses slot 0, which must be idle. Results do not depend on the choice. No
 * counterpart in the reference (its tunables are compile-time constants,
 * kernels.cu:52-89). */
warpjq_status warpjq_amd_autotune(warpjq_ctx *ctx, const uint8_t *d_bytes, uint64_t n_bytes, uint32_t *claim_out,
                                  uint32_t *arena_out, float *best_ms_out, char *err, size_t err_len);

/* Group table slots in use for a grouped context (a power of two, 4,096 up
 * to what the context allocated, 2^16..2^22 by chunk size). Fewer slots keep
 * probes of low-cardinality keys in fewer cache lines; a chunk with more
 * groups than fit sets group_overflow and the host splits it, so results do
 * not depend on it. warpjq_amd_autotune picks it for grouped contexts. No
 * counterpart in the reference (its table is a fixed 65,536 slots,
 * kernels.cu:57-60). */
warpjq_status warpjq_amd_set_group_table(warpjq_ctx *ctx, uint32_t slots);
uint32_t warpjq_amd_group_table(const warpjq_ctx *ctx);

/* Keep up to `per_slot` captured kernel chains (hipGraph) per slot: a
 * resident submit (warpjq_amd_submit_device / _spans) whose span table
 * matches an earlier one under the same program, claim and arena replays it
 * with one launch. 0 (the default) turns capture off. A serving knob for
 * queries repeated over the same data; results do not depend on it. No
 * counterpart in the reference. */
warpjq_status warpjq_amd_set_graphs(warpjq_ctx *ctx, uint32_t per_slot);

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

/* The same over several inputs read as one stream, as the reference reads a
 * list of files (Input::Chained, chunk.rs:46-57, :198-203): aggregates total
 * across files, a chunk never crosses a file, each file's BOM is dropped and
 * its last line is complete at its end, line numbers continue. Small files
 * are packed into one slot and run as one launch (warpjq_amd_submit_pieces).
 * Same arguments and return codes as warpjq_amd_run_buffer. */
int warpjq_amd_run_files(const char *query, const uint8_t *const *files, const uint64_t *lens, uint32_t n_files,
                         int format, int on_invalid, uint64_t chunk_bytes, uint64_t max_line_bytes, uint8_t **out,
                         uint64_t *out_len, warpjq_amd_stats *stats, char *err, size_t err_len);

/* The same over a CSV input, which the device transcodes into NDJSON
 * records first: the header line names the columns and reads as a blank
 * line; every later line is the object {"name":value,...} (an empty unquoted
 * field null, an unquoted JSON number a number, anything else a string; ""
 * inside quotes is a quote; a trailing CR dropped); a line with bad quoting
 * or another field count than the header's is one malformed line. No
 * counterpart in the reference, which reads NDJSON only. Same arguments and
 * return codes as warpjq_amd_run_buffer; stats->kernel_ms includes the
 * transcoding kernels. */
int warpjq_amd_run_csv(const char *query, const uint8_t *data, uint64_t n, int format, int on_invalid,
                       uint64_t chunk_bytes, uint64_t max_line_bytes, uint8_t **out, uint64_t *out_len,
                       warpjq_amd_stats *stats, char *err, size_t err_len);

/* The same over files named by path, as the reference's CLI opens its inputs
 * (Input::open / Input::stdin / Input::chain, chunk.rs:60-205; run.rs:67-83):
 * regular files are mapped read-only, "-" (or n_paths == 0) reads stdin and
 * pipes are read as a stream with the partial last line carried into the
 * next read (stream_chunks, chunk.rs:292-387); empty files add nothing. All
 * paths are one stream, as in warpjq_amd_run_files. Returns 3 with "could
 * not open `path`: ..." when a path cannot be opened, else the codes of
 * warpjq_amd_run_buffer. */
int warpjq_amd_run_paths(const char *query, const char *const *paths, uint32_t n_paths, int format, int on_invalid,
                         uint64_t chunk_bytes, uint64_t max_line_bytes, uint8_t **out, uint64_t *out_len,
                         warpjq_amd_stats *stats, char *err, size_t err_len);

void warpjq_amd_free(void *p);

/* The served form. A server keeps datasets resident in HBM within
 * `hbm_budget` bytes (loaded on first use, least recently used evicted when
 * no running request holds them) and runs queued requests on `workers`
 * threads, each with its own device contexts and streams: requests overlap on
 * the GPU and a worker takes the next admissible request as soon as it
 * finishes one. A request's result is exactly warpjq_amd_run_files' over its
 * datasets in order. No counterpart in the reference. */
typedef struct warpjq_amd_server warpjq_amd_server;
typedef struct {
  uint64_t budget, reserved, loads, evictions, completed, resident;
  double p50_ms, p99_ms, max_ms; /* arrival to result, completed requests */
  uint64_t engines_opened, engines_reprogrammed; /* contexts created / reused for another query */
} warpjq_amd_server_stats;

/* chunk_bytes: the device chunk of the server's contexts (0: 64 MiB, at most
 * 1 GiB); datasets are resident, so chunks past the reference's 64 MiB only
 * cut launches per request. Each dataset keeps a newline index beside it in
 * device memory (one byte per 64-byte block, counted in the budget) from
 * which row requests count their lines. */
warpjq_amd_server *warpjq_amd_server_create(uint64_t hbm_budget, uint32_t workers, uint64_t chunk_bytes, char *err,
                                            size_t err_len);
void warpjq_amd_server_destroy(warpjq_amd_server *s);
/* Registers host bytes that stay valid and unchanged for the server's life
 * (they back row chunking and overflow splits). Returns the dataset id. */
int64_t warpjq_amd_server_add_dataset(warpjq_amd_server *s, const uint8_t *bytes, uint64_t n);
/* Queues `query` over the datasets, in order, read as one stream. Returns
 * the request id, or -1 on a bad argument. */
int64_t warpjq_amd_server_submit(warpjq_amd_server *s, const char *query, const uint32_t *datasets,
                                 uint32_t n_datasets, int format, int on_invalid);
/* Blocks until request `id` is done; then as warpjq_amd_run_buffer returns,
 * plus its latency from submit to result. */
int warpjq_amd_server_wait(warpjq_amd_server *s, int64_t id, uint8_t **out, uint64_t *out_len,
                           warpjq_amd_stats *stats, double *latency_ms, char *err, size_t err_len);
void warpjq_amd_server_get_stats(warpjq_amd_server *s, warpjq_amd_server_stats *out);

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

/* The same for the nested schema (~330 B/line, objects six deep, an escaped
 * key, .user drawn from 1M values with Zipf s = 1.1, .bytes an integer on
 * 70% of lines / a %.17g double on 28% / null on 2%, .resp.latency; 0.1% of
 * lines malformed, counted in *n_malformed). No counterpart in the reference. */
int warpjq_amd_gen_nested(uint64_t seed, uint64_t target_bytes, uint32_t threads, uint64_t pad, uint8_t **out,
                          uint64_t *n_bytes, uint64_t *n_lines, uint64_t *n_malformed);

/* An ungrouped aggregate (count, sum, min, max, avg, with any filter) over
 * n_bytes already in device memory that allocates nothing: the result of
 * warpjq_ctx_create + warpjq_amd_submit_device + warpjq_wait for the same
 * program and bytes, with every device buffer those kernels use besides the
 * input kept in a fixed array that each GPU holds for the life of the
 * process (8 MiB; warpjq_amd_noalloc_max_bytes() gives the largest input it
 * serves, about 2.7 GiB; larger inputs are refused with
 * WARPJQ_ERR_INVALID_ARG). No device memory is allocated or freed per call;
 * the host side only builds the program's small tables. Calls on one GPU
 * share that array and must not overlap (concurrent calls race); the
 * caller's device must be current. Declined lines are counted in
 * n_fallback without their offsets. `ext` (optional) receives the exact sum
 * parts. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * hosts that forbid allocation after start-up). */
uint64_t warpjq_amd_noalloc_max_bytes(void);
warpjq_status warpjq_amd_aggregate_device_noalloc(const warpjq_program *prog, const uint8_t *d_bytes, uint64_t n_bytes,
                                                  warpjq_chunk_result *out, warpjq_amd_agg_ext *ext, char *err,
                                                  size_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* WARPJQ_AMD_H */
