/* warpjq device ABI, implemented for AMD Instinct MI355X (gfx950).
 *
 * This header declares the exact C boundary that warpjq's host code binds
 * (crates/warpjq-core/cuda/warpjq_kernels.h:1-280 in the reference, mirrored
 * by crates/warpjq-core/src/gpu/ffi.rs:47-184). Every struct keeps the
 * reference's field order, widths and padding so that the caller's
 * warpjq_abi_check (ffi.rs:222-242) sees sizes 20/8/32/8/48/56 and a
 * `warpjq_chunk_result` filled by this library reads identically on the
 * caller's side.
 *
 * Conventions (unchanged from the reference):
 *  - every pointer passed in is borrowed for the duration of the call;
 *  - the context owns all device memory and all pinned host memory;
 *  - result pointers stay valid until the next submit on the same slot;
 *  - errors are a warpjq_status plus "op: detail" text in `err`;
 *  - capacity problems are not errors: chunk_overflow / group_overflow ask
 *    the host to redo the chunk, FALLBACK lines ask it to redo single lines.
 */
#ifndef WARPJQ_KERNELS_H
#define WARPJQ_KERNELS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes (kernels.h:25-32). WARPJQ_ERR_CUDA keeps its name for ABI
 * compatibility; here it reports a HIP runtime error. */
typedef enum {
  WARPJQ_OK = 0,
  WARPJQ_ERR_CUDA = 1,
  WARPJQ_ERR_NO_DEVICE = 2,
  WARPJQ_ERR_OOM = 3,
  WARPJQ_ERR_ABI = 4,
  WARPJQ_ERR_INVALID_ARG = 5
} warpjq_status;

/* Per-line outcome (kernels.h:35-41). */
typedef enum {
  WARPJQ_LINE_OK = 0,
  WARPJQ_LINE_INVALID = 1,
  WARPJQ_LINE_TYPE_ERROR = 2,
  WARPJQ_LINE_FALLBACK = 3,
  WARPJQ_LINE_BLANK = 4
} warpjq_line_status;

/* Program-table enums (kernels.h:45-86; query/mod.rs:24-37, :202-207). */
enum { WARPJQ_STEP_KEY = 0, WARPJQ_STEP_INDEX = 1 };
enum {
  WARPJQ_LIT_NULL = 0,
  WARPJQ_LIT_FALSE = 1,
  WARPJQ_LIT_TRUE = 2,
  WARPJQ_LIT_NUM = 3,
  WARPJQ_LIT_STR = 4
};
enum {
  WARPJQ_COND_CMP = 0,
  WARPJQ_COND_TRUTHY = 1,
  WARPJQ_COND_AND = 2,
  WARPJQ_COND_OR = 3,
  WARPJQ_COND_NOT = 4
};
enum {
  WARPJQ_OP_EQ = 0,
  WARPJQ_OP_NE = 1,
  WARPJQ_OP_LT = 2,
  WARPJQ_OP_LE = 3,
  WARPJQ_OP_GT = 4,
  WARPJQ_OP_GE = 5
};
enum {
  WARPJQ_OUT_PASSTHROUGH = 0,
  WARPJQ_OUT_PATH = 1,
  WARPJQ_OUT_PROJECT = 2,
  WARPJQ_OUT_AGG = 3
};
enum {
  WARPJQ_AGG_COUNT = 0,
  WARPJQ_AGG_SUM = 1,
  WARPJQ_AGG_MIN = 2,
  WARPJQ_AGG_MAX = 3,
  WARPJQ_AGG_AVG = 4
};

/* One path hop (kernels.h:88-94; GpuStep query/mod.rs:39-47). 20 bytes. */
typedef struct {
  uint32_t kind;
  uint32_t index;
  uint32_t key_off;
  uint32_t key_len;
  uint32_t key_hash; /* FNV-1a 32 of the decoded key */
} warpjq_step;

/* A path is a run of steps (kernels.h:96-99). 8 bytes. */
typedef struct {
  uint32_t step_off;
  uint32_t step_count;
} warpjq_path;

/* `.path <op> literal` (kernels.h:101-109). 32 bytes. */
typedef struct {
  uint32_t path;
  uint32_t op;
  uint32_t lit_kind;
  uint32_t lit_off;
  uint32_t lit_len;
  uint32_t _pad;
  double lit_num;
} warpjq_cmp;

/* One RPN instruction of the filter (kernels.h:111-114). 8 bytes. */
typedef struct {
  uint32_t op;
  uint32_t arg;
} warpjq_cond_op;

/* The lowered query (kernels.h:117-153; built by gpu/lower.rs:38-160). */
typedef struct {
  const warpjq_step *steps;
  uint32_t n_steps;
  const warpjq_path *paths;
  uint32_t n_paths;
  const warpjq_cmp *cmps;
  uint32_t n_cmps;
  const warpjq_cond_op *cond_rpn;
  uint32_t n_cond;
  const uint8_t *blob;
  uint32_t n_blob;

  const uint32_t *needed_paths;
  uint32_t n_needed;
  const uint32_t *slot_of_path;

  uint32_t output_kind;
  uint32_t output_path;
  uint32_t agg_kind;
  uint32_t agg_path;
  uint32_t group_path;
  uint32_t has_filter;

  uint32_t n_fields;
  const uint32_t *field_paths;
  const uint32_t *prefix_off;
  const uint32_t *prefix_len;
  uint32_t suffix_off;
  uint32_t suffix_len;
  uint32_t csv_mode;
} warpjq_program;

/* Ungrouped aggregate partial for one chunk (kernels.h:159-167). 48 bytes. */
typedef struct {
  uint64_t count;
  uint64_t numeric;
  double sum;
  double min;
  double max;
  uint32_t saw_non_numeric;
  uint32_t _pad;
} warpjq_agg_partial;

/* One occupied group (kernels.h:170-180). key_off is chunk-relative; a
 * string key's bytes exclude the quotes. 56 bytes. */
typedef struct {
  uint32_t key_off;
  uint32_t key_len;
  uint32_t key_kind;
  uint32_t _pad;
  uint64_t count;
  uint64_t numeric;
  double sum;
  double min;
  double max;
} warpjq_group;

/* Everything one chunk produced (kernels.h:182-230). */
typedef struct {
  uint64_t n_lines; /* includes blank lines */
  uint64_t n_blank;
  uint64_t n_invalid;
  uint64_t n_type_error;
  uint64_t n_fallback;
  uint64_t n_selected;

  const uint8_t *out_bytes;
  uint64_t out_len;
  const uint32_t *out_line_idx; /* ascending */
  const uint64_t *out_row_off;  /* n_selected + 1 entries */

  const uint32_t *fallback_idx; /* unsorted */
  const uint32_t *fallback_off;
  const uint32_t *fallback_len;
  uint32_t chunk_overflow;
  uint32_t first_invalid; /* lowest invalid line index, or 0xFFFFFFFF */

  warpjq_agg_partial agg;

  const warpjq_group *groups;
  uint32_t n_groups;
  uint32_t group_overflow;
} warpjq_chunk_result;

typedef struct warpjq_ctx warpjq_ctx;

/* kernels.h:238 — fails when no usable gfx950 device is present. */
warpjq_status warpjq_probe(char *err, size_t err_len);

/* kernels.h:241 — compares the caller's six struct sizes with ours. */
warpjq_status warpjq_abi_check(const uint64_t *sizes, size_t n);

/* kernels.h:243-245 — uploads the program, allocates n_slots slots of
 * max_chunk_bytes pinned staging plus device buffers. */
warpjq_status warpjq_ctx_create(const warpjq_program *prog,
                                uint64_t max_chunk_bytes, uint32_t n_slots,
                                warpjq_ctx **out, char *err, size_t err_len);

/* kernels.h:247 */
void warpjq_ctx_destroy(warpjq_ctx *ctx);

/* kernels.h:250-251 — pinned staging buffer of `slot`, and its size. */
uint8_t *warpjq_slot_buffer(warpjq_ctx *ctx, uint32_t slot);
uint64_t warpjq_slot_capacity(const warpjq_ctx *ctx);

/* kernels.h:260 — most lines one submitted chunk may hold. */
uint64_t warpjq_max_lines(const warpjq_ctx *ctx);

/* kernels.h:265-266 — queues H2D plus kernels for `slot`; asynchronous. */
warpjq_status warpjq_submit(warpjq_ctx *ctx, uint32_t slot, uint64_t n_bytes,
                            char *err, size_t err_len);

/* kernels.h:270-271 — blocks until `slot` is done and fills `out`. */
warpjq_status warpjq_wait(warpjq_ctx *ctx, uint32_t slot,
                          warpjq_chunk_result *out, char *err, size_t err_len);

/* kernels.h:274 */
warpjq_status warpjq_device_name(char *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* WARPJQ_KERNELS_H */
