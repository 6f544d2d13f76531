// Internal interface of the host driver (host.cpp), shared with the
// serving loop (serve.cpp). Not exported.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/warpjq_amd.h"
#include "../../include/warpjq_kernels.h"
#include "query.h"

namespace wjqd {

enum Format { F_NDJSON = 0, F_CSV = 1, F_COUNT = 2 };
enum OnInvalid { OI_WARN = 0, OI_SKIP = 1, OI_ABORT = 2 };

struct Chunk {
  uint64_t off, len;
};

// One submit's input: newline-complete pieces in stream order (a chunk of
// one buffer, or several small files packed into one slot).
struct Piece {
  const uint8_t *p;  // host bytes (always: overflow splits and row cuts need them)
  uint64_t len;
  const uint8_t *d = nullptr;  // the same bytes resident in device memory, or null
  bool big = false;  // one line longer than the engine's chunk: run alone on a context sized for it
};
using Batch = std::vector<Piece>;

// Where a run's batches come from. engine_run sets cap / max_lines from the
// context before the first next(); the bytes of a batch stay valid until
// the batch after the next one has been taken (a run keeps at most two
// batches on the device and redoes an overflowed one before taking more).
struct Source {
  uint64_t cap = 0, max_lines = 0;
  const uint8_t *const *dfiles = nullptr;  // device copies of the inputs, or null
  virtual ~Source() = default;
  // false at the end of the input, or on an error (err set)
  virtual bool next(Batch &b, std::string &err) = 0;
};

// Batches over a list of in-memory inputs read as one stream, the way the
// reference chains files (chunk.rs:46-57, :198-203): a chunk never crosses
// an input, each input's BOM is dropped (chunk.rs:210-212) and its last line
// is complete at its end. Whole small inputs share a batch; a large one is
// cut into newline-aligned chunks (gpu/mod.rs:697-810). A line longer than
// the chunk but within max_line comes alone, as a batch of one `big` piece
// (gpu/mod.rs:411-458, :824-857).
struct Batcher : Source {
  const uint8_t *const *files;
  const uint64_t *lens;
  uint32_t n_files;
  uint64_t max_line;
  uint32_t max_pieces;
  uint32_t fi = 0;
  uint64_t pos = 0;
  bool started = false;

  Batcher(const uint8_t *const *files_, const uint64_t *lens_, uint32_t n_files_, uint64_t max_line_,
          uint32_t max_pieces_)
      : files(files_), lens(lens_), n_files(n_files_), max_line(max_line_), max_pieces(max_pieces_) {}
  bool next(Batch &b, std::string &err) override;
};

// Batches over files named by path (Input::open / Input::stdin / Input::
// chain, chunk.rs:60-205): regular files are mapped read-only and batched as
// above; "-", pipes and other unmappable inputs are read as a stream, the
// partial last line carried into the next read (chunk.rs:292-387). One
// stream over the whole list: chunks never cross an input.
struct PathSource : Source {
  explicit PathSource(uint64_t max_line);
  ~PathSource() override;
  // Opens every path (nullptr or "-" = stdin). False with err set when one
  // cannot be opened or mapped; nothing is read yet.
  bool open(const char *const *paths, uint32_t n, std::string &err);
  bool next(Batch &b, std::string &err) override;

 private:
  struct Part;
  std::vector<Part *> parts_;
  size_t cur_ = 0;
  uint64_t max_line_;
};

// A lowered query with its device context, reusable run after run (one
// thread at a time).
struct Engine {
  wjq::Program prog;
  wjq::Lowered low;
  warpjq_program view{};
  warpjq_ctx *ctx = nullptr;
  uint64_t cap = 0, max_lines = 0;
  bool csv = false;
  // largest chunk engine_open accepts: the reference's 64 MiB device chunk
  // (gpu/mod.rs:57); the served form, whose inputs are already resident,
  // may raise it to cut launches per request
  uint64_t cap_limit = 64ull << 20;
};

// 0 ok, 1 query error, 3 other error (message in err).
int engine_open(Engine &g, const char *query, int format, uint64_t chunk_bytes, char *err, size_t err_len);
// Points an open engine at another query whose output needs the same device
// buffers, keeping its context: 0 ok, 1 query error, 2 not compatible (open
// another engine), 3 other error.
int engine_reprogram(Engine &g, const char *query, int format, char *err, size_t err_len);
void engine_close(Engine &g);
// Runs every batch `src` yields; the return codes of warpjq_amd_run_buffer.
int engine_run(Engine &g, Source &src, int format, int on_invalid, uint8_t **out, uint64_t *out_len,
               warpjq_amd_stats *stats, char *err, size_t err_len);

// Newline indices of resident datasets (warpjq_device.hip): newlines per
// 64-byte block of the dataset, one byte each, kept beside it in device
// memory. Row outputs over a span inside a registered dataset count each
// claim's lines from the index instead of reading the whole claim.
uint64_t nlindex_bytes(uint64_t n);

// Host bytes from malloc, freed with the object.
struct HostBytes {
  uint8_t *p = nullptr;
  uint64_t n = 0;
  HostBytes() = default;
  HostBytes(const HostBytes &) = delete;
  HostBytes &operator=(const HostBytes &) = delete;
  ~HostBytes() { free(p); }
};

// CSV input (csv.hip): the NDJSON the records of data[0..n) read as, built
// on the device (the header as a blank line), in device memory (*dev_out,
// 256 readable bytes past its end) and copied to the host (out);
// *records = lines after the header; *kernel_ms = the transcoding kernels'
// time. The caller frees *dev_out with csv_free, also when false is
// returned. Calls from several threads take turns.
bool csv_transcode(const uint8_t *data, uint64_t n, HostBytes &out, uint8_t **dev_out, uint64_t *records,
                   float *kernel_ms, std::string &err);
void csv_free(uint8_t *dev);
// Builds the index of d[0..n) into idx (device memory, nlindex_bytes(n)
// bytes); waits for it.
bool nlindex_build(const uint8_t *d, uint64_t n, uint8_t *idx, std::string &err);
void nlindex_register(const uint8_t *d, uint64_t n, const uint8_t *idx);
void nlindex_unregister(const uint8_t *d);
// The index of the registered dataset holding d[0..n), and d's offset in it
// (null when none does).
void nlindex_find(const uint8_t *d, uint64_t n, const uint8_t **idx, uint64_t *off);

}  // namespace wjqd
