// Query front end for the jq subset warpjq accepts: text -> Program -> the
// flat tables of warpjq_program.
//
// Semantics follow the reference's front end:
//   lexer   crates/warpjq-core/src/query/lex.rs:56-317
//   parser  crates/warpjq-core/src/query/parse.rs:59-631
//   IR      crates/warpjq-core/src/query/ir.rs:11-252
//   tables  crates/warpjq-core/src/query/mod.rs:15-198
//   lower   crates/warpjq-core/src/gpu/lower.rs:38-160
// This is host code that runs once per query; it is shared by the GPU host
// driver and by the CPU restatement used to check it.
#pragma once

#include <cstdint>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "../../include/warpjq_kernels.h"

namespace wjq {

using PathId = uint16_t;

struct Step {
  bool is_index = false;
  std::string key;     // decoded UTF-8, for keys
  uint32_t index = 0;  // for indices
  bool operator==(const Step &o) const {
    return is_index == o.is_index && key == o.key && index == o.index;
  }
};

struct Path {
  std::vector<Step> steps;
  bool is_identity() const { return steps.empty(); }
  bool operator==(const Path &o) const { return steps == o.steps; }
  // jq spelling: `.`, `.a.b[0]`, `."odd key"` (ir.rs:32-46).
  std::string display() const;
};

enum CmpOp : uint32_t { OP_EQ = 0, OP_NE = 1, OP_LT = 2, OP_LE = 3, OP_GT = 4, OP_GE = 5 };

// jq's three-way result applied to an operator (ir.rs:66-76). `ord` is -1/0/1.
inline bool cmp_accepts(CmpOp op, int ord) {
  switch (op) {
    case OP_EQ: return ord == 0;
    case OP_NE: return ord != 0;
    case OP_LT: return ord < 0;
    case OP_LE: return ord <= 0;
    case OP_GT: return ord > 0;
    case OP_GE: return ord >= 0;
  }
  return false;
}
const char *cmp_str(CmpOp op);

struct Literal {
  enum Kind { NUL, BOOL, NUM, STR } kind = NUL;
  bool b = false;
  double num = 0.0;
  std::string str;  // decoded
  int rank() const {  // ir.rs:100-108
    switch (kind) {
      case NUL: return 0;
      case BOOL: return b ? 2 : 1;
      case NUM: return 3;
      case STR: return 4;
    }
    return 0;
  }
};

struct Cond {
  enum Type { CMP, TRUTHY, AND, OR, NOT } type = CMP;
  PathId path = 0;
  CmpOp op = OP_EQ;
  Literal lit;
  std::unique_ptr<Cond> a, b;
  // ir.rs:128-137 — every path the condition touches, left to right.
  template <class F>
  void visit_paths(F &&f) const {
    switch (type) {
      case CMP:
      case TRUTHY: f(path); break;
      case AND:
      case OR: a->visit_paths(f); b->visit_paths(f); break;
      case NOT: a->visit_paths(f); break;
    }
  }
};

enum class AggKind : uint32_t { Count = 0, Sum = 1, Min = 2, Max = 3, Avg = 4 };
const char *agg_name(AggKind k);

struct Output {
  enum Kind { PASSTHROUGH, PATH, PROJECT, AGG } kind = PASSTHROUGH;
  PathId path = 0;                                     // PATH
  std::vector<std::pair<std::string, PathId>> fields;  // PROJECT
  AggKind agg = AggKind::Count;                        // AGG
  std::optional<PathId> arg;                           // AGG
};

struct Program {
  std::vector<Path> paths;
  std::unique_ptr<Cond> filter;
  std::optional<PathId> group_by;
  Output output;
  std::string source;

  const Path &path(PathId id) const { return paths[id]; }
  bool is_aggregate() const { return output.kind == Output::AGG; }
  std::vector<PathId> required_paths() const;                   // ir.rs:206-234
  std::optional<std::vector<std::string>> csv_headers() const;  // ir.rs:237-251
};

struct QueryError {
  std::string message;
  size_t offset = 0;
  std::string hint;
  std::string render(const std::string &src) const;
};

// Returns false and fills `err` on rejection.
bool parse(const std::string &src, Program &out, QueryError &err);

// FNV-1a 32 (query/mod.rs:15-22).
inline uint32_t key_hash(const uint8_t *p, size_t n) {
  uint32_t h = 0x811c9dc5u;
  for (size_t i = 0; i < n; i++) {
    h ^= p[i];
    h *= 0x01000193u;
  }
  return h;
}

// The flat tables plus the projection prefix blob (query/mod.rs:81-198 and
// gpu/lower.rs:17-127). `view()` returns the C struct pointing into `this`.
struct Lowered {
  std::vector<warpjq_step> steps;
  std::vector<warpjq_path> paths;
  std::vector<warpjq_cmp> cmps;
  std::vector<warpjq_cond_op> cond_rpn;
  std::vector<uint8_t> blob;
  uint32_t cond_stack_depth = 0;
  std::vector<uint32_t> needed_paths, slot_of_path, field_paths, prefix_off, prefix_len;
  uint32_t suffix_off = 0, suffix_len = 0;
  uint32_t output_kind = WARPJQ_OUT_PASSTHROUGH, output_path = 0xFFFFFFFFu;
  uint32_t agg_kind = WARPJQ_AGG_COUNT, agg_path = 0xFFFFFFFFu, group_path = 0xFFFFFFFFu;
  uint32_t has_filter = 0, csv_mode = 0;
  warpjq_program view() const;
};

void lower(const Program &p, bool csv, Lowered &out);

// JSON string escaping of output.rs:24-42, shared by lowering and the host
// writer.
void write_json_string(std::string &out, const uint8_t *p, size_t n);

}  // namespace wjq
