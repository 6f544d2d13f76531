// Host-side digestion of a warpjq_program into the per-slot tables and masks
// the kernels read (jscan.h DProg). Used at context creation.
#pragma once

#include <string>
#include <vector>

#include "jscan.h"

namespace wjqd {

struct DProgHost {
  std::vector<SlotPath> spath;
  std::vector<uint32_t> field_slot;
  std::vector<warpjq_cond_op> cond;  // the filter RPN the kernels run (see order_rpn)
  DProg d{};  // host pointers into the caller's warpjq_program and into this
  uint32_t ns = 1;  // kernel template width: 1, 2, 4, 8 or 16 slots
};

// The kernels keep the filter's boolean stack in one 64-bit word
// (eval_cond), so the RPN they run must never hold more than 64 values.
#define WJQ_COND_STACK 64u

// Replays an RPN filter: false when it is malformed (an operator without
// its operands, or not exactly one value left); `depth` gets the largest
// stack it reaches.
inline bool rpn_depth(const warpjq_cond_op *c, uint32_t n, uint32_t &depth) {
  uint32_t sp = 0;
  depth = 0;
  for (uint32_t k = 0; k < n; k++) {
    switch (c[k].op) {
      case WARPJQ_COND_CMP:
      case WARPJQ_COND_TRUTHY: sp++; break;
      case WARPJQ_COND_NOT:
        if (sp < 1) return false;
        break;
      case WARPJQ_COND_AND:
      case WARPJQ_COND_OR:
        if (sp < 2) return false;
        sp--;
        break;
      default: return false;
    }
    if (sp > depth) depth = sp;
  }
  return n == 0 || sp == 1;
}

// Re-emits a well-formed RPN with the deeper operand of every and/or first
// (Sethi-Ullman order), so the stack grows with the logarithm of the number
// of operands rather than with their nesting. The kernels evaluate every
// operand (no short circuit) and and/or are commutative, so the value of the
// filter for every line is unchanged.
inline void order_rpn(const warpjq_cond_op *c, uint32_t n, std::vector<warpjq_cond_op> &out) {
  struct Node {
    warpjq_cond_op op;
    int32_t a, b;
    uint32_t need;  // stack slots its evaluation takes
  };
  std::vector<Node> nodes;
  std::vector<int32_t> st;
  nodes.reserve(n);
  for (uint32_t k = 0; k < n; k++) {
    Node x{c[k], -1, -1, 1};
    if (c[k].op == WARPJQ_COND_NOT) {
      x.a = st.back();
      st.pop_back();
      x.need = nodes[x.a].need;
    } else if (c[k].op == WARPJQ_COND_AND || c[k].op == WARPJQ_COND_OR) {
      x.b = st.back();
      st.pop_back();
      x.a = st.back();
      st.pop_back();
      const uint32_t na = nodes[x.a].need, nb = nodes[x.b].need;
      x.need = na == nb ? na + 1 : (na > nb ? na : nb);
    }
    nodes.push_back(x);
    st.push_back((int32_t)nodes.size() - 1);
  }
  out.clear();
  if (st.empty()) return;
  // iterative post-order, deeper child first
  std::vector<std::pair<int32_t, int>> work{{st.back(), 0}};
  while (!work.empty()) {
    auto &[id, stage] = work.back();
    const Node &x = nodes[id];
    if (x.a < 0) {
      out.push_back(x.op);
      work.pop_back();
    } else if (x.b < 0) {
      if (stage == 0) {
        stage = 1;
        work.push_back({x.a, 0});
      } else {
        out.push_back(x.op);
        work.pop_back();
      }
    } else {
      const bool b_first = nodes[x.b].need > nodes[x.a].need;
      if (stage < 2) {
        const int32_t next = (stage == 0) == b_first ? x.b : x.a;
        stage++;
        work.push_back({next, 0});
      } else {
        out.push_back(x.op);
        work.pop_back();
      }
    }
  }
}

// Returns false (with a message) for a program the kernels cannot host.
inline bool build_dprog(const warpjq_program &p, DProgHost &h, std::string &err) {
  const uint32_t NONE = 0xFFFFFFFFu;
  if (p.n_needed > 16) {
    err = "more distinct field paths than the kernel slot table holds";
    return false;
  }
  if (p.output_kind > WARPJQ_OUT_AGG) {
    err = "unknown output kind";
    return false;
  }
  h.spath.assign(p.n_needed, SlotPath{});
  uint32_t fused = 0;
  for (uint32_t s = 0; s < p.n_needed; s++) {
    uint32_t pid = p.needed_paths[s];
    if (pid >= p.n_paths) {
      err = "needed path out of range";
      return false;
    }
    const warpjq_path &pp = p.paths[pid];
    if (pp.step_count == 0 || (uint64_t)pp.step_off + pp.step_count > p.n_steps) {
      err = "path steps out of range";
      return false;
    }
    SlotPath sp{pp.step_count, pp.step_off, 0, ~0u, {0, 0, 0, 0}};
    bool ok = pp.step_count <= WJQ_FUSED_STEPS;
    for (uint32_t k = 0; k < pp.step_count; k++) {
      const warpjq_step &st = p.steps[pp.step_off + k];
      if (st.kind == WARPJQ_STEP_INDEX && st.index >= 0xFFFFu) ok = false;
      if (st.kind == WARPJQ_STEP_KEY && (uint64_t)st.key_off + st.key_len > p.n_blob) {
        err = "key out of range";
        return false;
      }
    }
    sp.fused = ok ? 1 : 0;
    if (ok) fused |= 1u << s;
    h.spath[s] = sp;
  }
  auto slot_of = [&](uint32_t pid) -> uint32_t {
    if (pid == NONE || pid >= p.n_paths) return NONE;
    return p.slot_of_path[pid];
  };
  uint32_t filter_mask = 0;
  for (uint32_t k = 0; k < p.n_cond; k++) {
    const warpjq_cond_op &o = p.cond_rpn[k];
    uint32_t s = NONE;
    if (o.op == WARPJQ_COND_CMP) {
      if (o.arg >= p.n_cmps) {
        err = "comparison index out of range";
        return false;
      }
      s = slot_of(p.cmps[o.arg].path);
      const warpjq_cmp &c = p.cmps[o.arg];
      if (c.lit_kind == WARPJQ_LIT_STR && (uint64_t)c.lit_off + c.lit_len > p.n_blob) {
        err = "literal out of range";
        return false;
      }
    } else if (o.op == WARPJQ_COND_TRUTHY) {
      s = slot_of(o.arg);
    } else {
      continue;
    }
    if (s == NONE || s >= p.n_needed) {
      err = "filter path has no slot";
      return false;
    }
    filter_mask |= 1u << s;
  }
  uint32_t depth = 0;
  if (!rpn_depth(p.cond_rpn, p.n_cond, depth)) {
    err = "malformed filter program";
    return false;
  }
  if (depth <= WJQ_COND_STACK) {
    h.cond.assign(p.cond_rpn, p.cond_rpn + p.n_cond);
  } else {
    order_rpn(p.cond_rpn, p.n_cond, h.cond);
    if (!rpn_depth(h.cond.data(), (uint32_t)h.cond.size(), depth) || depth > WJQ_COND_STACK) {
      err = "filter needs a deeper stack than the kernels keep";
      return false;
    }
  }
  uint32_t all = p.n_needed == 32 ? ~0u : ((1u << p.n_needed) - 1);
  h.field_slot.assign(p.n_fields, 0);
  for (uint32_t f = 0; f < p.n_fields; f++) {
    uint32_t s = slot_of(p.field_paths[f]);
    if (s == NONE || s >= p.n_needed) {
      err = "projected path has no slot";
      return false;
    }
    h.field_slot[f] = s;
  }
  DProg &d = h.d;
  d.steps = p.steps;
  d.cmps = p.cmps;
  d.cond = h.cond.data();
  d.blob = p.blob;
  d.spath = h.spath.data();
  d.slot_of_path = p.slot_of_path;
  d.field_slot = h.field_slot.data();
  d.prefix_off = p.prefix_off;
  d.prefix_len = p.prefix_len;
  d.n_slots = p.n_needed;
  d.n_cond = p.n_cond;
  d.fused_mask = fused;
  uint32_t lens0 = 0;
  for (uint32_t q = 0; q < p.n_needed; q++) {
    if (!((fused >> q) & 1)) continue;
    const warpjq_step &st = p.steps[h.spath[q].step_off];
    if (st.kind == WARPJQ_STEP_KEY) lens0 |= 1u << (st.key_len < 31u ? st.key_len : 31u);
    else lens0 |= 1u << 31;
  }
  d.key_lens0 = lens0;
  // keys a minified flat line can hold verbatim, compared as words
  uint32_t plain = 0;
  for (uint32_t q = 0; q < p.n_needed; q++) {
    if (!((fused >> q) & 1)) continue;
    const warpjq_step &st = p.steps[h.spath[q].step_off];
    if (st.kind != WARPJQ_STEP_KEY || st.key_len > 16) continue;
    bool ok = true;
    for (uint32_t k = 0; k < st.key_len; k++) {
      const uint8_t c = p.blob[st.key_off + k];
      if (c == '"' || c == '\\' || c < 0x20) ok = false;
    }
    if (!ok) continue;
    SlotPath &sp = h.spath[q];
    sp.klen = st.key_len;
    for (uint32_t k = 0; k < st.key_len; k++) sp.kw[k >> 2] |= (uint32_t)p.blob[st.key_off + k] << (8 * (k & 3));
    plain |= 1u << q;
  }
  d.plain_keys = plain;
  d.filter_mask = filter_mask;
  d.post_mask = all & ~filter_mask;
  d.has_filter = p.has_filter;
  d.output_kind = p.output_kind;
  d.output_slot = p.output_kind == WARPJQ_OUT_PATH ? slot_of(p.output_path) : NONE;
  if (p.output_kind == WARPJQ_OUT_PATH && d.output_slot == NONE) {
    err = "output path has no slot";
    return false;
  }
  d.agg_kind = p.agg_kind;
  d.agg_slot = p.output_kind == WARPJQ_OUT_AGG ? slot_of(p.agg_path) : NONE;
  d.group_slot = p.output_kind == WARPJQ_OUT_AGG ? slot_of(p.group_path) : NONE;
  d.n_fields = p.n_fields;
  d.suffix_off = p.suffix_off;
  d.suffix_len = p.suffix_len;
  d.csv_mode = p.csv_mode;
  uint32_t n = p.n_needed;
  h.ns = n <= 1 ? 1 : n <= 2 ? 2 : n <= 4 ? 4 : n <= 8 ? 8 : 16;
  return true;
}

}  // namespace wjqd
