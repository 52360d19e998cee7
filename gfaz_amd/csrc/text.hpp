// P- and W-line step text: formatting decoded traversals and parsing step lists.
#pragma once

#include "ctx.hpp"

#include <vector>

namespace gfz {

constexpr int kTextP = 0;  // "12+,13-"
constexpr int kTextW = 1;  // ">12<13"

// Layout of the formatted lines: every step's and line's byte position.
struct FormatPlan {
  int kind = 0;
  const int32_t *nodes = nullptr;
  uint64_t n = 0, n_trav = 0, n_tiles = 0, total = 0;
  DBuf<uint32_t> heads;       // traversal-start bitmap
  DBuf<uint64_t> tstart;      // n_trav + 1 step offsets
  DBuf<uint64_t> tile_off;    // step-text bytes before each tile (n_tiles + 1)
  DBuf<uint64_t> tpos;        // step-text bytes before each traversal (n_trav + 1)
  std::vector<uint64_t> line_off, head_len;  // host
};

// Sizes the lines head_t + steps_t + tail_t of traversals d_nodes (device;
// lengths host, n_trav) with heads/tails of the given offsets (host, n_trav + 1
// each); plan.total is the text size. d_nodes must stay valid for format_write.
void format_plan(Ctx &ctx, int kind, const int32_t *d_nodes, uint64_t n,
                 const uint32_t *lengths, uint64_t n_trav, const uint64_t *head_off,
                 const uint64_t *tail_off, FormatPlan &plan);
// Writes the planned lines back to back into d_out (device, out_cap bytes);
// heads/tails: host bytes. Synchronises the ctx stream.
void format_write(Ctx &ctx, const FormatPlan &plan, const uint8_t *heads_txt,
                  const uint64_t *head_off, const uint8_t *tails_txt, const uint64_t *tail_off,
                  uint8_t *d_out, uint64_t out_cap);

struct ParsePlan {
  int kind = 0;
  const uint8_t *txt = nullptr;
  uint64_t n_bytes = 0, n_lines = 0, n_tiles = 0, total = 0;
  DBuf<uint64_t> fb, fe;     // field begins and ends (n_lines each)
  DBuf<uint64_t> tile_pos;   // kept steps before each tile (n_tiles + 1)
  DBuf<uint64_t> first;      // first field beginning in or after each tile (n_tiles + 1)
};

// Counts the steps of the step fields d_txt[field_begin[l] .. field_end[l]) (device
// text; host ranges, ordered and disjoint; bytes outside them are ignored):
// plan.total, and per line in lengths (host, n_lines, may be null). Throws the
// reference's parse_uint32 messages on a malformed name. d_txt must stay valid for
// parse_write.
void parse_plan(Ctx &ctx, int kind, const uint8_t *d_txt, uint64_t n_bytes,
                const uint64_t *field_begin, const uint64_t *field_end, uint64_t n_lines,
                uint32_t *lengths, ParsePlan &plan);
// Writes the planned steps into d_nodes (device, nodes_cap). Synchronises.
void parse_write(Ctx &ctx, const ParsePlan &plan, int32_t *d_nodes, uint64_t nodes_cap);

}  // namespace gfz
