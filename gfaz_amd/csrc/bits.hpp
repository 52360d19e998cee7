// Bit-parallel greedy pair selection, shared by the device kernels and the
// host-side unit test (compiled with g++ there, so no HIP types here).
//
// The reference encoder (src/grammar/path_encoder.cpp:29-83) scans a traversal
// left to right with a stack and replaces the top two symbols whenever they form
// a rule. Because a freshly made rule symbol can never pair with the next input
// symbol in the same round (all keys of a round are built from older symbols),
// the encoder is exactly
//     pick[i] = cand[i] && !pick[i-1],   pick[-1] = 0 at a traversal start,
// and cand is 0 at the last element of every traversal, so the recurrence may
// run straight across traversal boundaries. Inside a run of candidates starting
// at s, the picks are s, s+2, s+4, ...; the carry into a 64-position word is the
// single bit pick[-1].
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define GFZ_HD __host__ __device__
#else
#define GFZ_HD
#endif

namespace gfz {

// Picks of one 64-candidate word given pick[-1] = p_in.
GFZ_HD inline uint64_t word_picks(uint64_t c, uint32_t p_in) {
  const uint64_t E = 0x5555555555555555ull, O = ~E;
  // Run starts: a candidate whose predecessor is not a candidate; with
  // p_in = 1 the run touching bit 0 continues one that began at bit -1.
  const uint64_t starts = c & ~((c << 1) | (uint64_t)(p_in & 1));
  const uint64_t se = starts & E;
  // Adding a run's start bit carries through the run and clears it, so
  // c & ~(c + se) is the union of runs that start at an even position.
  const uint64_t runs_even = c & ~(c + se);
  const uint64_t runs_odd = c & ~runs_even;
  return (runs_even & E) | (runs_odd & O);
}

// Transfer function of a segment of the pick recurrence: for an incoming
// pick bit p, the outgoing pick bit o[p] and a count n[p]. Packed in 32 bits
// for counts below 2^15: bit0 = o[0], bit1 = o[1], bits 2..16 = n[0],
// bits 17..31 = n[1].
GFZ_HD inline uint32_t fn_pack(uint32_t o0, uint32_t n0, uint32_t o1, uint32_t n1) {
  return (o0 & 1) | ((o1 & 1) << 1) | (n0 << 2) | (n1 << 17);
}
GFZ_HD inline uint32_t fn_out(uint32_t f, uint32_t p) { return (f >> p) & 1; }
GFZ_HD inline uint32_t fn_cnt(uint32_t f, uint32_t p) {
  return p ? (f >> 17) & 0x7FFF : (f >> 2) & 0x7FFF;
}
// Apply f then g.
GFZ_HD inline uint32_t fn_then(uint32_t f, uint32_t g) {
  const uint32_t a0 = fn_out(f, 0), a1 = fn_out(f, 1);
  return fn_pack(fn_out(g, a0), fn_cnt(f, 0) + fn_cnt(g, a0), fn_out(g, a1),
                 fn_cnt(f, 1) + fn_cnt(g, a1));
}
GFZ_HD inline uint32_t fn_identity() { return fn_pack(0, 0, 1, 0); }

// Transfer function of one candidate word: counts are picks.
GFZ_HD inline uint32_t word_fn(uint64_t c) {
  const uint64_t p0 = word_picks(c, 0), p1 = word_picks(c, 1);
  return fn_pack((uint32_t)(p0 >> 63), (uint32_t)__builtin_popcountll(p0),
                 (uint32_t)(p1 >> 63), (uint32_t)__builtin_popcountll(p1));
}

// Wide form for scans across tiles: o[p] and a 64-bit survivor count s[p].
struct FnWide {
  uint64_t s0, s1;
  uint32_t o0, o1;
};
GFZ_HD inline FnWide fnw_then(const FnWide &f, const FnWide &g) {
  FnWide h;
  h.o0 = f.o0 ? g.o1 : g.o0;
  h.s0 = f.s0 + (f.o0 ? g.s1 : g.s0);
  h.o1 = f.o1 ? g.o1 : g.o0;
  h.s1 = f.s1 + (f.o1 ? g.s1 : g.s0);
  return h;
}
GFZ_HD inline FnWide fnw_identity() { return FnWide{0, 0, 0, 1}; }

}  // namespace gfz
