// Lane program: the product graph of a depth <= 3 model as one wave-uniform instruction
// stream executed by one atom per lane (kernel in lane_kernel.hip).
//
// Semantics are those of the flat loop of src/force/sus2_sh.cu:4799-4846: forward
// M[t] += c M[l] M[r], E = sum_s w_s M[map[s]], reverse through the same products.
//
// For a graph whose products are at most two layers deep (every standard graph up to body
// order 5: layer 1 couples basic tensors, layer 2 contracts them into scalars) the
// per-atom energy is a quadratic form in the layer-<=1 nodes:
//   E' = sum_{layer-2 p} k_p v_l v_r + sum_n d_n v_n,   k_p = w_{t_p} c_p,  d_n = sum w_s [map_s = n]
// Its layer-2 products couple only nodes of one (L, M) class, so the layer-<=1 nodes split
// into small connected components (union-find over the layer-2 products).  A component is
// evaluated on its own:
//   FWD  v_n = sum c b_l b_r                   (layer-1 nodes; basic nodes alias b)
//   ADJ  g_n = d_n + sum_{p contains n} k_p v_other   (the dE'/dv_n of the quadratic form)
//   ENG  E' += 1/2 sum_n v_n (g_n + d_n)       (exact for a quadratic form)
//   BWD  G_b[q] += sum c g_n b_other  (+ g_q for a basic node q)
// Components are packed into batches whose nodes fit the per-lane value region; each batch
// runs as four segments (FWD, ADJ, ENG, BWD).  Inside a segment rows are independent: they
// are spread longest-first over `waves` x `nslot` accumulators -- `waves` waves share one
// 64-atom tile and split the rows, each wave's rows interleaved as groups of nslot
// instructions so that one group issues nslot independent LDS reads per operand.
//
// Three-layer graphs (body order 6): the layer-2 nodes Y that feed live layer-3 products
// join the quadratic form as nodes (a layer-3 product couples a Y node with a node of the
// same (L, M) class), so the form is quadratic in the layer-<=1 and Y values.  A Y value is
// computed in the FWD segment of its batch from its layer-1 operands P, whose values are
// computed once, before the first batch, into persistent slots.  The total adjoint of a P
// node is its quadratic-form adjoint (added to a persistent slot by the BWD segment of its
// batch) plus the part through Y (a segment after each batch holding Y nodes: g_P += c g_Y
// v_other, basic operands straight into the basic adjoints); one segment after the last
// batch carries the P adjoints to the basic adjoints.
// Per-lane slot map: [0, B) basics, [B, 2B) basic adjoints, 2B the constant 1, (2B + 1 a
// write-only trash slot, only when a store segment has an accumulator without rows), then
// the P values and P adjoints (three-layer graphs), [v_base, v_base + cap_v) component
// values, [g_base, g_base + cap_g) component adjoints.
// Instruction word: x | y << 10 | t << 20 | end << 30 with end 0 (none), 1 (slot t = acc),
// 2 (slot t += acc), 3 (site energy += acc).  Binary segments add c X[x] X[y], unary
// segments add c X[x].  A row much longer than a segment's mean bin load is split between
// accumulators 0 and 1 of one wave (alternate terms) and combined at its end.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "model.hpp"

namespace sus2sh {

struct LaneProgram {
  bool ok = false;
  std::string reason;  // why the model cannot use the lane program (when !ok)
  int basic_count = 0;
  int nslot = 8;
  int waves = 1;  // waves that share one 64-atom tile and split every segment's rows
  int atoms = 1;  // atoms per lane of the kernel that runs it (its LDS column width / 64)
  int cap_v = 0, cap_g = 0;  // value / adjoint region sizes
  int slots = 0;  // per-lane slots (LDS = slots * 64 * sizeof(Real) per wave)
  int ones_slot = 0, trash_slot = 0, v_base = 0, g_base = 0;
  int components = 0, largest_component = 0, batches = 0;
  long long useful_terms = 0;  // instructions that are not padding
  // segments: {kind (0 binary, 1 unary), first instruction, groups}
  std::vector<int> seg_kind, seg_begin, seg_groups;
  // 1: the waves of a tile synchronise before the segment; 0: it runs in the same epoch as
  // the segment before it (no slot written in an epoch is read in it or written twice)
  std::vector<uint8_t> seg_sync;
  std::vector<uint32_t> ins;
  std::vector<double> coeff;
  // device form of the row ends: slot each instruction's group writes its accumulator to
  // (store segments: the row's target; add segments: the row's target at its end) and
  // whether the accumulator is kept after the group (0 at a row end, except the last row of
  // a store accumulator, whose sum is rewritten unchanged by the padding after it)
  std::vector<uint16_t> write_slot;
  std::vector<uint8_t> keep;
  // 1 on accumulator 0 of the group that ends a split row: accumulator 1 (the row's other
  // half, in the same wave) is added to accumulator 0 before the row end, then cleared
  std::vector<uint8_t> combine;
  bool needs_trash = false;  // a store segment has an accumulator without rows
};

constexpr uint32_t kLaneEndNone = 0, kLaneEndStore = 1, kLaneEndAdd = 2, kLaneEndEnergy = 3;
// segment kinds: every row of a segment ends the same way
constexpr int kLaneBinaryStore = 0, kLaneUnaryStore = 1, kLaneBinaryEnergy = 2, kLaneBinaryAdd = 3;
constexpr int kLaneMaxSlot = 1023;

// Build the lane program with at most `max_slots` per-lane slots (the value / adjoint
// region sizes follow from it).  Never throws for an unsuitable graph: returns
// ok = false with the reason instead.
LaneProgram compile_lane_program(const Model& model, int max_slots, int nslot = 8,
                                 int waves = 1);

// Host execution of a lane program for one atom (test and self-check support):
// returns E' and writes dE'/db into grad_b (B values).
double run_lane_program_host(const LaneProgram& p, const double* basics, double* grad_b);

// Device encoding of a lane program: one record of rec_words() 32-bit words per group of
// nslot instructions, wave-major inside a segment (record (w * groups + g) of the segment),
// every word used so that a batch of records is a few wide scalar loads:
//   words [0, nslot)          x word offsets in the lane column layout (slot * col, col =
//                             64 x atoms per lane)
//   words [nslot, 2 nslot)    y word offsets
//   words [2 nslot, 3 nslot)  target word offsets: where an accumulator goes at its row's
//                             end (stored by store segments, added by add segments; energy
//                             segments sum their accumulators once, at the segment's end)
//   word  3 nslot             row-end mask: bit j = instruction j ends its row; bit nslot =
//                             a row split between accumulators 0 and 1 ends (1 joins 0 first)
//   then nslot coefficients (Real) from word header_words(): 8-byte coefficients start on
//   an even word, with one padding word when 3 nslot + 1 is odd.
constexpr int lane_header_words(int ns, int real_bytes)
{
  return real_bytes == 8 ? (3 * ns + 1 + 1) / 2 * 2 : 3 * ns + 1;
}
constexpr int lane_record_words(int ns, int real_bytes)
{
  return lane_header_words(ns, real_bytes) + ns * real_bytes / 4;
}

struct LaneRecords {
  int nslot = 0, real_bytes = 0;
  int col = 64;  // LDS words per slot row (64 x atoms per lane)
  int header_words() const { return lane_header_words(nslot, real_bytes); }
  int rec_words() const { return lane_record_words(nslot, real_bytes); }
  std::vector<uint32_t> words;
  std::vector<int> seg_first;  // first record of each segment
};
LaneRecords encode_lane_records(const LaneProgram& p, int real_bytes, int col = 64);

// Host execution of the device encoding (same arithmetic order as the kernel, in double).
double run_lane_records_host(const LaneProgram& p, const LaneRecords& r, const double* basics,
                             double* grad_b);

// Host execution of the flat graph for one atom (same outputs), the reference the lane
// program is checked against when it is compiled.
double run_flat_graph_host(const Model& model, const double* basics, double* grad_b);

}  // namespace sus2sh
