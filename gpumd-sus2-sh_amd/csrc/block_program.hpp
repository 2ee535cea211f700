// Block program: the product graph of a depth <= 2 model whose layer-1 nodes are complete real
// Clebsch-Gordan couplings of basic tensors (degree <= 3) and whose layer-2 nodes contract two
// tensors of one degree L over all 2L+1 components with one coefficient -- every standard
// graph up to body order 5 (src/force/sus2_sh.cu:1227-1706) -- run by the block graph kernel
// (block_kernel.hip), one atom per lane.
//
// Semantics are those of the flat loop of src/force/sus2_sh.cu:4799-4846.  With tensors
// T (basic tensors b_(k,l), layer-1 tensors P = (b_A (x) b_B)_L, components M = -L..L):
//   P[M] = sum_{m1,m2} C(lA m1; lB m2 | L M) b_A[m1] b_B[m2]            (FWD)
//   E'   = sum_L sum_{T,T' of degree L} K_TT' sum_M T[M] T'[M] + sum_T d_T T[0]
// so per tensor g_T[M] = dE'/dT[M] = sum_T' K'_TT' T'[M] + d_T [M = 0], K' = K + K^T,
// E' = 1/2 sum_T sum_M T[M] (g_T[M] + d_T [M = 0]), and the reverse pass adds
// W = sum_M g_P[M] C(. . | L M) into the basic adjoints: G_A += W b_B, G_B += W^T b_A
// (G_T += g_T for a basic tensor).
//
// The live layer-1 tensors are grouped by degree; whole degree classes -- or, when that makes
// the tile small enough for two workgroups per CU, the two component halves [0, L + 1) and
// [L + 1, 2L + 1) of the large classes of degree >= 2 (the quadratic form couples equal
// components only) -- are packed into batches that fit the per-lane value region.  A batch is two epochs: FWD (the batch's
// layer-1 tensors, split between the waves of the tile) and TENSOR (per quadratic-form tensor
// of the batch: g_T, its energy share and its reverse pass, the basic adjoints summed with
// LDS atomics).  Each wave's records of an epoch are sorted into runs of one block type, so
// the kernel dispatches once per run.
//
// Per-lane slot map: [0, B) basics, [B, B + cap) tensor values (Real); the basic adjoints
// sit in a separate region of B double slots (FP64 LDS atomics in both precisions).
// Device words of one (batch, wave, epoch): runs of  header = kind | count << 8  followed by
// count records.
//   FWD record (2 words):    v_off | a_off << 16,  b_off
//   TENSOR record:           v_off | npart << 16,  a_off | b_off << 16,  d (Real),
//                            npart x (partner v_off, K' (Real))
// The word array ends with 8 padding words.
// Offsets are LDS words (slot * 64; for a half, the slot of its first component); a Real takes
// 1 (FP32) or 2 (FP64) words.  Kind = t | part << 6 for CG block type t of cg_table.hpp, and
// kBlockBasicKind + 4 part + L for a basic tensor of degree L; part 0 is the whole tensor,
// 1 and 2 its lower and upper component halves.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "model.hpp"

namespace sus2sh {

constexpr int kBlockBasicKind = 192;

struct BlockProgram {
  bool ok = false;
  std::string reason;  // why the model cannot use the block program (when !ok)
  int basic_count = 0;
  int waves = 1;       // waves sharing one 64-atom tile
  int slots = 0;       // per-lane Real slots (basics and tensor values)
  int v_base = 0, cap = 0;
  int batches = 0;
  int units = 0;       // degree classes or class halves
  int tensors = 0;     // quadratic-form tensors
  int fwd_blocks = 0;  // layer-1 tensors computed
  long long fwd_terms = 0, adj_terms = 0, bwd_terms = 0;
  int real_bytes = 4;  // width of the coefficients in `words`
  // word ranges [fwd_begin, fwd_end, ten_begin, ten_end) of (batch, wave), 4 per entry
  std::vector<int> seg;
  std::vector<uint32_t> words;
};

// Build the block program for `waves` waves per tile with coefficients of real_bytes (4 or
// 8) and at most max_slots per-lane Real slots, halving large classes when that brings it to
// target_slots (<= 0: never).  Never throws for an unsuitable model: returns ok = false with
// the reason instead.  The program is checked against the flat graph on seeded basic vectors
// before it is returned.
BlockProgram compile_block_program(const Model& model, int waves, int real_bytes, int max_slots,
                                   int target_slots = 0);

// Host execution of the device words for one atom, in double (test and self-check support):
// returns E' and writes dE'/db into grad_b (B values).
double run_block_program_host(const BlockProgram& p, const double* basics, double* grad_b);

}  // namespace sus2sh
