// Compilation of the explicit product graph into the layered program run by the
// on-chip graph kernel.
//
// Semantics are those of the flat forward / energy / reverse loop
// (src/force/sus2_sh.cu:4799-4846) and of the layered row / back-row form
// (src/force/sus2_sh.cu:1913-2093):
//   * forward rows: one per live target, M[t] = sum c M[l] M[r], grouped by layer;
//   * terminal scalars (mapped nodes that feed no product) are never stored: their row
//     sum goes straight into the site energy and their adjoint is the constant w_t;
//   * reverse: per layer (descending) and per source node, G[src] += sum c G[t] M[o];
//     duplicate (target, other) terms of a source are merged (cf. sus2_sh.cu:8945-9121);
//   * nodes that cannot reach any scalar are pruned.
// Adjoints are computed without the species factor sigma (G' = G / sigma) and scaled
// once at the end, because the seeds are G[map[s]] = sigma * w_s.
#pragma once

#include <cstdint>
#include <vector>

#include "model.hpp"

namespace sus2sh {

struct GraphProgram {
  int basic_count = 0;
  int slot_count = 0;  // stored nodes: basics occupy slots [0, basic_count)
  int layer_count = 0;
  std::vector<int> node_slot;  // [A] slot or -1

  // forward
  std::vector<int> fwd_layer_begin;  // [layer_count + 1] row offsets, layer 1..L
  std::vector<int> fwd_row_target;   // slot, or -1 for a terminal scalar row
  std::vector<int> fwd_row_begin;
  std::vector<int> fwd_row_count;
  std::vector<double> fwd_row_weight;  // w_t for terminal rows
  std::vector<int> fwd_row_node;       // node id (for diagnostics / parameter gradient)
  std::vector<uint32_t> fwd_term_idx;  // left_slot | right_slot << 16
  std::vector<double> fwd_term_coeff;

  // site energy contributions of stored scalar slots
  std::vector<int> energy_slot;
  std::vector<double> energy_weight;

  // adjoint seeds per slot (sum of w_s over mapped scalars)
  std::vector<double> seed;

  // reverse, processed layer by layer from layer_count down to 1
  std::vector<int> bwd_layer_begin;  // [layer_count + 1] chunk offsets, layer 1..L
  std::vector<int> bwd_chunk_src;
  std::vector<int> bwd_chunk_begin;
  std::vector<int> bwd_chunk_count;
  std::vector<int> bwd_chunk_atomic;  // 1 when another chunk of the same layer adds to src
  std::vector<uint32_t> bwd_term_idx;  // other_slot | target_slot << 16 (0xFFFF: constant)
  std::vector<double> bwd_term_coeff;  // c, or c * w_t for a terminal target

  // parameter-gradient support: scalar s -> (slot or -1, terminal row or -1), and for each
  // forward row the scalars mapped to its (terminal) node, CSR [rows + 1] / [entries]
  std::vector<int> scalar_slot;
  std::vector<int> scalar_row;
  std::vector<int> row_scalar_begin;
  std::vector<int> row_scalar_list;
  std::vector<int> stored_scalar_index;  // scalars whose node has a slot
  std::vector<int> stored_scalar_slot;

  // Per-thread instruction streams (threads_per_block lanes), interleaved so that
  // instruction k of thread t sits at stream_begin[layer] + k * threads + t.
  //   forward:  a < 0xFFFE: acc += c V[a] V[b];  a == 0xFFFF: end of row -> V[b] = acc, or
  //             (b == 0xFFFF) e += c acc;  a == 0xFFFE: no-op
  //   backward: a < 0xFFFE: acc += c G[b] V[a]  (b = ones slot for a terminal target);
  //             a == 0xFFFF: flush acc into G[b] (c != 0: atomically);  a == 0xFFFE: no-op
  // Two assignments of rows/chunks to threads are compiled: [0] longest-first onto the
  // least-loaded thread (balanced), [1] round-robin in target order (neighbouring lanes
  // work on rows of the same tensor, so their LDS operands coincide).
  struct Streams {
    std::vector<int> fwd_begin;  // [layer_count + 1]
    std::vector<int> fwd_len;    // [layer_count] instructions per thread
    std::vector<uint32_t> fwd_ab;
    std::vector<double> fwd_c;
    std::vector<int> bwd_begin;
    std::vector<int> bwd_len;
    std::vector<uint32_t> bwd_ab;
    std::vector<double> bwd_c;
  };
  int threads = 0;
  Streams streams[2];
  int ones_slot = 0;  // adjoint slot that always holds 1 (== slot_count)

  int forward_terms() const { return static_cast<int>(fwd_term_idx.size()); }
  int backward_terms() const { return static_cast<int>(bwd_term_idx.size()); }
};

constexpr uint32_t kNoTarget = 0xFFFFu;
constexpr int kMaxSlots = 0xFFFF;

// store_mapped: every mapped scalar node gets a slot (no terminal rows fused into the site
// energy), so that a kernel can read scalar values and seed their adjoints per atom -- the
// gate layer's program, compiled for a model whose scalars are the gate scalars.
GraphProgram compile_graph(const Model& model, int chunk_terms = 8, int threads = 256,
                           bool store_mapped = false);

}  // namespace sus2sh
