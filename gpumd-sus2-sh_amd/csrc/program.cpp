#include "program.hpp"

#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>

namespace sus2sh {

GraphProgram compile_graph(const Model& m, int chunk_terms)
{
  const int A = m.moments_count;
  const int B = m.basic_count;
  const int T = static_cast<int>(m.products.size());
  GraphProgram g;
  g.basic_count = B;
  g.layer_count = m.layer_count;

  // --- liveness: a product is live when its target can reach a mapped scalar ---
  std::vector<char> live(A, 0), product_live(T, 0), consumed(A, 0), mapped(A, 0);
  for (int s = 0; s < m.scalar_count; ++s) {
    live[m.moment_mapping[s]] = 1;
    mapped[m.moment_mapping[s]] = 1;
  }
  for (int p = T - 1; p >= 0; --p) {
    const Product& q = m.products[p];
    if (!live[q.target]) continue;
    product_live[p] = 1;
    live[q.left] = 1;
    live[q.right] = 1;
    consumed[q.left] = 1;
    consumed[q.right] = 1;
  }

  // --- slots: basics first, then live consumed targets in node order ---
  g.node_slot.assign(A, -1);
  int slots = 0;
  for (int b = 0; b < B; ++b) g.node_slot[b] = slots++;
  for (int n = B; n < A; ++n)
    if (live[n] && consumed[n]) g.node_slot[n] = slots++;
  if (slots >= kMaxSlots) throw InputError("SUS2_SH graph too large for the on-chip program.");
  g.slot_count = slots;

  // terminal weights w_t (summed over scalars mapped to the same node)
  std::vector<double> node_weight(A, 0.0);
  for (int s = 0; s < m.scalar_count; ++s) node_weight[m.moment_mapping[s]] += m.moment_coeffs[s];

  g.seed.assign(slots, 0.0);
  for (int n = 0; n < A; ++n)
    if (g.node_slot[n] >= 0 && mapped[n]) g.seed[g.node_slot[n]] = node_weight[n];
  for (int n = 0; n < A; ++n)
    if (g.node_slot[n] >= 0 && mapped[n]) {
      g.energy_slot.push_back(g.node_slot[n]);
      g.energy_weight.push_back(node_weight[n]);
    }

  // --- forward rows per layer, longest rows first ---
  std::vector<std::vector<int>> terms_of(A);
  for (int p = 0; p < T; ++p)
    if (product_live[p]) terms_of[m.products[p].target].push_back(p);
  std::vector<std::vector<int>> rows_of_layer(m.layer_count + 1);
  for (int n = B; n < A; ++n)
    if (!terms_of[n].empty()) rows_of_layer[m.node_layer[n]].push_back(n);
  std::vector<int> node_row(A, -1);
  g.fwd_layer_begin.assign(m.layer_count + 1, 0);
  for (int layer = 1; layer <= m.layer_count; ++layer) {
    g.fwd_layer_begin[layer - 1] = static_cast<int>(g.fwd_row_target.size());
    std::vector<int>& rows = rows_of_layer[layer];
    std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) {
      return terms_of[a].size() > terms_of[b].size();
    });
    for (int n : rows) {
      node_row[n] = static_cast<int>(g.fwd_row_target.size());
      g.fwd_row_target.push_back(g.node_slot[n]);
      g.fwd_row_node.push_back(n);
      g.fwd_row_begin.push_back(static_cast<int>(g.fwd_term_idx.size()));
      g.fwd_row_count.push_back(static_cast<int>(terms_of[n].size()));
      g.fwd_row_weight.push_back(g.node_slot[n] < 0 ? node_weight[n] : 0.0);
      for (int p : terms_of[n]) {
        const Product& q = m.products[p];
        g.fwd_term_idx.push_back(static_cast<uint32_t>(g.node_slot[q.left]) |
                                 (static_cast<uint32_t>(g.node_slot[q.right]) << 16));
        g.fwd_term_coeff.push_back(q.coeff);
      }
    }
  }
  g.fwd_layer_begin[m.layer_count] = static_cast<int>(g.fwd_row_target.size());

  g.scalar_slot.assign(m.scalar_count, -1);
  g.scalar_row.assign(m.scalar_count, -1);
  for (int s = 0; s < m.scalar_count; ++s) {
    const int n = m.moment_mapping[s];
    g.scalar_slot[s] = g.node_slot[n];
    if (g.node_slot[n] < 0) g.scalar_row[s] = node_row[n];
  }
  {
    const int nrows = static_cast<int>(g.fwd_row_target.size());
    std::vector<std::vector<int>> per_row(nrows);
    for (int s = 0; s < m.scalar_count; ++s) {
      if (g.scalar_row[s] >= 0) per_row[g.scalar_row[s]].push_back(s);
      if (g.scalar_slot[s] >= 0) {
        g.stored_scalar_index.push_back(s);
        g.stored_scalar_slot.push_back(g.scalar_slot[s]);
      }
    }
    g.row_scalar_begin.assign(nrows + 1, 0);
    for (int r = 0; r < nrows; ++r) {
      g.row_scalar_begin[r] = static_cast<int>(g.row_scalar_list.size());
      g.row_scalar_list.insert(g.row_scalar_list.end(), per_row[r].begin(), per_row[r].end());
    }
    g.row_scalar_begin[nrows] = static_cast<int>(g.row_scalar_list.size());
  }

  // --- reverse: per layer, per source; merge duplicate (target, other) terms ---
  g.bwd_layer_begin.assign(m.layer_count + 1, 0);
  for (int layer = 1; layer <= m.layer_count; ++layer) {
    // source slot -> ordered map (target code, other slot) -> coefficient
    std::map<int, std::map<std::pair<uint32_t, int>, double>> by_source;
    for (int n : rows_of_layer[layer]) {
      const bool terminal = g.node_slot[n] < 0;
      const uint32_t tcode = terminal ? kNoTarget : static_cast<uint32_t>(g.node_slot[n]);
      const double scale = terminal ? node_weight[n] : 1.0;
      for (int p : terms_of[n]) {
        const Product& q = m.products[p];
        const int ls = g.node_slot[q.left], rs = g.node_slot[q.right];
        by_source[ls][{tcode, rs}] += q.coeff * scale;
        by_source[rs][{tcode, ls}] += q.coeff * scale;
      }
    }
    g.bwd_layer_begin[layer - 1] = static_cast<int>(g.bwd_chunk_src.size());
    struct Chunk {
      int src, begin, count, atomic;
    };
    std::vector<Chunk> chunks;
    for (const auto& kv : by_source) {
      const int src = kv.first;
      const int nterms = static_cast<int>(kv.second.size());
      const int nchunks = (nterms + chunk_terms - 1) / chunk_terms;
      const int begin = static_cast<int>(g.bwd_term_idx.size());
      for (const auto& e : kv.second) {
        g.bwd_term_idx.push_back(static_cast<uint32_t>(e.first.second) | (e.first.first << 16));
        g.bwd_term_coeff.push_back(e.second);
      }
      for (int c = 0; c < nchunks; ++c) {
        const int cb = begin + c * chunk_terms;
        const int cn = std::min(chunk_terms, nterms - c * chunk_terms);
        chunks.push_back({src, cb, cn, nchunks > 1 ? 1 : 0});
      }
    }
    std::stable_sort(chunks.begin(), chunks.end(),
                     [](const Chunk& a, const Chunk& b) { return a.count > b.count; });
    for (const Chunk& c : chunks) {
      g.bwd_chunk_src.push_back(c.src);
      g.bwd_chunk_begin.push_back(c.begin);
      g.bwd_chunk_count.push_back(c.count);
      g.bwd_chunk_atomic.push_back(c.atomic);
    }
  }
  g.bwd_layer_begin[m.layer_count] = static_cast<int>(g.bwd_chunk_src.size());
  return g;
}

}  // namespace sus2sh
