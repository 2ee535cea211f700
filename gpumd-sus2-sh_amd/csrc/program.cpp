#include "program.hpp"

#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>

namespace sus2sh {

namespace {

constexpr uint32_t kEnd = 0xFFFFu;
constexpr uint32_t kNop = 0xFFFEu;

struct Ins {
  uint32_t a, b;
  double c;
};

// Longest-processing-time assignment of work items (each a list of instructions) to
// `threads` streams, then interleaving with no-op padding.
void emit_streams(const std::vector<std::vector<Ins>>& items, int threads, bool balanced,
                  std::vector<uint32_t>& ab, std::vector<double>& cs, int* len)
{
  std::vector<std::vector<Ins>> per(threads);
  if (balanced) {
    std::vector<int> order(items.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = static_cast<int>(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return items[x].size() > items[y].size(); });
    std::vector<size_t> load(threads, 0);
    for (int it : order) {
      int best = 0;
      for (int t = 1; t < threads; ++t)
        if (load[t] < load[best]) best = t;
      per[best].insert(per[best].end(), items[it].begin(), items[it].end());
      load[best] += items[it].size();
    }
  } else {
    for (size_t it = 0; it < items.size(); ++it) {
      std::vector<Ins>& dst = per[it % threads];
      dst.insert(dst.end(), items[it].begin(), items[it].end());
    }
  }
  size_t L = 0;
  for (int t = 0; t < threads; ++t) L = std::max(L, per[t].size());
  L = (L + 3) / 4 * 4;  // the kernel consumes instructions in groups of four
  for (size_t k = 0; k < L; ++k)
    for (int t = 0; t < threads; ++t) {
      const Ins in = k < per[t].size() ? per[t][k] : Ins{kNop, 0, 0.0};
      ab.push_back((in.a & 0xFFFFu) | (in.b << 16));
      cs.push_back(in.c);
    }
  *len = static_cast<int>(L);
}

}  // namespace

GraphProgram compile_graph(const Model& m, int chunk_terms, int threads, bool store_mapped)
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
    if (live[n] && (consumed[n] || (store_mapped && mapped[n]))) g.node_slot[n] = slots++;
  if (slots + 1 >= static_cast<int>(kNoTarget) - 1)
    throw InputError("SUS2_SH graph too large for the on-chip program.");
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

  // --- per-thread instruction streams ---
  g.threads = threads;
  g.ones_slot = g.slot_count;
  // forward items in layer order; within a layer by ascending target (round-robin keeps the
  // rows of one tensor on neighbouring lanes)
  std::vector<std::vector<std::vector<Ins>>> fwd_items(m.layer_count), bwd_items(m.layer_count);
  for (int layer = 0; layer < m.layer_count; ++layer) {
    std::vector<int> rows;
    for (int r = g.fwd_layer_begin[layer]; r < g.fwd_layer_begin[layer + 1]; ++r) rows.push_back(r);
    std::sort(rows.begin(), rows.end(),
              [&](int a, int b) { return g.fwd_row_node[a] < g.fwd_row_node[b]; });
    for (int r : rows) {
      std::vector<Ins> it;
      for (int t = 0; t < g.fwd_row_count[r]; ++t) {
        const uint32_t idx = g.fwd_term_idx[g.fwd_row_begin[r] + t];
        it.push_back({idx & 0xFFFFu, idx >> 16, g.fwd_term_coeff[g.fwd_row_begin[r] + t]});
      }
      const int tgt = g.fwd_row_target[r];
      it.push_back({kEnd, tgt >= 0 ? static_cast<uint32_t>(tgt) : kEnd,
                    tgt >= 0 ? 0.0 : g.fwd_row_weight[r]});
      fwd_items[layer].push_back(std::move(it));
    }
    std::vector<int> chunks;
    for (int c = g.bwd_layer_begin[layer]; c < g.bwd_layer_begin[layer + 1]; ++c) chunks.push_back(c);
    std::stable_sort(chunks.begin(), chunks.end(), [&](int a, int b) {
      return g.bwd_chunk_src[a] != g.bwd_chunk_src[b] ? g.bwd_chunk_src[a] < g.bwd_chunk_src[b]
                                                      : g.bwd_chunk_begin[a] < g.bwd_chunk_begin[b];
    });
    for (int c : chunks) {
      std::vector<Ins> it;
      for (int t = 0; t < g.bwd_chunk_count[c]; ++t) {
        const uint32_t idx = g.bwd_term_idx[g.bwd_chunk_begin[c] + t];
        const uint32_t other = idx & 0xFFFFu, tgt = idx >> 16;
        it.push_back({other, tgt == kNoTarget ? static_cast<uint32_t>(g.ones_slot) : tgt,
                      g.bwd_term_coeff[g.bwd_chunk_begin[c] + t]});
      }
      it.push_back({kEnd, static_cast<uint32_t>(g.bwd_chunk_src[c]),
                    g.bwd_chunk_atomic[c] ? 1.0 : 0.0});
      bwd_items[layer].push_back(std::move(it));
    }
  }
  for (int v = 0; v < 2; ++v) {
    GraphProgram::Streams& st = g.streams[v];
    const bool balanced = v == 0;
    st.fwd_begin.assign(m.layer_count + 1, 0);
    st.fwd_len.assign(m.layer_count, 0);
    st.bwd_begin.assign(m.layer_count + 1, 0);
    st.bwd_len.assign(m.layer_count, 0);
    for (int layer = 0; layer < m.layer_count; ++layer) {
      st.fwd_begin[layer] = static_cast<int>(st.fwd_ab.size());
      emit_streams(fwd_items[layer], threads, balanced, st.fwd_ab, st.fwd_c, &st.fwd_len[layer]);
      st.bwd_begin[layer] = static_cast<int>(st.bwd_ab.size());
      emit_streams(bwd_items[layer], threads, balanced, st.bwd_ab, st.bwd_c, &st.bwd_len[layer]);
    }
    st.fwd_begin[m.layer_count] = static_cast<int>(st.fwd_ab.size());
    st.bwd_begin[m.layer_count] = static_cast<int>(st.bwd_ab.size());
  }
  return g;
}

}  // namespace sus2sh
