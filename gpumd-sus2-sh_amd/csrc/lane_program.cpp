#include "lane_program.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace sus2sh {

namespace {

struct Term {
  double c;
  int x, y;
};

struct Row {
  std::vector<Term> terms;
  int target;
  uint32_t end;
};

int find_root(std::vector<int>& parent, int x)
{
  while (parent[x] != x) {
    parent[x] = parent[parent[x]];
    x = parent[x];
  }
  return x;
}

// Longest-first assignment of independent rows to waves * nslot accumulators.  The
// instructions of wave w form groups of nslot (slot j feeds accumulator j); every wave of a
// segment has the same number of groups (padding: c = 0 on the constant slot).
// Layout: instruction j of group g of wave w at begin + (w * groups + g) * nslot + j.
void emit_segment(LaneProgram& p, int kind, std::vector<Row>& rows, bool sync = true)
{
  const int ns = p.nslot, bins = p.nslot * p.waves;
  const bool store = kind == kLaneBinaryStore || kind == kLaneUnaryStore;
  rows.erase(std::remove_if(rows.begin(), rows.end(), [](const Row& r) { return r.terms.empty(); }),
             rows.end());
  if (rows.empty()) return;
  std::vector<int> order(rows.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return rows[a].terms.size() > rows[b].terms.size(); });
  // rows longer than the mean bin load are split between accumulators 0 and 1 of a wave
  // (energy rows are already split into parts); split rows go first, one wave after another
  size_t total = 0;
  for (const Row& r : rows) total += r.terms.size();
  const size_t mean = (total + bins - 1) / bins;
  // per[j]: (row, part) with part 0 the whole row, 1 its even terms, 2 its odd terms
  std::vector<std::vector<std::pair<int, int>>> per(bins);
  std::vector<size_t> load(bins, 0);
  size_t first = 0;
  if (ns >= 2 && kind != kLaneBinaryEnergy) {
    int w = 0;
    for (; first < order.size(); ++first) {
      const size_t n = rows[order[first]].terms.size();
      if (n < 4 || n <= mean) break;
      per[w * ns].push_back({order[first], 1});
      per[w * ns + 1].push_back({order[first], 2});
      load[w * ns] += (n + 1) / 2;
      load[w * ns + 1] += (n + 1) / 2;
      w = (w + 1) % p.waves;
    }
  }
  for (size_t i = first; i < order.size(); ++i) {
    const int r = order[i];
    const int j = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
    per[j].push_back({r, 0});
    load[j] += rows[r].terms.size();
  }
  const size_t groups = *std::max_element(load.begin(), load.end());
  std::vector<std::vector<uint32_t>> words(bins);
  std::vector<std::vector<double>> cs(bins);
  std::vector<std::vector<uint16_t>> wt(bins);
  std::vector<std::vector<uint8_t>> kp(bins), cb(bins);
  const uint32_t one = static_cast<uint32_t>(p.ones_slot);
  for (int j = 0; j < bins; ++j) {
    // the last row of a bin keeps its sum: the padding after it leaves it unchanged; a
    // store bin with no row writes the trash slot
    uint16_t last_target = static_cast<uint16_t>(p.trash_slot);
    if (store && per[j].empty()) p.needs_trash = true;
    for (size_t ri = 0; ri < per[j].size(); ++ri) {
      const Row& row = rows[per[j][ri].first];
      const int part = per[j][ri].second;
      const bool last_row = ri + 1 == per[j].size();
      std::vector<const Term*> ts;
      for (size_t k = part == 2 ? 1 : 0; k < row.terms.size(); k += part == 0 ? 1 : 2)
        ts.push_back(&row.terms[k]);
      // both halves of a split row are (n + 1) / 2 long: the odd half of an odd row starts
      // with a zero term
      const size_t len = part == 0 ? row.terms.size() : (row.terms.size() + 1) / 2;
      for (size_t k = 0; k < len; ++k) {
        const size_t pad = len - ts.size();
        const bool last = k + 1 == len;
        if (k < pad) {
          words[j].push_back(one | (one << 10));
          cs[j].push_back(0.0);
        } else {
          const Term& t = *ts[k - pad];
          // the odd half of a split row never ends: accumulator 0 ends the row
          const uint32_t end = last && part != 2 ? row.end : kLaneEndNone;
          const uint32_t tgt = end != kLaneEndNone && row.end != kLaneEndEnergy
                                 ? static_cast<uint32_t>(row.target) : 0u;
          words[j].push_back(static_cast<uint32_t>(t.x) | (static_cast<uint32_t>(t.y) << 10) |
                             (tgt << 20) | (end << 30));
          cs[j].push_back(t.c);
          ++p.useful_terms;
        }
        wt[j].push_back(static_cast<uint16_t>(row.target));
        kp[j].push_back(last && part != 2 && !(store && last_row) ? 0 : 1);
        cb[j].push_back(last && part == 1 ? 1 : 0);
      }
      last_target = static_cast<uint16_t>(row.target);
    }
    while (words[j].size() < groups) {
      words[j].push_back(one | (one << 10));
      cs[j].push_back(0.0);
      wt[j].push_back(last_target);
      kp[j].push_back(1);
      cb[j].push_back(0);
    }
  }
  p.seg_kind.push_back(kind);
  p.seg_sync.push_back(sync || p.seg_kind.size() == 1 ? 1 : 0);
  p.seg_begin.push_back(static_cast<int>(p.ins.size()));
  p.seg_groups.push_back(static_cast<int>(groups));
  for (int w = 0; w < p.waves; ++w)
    for (size_t g = 0; g < groups; ++g)
      for (int j = 0; j < ns; ++j) {
        p.ins.push_back(words[w * ns + j][g]);
        p.coeff.push_back(cs[w * ns + j][g]);
        p.write_slot.push_back(wt[w * ns + j][g]);
        p.keep.push_back(kp[w * ns + j][g]);
        p.combine.push_back(cb[w * ns + j][g]);
      }
}

uint64_t lcg(uint64_t& s)
{
  s = s * 6364136223846793005ULL + 1442695040888963407ULL;
  return s >> 11;
}

}  // namespace

namespace {
LaneProgram compile_impl(const Model& m, int max_slots, int nslot, int waves, bool with_trash);
}

LaneProgram compile_lane_program(const Model& m, int max_slots, int nslot, int waves)
{
  LaneProgram p = compile_impl(m, max_slots, nslot, waves, false);
  if (!p.ok && p.needs_trash) p = compile_impl(m, max_slots, nslot, waves, true);
  return p;
}

namespace {

LaneProgram compile_impl(const Model& m, int max_slots, int nslot, int waves, bool with_trash)
{
  LaneProgram p;
  p.nslot = nslot;
  p.waves = waves;
  const int A = m.moments_count, B = m.basic_count;
  p.basic_count = B;
  if (m.layer_count > 3) {
    p.reason = "product graph deeper than three layers";
    return p;
  }
  const std::vector<int>& layer = m.node_layer;

  // direct weights d_n and the structural "mapped" flag
  std::vector<double> d(A, 0.0);
  std::vector<char> mapped(A, 0);
  for (int s = 0; s < m.scalar_count; ++s) {
    d[m.moment_mapping[s]] += m.moment_coeffs[s];
    mapped[m.moment_mapping[s]] = 1;
  }
  const int T = static_cast<int>(m.products.size());
  // Y: layer-2 nodes that are operands of live layer-3 products (three-layer graphs only)
  std::vector<char> is_y(A, 0);
  for (int q = 0; q < T; ++q) {
    const Product& pr = m.products[q];
    if (layer[pr.target] != 3 || !mapped[pr.target]) continue;
    if (layer[pr.left] == 2) is_y[pr.left] = 1;
    if (layer[pr.right] == 2) is_y[pr.right] = 1;
  }
  std::vector<std::vector<int>> layer1_terms(A);  // products of each layer-1 target
  std::vector<std::vector<int>> y_terms(A);       // products of each Y node
  std::vector<int> l2;  // products of the quadratic form: live layer-2 and layer-3 products
  for (int q = 0; q < T; ++q) {
    const Product& pr = m.products[q];
    const int lt = layer[pr.target];
    if (lt == 1) {
      layer1_terms[pr.target].push_back(q);
    } else if (is_y[pr.target]) {
      y_terms[pr.target].push_back(q);
    } else if (mapped[pr.target]) {
      l2.push_back(q);
    }
  }
  // P: the layer-1 operands of the Y products.  Their values are computed once, before the
  // first batch, into persistent slots (the Y rows of any batch read them), and their
  // adjoints are collected in persistent slots: the quadratic-form part from their own
  // component's batch, the part through Y after each batch that holds Y nodes; their reverse
  // pass to the basic adjoints runs once, after the last batch.
  std::vector<int> p1;
  std::vector<int> pv_slot(A, -1), gp_slot(A, -1);
  for (int n = 0; n < A; ++n) {
    if (!is_y[n]) continue;
    for (int q : y_terms[n]) {
      for (int o : {m.products[q].left, m.products[q].right}) {
        if (layer[o] != 1 || pv_slot[o] >= 0) continue;
        pv_slot[o] = 0;
        p1.push_back(o);
      }
    }
  }
  std::sort(p1.begin(), p1.end());
  max_slots = std::min(max_slots, kLaneMaxSlot + 1);
  const int reserved = 2 * B + (with_trash ? 2 : 1);
  const int np = static_cast<int>(p1.size());
  if (max_slots - reserved - 2 * np < 2) {
    p.reason = "too many basic moments for the per-lane slot budget";
    return p;
  }
  p.ones_slot = 2 * B;
  p.trash_slot = with_trash ? 2 * B + 1 : 2 * B;  // without a trash slot no bin may be empty
  for (int i = 0; i < np; ++i) {
    pv_slot[p1[i]] = reserved + i;
    gp_slot[p1[i]] = reserved + np + i;
  }
  p.v_base = reserved + 2 * np;

  // component members: layer <= 1 and Y nodes that are mapped or feed a live product of the
  // quadratic form
  std::vector<char> member(A, 0);
  for (int n = 0; n < A; ++n)
    if (mapped[n] && (layer[n] <= 1 || is_y[n])) member[n] = 1;
  for (int q : l2) {
    member[m.products[q].left] = 1;
    member[m.products[q].right] = 1;
  }
  std::vector<int> parent(A);
  std::iota(parent.begin(), parent.end(), 0);
  for (int q : l2) {
    const int a = find_root(parent, m.products[q].left), b = find_root(parent, m.products[q].right);
    if (a != b) parent[a] = b;
  }
  std::vector<std::vector<int>> comps;
  {
    std::vector<int> comp_of(A, -1);
    for (int n = 0; n < A; ++n) {
      if (!member[n]) continue;
      const int r = find_root(parent, n);
      if (comp_of[r] < 0) {
        comp_of[r] = static_cast<int>(comps.size());
        comps.emplace_back();
      }
      comps[comp_of[r]].push_back(n);
    }
  }
  p.components = static_cast<int>(comps.size());
  // quadratic-form products by member node
  std::vector<std::vector<int>> l2_of(A);
  for (int q : l2) {
    l2_of[m.products[q].left].push_back(q);
    if (m.products[q].right != m.products[q].left) l2_of[m.products[q].right].push_back(q);
  }

  // region sizes: the free slots are split between values and adjoints (basic and P nodes
  // keep their values in their own slots)
  auto v_need = [&](const std::vector<int>& c) {
    int k = 0;
    for (int n : c) k += n >= B && pv_slot[n] < 0 ? 1 : 0;
    return k;
  };
  int nv_max = 0, ng_max = 0;
  for (const std::vector<int>& c : comps) {
    nv_max = std::max(nv_max, v_need(c));
    ng_max = std::max(ng_max, static_cast<int>(c.size()));
  }
  p.largest_component = ng_max;
  const int free_slots = max_slots - p.v_base;
  if (nv_max + ng_max > free_slots) {
    p.reason = "a graph component exceeds the per-lane slot budget";
    return p;
  }
  // split in proportion to the total needs (the batches are then filled evenly), but never
  // below the largest component's need
  long long nv_all = 0, ng_all = 0;
  for (const std::vector<int>& c : comps) {
    nv_all += v_need(c);
    ng_all += static_cast<long long>(c.size());
  }
  p.cap_v = ng_max > 0 ? std::max(nv_max, static_cast<int>(free_slots * nv_all / (nv_all + ng_all)))
                       : 0;
  if (free_slots - p.cap_v < ng_max) p.cap_v = free_slots - ng_max;
  p.cap_g = free_slots - p.cap_v;
  p.g_base = p.v_base + p.cap_v;
  p.slots = p.g_base + p.cap_g;

  // batches: first fit, largest components first
  std::vector<int> order(comps.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return comps[a].size() > comps[b].size(); });
  std::vector<std::vector<int>> batches;
  std::vector<int> bv, bg;
  for (int c : order) {
    const int nv = v_need(comps[c]), ng = static_cast<int>(comps[c].size());
    size_t b = 0;
    while (b < batches.size() && (bv[b] + nv > p.cap_v || bg[b] + ng > p.cap_g)) ++b;
    if (b == batches.size()) {
      batches.emplace_back();
      bv.push_back(0);
      bg.push_back(0);
    }
    batches[b].push_back(c);
    bv[b] += nv;
    bg[b] += ng;
  }
  p.batches = static_cast<int>(batches.size());

  // reverse terms of a layer-1 node n with adjoint slot gs into the basic adjoints
  auto add_layer1_back = [&](std::vector<std::vector<Term>>& gb, int n, int gs) {
    for (int q : layer1_terms[n]) {
      const Product& pr = m.products[q];
      if (pr.left == pr.right) {
        gb[pr.left].push_back({2.0 * pr.coeff, gs, pr.left});
      } else {
        gb[pr.left].push_back({pr.coeff, gs, pr.right});
        gb[pr.right].push_back({pr.coeff, gs, pr.left});
      }
    }
  };
  auto basic_rows = [&](std::vector<std::vector<Term>>& gb, std::vector<Row>& rows) {
    for (int q = 0; q < B; ++q) {
      if (gb[q].empty()) continue;
      Row r;
      r.target = B + q;
      r.end = kLaneEndAdd;
      r.terms = std::move(gb[q]);
      rows.push_back(std::move(r));
    }
  };

  // the operands of Y products are basic or P nodes, whose values sit in fixed slots
  auto operand_slot = [&](int o) { return o < B ? o : pv_slot[o]; };
  std::vector<int> vslot(A, -1), gslot(A, -1);
  for (int n : p1) vslot[n] = pv_slot[n];
  if (np > 0) {
    // P adjoints start at zero; P values from the basics (one epoch: disjoint slots)
    std::vector<Row> pre;
    for (int n : p1) {
      Row z;
      z.target = gp_slot[n];
      z.end = kLaneEndStore;
      z.terms.push_back({0.0, p.ones_slot, p.ones_slot});
      pre.push_back(std::move(z));
      Row r;
      r.target = pv_slot[n];
      r.end = kLaneEndStore;
      for (int q : layer1_terms[n])
        r.terms.push_back({m.products[q].coeff, m.products[q].left, m.products[q].right});
      if (r.terms.empty()) r.terms.push_back({0.0, p.ones_slot, p.ones_slot});
      pre.push_back(std::move(r));
    }
    emit_segment(p, kLaneBinaryStore, pre);
  }
  for (const std::vector<int>& batch : batches) {
    std::vector<int> nodes;
    for (int c : batch) nodes.insert(nodes.end(), comps[c].begin(), comps[c].end());
    int nv = 0, ng = 0;
    bool has_y = false;
    for (int n : nodes) {
      if (n >= B && pv_slot[n] < 0) vslot[n] = p.v_base + nv++;
      if (n < B) vslot[n] = n;
      gslot[n] = p.g_base + ng++;
      has_y = has_y || is_y[n];
    }
    // FWD: layer-1 values and Y values (Y rows read the basic and P slots)
    std::vector<Row> fwd;
    for (int n : nodes) {
      if (n < B || pv_slot[n] >= 0) continue;
      Row r;
      r.target = vslot[n];
      r.end = kLaneEndStore;
      if (is_y[n]) {
        for (int q : y_terms[n])
          r.terms.push_back({m.products[q].coeff, operand_slot(m.products[q].left),
                             operand_slot(m.products[q].right)});
      } else {
        for (int q : layer1_terms[n])
          r.terms.push_back({m.products[q].coeff, m.products[q].left, m.products[q].right});
      }
      if (r.terms.empty()) r.terms.push_back({0.0, p.ones_slot, p.ones_slot});
      fwd.push_back(std::move(r));
    }
    emit_segment(p, kLaneBinaryStore, fwd);
    // ADJ: g_n = d_n + sum k_p v_other
    std::vector<Row> adj;
    for (int n : nodes) {
      Row r;
      r.target = gslot[n];
      r.end = kLaneEndStore;
      if (mapped[n]) r.terms.push_back({d[n], p.ones_slot, 0});
      for (int q : l2_of[n]) {
        const Product& pr = m.products[q];
        const double k = d[pr.target] * pr.coeff;
        if (pr.left == pr.right) {
          r.terms.push_back({2.0 * k, vslot[n], 0});
        } else {
          const int other = pr.left == n ? pr.right : pr.left;
          r.terms.push_back({k, vslot[other], 0});
        }
      }
      if (r.terms.empty()) r.terms.push_back({0.0, p.ones_slot, 0});
      adj.push_back(std::move(r));
    }
    emit_segment(p, kLaneUnaryStore, adj);
    // ENG + BWD
    std::vector<Row> eb;
    {
      std::vector<Term> et;
      for (int n : nodes) {
        et.push_back({0.5, vslot[n], gslot[n]});
        if (mapped[n]) et.push_back({0.5 * d[n], vslot[n], p.ones_slot});
      }
      const int parts =
        std::max(1, std::min<int>(nslot * waves, static_cast<int>(et.size() + 3) / 4));
      for (int k = 0; k < parts; ++k) {
        Row r;
        r.target = 0;
        r.end = kLaneEndEnergy;
        for (size_t i = k; i < et.size(); i += parts) r.terms.push_back(et[i]);
        eb.push_back(std::move(r));
      }
    }
    emit_segment(p, kLaneBinaryEnergy, eb);
    eb.clear();
    std::vector<std::vector<Term>> gb(B);
    for (int n : nodes) {
      if (n < B) {
        gb[n].push_back({1.0, gslot[n], p.ones_slot});
      } else if (pv_slot[n] >= 0) {
        // a P node: its quadratic-form adjoint joins the persistent one
        Row r;
        r.target = gp_slot[n];
        r.end = kLaneEndAdd;
        r.terms.push_back({1.0, gslot[n], p.ones_slot});
        eb.push_back(std::move(r));
      } else if (!is_y[n]) {
        add_layer1_back(gb, n, gslot[n]);
      }
    }
    basic_rows(gb, eb);
    // BWD reads what ENG reads (values, adjoints) and adds into the basic and P adjoints,
    // which ENG does not read: both run in one epoch
    emit_segment(p, kLaneBinaryAdd, eb, false);
    if (has_y) {
      // the Y adjoints of this batch into the adjoints of their operands (P and basic
      // slots), in an epoch of its own (the BWD rows above add into the same slots)
      std::vector<std::vector<Term>> into(A);
      for (int n : nodes) {
        if (!is_y[n]) continue;
        for (int q : y_terms[n]) {
          const Product& pr = m.products[q];
          if (pr.left == pr.right) {
            into[pr.left].push_back({2.0 * pr.coeff, gslot[n], operand_slot(pr.left)});
          } else {
            into[pr.left].push_back({pr.coeff, gslot[n], operand_slot(pr.right)});
            into[pr.right].push_back({pr.coeff, gslot[n], operand_slot(pr.left)});
          }
        }
      }
      std::vector<Row> yb;
      for (int o = 0; o < A; ++o) {
        if (into[o].empty()) continue;
        Row r;
        r.target = o < B ? B + o : gp_slot[o];
        r.end = kLaneEndAdd;
        r.terms = std::move(into[o]);
        yb.push_back(std::move(r));
      }
      emit_segment(p, kLaneBinaryAdd, yb);
    }
  }
  if (np > 0) {
    // the P nodes' complete adjoints into the basic adjoints
    std::vector<std::vector<Term>> gb(B);
    for (int n : p1) add_layer1_back(gb, n, gp_slot[n]);
    std::vector<Row> rows;
    basic_rows(gb, rows);
    emit_segment(p, kLaneBinaryAdd, rows);
  }
  if (p.needs_trash && !with_trash) {
    p.reason = "a store segment has an empty accumulator: needs a trash slot";
    return p;
  }
  if (p.slots > kLaneMaxSlot + 1) {
    p.reason = "per-lane slot count exceeds the instruction encoding";
    return p;
  }
  // every slot the kernel touches lies inside the lane's column, and every segment covers
  // whole groups inside the instruction array
  const uint32_t nslots = static_cast<uint32_t>(p.slots);
  for (size_t s = 0; s < p.seg_kind.size(); ++s) {
    const size_t b = static_cast<size_t>(p.seg_begin[s]);
    const size_t e = b + static_cast<size_t>(p.seg_groups[s]) * nslot * waves;
    if (b % nslot != 0 || e > p.ins.size() || p.seg_groups[s] <= 0) {
      p.reason = "lane program segment table is inconsistent";
      return p;
    }
    for (size_t k = b; k < e; ++k) {
      const uint32_t w = p.ins[k];
      const uint32_t end = w >> 30;
      if ((w & 1023u) >= nslots || (p.seg_kind[s] != kLaneUnaryStore && ((w >> 10) & 1023u) >= nslots) ||
          ((end == kLaneEndStore || end == kLaneEndAdd) && ((w >> 20) & 1023u) >= nslots)) {
        p.reason = "lane program addresses a slot outside the lane's column";
        return p;
      }
    }
  }
  // waves of a tile run an epoch (segments between two barriers) concurrently: within an
  // epoch every stored / added slot is the target of one row only and is never read
  {
    std::vector<int> mark(p.slots, -1);
    int epoch = -1;
    for (size_t s = 0; s < p.seg_kind.size(); ++s) {
      if (p.seg_sync[s]) ++epoch;
      const size_t b = static_cast<size_t>(p.seg_begin[s]);
      const size_t e = b + static_cast<size_t>(p.seg_groups[s]) * nslot * waves;
      for (size_t k = b; k < e; ++k) {
        const uint32_t w = p.ins[k], end = w >> 30;
        if (end == kLaneEndStore || end == kLaneEndAdd) {
          const int t = static_cast<int>((w >> 20) & 1023u);
          if ((with_trash && t == p.trash_slot) || t == p.ones_slot || t < B) {
            p.reason = "lane program writes a reserved slot";
            return p;
          }
          if (mark[t] == epoch) {
            p.reason = "lane program writes one slot twice in an epoch";
            return p;
          }
          mark[t] = epoch;
        }
      }
      for (size_t k = b; k < e; ++k) {
        const uint32_t w = p.ins[k];
        if (mark[w & 1023u] == epoch ||
            (p.seg_kind[s] != kLaneUnaryStore && mark[(w >> 10) & 1023u] == epoch)) {
          p.reason = "lane program reads a slot its epoch writes";
          return p;
        }
      }
    }
    // reads of an earlier segment of the epoch against the writes of a later one
    std::vector<int> rmark(p.slots, -1);
    epoch = -1;
    for (size_t s = 0; s < p.seg_kind.size(); ++s) {
      if (p.seg_sync[s]) ++epoch;
      const size_t b = static_cast<size_t>(p.seg_begin[s]);
      const size_t e = b + static_cast<size_t>(p.seg_groups[s]) * nslot * waves;
      for (size_t k = b; k < e; ++k) {
        const uint32_t w = p.ins[k], end = w >> 30;
        if ((end == kLaneEndStore || end == kLaneEndAdd) && rmark[(w >> 20) & 1023u] == epoch) {
          p.reason = "lane program writes a slot its epoch reads";
          return p;
        }
      }
      for (size_t k = b; k < e; ++k) {
        const uint32_t w = p.ins[k];
        rmark[w & 1023u] = epoch;
        if (p.seg_kind[s] != kLaneUnaryStore) rmark[(w >> 10) & 1023u] = epoch;
      }
    }
  }
  if (p.ins.size() != p.coeff.size() || p.ins.size() > (1u << 30)) {
    p.reason = "lane program instruction table is inconsistent";
    return p;
  }

  // self-check against the flat graph on two seeded basic vectors: the program and its FP64
  // device encoding to 1e-9, the FP32 encoding (coefficients rounded to float) to 1e-5
  uint64_t seed = 0x5eed5eedULL;
  std::vector<double> b(B), g1(B), g2(B), g3(B), g4(B);
  const LaneRecords rec8 = encode_lane_records(p, 8), rec4 = encode_lane_records(p, 4);
  for (int trial = 0; trial < 2; ++trial) {
    for (int i = 0; i < B; ++i) b[i] = (static_cast<double>(lcg(seed) % 2000001) / 1e6 - 1.0) * 0.5;
    p.ok = true;
    const double e1 = run_lane_program_host(p, b.data(), g1.data());
    const double e2 = run_flat_graph_host(m, b.data(), g2.data());
    const double e3 = run_lane_records_host(p, rec8, b.data(), g3.data());
    const double e4 = run_lane_records_host(p, rec4, b.data(), g4.data());
    double scale = std::abs(e2), err = std::max(std::abs(e1 - e2), std::abs(e3 - e2));
    double err4 = std::abs(e4 - e2);
    for (int i = 0; i < B; ++i) {
      scale = std::max(scale, std::abs(g2[i]));
      err = std::max(err, std::max(std::abs(g1[i] - g2[i]), std::abs(g3[i] - g2[i])));
      err4 = std::max(err4, std::abs(g4[i] - g2[i]));
    }
    if (!(err <= 1e-9 * std::max(1.0, scale)) || !(err4 <= 1e-5 * std::max(1.0, scale))) {
      p.ok = false;
      p.reason = "lane program self-check failed";
      return p;
    }
  }
  p.ok = true;
  return p;
}

}  // namespace

double run_lane_program_host(const LaneProgram& p, const double* basics, double* grad_b)
{
  const int B = p.basic_count, ns = p.nslot;
  std::vector<double> X(p.slots, 0.0);
  for (int i = 0; i < B; ++i) X[i] = basics[i];
  X[p.ones_slot] = 1.0;
  double e = 0.0;
  std::vector<double> acc(ns), xv(ns), yv(ns);
  for (size_t s = 0; s < p.seg_kind.size(); ++s) {
    const bool bin = p.seg_kind[s] != kLaneUnaryStore;
    // waves run concurrently on the device: each wave's stream is replayed against a
    // snapshot of the slots taken at the start of the segment, and the writes of all waves
    // are applied afterwards (rows of one segment never read a slot the segment writes)
    const std::vector<double> X0 = X;
    for (int w = 0; w < p.waves; ++w) {
      std::fill(acc.begin(), acc.end(), 0.0);
      for (int g = 0; g < p.seg_groups[s]; ++g) {
        const size_t base = static_cast<size_t>(p.seg_begin[s]) +
                            (static_cast<size_t>(w) * p.seg_groups[s] + g) * ns;
        for (int j = 0; j < ns; ++j) {
          const uint32_t wd = p.ins[base + j];
          xv[j] = X0[wd & 1023u];
          yv[j] = bin ? X0[(wd >> 10) & 1023u] : 1.0;
        }
        for (int j = 0; j < ns; ++j) acc[j] += p.coeff[base + j] * xv[j] * yv[j];
        if (p.combine[base]) {
          acc[0] += acc[1];
          acc[1] = 0.0;
        }
        for (int j = 0; j < ns; ++j) {
          const uint32_t wd = p.ins[base + j];
          const uint32_t end = wd >> 30, t = (wd >> 20) & 1023u;
          if (end == kLaneEndStore) X[t] = acc[j];
          if (end == kLaneEndAdd) X[t] += acc[j];
          if (end == kLaneEndEnergy) e += acc[j];
          if (end != kLaneEndNone) acc[j] = 0.0;
        }
      }
    }
  }
  for (int i = 0; i < B; ++i) grad_b[i] = X[B + i];
  return e;
}

LaneRecords encode_lane_records(const LaneProgram& p, int real_bytes, int col_words)
{
  LaneRecords r;
  r.nslot = p.nslot;
  r.real_bytes = real_bytes;
  r.col = col_words;
  const int ns = p.nslot, hw = r.header_words(), rw = r.rec_words();
  const uint32_t col = static_cast<uint32_t>(col_words);  // words per slot row
  size_t nrec = 0;
  for (size_t s = 0; s < p.seg_kind.size(); ++s) nrec += static_cast<size_t>(p.seg_groups[s]) * p.waves;
  r.words.assign(nrec * rw, 0u);
  auto put_real = [&](uint32_t* at, double v) {
    if (real_bytes == 4) {
      const float f = static_cast<float>(v);
      std::memcpy(at, &f, 4);
    } else {
      std::memcpy(at, &v, 8);
    }
  };
  const int rs = real_bytes / 4;
  size_t rec = 0;
  for (size_t s = 0; s < p.seg_kind.size(); ++s) {
    r.seg_first.push_back(static_cast<int>(rec));
    const size_t groups = static_cast<size_t>(p.seg_groups[s]) * p.waves;
    for (size_t q = 0; q < groups; ++q, ++rec) {
      uint32_t* R = r.words.data() + rec * rw;
      const size_t base = static_cast<size_t>(p.seg_begin[s]) + q * ns;
      uint32_t mask = 0;
      for (int j = 0; j < ns; ++j) {
        const uint32_t w = p.ins[base + j];
        const uint32_t end = w >> 30;
        R[j] = (w & 1023u) * col;
        R[ns + j] = ((w >> 10) & 1023u) * col;
        R[2 * ns + j] = static_cast<uint32_t>(p.write_slot[base + j]) * col;
        if (end != kLaneEndNone) mask |= 1u << j;
        put_real(R + hw + j * rs, p.coeff[base + j]);
      }
      if (p.combine[base]) mask |= 1u << ns;
      R[3 * ns] = mask;
    }
  }
  return r;
}

double run_lane_records_host(const LaneProgram& p, const LaneRecords& r, const double* basics,
                             double* grad_b)
{
  const int B = p.basic_count, ns = p.nslot, hw = r.header_words(), rw = r.rec_words();
  const int rs = r.real_bytes / 4;
  const uint32_t col = static_cast<uint32_t>(r.col);
  auto get_real = [&](const uint32_t* at) {
    if (r.real_bytes == 4) {
      float f;
      std::memcpy(&f, at, 4);
      return static_cast<double>(f);
    }
    double d;
    std::memcpy(&d, at, 8);
    return d;
  };
  std::vector<double> X(p.slots, 0.0);
  for (int i = 0; i < B; ++i) X[i] = basics[i];
  X[p.ones_slot] = 1.0;
  double e = 0.0;
  std::vector<double> acc(ns), xv(ns), yv(ns);
  for (size_t s = 0; s < p.seg_kind.size(); ++s) {
    const int kind = p.seg_kind[s];
    const bool bin = kind != kLaneUnaryStore;
    const std::vector<double> X0 = X;  // waves of a segment read what earlier segments wrote
    for (int w = 0; w < p.waves; ++w) {
      std::fill(acc.begin(), acc.end(), 0.0);
      for (int g = 0; g < p.seg_groups[s]; ++g) {
        const uint32_t* R = r.words.data() +
                            (static_cast<size_t>(r.seg_first[s]) + static_cast<size_t>(w) * p.seg_groups[s] + g) * rw;
        for (int j = 0; j < ns; ++j) {
          xv[j] = X0[R[j] / col];
          yv[j] = bin ? X0[R[ns + j] / col] : 1.0;
          acc[j] += get_real(R + hw + j * rs) * xv[j] * yv[j];
        }
        const uint32_t mask = R[3 * ns];
        if (kind != kLaneBinaryEnergy && (mask & (1u << ns))) {
          acc[0] += acc[1];
          acc[1] = 0.0;
        }
        for (int j = 0; j < ns; ++j) {
          const uint32_t t = R[2 * ns + j] / col;
          if (kind == kLaneBinaryStore || kind == kLaneUnaryStore) {
            if (mask & (1u << j)) {
              X[t] = acc[j];
              acc[j] = 0.0;
            }
          } else if (kind == kLaneBinaryEnergy) {
            // energy terms are summed at the end of the segment
          } else if (mask & (1u << j)) {
            X[t] += acc[j];
            acc[j] = 0.0;
          }
        }
      }
      if (kind == kLaneBinaryEnergy)
        for (int j = 0; j < ns; ++j) e += acc[j];
    }
  }
  for (int i = 0; i < B; ++i) grad_b[i] = X[B + i];
  return e;
}

double run_flat_graph_host(const Model& m, const double* basics, double* grad_b)
{
  const int A = m.moments_count, B = m.basic_count;
  std::vector<double> M(A, 0.0), G(A, 0.0);
  for (int i = 0; i < B; ++i) M[i] = basics[i];
  for (const Product& q : m.products) M[q.target] += q.coeff * M[q.left] * M[q.right];
  double e = 0.0;
  for (int s = 0; s < m.scalar_count; ++s) {
    e += m.moment_coeffs[s] * M[m.moment_mapping[s]];
    G[m.moment_mapping[s]] += m.moment_coeffs[s];
  }
  for (size_t k = m.products.size(); k-- > 0;) {
    const Product& q = m.products[k];
    const double gt = q.coeff * G[q.target];
    G[q.left] += gt * M[q.right];
    G[q.right] += gt * M[q.left];
  }
  for (int i = 0; i < B; ++i) grad_b[i] = G[i];
  return e;
}

}  // namespace sus2sh
