// Block program compiler and its host executor (semantics: block_program.hpp).
#include "block_program.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "cg_table.hpp"
#include "lane_program.hpp"

namespace sus2sh {

namespace {

struct CgView {
  int n;
  const CgEntry* e;
};

// run-time view of the compile-time coupling tables
const CgView& cg_view(int t)
{
  static const CgView v[kCgTypes] = {
#define SUS2SH_CG_VIEW(T, l1, l2, L) {CgTable<T>::n, CgTable<T>::e},
    SUS2SH_CG_TYPES(SUS2SH_CG_VIEW)
#undef SUS2SH_CG_VIEW
  };
  return v[t];
}

int cg_type(int l1, int l2, int L)
{
  for (int t = 0; t < kCgTypes; ++t)
    if (kCgL1[t] == l1 && kCgL2[t] == l2 && kCgLL[t] == L) return t;
  return -1;
}

struct Tensor {
  int L = 0;
  bool basic = false;
  int first = -1;        // basic tensors: index of the m = -l basic
  int ta = -1, tb = -1;  // layer-1 tensors: operand (basic) tensors, ta <= tb
  int type = -1;         // layer-1 tensors: CG block type of (L(ta), L(tb)) -> L
  double d = 0.0;        // direct scalar weight (degree 0 only)
  std::map<int, double> partners;  // tensor -> K'
  int slot = -1;         // value slot (basic tensors: their basic slot)
  int batch = -1;
};

uint64_t lcg(uint64_t& s)
{
  s = s * 6364136223846793005ULL + 1442695040888963407ULL;
  return s >> 33;
}

void put_real(std::vector<uint32_t>& w, double v, int real_bytes)
{
  if (real_bytes == 4) {
    const float f = static_cast<float>(v);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    w.push_back(u);
  } else {
    uint32_t u[2];
    std::memcpy(u, &v, 8);
    w.push_back(u[0]);
    w.push_back(u[1]);
  }
}

double get_real(const uint32_t* w, int real_bytes)
{
  if (real_bytes == 4) {
    float f;
    std::memcpy(&f, w, 4);
    return f;
  }
  double d;
  std::memcpy(&d, w, 8);
  return d;
}

// longest-first assignment of items (cost) to `bins`
std::vector<std::vector<int>> balance(const std::vector<std::pair<double, int>>& items, int bins)
{
  std::vector<std::pair<double, int>> s = items;
  std::stable_sort(s.begin(), s.end(),
                   [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
                     return a.first > b.first;
                   });
  std::vector<std::vector<int>> out(bins);
  std::vector<double> load(bins, 0.0);
  for (const auto& it : s) {
    const int b = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
    out[b].push_back(it.second);
    load[b] += it.first;
  }
  return out;
}

}  // namespace

BlockProgram compile_block_program(const Model& m, int waves, int real_bytes, int max_slots,
                                   int target_slots)
{
  BlockProgram p;
  p.waves = waves;
  p.real_bytes = real_bytes;
  const int A = m.moments_count, B = m.basic_count;
  p.basic_count = B;
  auto fail = [&](const char* why) {
    p.ok = false;
    p.reason = why;
    return p;
  };
  if (waves < 1 || (real_bytes != 4 && real_bytes != 8)) return fail("invalid block program shape");
  if (m.layer_count > 2) return fail("product graph deeper than two layers");
  const std::vector<int>& layer = m.node_layer;

  // basic tensors: runs of 2l + 1 basics of one radial function mu, m = -l .. l
  std::vector<Tensor> ten;
  std::vector<int> comp_t(A, -1), comp_m(A, -1);  // node -> (tensor, M + L)
  for (int i = 0; i < B;) {
    const int l = m.basic_l[i];
    if (l > kCgLMax) return fail("basic tensor of degree above 3");
    if (i + 2 * l + 1 > B) return fail("basic moments are not whole tensors");
    for (int j = 0; j <= 2 * l; ++j)
      if (m.basic_l[i + j] != l || m.basic_m[i + j] != j - l || m.basic_mu[i + j] != m.basic_mu[i])
        return fail("basic moments are not whole tensors in m order");
    Tensor t;
    t.L = l;
    t.basic = true;
    t.first = i;
    t.slot = i;
    for (int j = 0; j <= 2 * l; ++j) {
      comp_t[i + j] = static_cast<int>(ten.size());
      comp_m[i + j] = j;
    }
    ten.push_back(t);
    i += 2 * l + 1;
  }

  std::vector<std::vector<int>> terms_of(A);
  for (size_t q = 0; q < m.products.size(); ++q) terms_of[m.products[q].target].push_back(static_cast<int>(q));

  // layer-1 nodes: one component of a CG coupling of two basic tensors
  std::map<std::tuple<int, int, int>, int> ptensor;
  for (int n = B; n < A; ++n) {
    if (layer[n] != 1 || terms_of[n].empty()) continue;
    const Product& q0 = m.products[terms_of[n][0]];
    int ta = comp_t[q0.left], tb = comp_t[q0.right];
    if (ta < 0 || tb < 0) return fail("layer-1 product of non-basic nodes");
    if (ta > tb) std::swap(ta, tb);
    const int la = ten[ta].L, lb = ten[tb].L, na = 2 * la + 1, nb = 2 * lb + 1;
    std::vector<double> C(static_cast<size_t>(na) * nb, 0.0);
    for (int qi : terms_of[n]) {
      const Product& q = m.products[qi];
      const int tx = comp_t[q.left], ty = comp_t[q.right];
      if (tx == ta && ty == tb)
        C[comp_m[q.left] * nb + comp_m[q.right]] += q.coeff;
      else if (tx == tb && ty == ta)
        C[comp_m[q.right] * nb + comp_m[q.left]] += q.coeff;
      else
        return fail("a layer-1 node mixes tensor pairs");
    }
    const bool same = ta == tb;
    if (same)
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < i; ++j) C[i * nb + j] = C[j * nb + i] = 0.5 * (C[i * nb + j] + C[j * nb + i]);
    double cmax = 0.0;
    for (double c : C) cmax = std::max(cmax, std::abs(c));
    if (cmax == 0.0) return fail("a layer-1 node is identically zero");
    int fL = -1, fM = -1, ft = -1;
    std::vector<double> R(C.size());
    for (int L = std::abs(la - lb); L <= la + lb && fL < 0; ++L) {
      const int t = cg_type(la, lb, L);
      if (t < 0) continue;
      for (int M = 0; M <= 2 * L && fL < 0; ++M) {
        std::fill(R.begin(), R.end(), 0.0);
        const CgView& v = cg_view(t);
        for (int k = 0; k < v.n; ++k)
          if (v.e[k].M == M) R[v.e[k].m1 * nb + v.e[k].m2] = v.e[k].c;
        if (same)
          for (int i = 0; i < na; ++i)
            for (int j = 0; j < i; ++j) R[i * nb + j] = R[j * nb + i] = 0.5 * (R[i * nb + j] + R[j * nb + i]);
        bool match = true;
        for (size_t k = 0; k < C.size() && match; ++k)
          match = std::abs(C[k] - R[k]) <= 1e-12 * std::max(1.0, std::abs(R[k]));
        if (match) {
          fL = L;
          fM = M;
          ft = t;
        }
      }
    }
    if (fL < 0) return fail("a layer-1 node is not a real Clebsch-Gordan coupling of two basic tensors");
    const auto key = std::make_tuple(ta, tb, fL);
    auto it = ptensor.find(key);
    int idx;
    if (it == ptensor.end()) {
      Tensor t;
      t.L = fL;
      t.ta = ta;
      t.tb = tb;
      t.type = ft;
      idx = static_cast<int>(ten.size());
      ten.push_back(t);
      ptensor.emplace(key, idx);
    } else {
      idx = it->second;
    }
    comp_t[n] = idx;
    comp_m[n] = fM;
  }

  // energy weights: direct (degree-0 nodes) and layer-2 contractions -> K'
  std::vector<double> w(A, 0.0);
  std::vector<char> mapped(A, 0);
  for (int s = 0; s < m.scalar_count; ++s) {
    w[m.moment_mapping[s]] += m.moment_coeffs[s];
    mapped[m.moment_mapping[s]] = 1;
  }
  for (int n = 0; n < A; ++n) {
    if (!mapped[n] || layer[n] > 1) continue;
    if (comp_t[n] < 0) return fail("a mapped node is not a tensor component");
    Tensor& t = ten[comp_t[n]];
    if (t.L != 0) return fail("a scalar maps to a component of a tensor of degree above 0");
    t.d += w[n];
  }
  for (int n = B; n < A; ++n) {
    if (layer[n] != 2 || !mapped[n]) continue;
    std::map<std::pair<int, int>, std::vector<double>> pc;
    for (int qi : terms_of[n]) {
      const Product& q = m.products[qi];
      const int tx = comp_t[q.left], ty = comp_t[q.right];
      if (tx < 0 || ty < 0) return fail("a layer-2 operand is not a tensor component");
      if (ten[tx].L != ten[ty].L || comp_m[q.left] != comp_m[q.right])
        return fail("a layer-2 product couples different components");
      std::vector<double>& v = pc[std::minmax(tx, ty)];
      if (v.empty()) v.assign(2 * ten[tx].L + 1, 0.0);
      v[comp_m[q.left]] += q.coeff;
    }
    for (const auto& kv : pc) {
      const std::vector<double>& v = kv.second;
      for (double c : v)
        if (std::abs(c - v[0]) > 1e-12 * std::max(1.0, std::abs(v[0])))
          return fail("a layer-2 node is not a uniform contraction over components");
      const double k = w[n] * v[0];
      if (k == 0.0) continue;
      const int a = kv.first.first, b = kv.first.second;
      if (a == b) {
        ten[a].partners[a] += 2.0 * k;
      } else {
        ten[a].partners[b] += k;
        ten[b].partners[a] += k;
      }
    }
  }

  // live tensors and degree classes
  std::vector<int> live;
  for (int i = 0; i < static_cast<int>(ten.size()); ++i) {
    for (auto it = ten[i].partners.begin(); it != ten[i].partners.end();)
      it = it->second == 0.0 ? ten[i].partners.erase(it) : std::next(it);
    if (ten[i].d != 0.0 || !ten[i].partners.empty()) live.push_back(i);
  }
  if (live.empty()) return fail("the product graph has no live tensor");
  for (const Tensor& t : ten)
    if (t.partners.size() >= (1u << 16)) return fail("a tensor has too many partners");
  constexpr int NC = 2 * kCgLMax + 1;
  std::vector<int> class_size(NC, 0);
  std::vector<char> class_used(NC, 0);
  for (int i : live) {
    class_used[ten[i].L] = 1;
    if (!ten[i].basic) class_size[ten[i].L] += 2 * ten[i].L + 1;
  }

  // Units: a degree class, or one of its two component halves [0, L + 1), [L + 1, 2L + 1)
  // (the quadratic form couples equal components only, so a half is evaluated on its own).
  // Whole classes when they fit target_slots; else the classes of degree >= 2 that exceed it
  // are halved, when that makes every unit fit; else whole classes up to max_slots.
  struct Unit {
    int L, part, M0, M1, size, batch;
  };
  auto make_units = [&](int cap_limit) {
    std::vector<Unit> u;
    for (int L = 0; L < NC; ++L) {
      if (!class_used[L]) continue;
      if (cap_limit > 0 && class_size[L] > cap_limit && L >= 2) {
        const int w0 = L + 1, w1 = L;
        u.push_back(Unit{L, 1, 0, w0, class_size[L] / (2 * L + 1) * w0, -1});
        u.push_back(Unit{L, 2, w0, 2 * L + 1, class_size[L] / (2 * L + 1) * w1, -1});
      } else {
        u.push_back(Unit{L, 0, 0, 2 * L + 1, class_size[L], -1});
      }
    }
    return u;
  };
  int cap = 1;
  for (int sz : class_size) cap = std::max(cap, sz);
  std::vector<Unit> units = make_units(0);
  if (target_slots > 0 && B + cap > target_slots) {
    std::vector<Unit> halves = make_units(target_slots - B);
    int hcap = 1;
    for (const Unit& un : halves) hcap = std::max(hcap, un.size);
    if (B + hcap <= target_slots) {
      units = halves;
      cap = hcap;
    }
  }
  // first-fit decreasing of the units into batches of `cap` values
  std::vector<int> order(units.size());
  for (size_t k = 0; k < units.size(); ++k) order[k] = static_cast<int>(k);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return units[a].size > units[b].size; });
  std::vector<int> batch_fill;
  for (int k : order) {
    int b = -1;
    for (int j = 0; j < static_cast<int>(batch_fill.size()) && b < 0; ++j)
      if (batch_fill[j] + units[k].size <= cap) b = j;
    if (b < 0) {
      b = static_cast<int>(batch_fill.size());
      batch_fill.push_back(0);
    }
    units[k].batch = b;
    batch_fill[b] += units[k].size;
  }
  p.batches = static_cast<int>(batch_fill.size());
  p.cap = cap;
  p.v_base = B;
  p.slots = B + cap;
  p.units = static_cast<int>(units.size());
  if (p.slots > max_slots || p.slots > 1024) return fail("block program exceeds the per-lane slot budget");
  // value slot of (tensor, unit): layer-1 tensors get the unit's width in its batch; a basic
  // tensor's components M0.. sit at its basic slot + M0
  std::vector<int> fill(p.batches, 0);
  std::map<std::pair<int, int>, int> slot_of;
  for (size_t k = 0; k < units.size(); ++k) {
    const Unit& un = units[k];
    for (int i : live) {
      const Tensor& t = ten[i];
      if (t.L != un.L) continue;
      if (t.basic) {
        slot_of[{i, static_cast<int>(k)}] = t.slot + un.M0;
      } else {
        slot_of[{i, static_cast<int>(k)}] = p.v_base + fill[un.batch];
        fill[un.batch] += un.M1 - un.M0;
      }
    }
  }

  // records: per batch, the FWD blocks and the quadratic-form tensors of its units, balanced
  // over the waves
  const int col = 64;
  p.seg.assign(static_cast<size_t>(4) * p.batches * waves, 0);
  for (int b = 0; b < p.batches; ++b) {
    // items: (cost, (tensor, unit))
    std::vector<std::pair<double, int>> fwd_items, ten_items;
    std::vector<std::pair<int, int>> item;
    for (size_t k = 0; k < units.size(); ++k) {
      const Unit& un = units[k];
      if (un.batch != b) continue;
      const double frac = static_cast<double>(un.M1 - un.M0) / (2 * un.L + 1);
      for (int i : live) {
        const Tensor& t = ten[i];
        if (t.L != un.L) continue;
        const int nl = un.M1 - un.M0;
        const int id = static_cast<int>(item.size());
        item.emplace_back(i, static_cast<int>(k));
        if (!t.basic) {
          const double e = kCgEntries[t.type] * frac;
          fwd_items.emplace_back(2.0 * e + 2 * ten[t.ta].L + 2 * ten[t.tb].L + 2 + nl, id);
          p.fwd_terms += static_cast<long long>(std::lround(e));
          p.bwd_terms += static_cast<long long>(std::lround(2 * e));
          ++p.fwd_blocks;
        }
        const double bwd = t.basic ? nl : 3.0 * kCgEntries[t.type] * frac + 2 * ten[t.ta].L + 2 * ten[t.tb].L + 2;
        ten_items.emplace_back(2.0 * nl * static_cast<double>(t.partners.size()) + 2.0 * nl + bwd, id);
        p.adj_terms += static_cast<long long>(nl) * static_cast<long long>(t.partners.size());
        ++p.tensors;
      }
    }
    const auto fwd_w = balance(fwd_items, waves), ten_w = balance(ten_items, waves);
    auto kind_of = [&](int id) {
      const Tensor& t = ten[item[id].first];
      const int part = units[item[id].second].part;
      return t.basic ? kBlockBasicKind + 4 * part + t.L : t.type | part << 6;
    };
    for (int wv = 0; wv < waves; ++wv) {
      int* sg = &p.seg[(static_cast<size_t>(b) * waves + wv) * 4];
      // FWD runs
      std::vector<int> list = fwd_w[wv];
      std::stable_sort(list.begin(), list.end(), [&](int x, int y) { return kind_of(x) < kind_of(y); });
      sg[0] = static_cast<int>(p.words.size());
      for (size_t k = 0; k < list.size();) {
        size_t e = k;
        while (e < list.size() && kind_of(list[e]) == kind_of(list[k])) ++e;
        p.words.push_back(static_cast<uint32_t>(kind_of(list[k])) | static_cast<uint32_t>(e - k) << 8);
        for (size_t r = k; r < e; ++r) {
          const Tensor& t = ten[item[list[r]].first];
          const int vs = slot_of[item[list[r]]];
          p.words.push_back(static_cast<uint32_t>(vs * col) | static_cast<uint32_t>(ten[t.ta].slot * col) << 16);
          p.words.push_back(static_cast<uint32_t>(ten[t.tb].slot * col));
        }
        k = e;
      }
      sg[1] = static_cast<int>(p.words.size());
      // TENSOR runs
      list = ten_w[wv];
      std::stable_sort(list.begin(), list.end(), [&](int x, int y) { return kind_of(x) < kind_of(y); });
      sg[2] = static_cast<int>(p.words.size());
      for (size_t k = 0; k < list.size();) {
        size_t e = k;
        while (e < list.size() && kind_of(list[e]) == kind_of(list[k])) ++e;
        p.words.push_back(static_cast<uint32_t>(kind_of(list[k])) | static_cast<uint32_t>(e - k) << 8);
        for (size_t r = k; r < e; ++r) {
          const int ti = item[list[r]].first, uk = item[list[r]].second;
          const Tensor& t = ten[ti];
          const int vs = slot_of[item[list[r]]];
          p.words.push_back(static_cast<uint32_t>(vs * col) | static_cast<uint32_t>(t.partners.size()) << 16);
          const int ao = t.basic ? t.slot : ten[t.ta].slot, bo = t.basic ? t.slot : ten[t.tb].slot;
          p.words.push_back(static_cast<uint32_t>(ao * col) | static_cast<uint32_t>(bo * col) << 16);
          put_real(p.words, t.d, real_bytes);
          for (const auto& pr : t.partners) {
            p.words.push_back(static_cast<uint32_t>(slot_of[{pr.first, uk}] * col));
            put_real(p.words, pr.second, real_bytes);
          }
        }
        k = e;
      }
      sg[3] = static_cast<int>(p.words.size());
    }
  }


  // padding: a record loop may fetch the two header words after its last record
  for (int k = 0; k < 8; ++k) p.words.push_back(0u);

  // self-check against the flat graph on two seeded basic vectors (FP32 coefficients: 1e-5)
  uint64_t seed = 0xb10cb10cULL;
  std::vector<double> bv(B), g1(B), g2(B);
  p.ok = true;
  for (int trial = 0; trial < 2; ++trial) {
    for (int i = 0; i < B; ++i) bv[i] = (static_cast<double>(lcg(seed) % 2000001) / 1e6 - 1.0) * 0.5;
    const double e1 = run_block_program_host(p, bv.data(), g1.data());
    const double e2 = run_flat_graph_host(m, bv.data(), g2.data());
    double scale = std::abs(e2), err = std::abs(e1 - e2);
    for (int i = 0; i < B; ++i) {
      scale = std::max(scale, std::abs(g2[i]));
      err = std::max(err, std::abs(g1[i] - g2[i]));
    }
    const double tol = real_bytes == 8 ? 1e-9 : 1e-5;
    if (!(err <= tol * std::max(1.0, scale))) return fail("block program self-check failed");
  }
  return p;
}

double run_block_program_host(const BlockProgram& p, const double* basics, double* grad_b)
{
  const int B = p.basic_count, col = 64, rw = p.real_bytes / 4;
  std::vector<double> X(static_cast<size_t>(p.slots), 0.0), G(static_cast<size_t>(B), 0.0);
  for (int i = 0; i < B; ++i) X[i] = basics[i];
  double e = 0.0;
  const uint32_t* W = p.words.data();
  // component range of a part of degree L
  auto range = [](int L, int part, int& M0, int& M1) {
    M0 = part == 2 ? L + 1 : 0;
    M1 = part == 1 ? L + 1 : 2 * L + 1;
  };
  for (int b = 0; b < p.batches; ++b) {
    for (int wv = 0; wv < p.waves; ++wv) {
      const int* sg = &p.seg[(static_cast<size_t>(b) * p.waves + wv) * 4];
      for (int k = sg[0]; k < sg[1];) {
        const int kind = static_cast<int>(W[k] & 255u), cnt = static_cast<int>(W[k] >> 8);
        ++k;
        const int t = kind & 63;
        int M0, M1;
        range(kCgLL[t], kind >> 6, M0, M1);
        const CgView& v = cg_view(t);
        for (int r = 0; r < cnt; ++r, k += 2) {
          const int vo = static_cast<int>(W[k] & 0xFFFFu) / col, ao = static_cast<int>(W[k] >> 16) / col;
          const int bo = static_cast<int>(W[k + 1]) / col;
          for (int M = M0; M < M1; ++M) X[vo + M - M0] = 0.0;
          for (int i = 0; i < v.n; ++i)
            if (v.e[i].M >= M0 && v.e[i].M < M1)
              X[vo + v.e[i].M - M0] += v.e[i].c * X[ao + v.e[i].m1] * X[bo + v.e[i].m2];
        }
      }
    }
    for (int wv = 0; wv < p.waves; ++wv) {
      const int* sg = &p.seg[(static_cast<size_t>(b) * p.waves + wv) * 4];
      for (int k = sg[2]; k < sg[3];) {
        const int kind = static_cast<int>(W[k] & 255u), cnt = static_cast<int>(W[k] >> 8);
        ++k;
        const bool basic = kind >= kBlockBasicKind;
        const int L = basic ? (kind - kBlockBasicKind) & 3 : kCgLL[kind & 63];
        int M0, M1;
        range(L, basic ? (kind - kBlockBasicKind) >> 2 : kind >> 6, M0, M1);
        const int nl = M1 - M0;
        for (int r = 0; r < cnt; ++r) {
          const int vo = static_cast<int>(W[k] & 0xFFFFu) / col, np = static_cast<int>(W[k] >> 16);
          const int ao = static_cast<int>(W[k + 1] & 0xFFFFu) / col, bo = static_cast<int>(W[k + 1] >> 16) / col;
          const double d = get_real(W + k + 2, p.real_bytes);
          k += 2 + rw;
          double g[2 * 2 * kCgLMax + 1] = {};
          g[0] = L == 0 ? d : 0.0;
          for (int j = 0; j < np; ++j, k += 1 + rw) {
            const int po = static_cast<int>(W[k]) / col;
            const double c = get_real(W + k + 1, p.real_bytes);
            for (int M = 0; M < nl; ++M) g[M] += c * X[po + M];
          }
          double en = L == 0 ? d * X[vo] : 0.0;
          for (int M = 0; M < nl; ++M) en += X[vo + M] * g[M];
          e += 0.5 * en;
          if (basic) {
            for (int M = 0; M < nl; ++M) G[vo + M] += g[M];
          } else {
            const CgView& v = cg_view(kind & 63);
            for (int i = 0; i < v.n; ++i) {
              if (v.e[i].M < M0 || v.e[i].M >= M1) continue;
              const double cg = v.e[i].c * g[v.e[i].M - M0];
              G[ao + v.e[i].m1] += cg * X[bo + v.e[i].m2];
              G[bo + v.e[i].m2] += cg * X[ao + v.e[i].m1];
            }
          }
        }
      }
    }
  }
  for (int i = 0; i < B; ++i) grad_b[i] = G[i];
  return e;
}

}  // namespace sus2sh
