// Host model layer: `.mtp` reader, CG algebra, standard graph generator, metadata.
// Reference semantics are cited per function (src/force/sus2_sh.cu).
#include "model.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstdint>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <tuple>
#include <unordered_map>

namespace sus2sh {

// ---------------------------------------------------------------------------
// Clebsch-Gordan coefficients (Racah closed form; sus2_sh.cu:1107-1142)
// ---------------------------------------------------------------------------
namespace {

double factorial(int n)
{
  if (n < 0) return 0.0;
  double v = 1.0;
  for (int i = 2; i <= n; ++i) v *= static_cast<double>(i);
  return v;
}

bool triangle(int a, int b, int c) { return std::abs(a - b) <= c && c <= a + b; }

}  // namespace

double clebsch_gordan(int j1, int m1, int j2, int m2, int j, int m)
{
  if (m != m1 + m2 || !triangle(j1, j2, j)) return 0.0;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m) > j) return 0.0;
  const double norm = std::sqrt((2.0 * j + 1.0) * factorial(j1 + j2 - j) *
                                factorial(j1 - j2 + j) * factorial(j2 - j1 + j) /
                                factorial(j1 + j2 + j + 1));
  const double mfac = std::sqrt(factorial(j1 + m1) * factorial(j1 - m1) * factorial(j2 + m2) *
                                factorial(j2 - m2) * factorial(j + m) * factorial(j - m));
  // summation index k over all values with non-negative factorial arguments
  const int kmin = std::max({0, j2 - j - m1, j1 + m2 - j});
  const int kmax = std::min({j1 + j2 - j, j1 - m1, j2 + m2});
  double sum = 0.0;
  for (int k = kmin; k <= kmax; ++k) {
    const double den = factorial(k) * factorial(j1 + j2 - j - k) * factorial(j1 - m1 - k) *
                       factorial(j2 + m2 - k) * factorial(j - j2 + m1 + k) *
                       factorial(j - j1 - m2 + k);
    sum += ((k & 1) ? -1.0 : 1.0) / den;
  }
  return norm * mfac * sum;
}

namespace {

// Row `real_m` of the unitary map from complex Y_l^{m} to the real harmonic Y_{l,real_m}
// (no Condon-Shortley phase; sus2_sh.cu:1149-1170).
std::complex<double> real_from_complex(int real_m, int complex_m)
{
  const double h = 1.0 / std::sqrt(2.0);
  if (real_m == 0) return complex_m == 0 ? 1.0 : 0.0;
  const int a = std::abs(real_m);
  if (std::abs(complex_m) != a) return 0.0;
  const double sgn = (a & 1) ? -1.0 : 1.0;
  if (real_m > 0) return complex_m == a ? std::complex<double>(h * sgn, 0.0)
                                        : std::complex<double>(h, 0.0);
  return complex_m == a ? std::complex<double>(0.0, -h * sgn) : std::complex<double>(0.0, h);
}

}  // namespace

double real_cg(int l1, int m1, int l2, int m2, int L, int M)
{
  std::complex<double> acc(0.0, 0.0);
  for (int cM = -L; cM <= L; ++cM) {
    const std::complex<double> uo = std::conj(real_from_complex(M, cM));
    if (std::abs(uo) == 0.0) continue;
    for (int c1 = -l1; c1 <= l1; ++c1) {
      const std::complex<double> u1 = real_from_complex(m1, c1);
      if (std::abs(u1) == 0.0) continue;
      for (int c2 = -l2; c2 <= l2; ++c2) {
        const std::complex<double> u2 = real_from_complex(m2, c2);
        if (std::abs(u2) == 0.0) continue;
        const double cg = clebsch_gordan(l1, c1, l2, c2, L, cM);
        if (std::abs(cg) < 1.0e-14) continue;
        acc += uo * cg * u1 * u2;
      }
    }
  }
  static const std::complex<double> phase[4] = {
    {1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}, {0.0, -1.0}};
  acc *= phase[(l1 + l2 - L) & 3];
  if (std::abs(acc.imag()) > 1.0e-10) {
    throw InputError("SUS2_SH standard graph produced a non-real CG coefficient.");
  }
  return acc.real();
}

// ---------------------------------------------------------------------------
// Standard real-CG graph generator (sus2_sh.cu:1227-1706)
// ---------------------------------------------------------------------------
namespace {

constexpr double kCoeffDrop = 1.0e-12;  // sus2_sh.cu:1320,1342,1383,1391,1634,1675

struct QEntry {
  int l;
  int k;
};

struct TensorNode {
  int l = 0;
  int layer = 0;
  bool zero = false;
  std::vector<int> nodes;  // 2l+1 node ids, component index m + l
};

struct CGEntry {
  int rm1, rm2, rM;
  double c;
};

struct ScalarSpec {
  int body;
  int q[5];
  int inter_l;
};

class GraphGenerator {
 public:
  GraphGenerator(int l_max, int k_max, const std::vector<int>& body_l_max)
    : l_max_(l_max), k_max_(k_max), body_l_max_(body_l_max)
  {
    // q list: l = L..0, then k = K-1..0 (sus2_sh.cu:1451-1465)
    for (int l = l_max_; l >= 0; --l)
      for (int k = k_max_ - 1; k >= 0; --k) q_.push_back({l, k});
    basic_tensor_of_q_.assign(q_.size(), -1);
  }

  StandardGraph run(int body_order)
  {
    collect_specs(body_order);
    for (size_t q = 0; q < q_.size(); ++q)
      if (required_[q]) basic_tensor(static_cast<int>(q));
    for (const ScalarSpec& spec : specs_) {
      const int node = build_scalar(spec);
      if (node >= 0) {
        out_.scalars.push_back(node);
        out_.scalar_body.push_back(spec.body);
      }
    }
    std::vector<Product> kept;
    kept.reserve(products_.size());
    for (const Product& p : products_)
      if (std::abs(p.coeff) >= kCoeffDrop) kept.push_back(p);
    out_.products.swap(kept);
    out_.node_count = node_count_;
    for (const TensorNode& t : tensors_) out_.max_layer = std::max(out_.max_layer, t.layer);
    return out_;
  }

 private:
  const std::vector<CGEntry>& cg_table(int l1, int l2, int L)
  {
    const int key = (l1 * 64 + l2) * 64 + L;
    auto it = cg_cache_.find(key);
    if (it != cg_cache_.end()) return it->second;
    std::vector<CGEntry> table;
    for (int a = -l1; a <= l1; ++a)
      for (int b = -l2; b <= l2; ++b)
        for (int c = -L; c <= L; ++c) {
          const double v = real_cg(l1, a, l2, b, L, c);
          if (std::abs(v) >= kCoeffDrop) table.push_back({a, b, c, v});
        }
    return cg_cache_.emplace(key, std::move(table)).first->second;
  }

  int new_node() { return node_count_++; }

  int basic_tensor(int q)
  {
    if (basic_tensor_of_q_[q] >= 0) return basic_tensor_of_q_[q];
    TensorNode t;
    t.l = q_[q].l;
    t.layer = 0;
    for (int m = -t.l; m <= t.l; ++m) {
      t.nodes.push_back(new_node());
      out_.basic_k.push_back(q_[q].k);
      out_.basic_l.push_back(t.l);
      out_.basic_m.push_back(m);
    }
    tensors_.push_back(t);
    basic_tensor_of_q_[q] = static_cast<int>(tensors_.size()) - 1;
    return basic_tensor_of_q_[q];
  }

  void add_product(int left, int right, int target, double coeff)
  {
    if (std::abs(coeff) < kCoeffDrop) return;
    if (right < left) std::swap(left, right);
    const uint64_t key = (static_cast<uint64_t>(left) << 42) |
                         (static_cast<uint64_t>(right) << 21) | static_cast<uint64_t>(target);
    auto it = product_index_.find(key);
    if (it != product_index_.end()) {
      products_[it->second].coeff += coeff;
      return;
    }
    product_index_.emplace(key, products_.size());
    products_.push_back({left, right, target, coeff});
  }

  // Couple tensors a (x) b -> L, memoized on the ordered (a, b, L) identity
  // (sus2_sh.cu:1292-1449).
  int couple(int a, int b, int L)
  {
    const auto key = std::make_tuple(a, b, L);
    auto found = coupled_.find(key);
    if (found != coupled_.end()) return found->second;

    const TensorNode A = tensors_[a];
    const TensorNode B = tensors_[b];
    TensorNode out;
    out.l = L;
    out.layer = 1 + std::max(A.layer, B.layer);
    if (A.zero || B.zero) {
      out.zero = true;
      return remember(key, out);
    }
    struct Local {
      int left, right, comp;
      double c;
    };
    std::vector<Local> local;
    std::map<std::tuple<int, int, int>, size_t> local_index;
    for (const CGEntry& e : cg_table(A.l, B.l, L)) {
      int left = A.nodes[e.rm1 + A.l];
      int right = B.nodes[e.rm2 + B.l];
      if (right < left) std::swap(left, right);
      const int comp = e.rM + L;
      const auto lk = std::make_tuple(left, right, comp);
      auto it = local_index.find(lk);
      if (it != local_index.end()) {
        local[it->second].c += e.c;
      } else {
        local_index.emplace(lk, local.size());
        local.push_back({left, right, comp, e.c});
      }
    }
    std::vector<Local> kept;
    for (const Local& x : local)
      if (std::abs(x.c) >= kCoeffDrop) kept.push_back(x);
    if (kept.empty()) {
      out.zero = true;
      return remember(key, out);
    }
    for (int M = -L; M <= L; ++M) out.nodes.push_back(new_node());
    const int id = remember(key, out);
    for (const Local& x : kept) add_product(x.left, x.right, tensors_[id].nodes[x.comp], x.c);
    return id;
  }

  int remember(const std::tuple<int, int, int>& key, const TensorNode& t)
  {
    tensors_.push_back(t);
    const int id = static_cast<int>(tensors_.size()) - 1;
    coupled_.emplace(key, id);
    return id;
  }

  bool allowed(int q, int body) const { return q_[q].l <= body_l_max_[body]; }

  // Tuples of q indices with non-increasing l and k (sus2_sh.cu:1472-1494).
  template <typename F>
  void for_each_tuple(int count, int body, F&& fn)
  {
    std::vector<int> tuple;
    std::function<void(int, int, int)> rec = [&](int pos, int max_l, int max_k) {
      if (pos == count) {
        fn(tuple);
        return;
      }
      for (int q = 0; q < static_cast<int>(q_.size()); ++q) {
        if (!allowed(q, body) || q_[q].l > max_l || q_[q].k > max_k) continue;
        tuple.push_back(q);
        rec(pos + 1, q_[q].l, q_[q].k);
        tuple.pop_back();
      }
    };
    rec(0, l_max_, k_max_ - 1);
  }

  void add_spec(int body, const std::vector<int>& qs, int inter_l)
  {
    ScalarSpec s{};
    s.body = body;
    for (int i = 0; i < 5; ++i) s.q[i] = i < static_cast<int>(qs.size()) ? qs[i] : -1;
    s.inter_l = inter_l;
    specs_.push_back(s);
    for (int q : qs) required_[q] = 1;
  }

  // Scalar specifications for body orders 2..6 (sus2_sh.cu:1523-1587).
  void collect_specs(int body_order)
  {
    required_.assign(q_.size(), 0);
    specs_.clear();
    if (body_order >= 2)
      for_each_tuple(1, 2, [&](const std::vector<int>& t) {
        if (q_[t[0]].l == 0) add_spec(2, t, 0);
      });
    if (body_order >= 3)
      for_each_tuple(2, 3, [&](const std::vector<int>& t) {
        if (q_[t[0]].l == q_[t[1]].l) add_spec(3, t, 0);
      });
    if (body_order >= 4)
      for_each_tuple(3, 4, [&](const std::vector<int>& t) {
        const int l0 = q_[t[0]].l, l1 = q_[t[1]].l, l2 = q_[t[2]].l;
        if (((l0 + l1 + l2) & 1) == 0 && triangle(l0, l1, l2)) add_spec(4, t, l2);
      });
    if (body_order >= 5)
      for_each_tuple(4, 5, [&](const std::vector<int>& t) {
        const int l0 = q_[t[0]].l, l1 = q_[t[1]].l, l2 = q_[t[2]].l, l3 = q_[t[3]].l;
        if ((l0 + l1 + l2 + l3) & 1) return;
        const int lo = std::max(std::abs(l0 - l1), std::abs(l2 - l3));
        const int hi = std::min(l0 + l1, l2 + l3);
        for (int L = lo; L <= hi; ++L) add_spec(5, t, L);
      });
    if (body_order >= 6)
      for_each_tuple(5, 6, [&](const std::vector<int>& t) {
        const int l0 = q_[t[0]].l, l1 = q_[t[1]].l, l2 = q_[t[2]].l, l3 = q_[t[3]].l,
                  l4 = q_[t[4]].l;
        if ((l0 + l1 + l2 + l3 + l4) & 1) return;
        const int lo = std::max(std::abs(l0 - l1), std::abs(l2 - l3));
        const int hi = std::min(l0 + l1, l2 + l3);
        for (int L = lo; L <= hi; ++L)
          if (triangle(L, l4, L)) add_spec(6, t, L);
      });
  }

  int scalar_node(int tensor) const
  {
    return tensors_[tensor].zero ? -1 : tensors_[tensor].nodes[0];
  }

  // sus2_sh.cu:1589-1629
  int build_scalar(const ScalarSpec& s)
  {
    const int t0 = basic_tensor(s.q[0]);
    if (s.body == 2) return tensors_[t0].nodes[0];
    const int t1 = basic_tensor(s.q[1]);
    if (s.body == 3) return scalar_node(couple(t0, t1, 0));
    const int t2 = basic_tensor(s.q[2]);
    if (s.body == 4) {
      const int pair = couple(t0, t1, s.inter_l);
      if (tensors_[pair].zero) return -1;
      return scalar_node(couple(pair, t2, 0));
    }
    const int t3 = basic_tensor(s.q[3]);
    const int left = couple(t0, t1, s.inter_l);
    const int right_pair = couple(t2, t3, s.inter_l);
    if (s.body == 6) {
      const int t4 = basic_tensor(s.q[4]);
      if (tensors_[left].zero || tensors_[right_pair].zero) return -1;
      const int right = couple(right_pair, t4, s.inter_l);
      if (tensors_[right].zero) return -1;
      return scalar_node(couple(left, right, 0));
    }
    if (tensors_[left].zero || tensors_[right_pair].zero) return -1;
    return scalar_node(couple(left, right_pair, 0));
  }

  int l_max_;
  int k_max_;
  std::vector<int> body_l_max_;
  std::vector<QEntry> q_;
  std::vector<int> basic_tensor_of_q_;
  std::vector<TensorNode> tensors_;
  std::map<std::tuple<int, int, int>, int> coupled_;
  std::unordered_map<uint64_t, size_t> product_index_;
  std::vector<Product> products_;
  std::vector<char> required_;
  std::vector<ScalarSpec> specs_;
  std::unordered_map<int, std::vector<CGEntry>> cg_cache_;
  int node_count_ = 0;
  StandardGraph out_;
};

}  // namespace

StandardGraph build_standard_graph(int l_max, int k_max, int body_order,
                                   const std::vector<int>& body_l_max)
{
  if (l_max < 0 || l_max > kMaxL || k_max <= 0 || body_order < 2 || body_order > 6 ||
      static_cast<int>(body_l_max.size()) < 7) {
    throw InputError("SUS2_SH standard graph: invalid generator arguments.");
  }
  GraphGenerator gen(l_max, k_max, body_l_max);
  return gen.run(body_order);
}

// ---------------------------------------------------------------------------
// Text helpers (sus2_sh.cu:269-477)
// ---------------------------------------------------------------------------
namespace {

[[noreturn]] void fail(const std::string& msg) { throw InputError(msg); }

std::string trim(const std::string& s)
{
  size_t b = 0, e = s.size();
  while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
  while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
  return s.substr(b, e - b);
}

size_t find_required(const std::string& text, const std::string& token, size_t from = 0)
{
  const size_t pos = text.find(token, from);
  if (pos == std::string::npos) fail("Missing token in SUS2-SH model file: " + token);
  return pos;
}

bool has_token(const std::string& text, const std::string& token)
{
  return text.find(token) != std::string::npos;
}

// Value after the FIRST textual occurrence of `token`, up to the end of that line.
std::string string_after(const std::string& text, const std::string& token)
{
  const size_t pos = find_required(text, token) + token.size();
  const size_t eol = text.find('\n', pos);
  return trim(text.substr(pos, eol == std::string::npos ? std::string::npos : eol - pos));
}

int int_after(const std::string& text, const std::string& token)
{
  const std::string v = string_after(text, token);
  try {
    return std::stoi(v);
  } catch (const std::exception&) {
    fail("Invalid integer in SUS2-SH model file for " + token + ": " + v);
  }
}

double double_after(const std::string& text, const std::string& token)
{
  const std::string v = string_after(text, token);
  try {
    return std::stod(v);
  } catch (const std::exception&) {
    fail("Invalid number in SUS2-SH model file for " + token + ": " + v);
  }
}

bool bool_value(const std::string& v, bool* ok)
{
  *ok = true;
  if (v == "1" || v == "true" || v == "TRUE" || v == "yes" || v == "on") return true;
  if (v == "0" || v == "false" || v == "FALSE" || v == "no" || v == "off") return false;
  *ok = false;
  return false;
}

bool optional_bool_after(const std::string& text, const std::string& token, bool dflt)
{
  if (!has_token(text, token)) return dflt;
  const std::string v = string_after(text, token);
  bool ok = false;
  const bool b = bool_value(v, &ok);
  if (!ok) fail("Invalid boolean value in SUS2-SH model file for " + token + ": " + v);
  return b;
}

// `{ ... }` block that starts at the first '{' after the first occurrence of token.
std::string braced_after(const std::string& text, const std::string& token)
{
  const size_t open = find_required(text, "{", find_required(text, token));
  int depth = 0;
  for (size_t i = open; i < text.size(); ++i) {
    if (text[i] == '{') {
      ++depth;
    } else if (text[i] == '}') {
      if (--depth == 0) return text.substr(open, i - open + 1);
    }
  }
  fail("Unterminated braced section in SUS2-SH model file: " + token);
}

// Groups `{...}` that sit directly inside the outer braces (depth 2).
std::vector<std::string> inner_groups(const std::string& text)
{
  std::vector<std::string> out;
  int depth = 0;
  size_t start = 0;
  for (size_t i = 0; i < text.size(); ++i) {
    if (text[i] == '{') {
      if (depth == 1) start = i;
      ++depth;
    } else if (text[i] == '}') {
      if (depth == 2) out.push_back(text.substr(start, i - start + 1));
      --depth;
    }
  }
  return out;
}

// All top-level `{...}` groups of a text span (depth 0 -> 1).
std::vector<std::string> outer_groups(const std::string& text)
{
  std::vector<std::string> out;
  int depth = 0;
  size_t start = 0;
  for (size_t i = 0; i < text.size(); ++i) {
    if (text[i] == '{') {
      if (depth == 0) start = i;
      ++depth;
    } else if (text[i] == '}') {
      if (--depth == 0) out.push_back(text.substr(start, i - start + 1));
    }
  }
  return out;
}

template <typename T>
std::vector<T> numbers(const std::string& text)
{
  std::string s = text;
  for (char& c : s)
    if (c == '{' || c == '}' || c == ',') c = ' ';
  std::istringstream in(s);
  std::vector<T> out;
  T v{};
  while (in >> v) out.push_back(v);
  return out;
}

RadialKind radial_kind_of(const std::string& type)
{
  if (type == "RBChebyshev_sss" || type == "RBChebyshev_sss_lmp") return RadialKind::ChebyshevSSS;
  if (type == "RBChebyshev_sss_rational" || type == "RBChebyshev_sss_rational_lmp")
    return RadialKind::ChebyshevSSSRational;
  if (type == "RBLaguerre_log1p" || type == "RBLaguerre_log1p_lmp")
    return RadialKind::LaguerreLog1p;
  fail("Unsupported SUS2_SH radial_basis_type in GPUMD: " + type +
       ". Supported now: RBChebyshev_sss[_lmp], RBChebyshev_sss_rational[_lmp], "
       "RBLaguerre_log1p[_lmp].");
}

std::string read_file(const std::string& path)
{
  std::ifstream in(path);
  if (!in) fail("Failed to open SUS2-SH model file: " + path);
  std::ostringstream ss;
  ss << in.rdbuf();
  return ss.str();
}

}  // namespace

// ---------------------------------------------------------------------------
// Explicit metadata (sus2_sh.cu:1872-2103)
// ---------------------------------------------------------------------------
namespace {

bool matches_standard(const Model& m)
{
  StandardGraph g = build_standard_graph(m.l_max, m.k_max, m.body_order, m.body_l_max);
  if (g.node_count != m.moments_count || static_cast<int>(g.basic_k.size()) != m.basic_count ||
      g.products.size() != m.products.size() || g.scalars.size() != m.moment_mapping.size())
    return false;
  for (int i = 0; i < m.basic_count; ++i) {
    if (m.basic_mu[i] != g.basic_k[i] * (m.l_max + 1) + g.basic_l[i] ||
        m.basic_l[i] != g.basic_l[i] || m.basic_m[i] != g.basic_m[i])
      return false;
  }
  for (size_t p = 0; p < m.products.size(); ++p) {
    const Product& e = g.products[p];
    const Product& a = m.products[p];
    const double tol = 2.0e-10 * (1.0 + std::abs(e.coeff));
    if (a.left != e.left || a.right != e.right || a.target != e.target ||
        std::abs(a.coeff - e.coeff) > tol)
      return false;
  }
  for (size_t i = 0; i < g.scalars.size(); ++i)
    if (g.scalars[i] != m.moment_mapping[i]) return false;
  return true;
}

}  // namespace

void prepare_graph_metadata(Model& m)
{
  const int A = m.moments_count;
  const int B = m.basic_count;
  if (A <= 0 || B <= 0 || B > A) fail("SUS2_SH invalid explicit product graph dimensions.");

  std::vector<int> last_def(A, -1);
  std::vector<std::vector<int>> terms_of(A);
  for (size_t p = 0; p < m.products.size(); ++p) {
    const Product& q = m.products[p];
    if (q.left < 0 || q.left >= A || q.right < 0 || q.right >= A || q.target < 0 ||
        q.target >= A)
      fail("SUS2_SH explicit product graph index out of range.");
    if (q.target < B) fail("SUS2_SH explicit product graph writes into a basic moment.");
    terms_of[q.target].push_back(static_cast<int>(p));
    last_def[q.target] = static_cast<int>(p);
  }
  for (size_t p = 0; p < m.products.size(); ++p) {
    const Product& q = m.products[p];
    for (int src : {q.left, q.right}) {
      if (src >= B && last_def[src] < 0)
        fail("SUS2_SH explicit product graph references an undefined moment.");
    }
    for (int src : {q.left, q.right}) {
      if (src >= B && last_def[src] >= static_cast<int>(p))
        fail("SUS2_SH explicit product graph is not topologically ordered.");
    }
  }
  std::vector<int> targets;
  std::vector<char> defined(A, 0);
  for (int i = 0; i < B; ++i) defined[i] = 1;
  for (int t = B; t < A; ++t)
    if (!terms_of[t].empty()) {
      targets.push_back(t);
      defined[t] = 1;
    }
  std::sort(targets.begin(), targets.end(), [&](int a, int b) {
    return last_def[a] != last_def[b] ? last_def[a] < last_def[b] : a < b;
  });
  for (int s = 0; s < m.scalar_count; ++s) {
    const int node = m.moment_mapping[s];
    if (node < 0 || node >= A || !defined[node])
      fail("SUS2_SH alpha_moment_mapping references an undefined moment.");
  }

  m.node_layer.assign(A, 0);
  int max_layer = 0;
  for (int t : targets) {
    int src_layer = 0;
    for (int p : terms_of[t]) {
      src_layer = std::max(src_layer, m.node_layer[m.products[p].left]);
      src_layer = std::max(src_layer, m.node_layer[m.products[p].right]);
    }
    m.node_layer[t] = src_layer + 1;
    max_layer = std::max(max_layer, src_layer + 1);
  }
  m.layer_count = max_layer;

  std::vector<std::vector<int>> by_layer(max_layer + 1);
  for (int t : targets) by_layer[m.node_layer[t]].push_back(t);
  m.rows.clear();
  m.row_terms.clear();
  m.layer_row_offsets.assign(max_layer + 2, 0);
  for (int layer = 1; layer <= max_layer; ++layer) {
    m.layer_row_offsets[layer] = static_cast<int>(m.rows.size());
    std::sort(by_layer[layer].begin(), by_layer[layer].end());
    for (int t : by_layer[layer]) {
      GraphRow r;
      r.layer = layer;
      r.target = t;
      r.term_begin = static_cast<int>(m.row_terms.size());
      r.term_count = static_cast<int>(terms_of[t].size());
      m.rows.push_back(r);
      for (int p : terms_of[t]) m.row_terms.push_back(m.products[p]);
    }
  }
  m.layer_row_offsets[max_layer + 1] = static_cast<int>(m.rows.size());

  // Back rows: per layer, one row per source node (sus2_sh.cu:2052-2092).
  m.back_row_count = 0;
  m.back_term_count = 0;
  for (int layer = 1; layer <= max_layer; ++layer) {
    std::map<int, int> per_source;
    for (int r = m.layer_row_offsets[layer]; r < m.layer_row_offsets[layer + 1]; ++r) {
      for (int t = 0; t < m.rows[r].term_count; ++t) {
        const Product& q = m.row_terms[m.rows[r].term_begin + t];
        per_source[q.left] += 1;
        per_source[q.right] += 1;
      }
    }
    m.back_row_count += static_cast<int>(per_source.size());
    for (const auto& kv : per_source) m.back_term_count += kv.second;
  }
  m.standard_graph_matched = matches_standard(m);
}

// ---------------------------------------------------------------------------
// Two-layer mu gate: header keys (sus2_sh.cu:2153-2266, mode names :495-518) and the checks
// of the gate scalar list against sh_scalar_info and the graph (sus2_sh.cu:2412-2478).  The
// checks and their messages are the reference's, so that a gate file accepted or rejected
// there is accepted or rejected here.
// ---------------------------------------------------------------------------
static void parse_gate_header(const std::string& text, Model& m)
{
  const std::string mode =
    has_token(text, "two_layer_gate_mode =") ? string_after(text, "two_layer_gate_mode =") : "legacy";
  if (mode == "mu-body-linear-combo")
    m.gate_mode = kGateLinearCombo;
  else if (mode == "mu-scalar-full" || mode == "mu-body-order")
    m.gate_mode = kGateScalarFull;
  else if (mode != "legacy")
    fail("Unsupported SUS2_SH two_layer_gate_mode in GPUMD: " + mode);
  if (m.gate_mode == 0)
    fail("GPUMD SUS2_SH mu-body-order branch requires two_layer_gate_mode = "
         "mu-body-linear-combo or mu-scalar-full; legacy gate models are not supported here.");
  if (optional_bool_after(text, "two_layer_gate_include_one_body =", false))
    fail("GPUMD SUS2_SH gate backend requires two_layer_gate_include_one_body = false.");
  if (has_token(text, "two_layer_gate_scale_mode ="))
    fail("GPUMD SUS2_SH mu-body gate does not support two_layer_gate_scale_mode.");
  if (has_token(text, "two_layer_gate_bias ="))
    fail("GPUMD SUS2_SH mu-body gate does not support two_layer_gate_bias.");
  if (optional_bool_after(text, "two_layer_residual_enabled =", false))
    fail("GPUMD SUS2_SH mu-body gate does not support residual gate fields.");
  if (has_token(text, "two_layer_gate_site_mode =")) {
    const std::string site = string_after(text, "two_layer_gate_site_mode =");
    if (site != "neighbor" && site != "double")
      fail("Unsupported SUS2_SH two_layer_gate_site_mode in GPUMD: " + site);
    if (site != "neighbor")
      fail("GPUMD SUS2_SH mu-body gate currently supports two_layer_gate_site_mode = neighbor. "
           "Double-site gate will be added after the single-gate CUDA path is verified.");
  }
  m.gate_amplitude = has_token(text, "two_layer_gate_tanh_amplitude =")
                       ? double_after(text, "two_layer_gate_tanh_amplitude =")
                       : 0.8;
  if (!std::isfinite(m.gate_amplitude) || m.gate_amplitude < 0.0 || m.gate_amplitude > 1.0)
    fail("SUS2_SH two_layer_gate_tanh_amplitude should be finite and in [0,1].");
  const std::string radial_mode = has_token(text, "two_layer_gate_radial_mode =")
                                    ? string_after(text, "two_layer_gate_radial_mode =")
                                    : std::string();
  if (radial_mode != "shared-radial")
    fail("GPUMD SUS2_SH gate backend currently requires shared-radial gate mode.");
  if (int_after(text, "two_layer_gate_body_order_max =") != m.k_max + 1)
    fail("SUS2_SH two_layer_gate_body_order_max should equal sh_k_max + 1.");
  m.gate_body_count = m.k_max;
  const int radial_count = int_after(text, "two_layer_gate_radial_coeff_count =");
  if (radial_count != m.radial_funcs_count * m.rb_size)
    fail("SUS2_SH two-layer gate radial coefficient count is inconsistent.");
  m.gate_radial_coeffs = numbers<double>(braced_after(text, "two_layer_gate_radial_coeffs ="));
  if (static_cast<int>(m.gate_radial_coeffs.size()) != radial_count)
    fail("SUS2_SH two-layer gate radial coeff list has wrong size.");
  if (has_token(text, "two_layer_gate_additive_coeff_count =")) {
    const int n = int_after(text, "two_layer_gate_additive_coeff_count =");
    if (n != m.species_count)
      fail("SUS2_SH two-layer gate additive coefficient count is inconsistent.");
    m.gate_additive = numbers<double>(braced_after(text, "two_layer_gate_additive_coeffs ="));
    if (static_cast<int>(m.gate_additive.size()) != n)
      fail("SUS2_SH two-layer gate additive coeff list has wrong size.");
  } else {
    m.gate_additive.assign(m.species_count, 1.0);
  }
  const int weight_count = int_after(text, "two_layer_gate_weight_count =");
  if (weight_count <= 0) fail("SUS2_SH two-layer gate should contain at least one scalar weight.");
  m.gate_scalar_indices = numbers<int>(braced_after(text, "two_layer_gate_scalar_indices ="));
  m.gate_weights = numbers<double>(braced_after(text, "two_layer_gate_weights ="));
  const int Qg = static_cast<int>(m.gate_scalar_indices.size());
  if (Qg <= 0) fail("SUS2_SH two-layer gate scalar index list is empty.");
  const int expected = m.gate_mode == kGateScalarFull ? m.radial_funcs_count * Qg : Qg;
  if (weight_count != expected || static_cast<int>(m.gate_weights.size()) != expected)
    fail("SUS2_SH two-layer gate scalar index/weight dimensions are inconsistent.");
  if (m.gate_mode == kGateLinearCombo) {
    const int n = int_after(text, "two_layer_gate_body_mix_weight_count =");
    if (n != m.radial_funcs_count * m.gate_body_count)
      fail("SUS2_SH two-layer gate body-mix weight count is inconsistent.");
    m.gate_body_mix = numbers<double>(braced_after(text, "two_layer_gate_body_mix_weights ="));
    if (static_cast<int>(m.gate_body_mix.size()) != n)
      fail("SUS2_SH two-layer gate body-mix weight list has wrong size.");
  } else if (has_token(text, "two_layer_gate_body_mix_weights =") ||
             has_token(text, "two_layer_gate_body_mix_weight_count =")) {
    fail("SUS2_SH mu-scalar-full gate should not contain body-mix weights.");
  }
  if (!has_token(text, "sh_scalar_info_count ="))
    fail("GPUMD SUS2_SH mu-body gate requires sh_scalar_info metadata.");
}

static void check_gate_scalars(Model& m)
{
  const int Qg = static_cast<int>(m.gate_scalar_indices.size());
  m.gate_moment_index.assign(Qg, 0);
  m.gate_body_id.assign(Qg, 0);
  std::vector<char> present(m.gate_body_count, 0);
  for (int q = 0; q < Qg; ++q) {
    const int s = m.gate_scalar_indices[q];
    if (s < 0 || s >= m.scalar_count) fail("SUS2_SH two-layer gate scalar index is out of range.");
    const int body = m.scalar_body_orders[s] - 2;
    if (body < 0 || body >= m.gate_body_count)
      fail("SUS2_SH two-layer gate scalar body order is outside the gate range.");
    m.gate_body_id[q] = body;
    present[body] = 1;
    const int node = m.moment_mapping[s];
    if (node < 0 || node >= m.moments_count)
      fail("SUS2_SH two-layer gate mapped moment index is out of range.");
    m.gate_moment_index[q] = node;
  }
  for (int b = 0; b < m.gate_body_count; ++b)
    if (!present[b]) fail("SUS2_SH two-layer gate scalar list misses a required body-order bucket.");
}

// ---------------------------------------------------------------------------
// Loader (sus2_sh.cu:2105-2482)
// ---------------------------------------------------------------------------
Model parse_model_text(const std::string& text)
{
  Model m;
  m.version = string_after(text, "version =");
  if (m.version != "1.1.0") fail("SUS2_SH currently supports only version = 1.1.0.");
  m.potential_tag = string_after(text, "potential_tag =");
  if (m.potential_tag != "SUS2-SH") fail("SUS2_SH selected for a non-SUS2-SH model.");
  m.scaling = double_after(text, "scaling =");
  m.scaling_map = string_after(text, "scaling_map =");
  m.species_count = int_after(text, "species_count =");
  m.l_max = int_after(text, "sh_l_max =");
  m.k_max = int_after(text, "sh_k_max =");
  m.body_order = int_after(text, "sh_body_order =");
  if (has_token(text, "sh_parity =") && string_after(text, "sh_parity =") != "even")
    fail("SUS2_SH GPUMD backend currently supports even parity models only.");
  m.radial_basis_type = string_after(text, "radial_basis_type =");
  m.radial_kind = radial_kind_of(m.radial_basis_type);
  m.max_dist = double_after(text, "max_dist =");
  m.zbl_enabled = optional_bool_after(text, "zbl_enabled =", false);
  if (m.zbl_enabled) {
    const auto optional_double = [&](const char* token, double dflt) {
      return has_token(text, token) ? double_after(text, token) : dflt;
    };
    m.zbl_inner = optional_double("zbl_inner =", 0.7);
    m.zbl_outer = optional_double("zbl_outer =", 1.4);
    // type-wise cutoffs: explicit flag, else implied by the presence of a factor
    m.zbl_typewise = has_token(text, "zbl_typewise_cutoff_enabled =")
                       ? optional_bool_after(text, "zbl_typewise_cutoff_enabled =", false)
                       : has_token(text, "zbl_typewise_cutoff_factor =");
    m.zbl_typewise_factor = optional_double("zbl_typewise_cutoff_factor =", 0.7);
    if (has_token(text, "zbl_atomic_numbers ="))
      m.zbl_atomic_numbers = numbers<int>(braced_after(text, "zbl_atomic_numbers ="));
  }
  m.rb_size = int_after(text, "radial_basis_size =");
  m.radial_funcs_count = int_after(text, "radial_funcs_count =");
  m.moments_count = int_after(text, "alpha_moments_count =");
  m.basic_count = int_after(text, "alpha_index_basic_count =");
  m.scalar_count = int_after(text, "alpha_scalar_moments =");
  m.gate_enabled = optional_bool_after(text, "two_layer_gate_enabled =", false);
  if (m.gate_enabled) parse_gate_header(text, m);

  if (m.scaling_map != "LK") fail("SUS2_SH requires scaling_map = LK.");
  if (m.l_max < 0 || m.l_max > kMaxL) fail("SUS2_SH supports sh_l_max in [0,6].");
  if (m.body_order < 2 || m.body_order > 6) fail("SUS2_SH supports sh_body_order in [2,6].");
  if (m.k_max <= 0 || m.radial_funcs_count != m.k_max * (m.l_max + 1))
    fail("SUS2_SH inconsistent sh_k_max/radial_funcs_count.");
  if (m.rb_size <= 0 || m.rb_size > kMaxRbSize || m.radial_funcs_count > kMaxRadialFuncs)
    fail("SUS2_SH GPU scratch limit exceeded: rb_size<=16, radial_funcs_count<=48.");
  if (m.basic_count <= 0 || m.basic_count > kMaxBasics)
    fail("SUS2_SH GPU scratch limit exceeded: alpha_index_basic_count<=320.");
  if (m.species_count <= 0) fail("SUS2_SH species_count should be positive.");

  m.body_l_max.assign(7, m.l_max);
  if (has_token(text, "sh_body_l_max =")) {
    const std::vector<int> v = numbers<int>(braced_after(text, "sh_body_l_max ="));
    const int expected = m.body_order >= 6 ? 5 : 4;
    if (static_cast<int>(v.size()) != expected) fail("Unexpected SUS2_SH sh_body_l_max size.");
    for (int body = 2; body <= 1 + expected; ++body) m.body_l_max[body] = v[body - 2];
  }
  for (int body = 2; body <= m.body_order; ++body)
    if (m.body_l_max[body] < 0 || m.body_l_max[body] > m.l_max)
      fail("SUS2_SH sh_body_l_max entry is out of range.");

  m.shift_coeffs = numbers<double>(braced_after(text, "shift_coeffs ="));
  m.scal_coeffs = numbers<double>(braced_after(text, "scal_coeffs ="));
  m.species_coeffs = numbers<double>(braced_after(text, "species_coeffs ="));
  m.moment_coeffs = numbers<double>(braced_after(text, "moment_coeffs ="));
  m.moment_mapping = numbers<int>(braced_after(text, "alpha_moment_mapping ="));

  const std::vector<int> braw = numbers<int>(braced_after(text, "alpha_index_basic ="));
  if (static_cast<int>(braw.size()) != 3 * m.basic_count)
    fail("Unexpected SUS2_SH alpha_index_basic size.");
  for (int i = 0; i < m.basic_count; ++i) {
    const int k = braw[3 * i], l = braw[3 * i + 1], mm = braw[3 * i + 2];
    if (k < 0 || k >= m.k_max || l < 0 || l > m.l_max || std::abs(mm) > l)
      fail("Invalid SUS2_SH alpha_index_basic entry.");
    m.basic_k.push_back(k);
    m.basic_l.push_back(l);
    m.basic_m.push_back(mm);
    m.basic_mu.push_back(k * (m.l_max + 1) + l);
  }

  const int product_count = int_after(text, "sh_product_count =");
  const std::vector<std::string> groups = inner_groups(braced_after(text, "sh_products ="));
  if (static_cast<int>(groups.size()) != product_count)
    fail("Unexpected SUS2_SH sh_products size.");
  m.products.resize(product_count);
  for (int p = 0; p < product_count; ++p) {
    const std::vector<double> v = numbers<double>(groups[p]);
    if (v.size() != 4) fail("Invalid SUS2_SH product entry.");
    Product& q = m.products[p];
    q.left = static_cast<int>(v[0]);
    q.right = static_cast<int>(v[1]);
    q.target = static_cast<int>(v[2]);
    q.coeff = v[3];
    if (q.left < 0 || q.left >= m.moments_count || q.right < 0 ||
        q.right >= m.moments_count || q.target < 0 || q.target >= m.moments_count)
      fail("SUS2_SH product index out of range.");
  }

  if (has_token(text, "sh_scalar_info_count =")) {
    const int n = int_after(text, "sh_scalar_info_count =");
    if (n != m.scalar_count)
      fail("SUS2_SH sh_scalar_info_count should equal alpha_scalar_moments.");
    const std::vector<std::string> info = inner_groups(braced_after(text, "sh_scalar_info ="));
    if (static_cast<int>(info.size()) != n) fail("Unexpected SUS2_SH sh_scalar_info size.");
    for (const std::string& g : info) {
      const std::vector<int> v = numbers<int>(g);
      if (v.empty()) fail("Invalid SUS2_SH sh_scalar_info entry.");
      if (v[0] < 2 || v[0] > m.body_order)
        fail("SUS2_SH sh_scalar_info body order is out of range.");
      m.scalar_body_orders.push_back(v[0]);
    }
  }

  m.radial_coeffs.assign(static_cast<size_t>(m.radial_funcs_count) * m.rb_size, 0.0);
  m.radial_type_coeffs.assign(m.species_count, 1.0);
  const size_t r0 = find_required(text, "radial_coeffs");
  const size_t r1 = find_required(text, "alpha_moments_count", r0);
  const std::vector<std::string> rgroups = outer_groups(text.substr(r0, r1 - r0));
  if (static_cast<int>(rgroups.size()) < m.radial_funcs_count)
    fail("Unexpected SUS2_SH radial_coeffs section.");
  for (int mu = 0; mu < m.radial_funcs_count; ++mu) {
    const std::vector<double> v = numbers<double>(rgroups[mu]);
    if (static_cast<int>(v.size()) != m.rb_size + m.species_count)
      fail("Unexpected SUS2_SH radial_coeffs row size.");
    for (int n = 0; n < m.rb_size; ++n) m.radial_coeffs[mu * m.rb_size + n] = v[n];
    if (mu == 0)
      for (int t = 0; t < m.species_count; ++t) m.radial_type_coeffs[t] = v[m.rb_size + t];
  }

  const size_t S = static_cast<size_t>(m.species_count);
  if (m.shift_coeffs.size() != S || m.species_coeffs.size() != S ||
      m.scal_coeffs.size() != 2 * S * S * m.radial_funcs_count ||
      static_cast<int>(m.moment_mapping.size()) != m.scalar_count ||
      static_cast<int>(m.moment_coeffs.size()) != m.scalar_count)
    fail("Unexpected SUS2_SH coefficient dimensions.");

  if (m.gate_enabled) check_gate_scalars(m);
  prepare_graph_metadata(m);
  return m;
}

Model load_model(const std::string& path) { return parse_model_text(read_file(path)); }

bool probe_model_file(const std::string& path)
{
  std::ifstream in(path);
  if (!in) return false;
  std::string first;
  in >> first;
  if (first != "MTP") return false;
  std::ostringstream ss;
  ss << in.rdbuf();
  return ss.str().find("potential_tag = SUS2-SH") != std::string::npos;
}

void build_radial_tables(const Model& m, std::vector<double>& coeffs,
                         std::vector<double>& scal_shift, const std::vector<double>* radial)
{
  const std::vector<double>& rc = radial ? *radial : m.radial_coeffs;
  const int S = m.species_count;
  const int Mr = m.radial_funcs_count;
  const int rb = m.rb_size;
  coeffs.assign(static_cast<size_t>(S) * S * Mr * rb, 0.0);
  scal_shift.assign(static_cast<size_t>(S) * S * Mr * 2, 0.0);
  for (int ti = 0; ti < S; ++ti)
    for (int tj = 0; tj < S; ++tj) {
      const int pair = ti * S + tj;
      const double scale = m.scaling * m.radial_type_coeffs[ti] * m.radial_type_coeffs[tj];
      for (int mu = 0; mu < Mr; ++mu) {
        const size_t so = static_cast<size_t>(2) * mu * S * S + static_cast<size_t>(S) * ti + tj;
        scal_shift[(static_cast<size_t>(pair) * Mr + mu) * 2 + 0] = m.scal_coeffs[so];
        scal_shift[(static_cast<size_t>(pair) * Mr + mu) * 2 + 1] =
          m.scal_coeffs[so + static_cast<size_t>(S) * S];
        for (int n = 0; n < rb; ++n)
          coeffs[(static_cast<size_t>(pair) * Mr + mu) * rb + n] =
            rc[static_cast<size_t>(mu) * rb + n] * scale;
      }
    }
}

bool has_full_basic_layout(const Model& m)
{
  if (m.basic_count != m.k_max * (m.l_max + 1) * (m.l_max + 1)) return false;
  int i = 0;
  for (int l = m.l_max; l >= 0; --l)
    for (int k = m.k_max - 1; k >= 0; --k)
      for (int mm = -l; mm <= l; ++mm, ++i)
        if (m.basic_k[i] != k || m.basic_l[i] != l || m.basic_m[i] != mm) return false;
  return true;
}

}  // namespace sus2sh
