// Host model layer of the SUS2-SH engine: `.mtp` reader, Clebsch-Gordan algebra,
// standard real-CG graph generator and the explicit layered graph metadata.
//
// Semantics follow the SUS2-SH GPUMD backend (src/force/sus2_sh.cu):
//   loader                       sus2_sh.cu:2105-2482
//   text helpers                 sus2_sh.cu:269-477
//   CG / real CG                 sus2_sh.cu:1090-1218
//   standard graph generator     sus2_sh.cu:1227-1706
//   graph match + rows/layers    sus2_sh.cu:1872-2093
//   radial tables                sus2_sh.cu:2484-2517
// Errors are reported by throwing sus2sh::InputError (the reference prints and
// calls exit(1) instead, sus2_sh.cu:269-273); the C ABI turns them into status codes.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace sus2sh {

struct InputError : public std::runtime_error {
  explicit InputError(const std::string& msg) : std::runtime_error(msg) {}
};

// Limits of this engine (the reference's own, src/force/sus2_sh.cu:29-33, 2273).
constexpr int kMaxL = 6;              // sh_l_max in [0,6]
constexpr int kMaxK = 6;              // per edge-kernel launch; larger k_max runs in k chunks
constexpr int kMaxRadialFuncs = 48;   // sus2_sh.cu:31
constexpr int kMaxRbSize = 16;        // sus2_sh.cu:32
constexpr int kMaxBasics = 320;       // sus2_sh.cu:33
constexpr int kMaxSpecies = 16;

constexpr int kGateLinearCombo = 1;  // two_layer_gate_mode = mu-body-linear-combo
constexpr int kGateScalarFull = 2;   // two_layer_gate_mode = mu-scalar-full (or mu-body-order)

enum class RadialKind : int {
  ChebyshevSSS = 0,          // xi = tanh(a (r - s) / 2)
  ChebyshevSSSRational = 1,  // xi = z / sqrt(1 + z^2), z = a (r - s) / 2 (README.md:55-56)
  LaguerreLog1p = 2,         // phi_n = (r - rc)^2 e^(-u/2) L_n(u), u = a log(1 + r / rho)
};

struct Product {
  int left = 0;
  int right = 0;
  int target = 0;
  double coeff = 0.0;
};

// One row of the explicit graph: target node and its product terms.
struct GraphRow {
  int layer = 0;
  int target = 0;
  int term_begin = 0;
  int term_count = 0;
};

struct Model {
  // scalar header
  std::string version;
  std::string potential_tag;
  std::string scaling_map;
  std::string radial_basis_type;
  RadialKind radial_kind = RadialKind::ChebyshevSSS;
  double scaling = 1.0;
  double max_dist = 0.0;
  int species_count = 0;
  int l_max = 0;
  int k_max = 0;
  int body_order = 0;
  std::vector<int> body_l_max;  // indexed by body order, size 7
  int rb_size = 0;
  int radial_funcs_count = 0;
  int moments_count = 0;
  int basic_count = 0;
  int scalar_count = 0;

  // coefficient arrays
  std::vector<double> shift_coeffs;        // [S]
  std::vector<double> species_coeffs;      // [S]
  std::vector<double> scal_coeffs;         // [2 S^2 M_r], per mu: [S^2 scal | S^2 shift]
  std::vector<double> moment_coeffs;       // [Q]
  std::vector<int> moment_mapping;         // [Q]
  std::vector<double> radial_coeffs;       // [M_r * rb]
  std::vector<double> radial_type_coeffs;  // [S], trailing numbers of the mu=0 row

  // basics: (mu, l, m) with mu = k (L+1) + l
  std::vector<int> basic_k, basic_l, basic_m, basic_mu;
  std::vector<Product> products;
  std::vector<int> scalar_body_orders;  // optional sh_scalar_info

  // ZBL pair term (sus2_sh.cu:2132-2147); cutoff tables are built at potential creation,
  // when the run.in symbols are known (zbl.hpp)
  bool zbl_enabled = false;
  double zbl_inner = 0.7;
  double zbl_outer = 1.4;
  bool zbl_typewise = false;
  double zbl_typewise_factor = 0.7;
  std::vector<int> zbl_atomic_numbers;  // optional metadata, one per species

  // two-layer mu gate (docs/mu-body-order-gate.md:25-63; loader sus2_sh.cu:2153-2266,
  // 2412-2478): a first graph pass on gate moments gives per-atom scalars s_q, combined into
  // h_{a,mu}; every main-moment edge i-j is weighted by G_{ij,mu} = 1 + A tanh(a_{t_j} h_{j,mu})
  bool gate_enabled = false;
  int gate_mode = 0;                          // kGateLinearCombo or kGateScalarFull
  double gate_amplitude = 0.8;                // A in [0, 1]
  std::vector<double> gate_radial_coeffs;     // [M_r * rb] (shared (a, s) and type coeffs)
  std::vector<double> gate_additive;          // [S] a_t
  std::vector<int> gate_scalar_indices;       // [Qg] scalar index q -> alpha scalar
  std::vector<double> gate_weights;           // linear combo [Qg]; scalar full [M_r][Qg]
  std::vector<double> gate_body_mix;          // linear combo [M_r][K]: c_{mu, body}
  int gate_body_count = 0;                    // body orders 2 .. k_max + 1
  std::vector<int> gate_moment_index;         // [Qg] graph node of scalar q
  std::vector<int> gate_body_id;              // [Qg] body order - 2

  // explicit metadata (sus2_sh.cu:1913-2093)
  bool standard_graph_matched = false;
  int layer_count = 0;
  std::vector<int> node_layer;       // [A]
  std::vector<GraphRow> rows;        // targets in layer order, ascending target id per layer
  std::vector<Product> row_terms;    // terms of each row, in file order
  std::vector<int> layer_row_offsets;  // [layer_count + 2]
  int back_row_count = 0;
  int back_term_count = 0;
};

// Standard real-CG graph (sus2_sh.cu:1231-1249).
struct StandardGraph {
  int node_count = 0;
  std::vector<int> basic_k, basic_l, basic_m;
  std::vector<Product> products;
  std::vector<int> scalars;
  std::vector<int> scalar_body;
  int max_layer = 0;
};

// --- Clebsch-Gordan algebra -------------------------------------------------
double clebsch_gordan(int j1, int m1, int j2, int m2, int j, int m);
// Real-basis coupling coefficient <l1 m1; l2 m2 | L M> for real spherical harmonics
// without Condon-Shortley phase, including the i^(l1+l2-L) phase (sus2_sh.cu:1186-1218).
double real_cg(int l1, int m1, int l2, int m2, int L, int M);

StandardGraph build_standard_graph(int l_max, int k_max, int body_order,
                                   const std::vector<int>& body_l_max);

// --- loader ------------------------------------------------------------------
Model parse_model_text(const std::string& text);
Model load_model(const std::string& path);
bool probe_model_file(const std::string& path);

// Fill explicit rows / layers / back-row counts and the standard-graph flag.
void prepare_graph_metadata(Model& model);

// Radial tables for pair p = t_i S + t_j: coefficient c[p][mu][n] and (a, s)[p][mu].
// `radial` replaces the model's radial_coeffs (the gate layer's own coefficients share the
// (a, s) and type coefficients, sus2_sh.cu:2484-2517 with its override argument).
void build_radial_tables(const Model& model, std::vector<double>& coeffs,
                         std::vector<double>& scal_shift,
                         const std::vector<double>* radial = nullptr);

// True when the basics are exactly the standard full (l desc, k desc, m asc) layout.
bool has_full_basic_layout(const Model& model);

}  // namespace sus2sh
