#include "zbl.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>

#include "model.hpp"

namespace sus2sh {

namespace {

// Element symbols H..Pu in atomic-number order (src/force/sus2_zbl_common.cuh:57-70).
const char* const kElements =
  "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca "
  "Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr Rb Sr Y Zr "
  "Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd "
  "Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg "
  "Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu";

// Type-wise covalent radii (A) for Z = 1..94, ten per row (src/force/sus2_zbl_common.cuh:
// 34-52; kept identical so that type-wise cutoffs match the reference).
constexpr double kCovalentRadius[94] = {
  0.426667, 0.613333, 1.6,      1.25333, 1.02667, 1.0,     0.946667, 0.84,    0.853333, 0.893333,
  1.86667,  1.66667,  1.50667,  1.38667, 1.46667, 1.36,    1.32,     1.28,    2.34667,  2.05333,
  1.77333,  1.62667,  1.61333,  1.46667, 1.42667, 1.38667, 1.33333,  1.32,    1.34667,  1.45333,
  1.49333,  1.45333,  1.53333,  1.46667, 1.52,    1.56,    2.52,     2.22667, 1.96,     1.85333,
  1.76,     1.65333,  1.53333,  1.50667, 1.50667, 1.44,    1.53333,  1.64,    1.70667,  1.68,
  1.68,     1.64,     1.76,     1.74667, 2.78667, 2.34667, 2.16,     1.96,    2.10667,  2.09333,
  2.08,     2.06667,  2.01333,  2.02667, 2.01333, 2.0,     1.98667,  1.98667, 1.97333,  2.04,
  1.94667,  1.82667,  1.74667,  1.64,    1.57333, 1.54667, 1.48,     1.49333, 1.50667,  1.76,
  1.73333,  1.73333,  1.81333,  1.74667, 1.84,    1.89333, 2.68,     2.41333, 2.22667,  2.10667,
  2.02667,  2.04,     2.05333,  2.06667};

[[noreturn]] void zbl_fail(const std::string& msg) { throw InputError(msg); }

}  // namespace

int zbl_atomic_number(const std::string& symbol)
{
  std::istringstream in(kElements);
  std::string s;
  for (int z = 1; in >> s; ++z)
    if (s == symbol) return z;
  return 0;
}

double zbl_covalent_radius(int z) { return z >= 1 && z <= 94 ? kCovalentRadius[z - 1] : 0.0; }

ZblTables build_zbl_tables(const Model& m, int nopts, const char** opts)
{
  const int S = m.species_count;
  ZblTables t;
  t.atomic_number = m.zbl_atomic_numbers;
  if (t.atomic_number.empty()) {  // from the element symbols after the model file
    if (nopts < S)
      zbl_fail("SUS2_SH ZBL model requires zbl_atomic_numbers metadata or element symbols "
               "after the model file.");
    for (int s = 0; s < S; ++s) {
      const std::string sym = opts[s] ? opts[s] : "";
      const int z = zbl_atomic_number(sym);
      if (z <= 0) zbl_fail("SUS2_SH ZBL cannot map element symbol to an atomic number: " + sym);
      t.atomic_number.push_back(z);
    }
  }
  if (static_cast<int>(t.atomic_number.size()) != S)
    zbl_fail("SUS2_SH ZBL metadata should provide one atomic number per species.");
  const auto error = [](const std::string& why) {
    zbl_fail("SUS2_SH ZBL metadata error: " + why);
  };
  // pair tables (src/force/sus2_zbl_common.cuh:92-168)
  const bool typewise = m.zbl_typewise;
  const double factor = typewise ? m.zbl_typewise_factor : 0.0;
  if (m.zbl_inner < 0.0) error("ZBL inner cutoff should be non-negative.");
  if (m.zbl_outer <= 0.0) error("ZBL outer cutoff should be positive.");
  if (typewise) {
    if (factor < 0.5) error("ZBL typewise cutoff factor should be at least 0.5.");
  } else if (m.zbl_outer <= m.zbl_inner) {
    error("ZBL cutoffs should satisfy 0 <= inner < outer.");
  }
  t.inner.assign(static_cast<size_t>(S) * S, 0.0);
  t.outer.assign(static_cast<size_t>(S) * S, 0.0);
  for (int a = 0; a < S; ++a) {
    for (int b = 0; b < S; ++b) {
      const int za = t.atomic_number[a], zb = t.atomic_number[b];
      for (int z : {za, zb})
        if (z <= 0 || (typewise && zbl_covalent_radius(z) <= 0.0))
          error("ZBL atomic numbers should be in [1, 94].");
      double inner = m.zbl_inner, outer = m.zbl_outer;
      if (typewise) {
        // outer = min(global outer, factor (r_a + r_b)), inner = 0
        inner = 0.0;
        outer = std::min(m.zbl_outer, factor * (zbl_covalent_radius(za) + zbl_covalent_radius(zb)));
      }
      if (outer <= inner) error("ZBL pair cutoffs should satisfy inner < outer.");
      t.inner[static_cast<size_t>(a) * S + b] = inner;
      t.outer[static_cast<size_t>(a) * S + b] = outer;
      t.outer_max = std::max(t.outer_max, outer);
    }
  }
  return t;
}

}  // namespace sus2sh
