// ZBL short-range pair term of SUS2-SH models (`zbl_enabled = true`).
//
// Reference: src/force/sus2_zbl_common.cuh (defaults :19-32, covalent radii :34-52,
// element symbols :57-70, type-wise outer cutoff :77-90, pair tables :92-168, pair energy
// :171-225), model keys src/force/sus2_sh.cu:2132-2147, atomic numbers from the run.in
// symbols and the neighbour cutoff :7824-7860, per-edge application :7522-7627.
//
// Per directed edge i -> j with r < outer(t_i, t_j): E_i += phi(r) / 2, and the pair force
// and virial follow from d(phi/2)/dr along the edge (the engine's edge-gradient convention).
// phi(r) = 14.3996454784255 Z_i Z_j / r * sum_k a_k exp(-b_k r / a) * sw(r), with the
// universal screening 1 / a = 2.134563 (Z_i^0.23 + Z_j^0.23) and the cosine switch
// sw = (1 + cos(pi (r - inner) / (outer - inner))) / 2 for r > inner.
#pragma once

#include <string>
#include <vector>

namespace sus2sh {

struct Model;

constexpr double kZblDefaultInner = 0.7;
constexpr double kZblDefaultOuter = 1.4;
constexpr double kZblDefaultTypewiseFactor = 0.7;

// 1..94 for a known element symbol, 0 otherwise.
int zbl_atomic_number(const std::string& symbol);
// Type-wise covalent radius used for the pair outer cutoff; 0 outside 1..94.
double zbl_covalent_radius(int atomic_number);

struct ZblTables {
  std::vector<int> atomic_number;  // [S]
  std::vector<double> inner;       // [S * S]
  std::vector<double> outer;       // [S * S]
  double outer_max = 0.0;
};

// Resolve the atomic numbers (file metadata, else the species symbols that follow the
// model file in run.in) and fill the pair cutoff tables, with the reference's checks and
// messages.  Throws InputError.
ZblTables build_zbl_tables(const Model& m, int nopts, const char** opts);

}  // namespace sus2sh
