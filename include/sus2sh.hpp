// sus2sh.hpp -- header-only C++ mirror of the GPUMD plugin class on top of the C ABI.
//
// `sus2sh::SUS2_SH` keeps the reference's constructor and compute() signatures
// (src/force/sus2_sh.cuh:12-25) so a GPUMD tree can wrap it in its Potential subclass.
// It is templated on the caller's Box / GPU_Vector types: anything
// with `cpu_h[18]`, `pbc_x/pbc_y/pbc_z` (Box) and `data()` / `size()` (vectors) works.
// Errors throw sus2sh::Error carrying the reference's message text; the reference prints
// it and calls exit(1) (src/force/sus2_sh.cu:269-273).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "sus2sh.h"

namespace sus2sh {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline bool is_sus2_sh_potential_file(const char* file) { return sus2sh_probe(file) == 1; }

class SUS2_SH {
 public:
  SUS2_SH(const char* file_potential, int num_atoms, int num_potential_options = 0,
          const char** potential_options = nullptr)
  {
    const int rc = sus2sh_create(file_potential, num_atoms, num_potential_options,
                                 potential_options, &h_);
    if (rc != SUS2SH_OK) throw Error(rc, sus2sh_last_error(nullptr));
    rc_ = sus2sh_cutoff(h_);
  }
  ~SUS2_SH() { sus2sh_destroy(h_); }
  SUS2_SH(const SUS2_SH&) = delete;
  SUS2_SH& operator=(const SUS2_SH&) = delete;

  double rc() const { return rc_; }

  // Raw-pointer form: SoA device arrays, outputs accumulated (+=).
  void compute(const sus2sh_box& box, int num_atoms, const int* type, const double* position,
               double* potential, double* force, double* virial, void* stream = nullptr)
  {
    const int rc = sus2sh_compute(h_, &box, num_atoms, type, position, potential, force, virial,
                                  stream);
    if (rc != SUS2SH_OK) throw Error(rc, sus2sh_last_error(h_));
  }

  // GPUMD form: compute(Box&, type, position, potential, force, virial).
  template <class BoxT, class IntVec, class DoubleVec>
  void compute(BoxT& box, const IntVec& type, const DoubleVec& position, DoubleVec& potential,
               DoubleVec& force, DoubleVec& virial)
  {
    sus2sh_box b;
    for (int k = 0; k < 18; ++k) b.h[k] = box.cpu_h[k];
    b.pbc[0] = box.pbc_x;
    b.pbc[1] = box.pbc_y;
    b.pbc[2] = box.pbc_z;
    compute(b, static_cast<int>(type.size()), type.data(), position.data(), potential.data(),
            force.data(), virial.data());
  }

  sus2sh_model_info info() const
  {
    sus2sh_model_info mi;
    sus2sh_get_model_info(h_, &mi);
    return mi;
  }

  sus2sh_handle* handle() { return h_; }

 private:
  sus2sh_handle* h_ = nullptr;
  double rc_ = 0.0;
};

}  // namespace sus2sh
