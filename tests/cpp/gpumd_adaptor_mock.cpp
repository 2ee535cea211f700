// Compiles the header-only GPUMD adaptor (include/sus2sh.hpp) against stand-ins for GPUMD's
// Box (cpu_h[18], pbc_x/y/z) and GPU_Vector (data(), size()), the members the reference's
// SUS2_SH::compute reads (src/force/sus2_sh.cuh:19-25), and drives the parts that run without
// a GPU: the file probe and the constructor's error path.
//
//   gpumd_adaptor_mock <model.mtp> <invalid.mtp> <not_a_model.txt>
// prints one line per check; exits 0 when every check holds.
#include <cstdio>
#include <string>
#include <vector>

#include "sus2sh.hpp"

namespace {

struct MockBox {
  double cpu_h[18];
  int pbc_x = 1, pbc_y = 1, pbc_z = 1;
};

template <typename T>
struct MockVector {
  std::vector<T> v;
  T* data() { return v.data(); }
  const T* data() const { return v.data(); }
  size_t size() const { return v.size(); }
};

// the GPUMD form of compute() is a template: taking its address instantiates it
using GpumdCompute = void (sus2sh::SUS2_SH::*)(MockBox&, const MockVector<int>&,
                                               const MockVector<double>&, MockVector<double>&,
                                               MockVector<double>&, MockVector<double>&);
const GpumdCompute kGpumdCompute =
  &sus2sh::SUS2_SH::compute<MockBox, MockVector<int>, MockVector<double>>;

}  // namespace

int main(int argc, char** argv)
{
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s model.mtp invalid.mtp not_a_model.txt\n", argv[0]);
    return 2;
  }
  int failures = 0;
  const bool good = sus2sh::is_sus2_sh_potential_file(argv[1]);
  const bool other = sus2sh::is_sus2_sh_potential_file(argv[3]);
  std::printf("probe model=%d not_a_model=%d\n", good ? 1 : 0, other ? 1 : 0);
  if (!good || other) ++failures;

  try {
    sus2sh::SUS2_SH pot(argv[2], 8);
    std::printf("invalid model accepted\n");
    ++failures;
  } catch (const sus2sh::Error& e) {
    std::printf("invalid model: code=%d message=%s\n", e.code(), e.what());
    if (e.code() == SUS2SH_OK || std::string(e.what()).empty()) ++failures;
  }

  const char* opts[] = {"Cu", "Zr", "sus2_float=1"};
  try {
    sus2sh::SUS2_SH pot(argv[1], 8, 3, opts);
    std::printf("created rc=%.3f\n", pot.rc());
  } catch (const sus2sh::Error& e) {
    // without a GPU the constructor reports the device error through the exception
    std::printf("create: code=%d message=%s\n", e.code(), e.what());
    if (e.code() == SUS2SH_OK) ++failures;
  }
  std::printf("gpumd compute form instantiated=%d\n", kGpumdCompute != nullptr ? 1 : 0);
  return failures == 0 ? 0 : 1;
}
