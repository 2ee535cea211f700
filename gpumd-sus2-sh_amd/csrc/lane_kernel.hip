// Lane graph kernel: the depth <= 2 product graph of one atom per lane, driven by a
// wave-uniform program (lane_program.hpp) -- forward values, site energy and the reverse
// pass down to the basic-moment adjoints (semantics: src/force/sus2_sh.cu:4799-4846).
//
// A workgroup is one wave of 64 atoms.  Each lane owns a column of the workgroup's LDS:
// slot s of lane i sits at word s * 64 + i, so a uniform slot index reads 64 consecutive
// words (no bank conflicts) and lanes never share data (no barriers).  Instruction words and
// coefficients are uniform across the wave and are fetched with scalar loads.  Each group
// of NS instructions feeds NS independent accumulators, so the NS (or 2 NS) LDS reads of a
// group are issued before the first FMA that needs them.
#include "kernels.hpp"

namespace sus2sh {

namespace {

constexpr int kLaneWave = 64;

template <typename Real, int NS>
__global__ __launch_bounds__(kLaneWave) void k_lane_graph(int N, const Real* __restrict__ basics,
                                                         const int* __restrict__ type_s,
                                                         LaneTable<Real> lt,
                                                         double* __restrict__ e_out,
                                                         Real* __restrict__ grad_out)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const unsigned lane = threadIdx.x;
  Real* X = reinterpret_cast<Real*>(smem_raw) + lane;  // slot s of this lane: X[s * 64]
  const unsigned atom = blockIdx.x * static_cast<unsigned>(kLaneWave) + lane;
  const bool active = atom < static_cast<unsigned>(N);
  const int B = lt.basics;

  for (int q = 0; q < B; ++q) {
    X[q * kLaneWave] = active ? basics[static_cast<size_t>(q) * N + atom] : Real(0);
    X[(B + q) * kLaneWave] = Real(0);
  }
  X[lt.ones_slot * kLaneWave] = Real(1);

  Real e = 0;
  for (int s = 0; s < lt.nseg; ++s) {
    const int4 seg = lt.segs[s];
    const uint32_t* __restrict__ I = lt.ins + seg.y;
    const Real* __restrict__ C = lt.coeff + seg.y;
    Real acc[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) acc[j] = Real(0);
    const bool binary = seg.x == 0;
    for (int g = 0; g < seg.z; ++g, I += NS, C += NS) {
      uint32_t w[NS];
      Real c[NS], xv[NS], yv[NS];
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        w[j] = I[j];
        c[j] = C[j];
      }
#pragma unroll
      for (int j = 0; j < NS; ++j) xv[j] = X[(w[j] & 1023u) * kLaneWave];
      if (binary) {
#pragma unroll
        for (int j = 0; j < NS; ++j) yv[j] = X[((w[j] >> 10) & 1023u) * kLaneWave];
#pragma unroll
        for (int j = 0; j < NS; ++j) acc[j] += (c[j] * xv[j]) * yv[j];
      } else {
#pragma unroll
        for (int j = 0; j < NS; ++j) acc[j] += c[j] * xv[j];
      }
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const uint32_t end = w[j] >> 30;
        if (end != 0u) {
          const uint32_t t = (w[j] >> 20) & 1023u;
          if (end == 1u)
            X[t * kLaneWave] = acc[j];
          else if (end == 2u)
            X[t * kLaneWave] += acc[j];
          else
            e += acc[j];
          acc[j] = Real(0);
        }
      }
    }
  }

  if (active) {
    const int t = type_s[atom];
    const double sg = lt.sigma[t];
    e_out[atom] = lt.shift[t] + sg + sg * static_cast<double>(e);
    const Real sgr = static_cast<Real>(sg);
    for (int q = 0; q < B; ++q)
      grad_out[static_cast<size_t>(q) * N + atom] = sgr * X[(B + q) * kLaneWave];
  }
}

}  // namespace

template <typename Real>
bool launch_lane_graph(int N, const Real* basics, const int* type_s, const LaneTable<Real>& lt,
                       double* e_s, Real* grads, hipStream_t s, size_t max_lds)
{
  if (N <= 0) return true;
  const size_t lds = static_cast<size_t>(lt.slots) * kLaneWave * sizeof(Real);
  if (lds > max_lds) return false;
  auto kern = &k_lane_graph<Real, 8>;
  if (lds > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
  const unsigned blocks = (static_cast<unsigned>(N) + kLaneWave - 1) / kLaneWave;
  kern<<<blocks, kLaneWave, lds, s>>>(N, basics, type_s, lt, e_s, grads);
  return true;
}

template bool launch_lane_graph<float>(int, const float*, const int*, const LaneTable<float>&,
                                       double*, float*, hipStream_t, size_t);
template bool launch_lane_graph<double>(int, const double*, const int*,
                                        const LaneTable<double>&, double*, double*, hipStream_t,
                                        size_t);

}  // namespace sus2sh
