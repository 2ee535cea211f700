// Lane graph kernel: the depth <= 2 product graph of one atom per lane, driven by a
// wave-uniform program (lane_program.hpp) -- forward values, site energy and the reverse
// pass down to the basic-moment adjoints (semantics: src/force/sus2_sh.cu:4799-4846).
//
// A workgroup owns a tile of 64 atoms.  Each lane owns a column of the tile's LDS: slot s of
// lane i sits at word s * 64 + i, so a uniform slot index reads 64 consecutive words (no
// bank conflicts).  The WPT waves of the workgroup serve the same 64 atoms: the program
// splits every segment's independent rows between them, so a wave runs one WPT-th of the
// instructions and the CU holds WPT times more waves for the same LDS.  A barrier separates
// segments (a segment reads what earlier segments wrote).  Instruction words and
// coefficients are uniform across a wave and are fetched with scalar loads; each group of NS
// instructions feeds NS independent accumulators, so a group issues its NS (or 2 NS) LDS
// reads before the first FMA that needs them.
#include <algorithm>

#include "kernels.hpp"

namespace sus2sh {

namespace {

constexpr int kLaneWave = 64;

template <typename Real, int NS, int WPT>
__global__ __launch_bounds__(kLaneWave * WPT) void k_lane_graph(int N, const Real* __restrict__ basics,
                                                               const int* __restrict__ type_s,
                                                               LaneTable<Real> lt,
                                                               double* __restrict__ e_out,
                                                               Real* __restrict__ grad_out)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const unsigned lane = threadIdx.x & (kLaneWave - 1);
  const int wave = static_cast<int>(threadIdx.x / kLaneWave);
  Real* X = reinterpret_cast<Real*>(smem_raw) + lane;  // slot s of this lane: X[s * 64]
  const unsigned atom = blockIdx.x * static_cast<unsigned>(kLaneWave) + lane;
  const bool active = atom < static_cast<unsigned>(N);
  const int B = lt.basics;

  for (int q = wave; q < B; q += WPT) {
    X[q * kLaneWave] = active ? basics[static_cast<size_t>(q) * N + atom] : Real(0);
    X[(B + q) * kLaneWave] = Real(0);
  }
  if (wave == 0) X[lt.ones_slot * kLaneWave] = Real(1);

  Real e = 0;
  for (int s = 0; s < lt.nseg; ++s) {
    __syncthreads();  // every wave runs every segment: uniform barrier
    const int4 seg = lt.segs[s];
    const size_t first = static_cast<size_t>(seg.y) + static_cast<size_t>(wave) * seg.z * NS;
    const uint32_t* __restrict__ I = lt.ins + first;
    const Real* __restrict__ C = lt.coeff + first;
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
  __syncthreads();

  // per-wave energies -> the (now unused) component-value region, summed by wave 0
  if (WPT > 1) {
    X[(lt.v_base + wave) * kLaneWave] = e;
    __syncthreads();
  }
  if (active) {
    if (wave == 0) {
      Real sum = e;
#pragma unroll
      for (int w = 1; w < WPT; ++w) sum += X[(lt.v_base + w) * kLaneWave];
      const int t = type_s[atom];
      const double sg = lt.sigma[t];
      e_out[atom] = lt.shift[t] + sg + sg * static_cast<double>(sum);
    }
    const Real sgr = static_cast<Real>(lt.sigma[type_s[atom]]);
    for (int q = wave; q < B; q += WPT)
      grad_out[static_cast<size_t>(q) * N + atom] = sgr * X[(B + q) * kLaneWave];
  }
}

template <typename Real, int NS, int WPT>
void launch_lane_shape(int N, const Real* basics, const int* type_s, const LaneTable<Real>& lt,
                       double* e_s, Real* grads, hipStream_t s, size_t lds)
{
  auto kern = &k_lane_graph<Real, NS, WPT>;
  if (lds > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
  const unsigned blocks = (static_cast<unsigned>(N) + kLaneWave - 1) / kLaneWave;
  kern<<<blocks, kLaneWave * WPT, lds, s>>>(N, basics, type_s, lt, e_s, grads);
}

}  // namespace

template <typename Real>
bool launch_lane_graph(int N, const Real* basics, const int* type_s, const LaneTable<Real>& lt,
                       double* e_s, Real* grads, hipStream_t s, size_t max_lds)
{
  if (N <= 0) return true;
  const size_t lds = static_cast<size_t>(lt.slots) * kLaneWave * sizeof(Real);
  if (lds > max_lds || lt.slots < lt.v_base + lt.waves) return false;
#define SUS2SH_LANE_CASE(NS, WPT)                                                    \
  if (lt.nslot == NS && lt.waves == WPT) {                                           \
    launch_lane_shape<Real, NS, WPT>(N, basics, type_s, lt, e_s, grads, s, lds);     \
    return true;                                                                     \
  }
  SUS2SH_LANE_CASE(4, 1)
  SUS2SH_LANE_CASE(4, 2)
  SUS2SH_LANE_CASE(4, 4)
  SUS2SH_LANE_CASE(8, 1)
  SUS2SH_LANE_CASE(8, 2)
  SUS2SH_LANE_CASE(8, 4)
#undef SUS2SH_LANE_CASE
  return false;
}

template bool launch_lane_graph<float>(int, const float*, const int*, const LaneTable<float>&,
                                       double*, float*, hipStream_t, size_t);
template bool launch_lane_graph<double>(int, const double*, const int*,
                                        const LaneTable<double>&, double*, double*, hipStream_t,
                                        size_t);

}  // namespace sus2sh
