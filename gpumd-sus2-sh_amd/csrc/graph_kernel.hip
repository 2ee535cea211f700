// Real-CG product graph on chip: forward rows layer by layer, site energy, and the
// reverse pass down to the basic-moment adjoints, for a tile of W atoms per workgroup.
//
// Semantics: the flat loop of src/force/sus2_sh.cu:4799-4846 (forward M[t] += c M[l] M[r],
// E_i = shift + sigma + sigma sum_s w_s M[map[s]], reverse G[l] += c G[t] M[r], ...),
// executed as the layered rows / back rows of src/force/sus2_sh.cu:1913-2093.
//
// Layout: the W atoms' values of one graph node are contiguous in LDS (vals[slot][W]), so a
// term reads both operands for all W atoms with two vector LDS loads and updates W
// accumulators; term metadata (packed slots + coefficient) is loaded once per W atoms.
// Rows of a layer are spread over the workgroup's threads, with a barrier between layers.
// Adjoints G' = G / sigma are seeded with w_s and scaled by sigma when written out.
#include "kernels.hpp"

namespace sus2sh {

namespace {

template <typename Real, int W>
__device__ __forceinline__ vec_t<Real, W> splat(Real x)
{
  vec_t<Real, W> v;
#pragma unroll
  for (int a = 0; a < W; ++a) v[a] = x;
  return v;
}

template <typename Real, int W, int NT>
__global__ __launch_bounds__(NT) void k_graph(int N, const Real* __restrict__ basics,
                                              const int* __restrict__ type_s, GraphTable<Real> gt,
                                              double* __restrict__ e_out,
                                              Real* __restrict__ grad_out, ParamGradOut pg)
{
  using V = vec_t<Real, W>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* vals = reinterpret_cast<V*>(smem_raw);
  V* grads = vals + gt.slots;
  V* red = grads + gt.slots;  // [NT / 64] wave partial energies
  Real* vals_s = reinterpret_cast<Real*>(vals);
  Real* grads_s = reinterpret_cast<Real*>(grads);

  const int tid = threadIdx.x;
  const int a0 = blockIdx.x * W;

  // basic moments of the W atoms -> slots [0, B)
  for (int idx = tid; idx < gt.basics * W; idx += NT) {
    const int b = idx / W, a = idx - b * W;
    const int atom = a0 + a;
    vals_s[idx] = atom < N ? basics[static_cast<size_t>(b) * N + atom] : Real(0);
  }
  for (int s = tid; s < gt.slots; s += NT) grads[s] = splat<Real, W>(gt.seed[s]);
  __syncthreads();

  // forward, layer by layer
  V e = splat<Real, W>(Real(0));
  for (int layer = 0; layer < gt.layers; ++layer) {
    const int re = gt.fwd_layer[layer + 1];
    for (int r = gt.fwd_layer[layer] + tid; r < re; r += NT) {
      const int4 row = gt.fwd_rows[r];
      V acc = splat<Real, W>(Real(0));
      for (int t = 0; t < row.z; ++t) {
        const Term<Real> term = gt.fwd_terms[row.y + t];
        const V x = vals[term.idx & 0xFFFFu];
        const V y = vals[term.idx >> 16];
        acc += term.c * x * y;
      }
      if (row.x >= 0) {
        vals[row.x] = acc;
      } else {
        e += gt.fwd_w[r] * acc;
      }
    }
    __syncthreads();
  }
  for (int k = tid; k < gt.n_e; k += NT) e += gt.e_w[k] * vals[gt.e_slot[k]];

  // block reduction of the per-atom site energies
#pragma unroll
  for (int a = 0; a < W; ++a) {
    Real v = e[a];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    e[a] = v;
  }
  const int wave = tid >> 6;
  if ((tid & 63) == 0) red[wave] = e;
  __syncthreads();
  if (tid < W) {
    const int atom = a0 + tid;
    if (atom < N) {
      double sum = 0.0;
      for (int w = 0; w < NT / 64; ++w) sum += static_cast<double>(reinterpret_cast<Real*>(red)[w * W + tid]);
      const int t = type_s[atom];
      const double sig = gt.sigma[t];
      e_out[atom] = gt.shift[t] + sig + sig * sum;
      if (pg.d_species != nullptr) {
        atomicAdd(pg.d_species + t, 1.0 + sum);
        atomicAdd(pg.d_shift + t, 1.0);
      }
    }
  }

  // reverse, layer by layer (descending)
  for (int layer = gt.layers - 1; layer >= 0; --layer) {
    const int ce = gt.bwd_layer[layer + 1];
    for (int c = gt.bwd_layer[layer] + tid; c < ce; c += NT) {
      const int4 ch = gt.bwd_chunks[c];
      V acc = splat<Real, W>(Real(0));
      for (int t = 0; t < ch.z; ++t) {
        const Term<Real> term = gt.bwd_terms[ch.y + t];
        const uint32_t tgt = term.idx >> 16;
        const V o = vals[term.idx & 0xFFFFu];
        if (tgt == 0xFFFFu) {
          acc += term.c * o;
        } else {
          acc += term.c * grads[tgt] * o;
        }
      }
      if (ch.w) {
#pragma unroll
        for (int a = 0; a < W; ++a) atomicAdd(grads_s + ch.x * W + a, acc[a]);
      } else {
        grads[ch.x] += acc;
      }
    }
    __syncthreads();
  }

  // sigma-scaled basic adjoints
  for (int idx = tid; idx < gt.basics * W; idx += NT) {
    const int b = idx / W, a = idx - b * W;
    const int atom = a0 + a;
    if (atom < N) {
      const Real sig = static_cast<Real>(gt.sigma[type_s[atom]]);
      grad_out[static_cast<size_t>(b) * N + atom] = sig * grads_s[idx];
    }
  }
}

template <typename Real, int W, int NT>
bool launch_tile(int N, const Real* basics, const int* type_s, const GraphTable<Real>& gt,
                 double* e_s, Real* grads, const ParamGradOut* pg, hipStream_t s, size_t max_lds)
{
  const size_t lds = graph_lds_bytes<Real>(W, gt.slots);
  if (lds > max_lds) return false;
  static bool attr_set = false;
  if (!attr_set && lds > 65536) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_graph<Real, W, NT>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(max_lds));
    attr_set = true;
  }
  ParamGradOut none{};
  const int blocks = (N + W - 1) / W;
  k_graph<Real, W, NT><<<blocks, NT, lds, s>>>(N, basics, type_s, gt, e_s, grads, pg ? *pg : none);
  return true;
}

}  // namespace

template <typename Real>
size_t graph_lds_bytes(int tile, int slots)
{
  return (2 * static_cast<size_t>(slots) + 4) * tile * sizeof(Real);
}

template <typename Real>
bool launch_graph(int tile, int N, const Real* basics, const int* type_s,
                  const GraphTable<Real>& gt, double* e_s, Real* grads, const ParamGradOut* pg,
                  hipStream_t s, size_t max_lds)
{
  if (N <= 0) return true;
  switch (tile) {
    case 1: return launch_tile<Real, 1, 256>(N, basics, type_s, gt, e_s, grads, pg, s, max_lds);
    case 2: return launch_tile<Real, 2, 256>(N, basics, type_s, gt, e_s, grads, pg, s, max_lds);
    case 4: return launch_tile<Real, 4, 256>(N, basics, type_s, gt, e_s, grads, pg, s, max_lds);
    case 8: return launch_tile<Real, 8, 256>(N, basics, type_s, gt, e_s, grads, pg, s, max_lds);
    case 16: return launch_tile<Real, 16, 256>(N, basics, type_s, gt, e_s, grads, pg, s, max_lds);
    default: return false;
  }
}

template size_t graph_lds_bytes<float>(int, int);
template size_t graph_lds_bytes<double>(int, int);
template bool launch_graph<float>(int, int, const float*, const int*, const GraphTable<float>&,
                                  double*, float*, const ParamGradOut*, hipStream_t, size_t);
template bool launch_graph<double>(int, int, const double*, const int*, const GraphTable<double>&,
                                   double*, double*, const ParamGradOut*, hipStream_t, size_t);

}  // namespace sus2sh
