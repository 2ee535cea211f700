// Two-layer mu gate (docs/mu-body-order-gate.md:25-63; per step src/force/sus2_sh.cu:
// 9821-10256): the gate layer's product graph on a tile of W atoms per workgroup.
//
//   mode 0 (forward, src/force/sus2_sh.cu:4565-4664): gate moments -> forward rows -> the gate
//     scalars s_q -> h_{a,mu} = sum_q weff[mu][q] s_q(a) (weff is the mu-scalar-full weight
//     matrix, or c_{mu,body(q)} w_q for mu-body-linear-combo, folded on the host), and per
//     atom and channel G = 1 + A tanh(a_t h), dG/dh = A a_t (1 - tanh^2);
//   mode 1 (reverse, :4666-4790, :10112-10140): forward rows again, the gate scalars' adjoints
//     seeded per atom with sum_mu adj[a][mu] weff[mu][q] (adj = dE/dh gathered by the gated
//     force kernel), back chunks down to the gate basic-moment adjoints.
//
// The tile state is the layout of the main graph's row kernel: the W atoms' values of a node
// are one W-vector in LDS (vals[slot][W], grads[slot][W]); forward rows and back chunks of a
// layer are spread over the 256 threads with a barrier between layers.  The program comes
// from compile_graph(..., store_mapped = true) on the gate scalars, so every gate scalar has a
// slot and no row is fused into an energy.
#include "kernels.hpp"

namespace sus2sh {

namespace {

constexpr int kGateThreads = 256;

template <typename Real, int W>
__device__ __forceinline__ vec_t<Real, W> gsplat(Real x)
{
  vec_t<Real, W> v;
#pragma unroll
  for (int a = 0; a < W; ++a) v[a] = x;
  return v;
}

template <typename Real, int W>
__global__ __launch_bounds__(kGateThreads) void k_gate_graph(
    int mode, int N, const Real* __restrict__ gbasics, const int* __restrict__ type_s,
    GraphTable<Real> gt, GateTable<Real> ga, Real* __restrict__ mult, Real* __restrict__ dmult,
    const Real* __restrict__ adj, Real* __restrict__ ggrad)
{
  using V = vec_t<Real, W>;
  constexpr int NT = kGateThreads;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* vals = reinterpret_cast<V*>(smem_raw);  // [slots + 1]
  V* grads = vals + gt.slots + 1;             // [slots + 1]
  Real* vals_s = reinterpret_cast<Real*>(vals);
  Real* grads_s = reinterpret_cast<Real*>(grads);
  const int tid = threadIdx.x;
  const int a0 = blockIdx.x * W;

  for (int idx = tid; idx < gt.basics * W; idx += NT) {
    const int b = idx / W, a = idx - b * W;
    const int atom = a0 + a;
    vals_s[idx] = atom < N ? gbasics[static_cast<size_t>(b) * N + atom] : Real(0);
  }
  if (mode == 1)
    for (int sl = tid; sl <= gt.slots; sl += NT) grads[sl] = gsplat<Real, W>(Real(0));
  __syncthreads();

  for (int layer = 0; layer < gt.layers; ++layer) {
    const int re = gt.fwd_layer[layer + 1];
    for (int r = gt.fwd_layer[layer] + tid; r < re; r += NT) {
      const int4 row = gt.fwd_rows[r];
      V acc = gsplat<Real, W>(Real(0));
      for (int t = 0; t < row.z; ++t) {
        const Term<Real> term = gt.fwd_terms[row.y + t];
        acc += term.c * vals[term.idx & 0xFFFFu] * vals[term.idx >> 16];
      }
      vals[row.x] = acc;  // every row of the gate program is stored
    }
    __syncthreads();
  }

  if (mode == 0) {
    for (int k = tid; k < W * ga.Mr; k += NT) {
      const int a = k / ga.Mr, mu = k - a * ga.Mr;
      const int atom = a0 + a;
      if (atom >= N) continue;
      Real h = 0;
      for (int q = 0; q < ga.Qg; ++q) h += ga.weff[mu * ga.Qg + q] * vals_s[ga.slot[q] * W + a];
      const Real ac = static_cast<Real>(ga.additive[type_s[atom]]);
      const Real amp = static_cast<Real>(ga.amplitude);
      Real t, sech2;
      dev_tanh_sech2<Real>(ac * h, t, sech2);
      mult[static_cast<size_t>(atom) * ga.Mr + mu] = Real(1) + amp * t;
      dmult[static_cast<size_t>(atom) * ga.Mr + mu] = amp * ac * sech2;
    }
    return;
  }

  // per-atom seeds of the gate scalars (several scalars may share a node: LDS atomics)
  for (int k = tid; k < W * ga.Qg; k += NT) {
    const int q = k / W, a = k - q * W;
    const int atom = a0 + a;
    if (atom >= N) continue;
    const Real* adj_a = adj + static_cast<size_t>(atom) * ga.Mr;
    Real seed = 0;
    for (int mu = 0; mu < ga.Mr; ++mu) seed += adj_a[mu] * ga.weff[mu * ga.Qg + q];
    atomicAdd(grads_s + ga.slot[q] * W + a, seed);
  }
  __syncthreads();

  for (int layer = gt.layers - 1; layer >= 0; --layer) {
    const int ce = gt.bwd_layer[layer + 1];
    for (int c = gt.bwd_layer[layer] + tid; c < ce; c += NT) {
      const int4 ch = gt.bwd_chunks[c];
      V acc = gsplat<Real, W>(Real(0));
      for (int t = 0; t < ch.z; ++t) {
        const Term<Real> term = gt.bwd_terms[ch.y + t];
        acc += term.c * grads[term.idx >> 16] * vals[term.idx & 0xFFFFu];
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

  for (int idx = tid; idx < gt.basics * W; idx += NT) {
    const int b = idx / W, a = idx - b * W;
    const int atom = a0 + a;
    if (atom < N) ggrad[static_cast<size_t>(b) * N + atom] = grads_s[idx];
  }
}

template <typename Real, int W>
bool launch_gate_tile(int mode, int N, const Real* gbasics, const int* type_s,
                      const GraphTable<Real>& gt, const GateTable<Real>& ga, Real* mult,
                      Real* dmult, const Real* adj, Real* ggrad, hipStream_t s, size_t max_lds)
{
  const size_t lds = gate_graph_lds_bytes(W, gt.slots, static_cast<int>(sizeof(Real)));
  if (lds > max_lds) return false;
  if (lds > 65536) allow_large_lds<&k_gate_graph<Real, W>>();
  const int tiles = (N + W - 1) / W;
  k_gate_graph<Real, W><<<tiles, kGateThreads, lds, s>>>(mode, N, gbasics, type_s, gt, ga, mult,
                                                         dmult, adj, ggrad);
  return true;
}

}  // namespace

size_t gate_graph_lds_bytes(int tile, int slots, int real_bytes)
{
  return 2 * (static_cast<size_t>(slots) + 1) * tile * static_cast<size_t>(real_bytes);
}

template <typename Real>
bool launch_gate_graph(int mode, int tile, int N, const Real* gbasics, const int* type_s,
                       const GraphTable<Real>& gt, const GateTable<Real>& ga, Real* mult,
                       Real* dmult, const Real* adj, Real* ggrad, hipStream_t s, size_t max_lds)
{
  if (N <= 0) return true;
  switch (tile) {
    case 1: return launch_gate_tile<Real, 1>(mode, N, gbasics, type_s, gt, ga, mult, dmult, adj, ggrad, s, max_lds);
    case 2: return launch_gate_tile<Real, 2>(mode, N, gbasics, type_s, gt, ga, mult, dmult, adj, ggrad, s, max_lds);
    case 4: return launch_gate_tile<Real, 4>(mode, N, gbasics, type_s, gt, ga, mult, dmult, adj, ggrad, s, max_lds);
    case 8: return launch_gate_tile<Real, 8>(mode, N, gbasics, type_s, gt, ga, mult, dmult, adj, ggrad, s, max_lds);
    case 16: return launch_gate_tile<Real, 16>(mode, N, gbasics, type_s, gt, ga, mult, dmult, adj, ggrad, s, max_lds);
    default: return false;
  }
}

template bool launch_gate_graph<float>(int, int, int, const float*, const int*,
                                       const GraphTable<float>&, const GateTable<float>&, float*,
                                       float*, const float*, float*, hipStream_t, size_t);
template bool launch_gate_graph<double>(int, int, int, const double*, const int*,
                                        const GraphTable<double>&, const GateTable<double>&,
                                        double*, double*, const double*, double*, hipStream_t,
                                        size_t);

}  // namespace sus2sh
