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
//
// With PG (parameter gradient) the workgroups are persistent: each walks tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... and keeps dE/dw_s, dE/dsigma, dE/dshift in LDS
// (double) until a single flush at the end; with pg.fwd_only it runs the forward pass and
// these sums only, beside a faster kernel that writes the energies and adjoints.
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

template <typename Real, int W>
__device__ __forceinline__ Real dot_w(const vec_t<Real, W>& a, const vec_t<Real, W>& b)
{
  Real s = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) s += a[i] * b[i];
  return s;
}

template <typename Real, int W, int NT, bool PG, bool VM>
__global__ __launch_bounds__(NT) void k_graph(int N, const Real* __restrict__ basics,
                                              const int* __restrict__ type_s, GraphTable<Real> gt,
                                              double* __restrict__ e_out,
                                              Real* __restrict__ grad_out, ParamGradOut pg)
{
  using V = vec_t<Real, W>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* vals = reinterpret_cast<V*>(smem_raw);  // [slots + 1]: the last slot holds 0
  V* grads = vals + gt.slots + 1;             // [slots + 1]: the last slot holds 1
  V* red = grads + gt.slots + 1;              // [NT / 64] wave partial energies
  V* sig = red + NT / 64;     // sigma of the tile's atoms (0 for padding lanes), PG only
  // PG: [Q] dE/dw, [S] dshift, [S] dsigma, 16-byte aligned after the tile state
  const size_t pw_off = (static_cast<size_t>(reinterpret_cast<unsigned char*>(sig + 1) - smem_raw) + 15) & ~size_t(15);
  double* pw = reinterpret_cast<double*>(smem_raw + pw_off);
  Real* vals_s = reinterpret_cast<Real*>(vals);
  Real* grads_s = reinterpret_cast<Real*>(grads);
  const int tid = threadIdx.x;
  const int tiles = (N + W - 1) / W;
  const uint32_t zero_slot = static_cast<uint32_t>(gt.slots);
  const uint32_t ones_slot = static_cast<uint32_t>(gt.slots);

  if (PG) {
    for (int k = tid; k < pg.Q + 2 * pg.S; k += NT) pw[k] = 0.0;
  }
  const bool reverse = !(PG && pg.fwd_only);

  for (int tile = blockIdx.x; tile < tiles; tile += (PG ? gridDim.x : tiles)) {
    const int a0 = tile * W;
    // basic moments of the W atoms -> slots [0, B)
    for (int idx = tid; idx < gt.basics * W; idx += NT) {
      const int b = idx / W, a = idx - b * W;
      const int atom = a0 + a;
      vals_s[idx] = atom < N ? basics[static_cast<size_t>(b) * N + atom] : Real(0);
    }
    for (int s = tid; s < gt.slots; s += NT) grads[s] = splat<Real, W>(gt.seed[s]);
    if (tid == 0) {
      grads[gt.slots] = splat<Real, W>(Real(1));
      vals[gt.slots] = splat<Real, W>(Real(0));
    }
    if (PG && tid < W) {
      const int atom = a0 + tid;
      reinterpret_cast<Real*>(sig)[tid] =
        atom < N ? static_cast<Real>(gt.sigma[type_s[atom]]) : Real(0);
    }
    __syncthreads();

    // forward, layer by layer
    V e = splat<Real, W>(Real(0));
    if (VM) {
      for (int layer = 0; layer < gt.layers; ++layer) {
        const int L = gt.fwd_slen[layer];
        const Term<Real>* ins = gt.fwd_ins + gt.fwd_sbegin[layer] + tid;
        V acc = splat<Real, W>(Real(0));
        Term<Real> buf[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) buf[j] = L > j ? ins[j * NT] : Term<Real>{};
        for (int k = 0; k < L; k += 4) {
          Term<Real> cur[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            cur[j] = buf[j];
            if (k + 4 + j < L) buf[j] = ins[(k + 4 + j) * NT];
          }
          // decode, then issue all LDS loads of the group before any FMA / store: END and
          // NOP instructions read the zero slot so the FMA path has no branch
          V x[4], y[4];
          uint32_t op[4], tg[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const uint32_t a = cur[j].idx & 0xFFFFu, b = cur[j].idx >> 16;
            op[j] = a;
            tg[j] = b;
            const uint32_t la = a < 0xFFFEu ? a : zero_slot;
            const uint32_t lb = a < 0xFFFEu ? b : zero_slot;
            x[j] = vals[la];
            y[j] = vals[lb];
          }
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            acc += (op[j] < 0xFFFEu ? cur[j].c : Real(0)) * x[j] * y[j];
            if (op[j] == 0xFFFFu) {
              if (tg[j] != 0xFFFFu)
                vals[tg[j]] = acc;
              else
                e += cur[j].c * acc;
              acc = splat<Real, W>(Real(0));
            }
          }
        }
        __syncthreads();
      }
    }
    for (int layer = 0; !VM && layer < gt.layers; ++layer) {
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
          if (PG) {
            const double m = static_cast<double>(dot_w<Real, W>(sig[0], acc));
            for (int k = pg.row_sbegin[r]; k < pg.row_sbegin[r + 1]; ++k) pw[pg.row_slist[k]] += m;
          }
        }
      }
      __syncthreads();
    }
    for (int k = tid; k < gt.n_e; k += NT) e += gt.e_w[k] * vals[gt.e_slot[k]];
    if (PG) {
      for (int k = tid; k < pg.n_stored; k += NT)
        pw[pg.stored_index[k]] += static_cast<double>(dot_w<Real, W>(sig[0], vals[pg.stored_slot[k]]));
    }

    // block reduction of the per-atom site energies
#pragma unroll
    for (int a = 0; a < W; ++a) {
      Real v = e[a];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      e[a] = v;
    }
    if ((tid & 63) == 0) red[tid >> 6] = e;
    __syncthreads();
    if (tid < W) {
      const int atom = a0 + tid;
      if (atom < N) {
        double sum = 0.0;
        for (int w = 0; w < NT / 64; ++w)
          sum += static_cast<double>(reinterpret_cast<Real*>(red)[w * W + tid]);
        const int t = type_s[atom];
        const double sg = gt.sigma[t];
        if (reverse) e_out[atom] = gt.shift[t] + sg + sg * sum;
        if (PG) {
          atomicAdd(pw + pg.Q + t, 1.0);
          atomicAdd(pw + pg.Q + pg.S + t, 1.0 + sum);
        }
      }
    }

    // reverse, layer by layer (descending)
    if (VM && reverse) {
      for (int layer = gt.layers - 1; layer >= 0; --layer) {
        const int L = gt.bwd_slen[layer];
        const Term<Real>* ins = gt.bwd_ins + gt.bwd_sbegin[layer] + tid;
        V acc = splat<Real, W>(Real(0));
        Term<Real> buf[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) buf[j] = L > j ? ins[j * NT] : Term<Real>{};
        for (int k = 0; k < L; k += 4) {
          Term<Real> cur[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            cur[j] = buf[j];
            if (k + 4 + j < L) buf[j] = ins[(k + 4 + j) * NT];
          }
          V x[4], y[4];
          uint32_t op[4], tg[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const uint32_t a = cur[j].idx & 0xFFFFu, b = cur[j].idx >> 16;
            op[j] = a;
            tg[j] = b;
            x[j] = grads[a < 0xFFFEu ? b : ones_slot];
            y[j] = vals[a < 0xFFFEu ? a : zero_slot];
          }
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            acc += (op[j] < 0xFFFEu ? cur[j].c : Real(0)) * x[j] * y[j];
            if (op[j] == 0xFFFFu) {
              if (cur[j].c != Real(0)) {
#pragma unroll
                for (int w = 0; w < W; ++w) atomicAdd(grads_s + tg[j] * W + w, acc[w]);
              } else {
                grads[tg[j]] += acc;
              }
              acc = splat<Real, W>(Real(0));
            }
          }
        }
        __syncthreads();
      }
    }
    for (int layer = gt.layers - 1; !VM && reverse && layer >= 0; --layer) {
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
    for (int idx = tid; reverse && idx < gt.basics * W; idx += NT) {
      const int b = idx / W, a = idx - b * W;
      const int atom = a0 + a;
      if (atom < N) {
        const Real sg = static_cast<Real>(gt.sigma[type_s[atom]]);
        grad_out[static_cast<size_t>(b) * N + atom] = sg * grads_s[idx];
      }
    }
    if (PG) __syncthreads();  // LDS tile state is rewritten by the next tile
  }

  if (PG) {
    __syncthreads();
    for (int k = tid; k < pg.Q; k += NT) {
      const double v = pw[k];
      if (v != 0.0) unsafeAtomicAdd(pg.d_moment + k, v);
    }
    for (int k = tid; k < pg.S; k += NT) {
      unsafeAtomicAdd(pg.d_shift + k, pw[pg.Q + k]);
      unsafeAtomicAdd(pg.d_species + k, pw[pg.Q + pg.S + k]);
    }
  }
}

template <typename Real, int W, int NT, bool PG, bool VM>
void launch_variant(int N, const Real* basics, const int* type_s, const GraphTable<Real>& gt,
                    double* e_s, Real* grads, const ParamGradOut& pg, hipStream_t s, size_t lds,
                    size_t max_lds, int blocks)
{
  auto kern = &k_graph<Real, W, NT, PG, VM>;
  if (lds > 65536) allow_large_lds<&k_graph<Real, W, NT, PG, VM>>();
  (void)max_lds;
  kern<<<blocks, NT, lds, s>>>(N, basics, type_s, gt, e_s, grads, pg);
}

template <typename Real, int W, int NT>
bool launch_tile(int mode, int N, const Real* basics, const int* type_s, const GraphTable<Real>& gt,
                 double* e_s, Real* grads, const ParamGradOut* pg, hipStream_t s, size_t max_lds,
                 int persistent_blocks)
{
  size_t lds = graph_lds_bytes<Real>(W, gt.slots);
  if (pg) lds = ((lds + 15) & ~size_t(15)) + sizeof(double) * (pg->Q + 2 * pg->S);
  if (lds > max_lds) return false;
  const int tiles = (N + W - 1) / W;
  if (pg) {
    const int blocks = std::max(1, std::min(tiles, persistent_blocks));
    launch_variant<Real, W, NT, true, false>(N, basics, type_s, gt, e_s, grads, *pg, s, lds,
                                             max_lds, blocks);
  } else if (mode == 0) {
    launch_variant<Real, W, NT, false, false>(N, basics, type_s, gt, e_s, grads, ParamGradOut{}, s,
                                              lds, max_lds, tiles);
  } else {
    launch_variant<Real, W, NT, false, true>(N, basics, type_s, gt, e_s, grads, ParamGradOut{}, s,
                                             lds, max_lds, tiles);
  }
  return true;
}

}  // namespace

template <typename Real>
size_t graph_lds_bytes(int tile, int slots)
{
  return (2 * static_cast<size_t>(slots) + 2 + 4 + 1) * tile * sizeof(Real);
}

template <typename Real>
bool launch_graph(int tile, int mode, int N, const Real* basics, const int* type_s,
                  const GraphTable<Real>& gt, double* e_s, Real* grads, const ParamGradOut* pg,
                  hipStream_t s, size_t max_lds, int persistent_blocks)
{
  if (N <= 0) return true;
  switch (tile) {
    case 1: return launch_tile<Real, 1, 256>(mode, N, basics, type_s, gt, e_s, grads, pg, s, max_lds, persistent_blocks);
    case 2: return launch_tile<Real, 2, 256>(mode, N, basics, type_s, gt, e_s, grads, pg, s, max_lds, persistent_blocks);
    case 4: return launch_tile<Real, 4, 256>(mode, N, basics, type_s, gt, e_s, grads, pg, s, max_lds, persistent_blocks);
    case 8: return launch_tile<Real, 8, 256>(mode, N, basics, type_s, gt, e_s, grads, pg, s, max_lds, persistent_blocks);
    case 16: return launch_tile<Real, 16, 256>(mode, N, basics, type_s, gt, e_s, grads, pg, s, max_lds, persistent_blocks);
    default: return false;
  }
}

template size_t graph_lds_bytes<float>(int, int);
template size_t graph_lds_bytes<double>(int, int);
template bool launch_graph<float>(int, int, int, const float*, const int*,
                                  const GraphTable<float>&, double*, float*, const ParamGradOut*,
                                  hipStream_t, size_t, int);
template bool launch_graph<double>(int, int, int, const double*, const int*,
                                   const GraphTable<double>&, double*, double*,
                                   const ParamGradOut*, hipStream_t, size_t, int);

}  // namespace sus2sh
