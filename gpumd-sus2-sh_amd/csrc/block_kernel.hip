// Block graph kernel: the depth <= 2 product graph of one atom per lane, organised by tensors
// (program: block_program.hpp; semantics src/force/sus2_sh.cu:4799-4846).
//
// A workgroup of WPT waves owns a tile of 64 atoms; slot s of lane i sits at LDS word
// s * 64 + i, so a uniform slot offset reads 64 consecutive words.  Per batch the waves first
// compute the batch's layer-1 tensors (FWD: one real Clebsch-Gordan block per tensor, the
// coupling pattern and its coefficients compile-time constants of cg_table.hpp, the operand
// components read once into registers), then, per quadratic-form tensor, its adjoint
// g = K' v (all 2L+1 components of a partner share one coefficient and one LDS base address),
// its share of the site energy 1/2 v . (g + d), and its reverse pass into the basic adjoints
// (W = sum_M g[M] C_M, G_A += W b_B, G_B += W^T b_A), which the waves add with LDS atomics.
// The basic adjoints are kept in double in both precisions: FP64 LDS atomic adds issue at the
// rate of other LDS operations on this GPU, FP32 ones are many times slower.
// The records of an epoch come in runs of one block type: the kernel dispatches once per run,
// and a record costs two to four scalar words plus its partners.
#include <utility>

#include "block_program.hpp"
#include "cg_table.hpp"
#include "kernels.hpp"

namespace sus2sh {

namespace {

constexpr int kCol = 64;

// read-only program words through the constant address space (scalar loads)
using cu32 = const __attribute__((address_space(4))) uint32_t;

template <typename Real>
__device__ __forceinline__ Real read_real(cu32* p);
template <>
__device__ __forceinline__ float read_real<float>(cu32* p)
{
  return __uint_as_float(p[0]);
}
template <>
__device__ __forceinline__ double read_real<double>(cu32* p)
{
  return __hiloint2double(static_cast<int>(p[1]), static_cast<int>(p[0]));
}

__device__ __forceinline__ void lds_add(double* p, double v)
{
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Compile-time flags of the entries of block type T with output component M in [M0, M1): the
// first such entry of each M, and the first of each (m1, m2) pair.
template <int T, int M0, int M1>
struct CgFirst {
  struct F {
    bool m[CgTable<T>::n];
    bool p[CgTable<T>::n];
  };
  static constexpr bool in(int i) { return CgTable<T>::e[i].M >= M0 && CgTable<T>::e[i].M < M1; }
  static constexpr F make()
  {
    F f{};
    for (int i = 0; i < CgTable<T>::n; ++i) {
      f.m[i] = in(i);
      f.p[i] = in(i);
      for (int j = 0; j < i; ++j) {
        if (!in(j)) continue;
        if (CgTable<T>::e[j].M == CgTable<T>::e[i].M) f.m[i] = false;
        if (CgTable<T>::e[j].m1 == CgTable<T>::e[i].m1 && CgTable<T>::e[j].m2 == CgTable<T>::e[i].m2)
          f.p[i] = false;
      }
    }
    return f;
  }
  static constexpr F v = make();
};

// o[M - M0] (+)= c a[m1] b[m2] for entry I of block type T (entries outside [M0, M1) skipped)
template <typename Real, int T, int M0, int M1, int I, int NA, int NB, int NO>
__device__ __forceinline__ void fwd_term(const Real (&a)[NA], const Real (&b)[NB], Real (&o)[NO])
{
  constexpr CgEntry e = CgTable<T>::e[I];
  if constexpr (e.M >= M0 && e.M < M1) {
    constexpr Real c = static_cast<Real>(e.c);
    const Real t = c * (a[e.m1] * b[e.m2]);
    if constexpr (CgFirst<T, M0, M1>::v.m[I])
      o[e.M - M0] = t;
    else
      o[e.M - M0] += t;
  }
}

template <typename Real, int T, int M0, int M1, int NA, int NB, int NO, int... I>
__device__ __forceinline__ void fwd_terms(const Real (&a)[NA], const Real (&b)[NB], Real (&o)[NO],
                                          std::integer_sequence<int, I...>)
{
  (fwd_term<Real, T, M0, M1, I>(a, b, o), ...);
}

// W[m1][m2] (+)= c g[M - M0] for entry I
template <typename Real, int T, int M0, int M1, int I, int NA, int NB, int NO>
__device__ __forceinline__ void w_term(const Real (&g)[NO], Real (&W)[NA][NB])
{
  constexpr CgEntry e = CgTable<T>::e[I];
  if constexpr (e.M >= M0 && e.M < M1) {
    constexpr Real c = static_cast<Real>(e.c);
    if constexpr (CgFirst<T, M0, M1>::v.p[I])
      W[e.m1][e.m2] = c * g[e.M - M0];
    else
      W[e.m1][e.m2] += c * g[e.M - M0];
  }
}

// for the first entry of each (m1, m2): ga[m1] += W b[m2], gb[m2] += W a[m1]
template <typename Real, int T, int M0, int M1, int I, int NA, int NB>
__device__ __forceinline__ void wab_term(const Real (&W)[NA][NB], const Real (&a)[NA],
                                         const Real (&b)[NB], Real (&ga)[NA], Real (&gb)[NB])
{
  if constexpr (CgFirst<T, M0, M1>::v.p[I]) {
    constexpr CgEntry e = CgTable<T>::e[I];
    ga[e.m1] += W[e.m1][e.m2] * b[e.m2];
    gb[e.m2] += W[e.m1][e.m2] * a[e.m1];
  }
}

template <typename Real, int T, int M0, int M1, int NA, int NB, int NO, int... I>
__device__ __forceinline__ void bwd_terms(const Real (&g)[NO], const Real (&a)[NA], const Real (&b)[NB],
                                          Real (&ga)[NA], Real (&gb)[NB], std::integer_sequence<int, I...>)
{
  Real W[NA][NB];
  (w_term<Real, T, M0, M1, I>(g, W), ...);
  (wab_term<Real, T, M0, M1, I>(W, a, b, ga, gb), ...);
}

// one layer-1 tensor (components [M0, M1)): v = (b_A (x) b_B)_L
template <typename Real, int T, int M0, int M1>
__device__ __forceinline__ void fwd_block(Real* X, uint32_t vo, uint32_t ao, uint32_t bo)
{
  constexpr int NA = 2 * kCgL1[T] + 1, NB = 2 * kCgL2[T] + 1, NO = M1 - M0;
  const Real* xa = X + ao;
  const Real* xb = X + bo;
  Real a[NA], b[NB], o[NO];
#pragma unroll
  for (int m = 0; m < NA; ++m) a[m] = xa[m * kCol];
#pragma unroll
  for (int m = 0; m < NB; ++m) b[m] = xb[m * kCol];
  fwd_terms<Real, T, M0, M1>(a, b, o, std::make_integer_sequence<int, CgTable<T>::n>{});
  Real* xo = X + vo;
#pragma unroll
  for (int M = 0; M < NO; ++M) xo[M * kCol] = o[M];
}

// g = d [M = 0 of a degree-0 tensor] + sum over partners K' v_partner (NL components);
// advances p past the partner list
template <typename Real, int NL, bool D>
__device__ __forceinline__ void adjoint(cu32*& p, uint32_t np, const Real* X, Real d, Real (&g)[NL])
{
  constexpr int PW = 1 + static_cast<int>(sizeof(Real)) / 4;
  constexpr int U = NL <= 5 ? 4 : 2;  // partners per scalar batch
#pragma unroll
  for (int M = 0; M < NL; ++M) g[M] = Real(0);
  if (D) g[0] = d;
  uint32_t j = 0;
  for (; j + U <= np; j += U) {
    uint32_t off[U];
    Real c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      off[u] = p[u * PW];
      c[u] = read_real<Real>(p + u * PW + 1);
    }
    p += U * PW;
    Real v[U][NL];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int M = 0; M < NL; ++M) v[u][M] = X[off[u] + M * kCol];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int M = 0; M < NL; ++M) g[M] += c[u] * v[u][M];
  }
  for (; j < np; ++j) {
    const uint32_t off = p[0];
    const Real c = read_real<Real>(p + 1);
    p += PW;
    Real v[NL];
#pragma unroll
    for (int M = 0; M < NL; ++M) v[M] = X[off + M * kCol];
#pragma unroll
    for (int M = 0; M < NL; ++M) g[M] += c * v[M];
  }
}

// 1/2 v . (g + d [M = 0]) into the lane's energy
template <typename Real, int NL, bool D>
__device__ __forceinline__ void energy(const Real* xv, Real d, const Real (&g)[NL], Real& e)
{
  Real v[NL];
#pragma unroll
  for (int M = 0; M < NL; ++M) v[M] = xv[M * kCol];
  Real s = D ? d * v[0] : Real(0);
#pragma unroll
  for (int M = 0; M < NL; ++M) s += v[M] * g[M];
  e += Real(0.5) * s;
}

// quadratic-form record of a layer-1 tensor of block type T (components [M0, M1))
template <typename Real, int T, int M0, int M1>
__device__ __forceinline__ void tensor_pair(cu32*& p, Real* X, double* G, Real& e)
{
  constexpr int NA = 2 * kCgL1[T] + 1, NB = 2 * kCgL2[T] + 1, NL = M1 - M0;
  constexpr bool D = kCgLL[T] == 0;
  const uint32_t w0 = p[0], w1 = p[1];
  const Real d = read_real<Real>(p + 2);
  p += 2 + static_cast<int>(sizeof(Real)) / 4;
  const uint32_t vo = w0 & 0xFFFFu, np = w0 >> 16, ao = w1 & 0xFFFFu, bo = w1 >> 16;
  Real g[NL];
  adjoint<Real, NL, D>(p, np, X, d, g);
  energy<Real, NL, D>(X + vo, d, g, e);
  const Real* xa = X + ao;
  const Real* xb = X + bo;
  Real a[NA], b[NB], ga[NA], gb[NB];
#pragma unroll
  for (int m = 0; m < NA; ++m) {
    a[m] = xa[m * kCol];
    ga[m] = Real(0);
  }
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    b[m] = xb[m * kCol];
    gb[m] = Real(0);
  }
  bwd_terms<Real, T, M0, M1>(g, a, b, ga, gb, std::make_integer_sequence<int, CgTable<T>::n>{});
  double* Ga = G + ao;
  double* Gb = G + bo;
#pragma unroll
  for (int m = 0; m < NA; ++m) lds_add(Ga + m * kCol, static_cast<double>(ga[m]));
#pragma unroll
  for (int m = 0; m < NB; ++m) lds_add(Gb + m * kCol, static_cast<double>(gb[m]));
}

// quadratic-form record of a basic tensor of degree L (components [M0, M1), the record's
// offset is that of component M0): its adjoint goes to G directly
template <typename Real, int L, int M0, int M1>
__device__ __forceinline__ void tensor_basic(cu32*& p, Real* X, double* G, Real& e)
{
  constexpr int NL = M1 - M0;
  const uint32_t w0 = p[0];
  const Real d = read_real<Real>(p + 2);
  p += 2 + static_cast<int>(sizeof(Real)) / 4;
  const uint32_t vo = w0 & 0xFFFFu, np = w0 >> 16;
  Real g[NL];
  adjoint<Real, NL, L == 0>(p, np, X, d, g);
  energy<Real, NL, L == 0>(X + vo, d, g, e);
  double* Gv = G + vo;
#pragma unroll
  for (int M = 0; M < NL; ++M) lds_add(Gv + M * kCol, static_cast<double>(g[M]));
}

// a run of `cnt` FWD records of block type T, part `part` (0 whole, 1 lower half, 2 upper;
// halves are compiled for FP32 only: an FP64 tile does not reach two workgroups per CU)
template <typename Real, int T>
__device__ __forceinline__ void fwd_run(uint32_t part, int cnt, cu32*& p, Real* X)
{
  constexpr int L = kCgLL[T];
  if constexpr (L >= 2 && sizeof(Real) == 4) {
    if (part == 1) {
      for (int r = 0; r < cnt; ++r, p += 2) fwd_block<Real, T, 0, L + 1>(X, p[0] & 0xFFFFu, p[0] >> 16, p[1]);
      return;
    }
    if (part == 2) {
      for (int r = 0; r < cnt; ++r, p += 2)
        fwd_block<Real, T, L + 1, 2 * L + 1>(X, p[0] & 0xFFFFu, p[0] >> 16, p[1]);
      return;
    }
  }
  for (int r = 0; r < cnt; ++r, p += 2) fwd_block<Real, T, 0, 2 * L + 1>(X, p[0] & 0xFFFFu, p[0] >> 16, p[1]);
}

template <typename Real, int T>
__device__ __forceinline__ void pair_run(uint32_t part, int cnt, cu32*& p, Real* X, double* G, Real& e)
{
  constexpr int L = kCgLL[T];
  if constexpr (L >= 2 && sizeof(Real) == 4) {
    if (part == 1) {
      for (int r = 0; r < cnt; ++r) tensor_pair<Real, T, 0, L + 1>(p, X, G, e);
      return;
    }
    if (part == 2) {
      for (int r = 0; r < cnt; ++r) tensor_pair<Real, T, L + 1, 2 * L + 1>(p, X, G, e);
      return;
    }
  }
  for (int r = 0; r < cnt; ++r) tensor_pair<Real, T, 0, 2 * L + 1>(p, X, G, e);
}

template <typename Real, int L>
__device__ __forceinline__ void basic_run(uint32_t part, int cnt, cu32*& p, Real* X, double* G, Real& e)
{
  if constexpr (L >= 2 && sizeof(Real) == 4) {
    if (part == 1) {
      for (int r = 0; r < cnt; ++r) tensor_basic<Real, L, 0, L + 1>(p, X, G, e);
      return;
    }
    if (part == 2) {
      for (int r = 0; r < cnt; ++r) tensor_basic<Real, L, L + 1, 2 * L + 1>(p, X, G, e);
      return;
    }
  }
  for (int r = 0; r < cnt; ++r) tensor_basic<Real, L, 0, 2 * L + 1>(p, X, G, e);
}

template <typename Real>
__device__ __forceinline__ void run_fwd(cu32* p, cu32* end, Real* X)
{
  while (p < end) {
    const uint32_t h = p[0];
    const int cnt = static_cast<int>(h >> 8);
    const uint32_t kind = h & 255u;
    ++p;
    switch (kind & 63u) {
#define SUS2SH_FWD_CASE(T, l1, l2, L) \
  case T:                             \
    fwd_run<Real, T>(kind >> 6, cnt, p, X); \
    break;
      SUS2SH_CG_TYPES(SUS2SH_FWD_CASE)
#undef SUS2SH_FWD_CASE
      default:
        break;
    }
  }
}

template <typename Real>
__device__ __forceinline__ void run_tensors(cu32* p, cu32* end, Real* X, double* G, Real& e)
{
  while (p < end) {
    const uint32_t h = p[0];
    const int cnt = static_cast<int>(h >> 8);
    const uint32_t kind = h & 255u;
    ++p;
    if (kind >= static_cast<uint32_t>(kBlockBasicKind)) {
      const uint32_t k = kind - kBlockBasicKind, part = k >> 2;
      switch (k & 3u) {
        case 0: basic_run<Real, 0>(part, cnt, p, X, G, e); break;
        case 1: basic_run<Real, 1>(part, cnt, p, X, G, e); break;
        case 2: basic_run<Real, 2>(part, cnt, p, X, G, e); break;
        default: basic_run<Real, 3>(part, cnt, p, X, G, e); break;
      }
      continue;
    }
    switch (kind & 63u) {
#define SUS2SH_TEN_CASE(T, l1, l2, L) \
  case T:                             \
    pair_run<Real, T>(kind >> 6, cnt, p, X, G, e); \
    break;
      SUS2SH_CG_TYPES(SUS2SH_TEN_CASE)
#undef SUS2SH_TEN_CASE
      default:
        break;
    }
  }
}

template <typename Real, int WPT>
__global__ __launch_bounds__(kCol * WPT) void k_block_graph(int N, const Real* __restrict__ basics,
                                                           const int* __restrict__ type_s,
                                                           BlockTable bt, double* __restrict__ e_out,
                                                           Real* __restrict__ grad_out)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const unsigned lane = threadIdx.x & (kCol - 1);
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kCol));
  const int B = bt.basics;
  // [B double slots: basic adjoints][Real slots: basics, tensor values]
  double* G = reinterpret_cast<double*>(smem_raw) + lane;
  Real* X = reinterpret_cast<Real*>(smem_raw + static_cast<size_t>(B) * kCol * sizeof(double)) + lane;
  const unsigned atom = blockIdx.x * static_cast<unsigned>(kCol) + lane;
  for (int q = wave; q < B; q += WPT) {
    X[q * kCol] = atom < static_cast<unsigned>(N) ? basics[static_cast<size_t>(q) * N + atom] : Real(0);
    G[q * kCol] = 0.0;
  }
  cu32* W = (cu32*)bt.words;
  Real e = Real(0);
  for (int b = 0; b < bt.batches; ++b) {
    cu32* sg = (cu32*)bt.seg + 4 * (b * WPT + wave);
    const uint32_t f0 = sg[0], f1 = sg[1], t0 = sg[2], t1 = sg[3];
    __syncthreads();  // basics loaded, or the previous batch's values no longer read
    run_fwd<Real>(W + f0, W + f1, X);
    __syncthreads();
    run_tensors<Real>(W + t0, W + t1, X, G, e);
  }
  __syncthreads();

  // per-wave energies -> the (now unused) value region, summed by wave 0
  if (WPT > 1) {
    X[(bt.v_base + wave) * kCol] = e;
    __syncthreads();
  }
  if (atom >= static_cast<unsigned>(N)) return;
  const int t = type_s[atom];
  if (wave == 0) {
    Real sum = e;
#pragma unroll
    for (int w = 1; w < WPT; ++w) sum += X[(bt.v_base + w) * kCol];
    const double sg = bt.sigma[t];
    e_out[atom] = bt.shift[t] + sg + sg * static_cast<double>(sum);
  }
  const Real sgr = static_cast<Real>(bt.sigma[t]);
  for (int q = wave; q < B; q += WPT)
    grad_out[static_cast<size_t>(q) * N + atom] = sgr * static_cast<Real>(G[q * kCol]);
}

template <typename Real, int WPT>
void launch_block_shape(int N, const Real* basics, const int* type_s, const BlockTable& bt,
                        double* e_s, Real* grads, hipStream_t s, size_t lds)
{
  if (lds > 65536) allow_large_lds<&k_block_graph<Real, WPT>>();
  const unsigned blocks = (static_cast<unsigned>(N) + kCol - 1) / kCol;
  k_block_graph<Real, WPT><<<blocks, kCol * WPT, lds, s>>>(N, basics, type_s, bt, e_s, grads);
}

}  // namespace

size_t block_graph_lds(const BlockTable& bt, int real_bytes)
{
  const int slots = bt.slots > bt.v_base + bt.waves ? bt.slots : bt.v_base + bt.waves;
  return static_cast<size_t>(bt.basics) * kCol * sizeof(double) +
         static_cast<size_t>(slots) * kCol * static_cast<size_t>(real_bytes);
}

template <typename Real>
bool launch_block_graph(int N, const Real* basics, const int* type_s, const BlockTable& bt,
                        double* e_s, Real* grads, hipStream_t s, size_t max_lds)
{
  if (N <= 0) return true;
  const size_t lds = block_graph_lds(bt, static_cast<int>(sizeof(Real)));
  if (lds > max_lds) return false;
  switch (bt.waves) {
    case 4: launch_block_shape<Real, 4>(N, basics, type_s, bt, e_s, grads, s, lds); return true;
    case 8: launch_block_shape<Real, 8>(N, basics, type_s, bt, e_s, grads, s, lds); return true;
    case 16: launch_block_shape<Real, 16>(N, basics, type_s, bt, e_s, grads, s, lds); return true;
    default: return false;
  }
}

template bool launch_block_graph<float>(int, const float*, const int*, const BlockTable&, double*,
                                        float*, hipStream_t, size_t);
template bool launch_block_graph<double>(int, const double*, const int*, const BlockTable&, double*,
                                         double*, hipStream_t, size_t);

}  // namespace sus2sh
