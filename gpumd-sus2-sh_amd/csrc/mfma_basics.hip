// Basic moments on the matrix cores (FP32 engine, uniform radial scaling, l_max <= 3).
//
// Every pair (t_i, t_j) scales one shared coefficient matrix c[mu][n] by a pair factor
// sigma_p = scaling * type_coeff[t_i] * type_coeff[t_j] (src/force/sus2_sh.cu:2484-2517), so
// with one (a, s) per pair (xi_p) the radial function is
//   R_mu(r) = sigma_p (r - rc)^2 sum_n c[mu][n] T_n(xi_p(r))   (src/force/sus2_sh.cu:3171-3220)
// and the basic moments of atom i factor into two small dense products:
//   C[n][lm]   = sum_j t_j[n] y_j[lm],    t_j[n] = sigma_p (r - rc)^2 T_n(xi_p(r_ij)),
//                                         y_j[lm] = N_l S_lm(r_hat_ij)    (T^T Y over edges)
//   B[mu][lm]  = sum_n c[mu][n] C[n][lm], basic (k, l, m) = B[k (l_max+1) + l][(l, m)].
// A wavefront works on 16 consecutive atoms (cell order): in each round its 64 lanes evaluate
// edge slot k0 + l / 16 of atom l % 16 (t_j and y_j; edge records of 16 consecutive atoms are
// 256 contiguous bytes, as in the vector kernel), write them to the wave's LDS tiles (per atom:
// n rows and lm rows, one column per edge slot), and then one MFMA per atom adds the round's
// edges to its 16x16 C tile: 16x16x4 FP32 MFMAs (rounds of 4 edges) or 16x16x16 BF16 MFMAs
// (rounds of 16 edges, BF16 operands, FP32 accumulation).  B = c C is four more FP32 MFMAs per
// atom whose B operand is C itself: MFMA output register r of lane group g holds row 4 g + r,
// so step r of the second product sums the rows {4 g + r} (a permutation of the n sum).  The
// moments go through LDS to coalesced row stores.
// Semantics: src/force/sus2_sh.cu:3486-3550, 3650-3724 (same moments as k_basics).
#include "kernels.hpp"

namespace sus2sh {

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef short s4v __attribute__((ext_vector_type(4)));

constexpr int kMfmaWaves = 4;  // waves per workgroup
constexpr int kGroup = 16;     // atoms per wave
// FP32 tile of one atom: 32 rows (16 n, 16 lm) x 4 edge slots, padded to 132 words so that the
// 16 atoms of a round's stores start in different banks
constexpr int kF32Atom = 132;
// BF16 tile of one atom: 32 rows x 16 edge slots (16 halves = 8 words per row), 260 words
constexpr int kBf16Atom = 260;
constexpr int kMaxBasics = 64;  // staging rows for the outputs

__device__ __forceinline__ short to_bf16(float x)
{
  // round to nearest even (finite inputs)
  const unsigned u = __float_as_uint(x);
  return static_cast<short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

template <bool BF16>
__host__ __device__ constexpr size_t wave_lds_bytes()
{
  return (BF16 ? kGroup * kBf16Atom : kGroup * kF32Atom) * 4 + kMaxBasics * kGroup * 4;
}

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// t_j[n] (n < NB) and y_j[lm] of one edge
template <int L, int NB>
__device__ __forceinline__ void edge_rows(const Edge<float>& ed, int ti, int S,
                                          const float* __restrict__ s_pair, int kind, float rcut,
                                          float (&t)[NB], float (&y)[(L + 1) * (L + 1)])
{
  const float r2 = ed.dx * ed.dx + ed.dy * ed.dy + ed.dz * ed.dz;
  const float inv = dev_rsqrt<float>(r2);
  const float r = r2 * inv;
  const int pair = ti * S + (ed.j >> kEdgeTypeShift);
  float xi, dxi;
  radial_map<float>(kind, s_pair[3 * pair], s_pair[3 * pair + 1], r, xi, dxi);
  const float w = s_pair[3 * pair + 2] * (r - rcut) * (r - rcut);
  float tm = 1.0f, tn = xi;
  t[0] = w;
  if (NB > 1) t[1] = w * xi;
#pragma unroll
  for (int n = 2; n < NB; ++n) {
    const float tp = 2.0f * xi * tn - tm;
    tm = tn;
    tn = tp;
    t[n] = w * tp;
  }
  float Sv[(L + 1) * (L + 1)];
  solid_values<float, L>(ed.dx * inv, ed.dy * inv, ed.dz * inv, Sv);
#pragma unroll
  for (int l = 0; l <= L; ++l) {
    const float nl = static_cast<float>(norm_of_l(l));
#pragma unroll
    for (int m = -l; m <= l; ++m) y[sh_index(l, m)] = nl * Sv[sh_index(l, m)];
  }
}

template <int L, int NB, bool BF16>
__global__ __launch_bounds__(64 * kMfmaWaves) void k_basics_mfma(EdgeLaunch<float> a,
                                                                 float* __restrict__ out)
{
  constexpr int NC = (L + 1) * (L + 1);
  static_assert(NC <= 16 && NB <= 16, "one 16x16 tile per atom");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int S = a.rt.S, pairs = S * S, Mr = a.K * (L + 1);
  // block table: per pair (a, s, sigma)
  float* s_pair = reinterpret_cast<float*>(smem_raw);
  for (int i = threadIdx.x; i < pairs; i += blockDim.x) {
    s_pair[3 * i] = a.rt.scal[static_cast<size_t>(i) * Mr * 2];
    s_pair[3 * i + 1] = a.rt.scal[static_cast<size_t>(i) * Mr * 2 + 1];
    s_pair[3 * i + 2] = a.pair_sigma[i];
  }
  const int wave = static_cast<int>(threadIdx.x) >> 6, lane = static_cast<int>(threadIdx.x) & 63;
  float* wl = reinterpret_cast<float*>(smem_raw + ((3 * pairs * sizeof(float) + 15) & ~size_t(15)) +
                                       wave * wave_lds_bytes<BF16>());
  float* stage = wl + (BF16 ? kGroup * kBf16Atom : kGroup * kF32Atom);  // [kMaxBasics][16]
  short* wb = reinterpret_cast<short*>(wl);
  const int g = lane >> 4, c16 = lane & 15;
  // rows n >= NB and lm >= NC stay zero
  if (BF16) {
    for (int i = lane; i < kGroup * kBf16Atom; i += 64) wl[i] = 0.0f;
  } else {
    for (int i = lane; i < kGroup * kF32Atom; i += 64) wl[i] = 0.0f;
  }
  __syncthreads();

  // A operands of B = c C: lane l, step r holds c[mu = l % 16][n = 4 (l / 16) + r]
  float cA[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int n = 4 * g + r;
    cA[r] = (c16 < Mr && n < a.rt.rb) ? a.mfma_rc[c16 * a.rt.rb + n] : 0.0f;
  }
  // basic index of lane l, register r: mu = 4 g + r, (l, m) = lm = l % 16 (-1: no basic)
  int obase[4];
  {
    int lm_l = 0;
    while ((lm_l + 1) * (lm_l + 1) <= c16) ++lm_l;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mu = 4 * g + r;
      obase[r] = -1;
      if (mu < Mr && c16 < NC && mu % (L + 1) == lm_l) {
        const int k = mu / (L + 1), m = c16 - lm_l * lm_l - lm_l;
        int off = 0;  // full (k, l, m) grid index: l = L..0, k = K-1..0, m = -l..l
        for (int lp = L; lp > lm_l; --lp) off += a.K * (2 * lp + 1);
        obase[r] = a.fmap[off + (a.K - 1 - k) * (2 * lm_l + 1) + (m + lm_l)];
      }
    }
  }

  const int N = a.N;
  const float rcut = a.rt.rc;
  const int kind = a.rt.kind;
  const int p0 = (xcd_block() * kMfmaWaves + wave) * kGroup;
  // lane l evaluates edge slots of atom p0 + l % 16
  const int p = p0 + c16;
  const bool valid = p < N;
  const int cnt = valid ? a.count[p] : 0;
  const int ti = valid ? a.type_s[p] : 0;
  int maxcnt = cnt;
  for (int off = 1; off < 64; off <<= 1) maxcnt = max(maxcnt, __shfl_xor(maxcnt, off, 64));

  f4v acc[kGroup];
#pragma unroll
  for (int q = 0; q < kGroup; ++q) acc[q] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
  // BF16 rounds: lane (g, atom) evaluates the 4 edges k0 + 4 g + v (v = 0..3), so that each row
  // of its 4 columns is one 8-byte store; FP32 rounds: edge k0 + g.  The edge record of the
  // next round is loaded while the current one is evaluated.
  constexpr int SLOTS = BF16 ? 16 : 4;   // edge slots per round
  constexpr int PER = BF16 ? 4 : 1;      // edges per lane per round
  Edge<float> nxt[PER];
#pragma unroll
  for (int v = 0; v < PER; ++v) {
    const int e = PER * g + v;
    nxt[v] = e < cnt ? a.nl[static_cast<size_t>(e) * N + p] : Edge<float>{};
  }
  for (int k0 = 0; k0 < maxcnt; k0 += SLOTS) {
    Edge<float> cur[PER];
#pragma unroll
    for (int v = 0; v < PER; ++v) {
      cur[v] = nxt[v];
      const int e = k0 + SLOTS + PER * g + v;
      if (e < cnt) nxt[v] = a.nl[static_cast<size_t>(e) * N + p];
    }
    if (BF16) {
      unsigned lo[NB + NC], hi[NB + NC];  // packed BF16 pairs: columns (0,1) and (2,3)
#pragma unroll
      for (int v = 0; v < PER; ++v) {
        float t[NB], y[NC];
#pragma unroll
        for (int n = 0; n < NB; ++n) t[n] = 0.0f;
#pragma unroll
        for (int i = 0; i < NC; ++i) y[i] = 0.0f;
        if (k0 + PER * g + v < cnt) edge_rows<L, NB>(cur[v], ti, S, s_pair, kind, rcut, t, y);
#pragma unroll
        for (int i = 0; i < NB + NC; ++i) {
          const unsigned h = static_cast<unsigned short>(to_bf16(i < NB ? t[i] : y[i - NB]));
          if (v == 0) lo[i] = h;
          else if (v == 1) lo[i] |= h << 16;
          else if (v == 2) hi[i] = h;
          else hi[i] |= h << 16;
        }
      }
      unsigned* at = reinterpret_cast<unsigned*>(wb + 2 * c16 * kBf16Atom);  // 8 words per row
#pragma unroll
      for (int i = 0; i < NB + NC; ++i) {
        const int row = i < NB ? i : 16 + (i - NB);
        *reinterpret_cast<uint2*>(at + row * 8 + 2 * g) = make_uint2(lo[i], hi[i]);
      }
    } else {
      float t[NB], y[NC];
#pragma unroll
      for (int n = 0; n < NB; ++n) t[n] = 0.0f;
#pragma unroll
      for (int i = 0; i < NC; ++i) y[i] = 0.0f;
      if (k0 + g < cnt) edge_rows<L, NB>(cur[0], ti, S, s_pair, kind, rcut, t, y);
      float* at = wl + c16 * kF32Atom;
#pragma unroll
      for (int n = 0; n < NB; ++n) at[n * 4 + g] = t[n];
#pragma unroll
      for (int i = 0; i < NC; ++i) at[(16 + i) * 4 + g] = y[i];
    }
    wave_sync();
    // atom q: C[n][lm] += sum over this round's slots of T[q][n][slot] Y[q][lm][slot]
#pragma unroll
    for (int q = 0; q < kGroup; ++q) {
      if (BF16) {
        const short* at = wb + 2 * q * kBf16Atom;
        const s4v av = *reinterpret_cast<const s4v*>(at + c16 * 16 + 4 * g);
        const s4v bv = *reinterpret_cast<const s4v*>(at + (16 + c16) * 16 + 4 * g);
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[q], 0, 0, 0);
      } else {
        const float* at = wl + q * kF32Atom;
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(at[c16 * 4 + g], at[(16 + c16) * 4 + g],
                                                      acc[q], 0, 0, 0);
      }
    }
    wave_sync();
  }
  // B[mu][lm] = sum_n c[mu][n] C[n][lm]; C row 4 g + r is acc[r] of lane group g
#pragma unroll
  for (int q = 0; q < kGroup; ++q) {
    f4v bm = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int r = 0; r < 4; ++r) bm = __builtin_amdgcn_mfma_f32_16x16x4f32(cA[r], acc[q][r], bm, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (obase[r] >= 0) stage[obase[r] * kGroup + q] = bm[r];
  }
  wave_sync();
  // rows of 16 consecutive atoms
  const int B = a.B;
  for (int i = lane; i < B * kGroup; i += 64) {
    const int b = i >> 4, q = i & 15;
    if (p0 + q < N) out[static_cast<size_t>(b) * N + p0 + q] = stage[i];
  }
}

template <int L, int NB, bool BF16>
void launch_mfma_shape(const EdgeLaunch<float>& a, float* basics, hipStream_t s)
{
  const int per_block = kMfmaWaves * kGroup;
  const int blocks = (a.N + per_block - 1) / per_block;
  const size_t lds = ((3 * static_cast<size_t>(a.rt.S) * a.rt.S * sizeof(float) + 15) & ~size_t(15)) +
                     kMfmaWaves * wave_lds_bytes<BF16>();
  k_basics_mfma<L, NB, BF16><<<blocks, 64 * kMfmaWaves, lds, s>>>(a, basics);
}

template <int L, bool BF16>
bool launch_mfma_nb(const EdgeLaunch<float>& a, float* basics, hipStream_t s)
{
  const int rb = a.rt.rb;
  if (rb <= 8)
    launch_mfma_shape<L, 8, BF16>(a, basics, s);
  else if (rb <= 12)
    launch_mfma_shape<L, 12, BF16>(a, basics, s);
  else if (rb <= 16)
    launch_mfma_shape<L, 16, BF16>(a, basics, s);
  else
    return false;
  return true;
}

}  // namespace

bool mfma_basics_supported(int l_max, int k_max, int rb, int S, bool uniform, bool gate)
{
  return uniform && !gate && l_max <= 3 && k_max * (l_max + 1) <= 16 && rb <= 16 && S <= 16 &&
         k_max * (l_max + 1) * (l_max + 1) <= kMaxBasics;
}

bool launch_basics_mfma(const EdgeLaunch<float>& a, float* basics, bool bf16, hipStream_t s)
{
  if (a.N <= 0) return true;
  if (!a.mfma_rc || !a.pair_sigma || !a.rt.uniform) return false;
  if (!mfma_basics_supported(a.L, a.K, a.rt.rb, a.rt.S, a.rt.uniform != 0, a.gate_m != nullptr))
    return false;
#define SUS2SH_MFMA_L(LL)                                                            \
  case LL:                                                                          \
    return bf16 ? launch_mfma_nb<LL, true>(a, basics, s) : launch_mfma_nb<LL, false>(a, basics, s);
  switch (a.L) {
    SUS2SH_MFMA_L(0)
    SUS2SH_MFMA_L(1)
    SUS2SH_MFMA_L(2)
    SUS2SH_MFMA_L(3)
    default: return false;
  }
#undef SUS2SH_MFMA_L
}

}  // namespace sus2sh
