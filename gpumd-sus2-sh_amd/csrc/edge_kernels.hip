// Per-edge kernels: basic moments and forces/virial.  Atoms in cell-sorted order, npl
// consecutive lanes per atom (1 by default; 2-8 give small structures more lanes than atoms):
// each lane walks its share of the atom's neighbours and keeps the full (k, l, m) accumulator
// grid in VGPRs, so nothing per-edge is ever stored; the npl partial grids are summed with
// lane shuffles.
//
//   basic moments   B_i[k,l,m] = sum_j R_{k(L+1)+l}(r_ij) Y_lm(r_ij)
//                   (src/force/sus2_sh.cu:3486-3550, 3650-3724)
//   forces/virial   g_ij = sum_b G_b (R'_mu r_hat Y_y + R_mu grad Y_y); F_i += g, F_j -= g,
//                   W_i -= r_ij (x) g   (src/force/sus2_sh.cu:5996-6035, 6347-6446)
// The backward pass recomputes R, R', Y and grad Y per edge instead of storing them.
// Radial tables are staged in LDS (pair type varies per lane).
//
// The kernels are instantiated for 2 precisions x 7 values of l_max; each (precision, l_max)
// pair is compiled as a unit of its own (SUS2SH_EDGE_REAL, SUS2SH_EDGE_L set by the build) so
// that the build runs them in parallel.  Compiled without them, this file holds only the
// dispatch on l_max.
#include "kernels.hpp"

namespace sus2sh {

#if defined(SUS2SH_EDGE_REAL) && defined(SUS2SH_EDGE_L)

namespace {

// Pair strides of the LDS radial tables: padded to 1 mod 32 words so that lanes reading the
// same coefficient of different pair tables hit different banks.
__host__ __device__ constexpr int padded_stride(int n) { return n + ((1 - n % 32) + 32) % 32; }

// Coefficient row length in LDS: NB (zero-padded) for the uniform kernels, rb otherwise.
__host__ __device__ constexpr int coeff_row(int NB, int rb) { return NB > 0 ? NB : rb; }

// Pair stride (in Reals) of the coefficient table.  The uniform kernels (NB > 0) read their
// 8-, 12- or 16-coefficient rows as 16-byte vectors, so each pair table starts 16-byte aligned, 4
// words past a multiple of 32 banks: the pair tables that the lanes of a wave read at once
// then sit in disjoint bank quads.  Scalar rows (NB = 0) keep the 1 mod 32 padding.
template <typename Real>
__host__ __device__ constexpr int coeff_stride(int NB, int n)
{
  if (NB == 0) return padded_stride(n);
  constexpr int wpr = static_cast<int>(sizeof(Real)) / 4;  // words per Real
  const int w = n * wpr;
  return (w + ((4 - w % 32) + 32) % 32) / wpr;
}

// Sum of v over the npl consecutive lanes of one atom (npl a power of two dividing 64):
// butterfly of xor shuffles, so every lane of the group ends with the same value.
template <typename Real>
__device__ __forceinline__ Real group_sum(Real v, int npl)
{
  for (int off = 1; off < npl; off <<= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// The FP32 uniform force kernel reads its coefficient rows from the rows shared by all pairs
// (scalar loads, SGPR operands: no per-edge LDS reads, and the kernel fits its 3-waves-per-SIMD
// register budget without spilling); the other kernels stage every pair's rows in LDS.
template <typename Real, int NB, bool FORCE>
constexpr bool shared_rows()
{
  return FORCE && NB > 0 && sizeof(Real) == sizeof(float);
}

// Bytes of the radial-table region at the start of an edge kernel's LDS: (a, s, sigma) per
// pair with shared rows, else every pair's coefficient rows (padded to the kernel's row
// length) and (a, s) per radial function.
template <typename Real>
__host__ __device__ constexpr size_t table_region_bytes(bool shared, int NB, int S, int Mr, int rb)
{
  return shared ? (3 * static_cast<size_t>(S) * S * sizeof(Real) + 15) & ~size_t(15)
                : ((static_cast<size_t>(S) * S *
                    (coeff_stride<Real>(NB, Mr * coeff_row(NB, rb)) + padded_stride(Mr * 2)) *
                    sizeof(Real)) + 15) & ~size_t(15);
}

// (a, s, sigma) of every pair for the shared-row kernels; ends with a barrier
template <typename Real>
__device__ __forceinline__ void stage_pairs(const RadialTable<Real>& rt, int Mr, Real* s_pair)
{
  for (int i = threadIdx.x; i < rt.S * rt.S; i += blockDim.x) {
    s_pair[3 * i] = rt.scal[static_cast<size_t>(i) * Mr * 2];
    s_pair[3 * i + 1] = rt.scal[static_cast<size_t>(i) * Mr * 2 + 1];
    s_pair[3 * i + 2] = rt.sigma[i];
  }
  __syncthreads();
}

template <typename Real>
using const_rows = const __attribute__((address_space(4))) Real;

template <typename Real>
__device__ __forceinline__ void stage_tables(const RadialTable<Real>& rt, int Mr, int row, int NB,
                                             Real* s_coeff, Real* s_scal)
{
  const int pairs = rt.S * rt.S;
  const int nc = Mr * row, ns = Mr * 2;
  const int cs = coeff_stride<Real>(NB, nc), ss = padded_stride(ns);
  for (int i = threadIdx.x; i < pairs * nc; i += blockDim.x) {
    const int p = i / nc, k = i - p * nc, mu = k / row, n = k - mu * row;
    s_coeff[p * cs + k] = n < rt.rb ? rt.coeff[(p * Mr + mu) * rt.rb + n] : Real(0);
  }
  for (int i = threadIdx.x; i < pairs * ns; i += blockDim.x) {
    const int p = i / ns;
    s_scal[p * ss + (i - p * ns)] = rt.scal[i];
  }
  __syncthreads();
}

// GATE: every edge's radial channel mu is weighted by the neighbour's gate multiplier
// G_{j,mu} (two-layer mu gate, src/force/sus2_sh.cu:4246-4310).
template <typename Real, int L, int K, int NB, bool GATE>
__global__ __launch_bounds__(256) void k_basics(EdgeLaunch<Real> a, Real* __restrict__ out)
{
  constexpr int MR = K * (L + 1);
  constexpr int NC = (L + 1) * (L + 1);
  constexpr int FG = K * NC;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* s_coeff = reinterpret_cast<Real*>(smem_raw);
  const int row = coeff_row(NB, a.rt.rb);
  const int cs = coeff_stride<Real>(NB, MR * row), ss = padded_stride(MR * 2);
  Real* s_scal = s_coeff + a.rt.S * a.rt.S * cs;
  stage_tables(a.rt, MR, row, NB, s_coeff, s_scal);

  const int N = a.N;
  // npl lanes per atom: lane `sub` of atom p takes the edges k = sub, sub + npl, ...
  const int npl = a.npl;
  const int p = xcd_block() * (static_cast<int>(blockDim.x) / npl) + static_cast<int>(threadIdx.x) / npl;
  const int sub = static_cast<int>(threadIdx.x) & (npl - 1);
  const bool valid = p < N;
  if (npl == 1 && !valid) return;
  Real acc[FG];
#pragma unroll
  for (int f = 0; f < FG; ++f) acc[f] = Real(0);
  const int ti = valid ? a.type_s[p] : 0;
  const int cnt = valid ? a.count[p] : 0;
  const int S = a.rt.S;
  const int rb = a.rt.rb;
  const Real rc = a.rt.rc;
  // the next edge is loaded while the current one is evaluated
  Edge<Real> next = sub < cnt ? a.nl[static_cast<size_t>(sub) * N + p] : Edge<Real>{};
  for (int k = sub; k < cnt; k += npl) {
    const Edge<Real> e = next;
    if (k + npl < cnt) next = a.nl[static_cast<size_t>(k + npl) * N + p];
    const Real r2 = e.dx * e.dx + e.dy * e.dy + e.dz * e.dz;
    const Real inv = dev_rsqrt<Real>(r2);
    const Real r = r2 * inv;
    const int pair = ti * S + (e.j >> kEdgeTypeShift);
    Real R[MR], dR[MR];
    eval_radial<Real, MR, false, NB>(s_coeff + pair * cs, s_scal + pair * ss, rb, rc,
                                     a.rt.kind, r, R, dR);
    if constexpr (GATE) {
      const Real* gm = a.gate_m + static_cast<size_t>(e.j & kEdgeIndexMask) * MR;
#pragma unroll
      for (int mu = 0; mu < MR; ++mu) R[mu] *= gm[mu];
    }
    Real Sv[NC];
    solid_values<Real, L>(e.dx * inv, e.dy * inv, e.dz * inv, Sv);
#pragma unroll
    for (int l = 0; l <= L; ++l) {
      const Real nl = Real(norm_of_l(l));
#pragma unroll
      for (int m = -l; m <= l; ++m) {
        const Real y = nl * Sv[sh_index(l, m)];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) acc[full_index<L, K>(l, kk, m)] += R[kk * (L + 1) + l] * y;
      }
    }
  }
  if (npl > 1) {
#pragma unroll
    for (int f = 0; f < FG; ++f) acc[f] = group_sum(acc[f], npl);
  }
  if (!valid) return;
#pragma unroll
  for (int f = 0; f < FG; ++f) {
    const int b = a.fmap[f];
    if (b >= 0 && (f & (npl - 1)) == sub) out[static_cast<size_t>(b) * N + p] = acc[f];
  }
}

// Adjoint of the radial coefficients for one edge: with A_mu = sum_m G_{k,l,m} Y_lm,
//   dE/dc^p_{mu n} += A_mu (r - rc)^2 T_n(xi_mu),
//   dE/da^p_mu     += A_mu (r - rc)^2 sum_n c_n T'_n(xi_mu) dxi/da   (same for s),
// accumulated per pair p in LDS (double), entry i of the pair at D[i * cp]: the caller's lane
// adds into its own one of cp interleaved copies, so that the lanes of a wave, which mostly
// hit the same few pairs, do not serialise on one address.  No reference counterpart
// (training path).
template <typename Real, int MR>
__device__ __forceinline__ void radial_param_adjoint(const Real* __restrict__ coeff,
                                                     const Real* __restrict__ scal, int rb,
                                                     Real rc, int kind, Real r,
                                                     const Real (&A)[MR], double* __restrict__ D,
                                                     double* __restrict__ Da,
                                                     double* __restrict__ Ds, int cp)
{
  const Real env = (r - rc) * (r - rc);
#pragma unroll
  for (int mu = 0; mu < MR; ++mu) {
    const Real a = scal[2 * mu], s = scal[2 * mu + 1];
    const Real z = Real(0.5) * a * (r - s);
    Real xi, dxdz;
    if (kind == 1) {
      const Real q = dev_rsqrt<Real>(Real(1) + z * z);
      xi = z * q;
      dxdz = q * q * q;
    } else {
      xi = dev_tanh<Real>(z);
      dxdz = Real(1) - xi * xi;
    }
    const Real* c = coeff + mu * rb;
    const Real w = A[mu] * env;
    Real t0 = 1, t1 = xi, d0 = 0, d1 = 1;
    atomicAdd(D + (mu * rb) * cp, static_cast<double>(w));
    Real dsum = 0;
    if (rb > 1) {
      atomicAdd(D + (mu * rb + 1) * cp, static_cast<double>(w * xi));
      dsum = c[1];
    }
    for (int n = 2; n < rb; ++n) {
      const Real t2 = Real(2) * xi * t1 - t0;
      const Real d2 = Real(2) * (t1 + xi * d1) - d0;
      atomicAdd(D + (mu * rb + n) * cp, static_cast<double>(w * t2));
      dsum += c[n] * d2;
      t0 = t1;
      t1 = t2;
      d0 = d1;
      d1 = d2;
    }
    atomicAdd(Da + mu * cp, static_cast<double>(w * dsum * Real(0.5) * (r - s) * dxdz));
    atomicAdd(Ds + mu * cp, static_cast<double>(w * dsum * Real(-0.5) * a * dxdz));
  }
}

// Position of sorted atom j in the LDS accumulators of a tile (-1: outside the windows):
// windows are disjoint and sorted by first atom; w_off is their prefix size.
__device__ __forceinline__ int window_slot(const int* __restrict__ w_first,
                                           const int* __restrict__ w_off, int count, int j)
{
  if (count <= 0) return -1;
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (w_first[mid] <= j)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int d = j - w_first[lo];
  return (d >= 0 && d < w_off[lo + 1] - w_off[lo]) ? w_off[lo] + d : -1;
}

// WIN: neighbour forces of the tile (kWinTile atoms) are summed in LDS over the tile's
// neighbour windows (launch_windows) and flushed with one global atomic per neighbour and
// component; a tile without windows (count < 0) falls back to per-edge global atomics.  The
// window accumulators are double in both precisions: FP64 LDS atomic adds issue at the rate
// of other LDS operations on this GPU, FP32 ones are many times slower.
// The FP32 windowed kernel with 8-coefficient rows is held to 168 VGPRs (3 waves per SIMD)
// for l_max <= 4; the l_max = 5, 6 grids would spill far more at that budget.
// GATE (two-layer mu gate, src/force/sus2_sh.cu:6077-6123, 6555-6693): the edge's radial
// channels are weighted by the neighbour's multiplier G_{j,mu}, and
// dE/dh_{j,mu} += dG/dh_{j,mu} R_mu sum_m G_{k,l,m} Y_lm is scattered to the neighbour.
// FIX (deterministic summation): f_s holds 64-bit fixed-point sums (fix_encode); the window
// accumulators are fixed point too, so every total is independent of the order of the adds.
template <typename Real, int L, int K, bool PG, int NB, bool WIN, bool GATE, bool FIX>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu((sizeof(Real) == 4 && WIN && NB == 8 && L <= 4) ? 3 : 1))) void k_forces(EdgeLaunch<Real> a, const Real* __restrict__ grads,
                                                double* __restrict__ f_s,
                                                double* __restrict__ v_s,
                                                double* __restrict__ pg_rad,
                                                const int* __restrict__ win)
{
  constexpr int MR = K * (L + 1);
  constexpr int NC = (L + 1) * (L + 1);
  constexpr int FG = K * NC;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int S = a.rt.S;
  const int rb = a.rt.rb;
  Real* s_coeff = reinterpret_cast<Real*>(smem_raw);
  const int row = coeff_row(NB, rb);
  const int cs = coeff_stride<Real>(NB, MR * row), ss = padded_stride(MR * 2);
  Real* s_scal = s_coeff + S * S * cs;
  // PG: per-pair adjoint accumulators [S*S][MR*rb + 2*MR] (double), 16-byte aligned
  constexpr bool SHARED = shared_rows<Real, NB, true>();
  const size_t pg_off = table_region_bytes<Real>(SHARED, NB, S, MR, rb);
  double* s_pg = reinterpret_cast<double*>(smem_raw + pg_off);
  const int pg_stride = MR * rb + 2 * MR;
  const int pg_cp = a.pg_copies;  // interleaved copies of the accumulators (power of two)
  if (PG) {
    for (int i = threadIdx.x; i < S * S * pg_stride * pg_cp; i += blockDim.x) s_pg[i] = 0.0;
  }
  // WIN: window table (first atom, prefix size) and the [3][total] accumulators
  int* w_first = reinterpret_cast<int*>(smem_raw + pg_off);
  int* w_off = w_first + kWinMax;
  double* acc_w = reinterpret_cast<double*>(smem_raw + pg_off + 16 * ((sizeof(int) * (2 * kWinMax + 1) + 15) / 16));
  unsigned long long* acc_q = reinterpret_cast<unsigned long long*>(acc_w);  // FIX
  unsigned long long* f_q = reinterpret_cast<unsigned long long*>(f_s);      // FIX
  int w_count = -1, w_total = 0;
  // npl lanes per atom (edge k to lane k mod npl); a workgroup covers `per_block` atoms of
  // the cell order, inside one window tile (kWinTile atoms, a multiple of per_block)
  const int npl = a.npl;
  const int per_block = static_cast<int>(blockDim.x) / npl;
  const int block = xcd_block();
  if (WIN) {
    const int* tw = win + static_cast<size_t>(block * per_block / kWinTile) * kWinStride;
    w_count = tw[0];
    w_total = tw[1];
    if (w_count >= 0) {
      if (threadIdx.x < w_count) w_first[threadIdx.x] = tw[2 + 2 * threadIdx.x];
      if (threadIdx.x == 0) {
        int o = 0;
        for (int w = 0; w < w_count; ++w) {
          w_off[w] = o;
          o += tw[3 + 2 * w];
        }
        w_off[w_count] = o;
      }
      for (int i = threadIdx.x; i < 3 * w_total; i += blockDim.x) {
        if constexpr (FIX)
          acc_q[i] = 0ull;
        else
          acc_w[i] = 0.0;
      }
    }
  }
  if constexpr (SHARED)
    stage_pairs(a.rt, MR, s_coeff);  // (a, s, sigma) per pair; ends with a barrier
  else
    stage_tables(a.rt, MR, row, NB, s_coeff, s_scal);  // ends with a barrier
  const_rows<Real>* shared = (const_rows<Real>*)a.rt.shared;

  const int N = a.N;
  const int p = block * per_block + static_cast<int>(threadIdx.x) / npl;
  const int sub = static_cast<int>(threadIdx.x) & (npl - 1);
  const bool valid = p < N;
  if (!PG && !WIN && npl == 1 && !valid) return;
  Real G[FG];
#pragma unroll
  for (int f = 0; f < FG; ++f) {
    const int b = a.fmap[f];
    G[f] = (b >= 0 && valid) ? grads[static_cast<size_t>(b) * N + p] : Real(0);
  }
  const int ti = valid ? a.type_s[p] : 0;
  const int cnt = valid ? a.count[p] : 0;
  const Real rc = a.rt.rc;
  Real fx = 0, fy = 0, fz = 0;
  Real w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  Edge<Real> next = sub < cnt ? a.nl[static_cast<size_t>(sub) * N + p] : Edge<Real>{};
  for (int k = sub; k < cnt; k += npl) {
    const Edge<Real> e = next;
    if (k + npl < cnt) next = a.nl[static_cast<size_t>(k + npl) * N + p];
    const Real r2 = e.dx * e.dx + e.dy * e.dy + e.dz * e.dz;
    const Real inv = dev_rsqrt<Real>(r2);
    const Real r = r2 * inv;
    const Real ux = e.dx * inv, uy = e.dy * inv, uz = e.dz * inv;
    const int pair = ti * S + (e.j >> kEdgeTypeShift);
    const int jn = e.j & kEdgeIndexMask;
    Real R[MR], dR[MR];
    if constexpr (SHARED)
      eval_radial_shared<Real, MR, true, NB>(shared, s_coeff[3 * pair + 2], s_coeff[3 * pair],
                                             s_coeff[3 * pair + 1], rc, a.rt.kind, r, R, dR);
    else
      eval_radial<Real, MR, true, NB>(s_coeff + pair * cs, s_scal + pair * ss, rb, rc,
                                    a.rt.kind, r, R, dR);
    Real Sv[NC];
    solid_values<Real, L>(ux, uy, uz, Sv);
    if constexpr (GATE) {
      const Real* gm = a.gate_m + static_cast<size_t>(jn) * MR;
      const Real* gd = a.gate_d + static_cast<size_t>(jn) * MR;
      Real* ga = a.gate_adj + static_cast<size_t>(jn) * MR;
#pragma unroll
      for (int l = 0; l <= L; ++l) {
        const Real nl = Real(norm_of_l(l));
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
          const int mu = kk * (L + 1) + l;
          Real proj = 0;
#pragma unroll
          for (int m = -l; m <= l; ++m) proj += G[full_index<L, K>(l, kk, m)] * Sv[sh_index(l, m)];
          const Real v = gd[mu] * R[mu] * proj * nl;
          if (v != Real(0)) unsafeAtomicAdd(ga + mu, v);
        }
      }
#pragma unroll
      for (int mu = 0; mu < MR; ++mu) {
        R[mu] *= gm[mu];
        dR[mu] *= gm[mu];
      }
    }
    // g = u * sum (V - l U / r) Y + (1/r) sum U N_l grad S,  with U = sum_k G R, V = sum_k G R'
    Real sv = 0, gx = 0, gy = 0, gz = 0;
#pragma unroll
    for (int l = 0; l <= L; ++l) {
      const Real nl = Real(norm_of_l(l));
#pragma unroll
      for (int m = -l; m <= l; ++m) {
        Real U = 0, V = 0;
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
          const Real gb = G[full_index<L, K>(l, kk, m)];
          U += gb * R[kk * (L + 1) + l];
          V += gb * dR[kk * (L + 1) + l];
        }
        const int y = sh_index(l, m);
        sv += (V - Real(l) * inv * U) * nl * Sv[y];
        add_solid_grad<Real, L>(l, m, U * nl, Sv, gx, gy, gz);
      }
    }
    if (PG) {
      Real A[MR];
#pragma unroll
      for (int l = 0; l <= L; ++l) {
        const Real nl = Real(norm_of_l(l));
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
          Real acc = 0;
#pragma unroll
          for (int m = -l; m <= l; ++m) acc += G[full_index<L, K>(l, kk, m)] * Sv[sh_index(l, m)];
          A[kk * (L + 1) + l] = acc * nl;
        }
      }
      double* Dp = s_pg + static_cast<size_t>(pair) * pg_stride * pg_cp + (threadIdx.x & (pg_cp - 1));
      radial_param_adjoint<Real, MR>(s_coeff + pair * cs, s_scal + pair * ss, rb, rc,
                                     a.rt.kind, r, A, Dp, Dp + MR * rb * pg_cp,
                                     Dp + (MR * rb + MR) * pg_cp, pg_cp);
    }
    const Real g0 = ux * sv + gx * inv;
    const Real g1 = uy * sv + gy * inv;
    const Real g2 = uz * sv + gz * inv;
    fx += g0;
    fy += g1;
    fz += g2;
    const int slot = (WIN && w_count >= 0) ? window_slot(w_first, w_off, w_count, jn) : -1;
    if constexpr (FIX) {
      const unsigned long long q0 = fix_encode(-static_cast<double>(g0), kFixPairMax, a.fix_flag);
      const unsigned long long q1 = fix_encode(-static_cast<double>(g1), kFixPairMax, a.fix_flag);
      const unsigned long long q2 = fix_encode(-static_cast<double>(g2), kFixPairMax, a.fix_flag);
      if (slot >= 0) {
        atomicAdd(acc_q + slot, q0);
        atomicAdd(acc_q + w_total + slot, q1);
        atomicAdd(acc_q + 2 * w_total + slot, q2);
      } else {
        atomicAdd(f_q + jn, q0);
        atomicAdd(f_q + N + jn, q1);
        atomicAdd(f_q + 2 * N + jn, q2);
      }
    } else if (slot >= 0) {
      atomicAdd(acc_w + slot, -static_cast<double>(g0));
      atomicAdd(acc_w + w_total + slot, -static_cast<double>(g1));
      atomicAdd(acc_w + 2 * w_total + slot, -static_cast<double>(g2));
    } else {
      unsafeAtomicAdd(f_s + jn, -static_cast<double>(g0));
      unsafeAtomicAdd(f_s + N + jn, -static_cast<double>(g1));
      unsafeAtomicAdd(f_s + 2 * N + jn, -static_cast<double>(g2));
    }
    w[0] -= e.dx * g0;
    w[1] -= e.dy * g1;
    w[2] -= e.dz * g2;
    w[3] -= e.dx * g1;
    w[4] -= e.dx * g2;
    w[5] -= e.dy * g2;
    w[6] -= e.dy * g0;
    w[7] -= e.dz * g0;
    w[8] -= e.dz * g1;
  }
  if (npl > 1) {
    fx = group_sum(fx, npl);
    fy = group_sum(fy, npl);
    fz = group_sum(fz, npl);
#pragma unroll
    for (int c = 0; c < 9; ++c) w[c] = group_sum(w[c], npl);
  }
  if (valid && sub == 0) {
    if constexpr (FIX) {
      atomicAdd(f_q + p, fix_encode(static_cast<double>(fx), kFixSumMax, a.fix_flag));
      atomicAdd(f_q + N + p, fix_encode(static_cast<double>(fy), kFixSumMax, a.fix_flag));
      atomicAdd(f_q + 2 * N + p, fix_encode(static_cast<double>(fz), kFixSumMax, a.fix_flag));
    } else {
      unsafeAtomicAdd(f_s + p, static_cast<double>(fx));
      unsafeAtomicAdd(f_s + N + p, static_cast<double>(fy));
      unsafeAtomicAdd(f_s + 2 * N + p, static_cast<double>(fz));
    }
    if (a.v_accumulate) {
#pragma unroll
      for (int c = 0; c < 9; ++c) v_s[static_cast<size_t>(c) * N + p] += static_cast<double>(w[c]);
    } else {
#pragma unroll
      for (int c = 0; c < 9; ++c) v_s[static_cast<size_t>(c) * N + p] = static_cast<double>(w[c]);
    }
  }
  if (WIN && w_count >= 0) {
    __syncthreads();
    for (int i = threadIdx.x; i < w_total; i += blockDim.x) {
      int lo = 0, hi = w_count - 1;  // window of slot i
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (w_off[mid] <= i)
          lo = mid;
        else
          hi = mid - 1;
      }
      const int j = w_first[lo] + (i - w_off[lo]);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if constexpr (FIX) {
          const unsigned long long v = acc_q[c * w_total + i];
          if (v != 0ull) atomicAdd(f_q + static_cast<size_t>(c) * N + j, v);
        } else {
          const double v = acc_w[c * w_total + i];
          if (v != 0.0) unsafeAtomicAdd(f_s + static_cast<size_t>(c) * N + j, v);
        }
      }
    }
  }
  if (PG) {
    __syncthreads();
    for (int i = threadIdx.x; i < S * S * pg_stride; i += blockDim.x) {
      double v = 0.0;
      for (int c = 0; c < pg_cp; ++c) v += s_pg[static_cast<size_t>(i) * pg_cp + c];
      if (v != 0.0) unsafeAtomicAdd(pg_rad + i, v);
    }
  }
}

template <typename Real>
size_t table_bytes(const EdgeLaunch<Real>& a, int NB, bool force)
{
  const bool shared = force && NB > 0 && sizeof(Real) == sizeof(float);
  return table_region_bytes<Real>(shared, NB, a.rt.S, a.K * (a.L + 1), a.rt.rb);
}

// NB of the uniform kernels: coefficient rows padded to 8, 12 or 16 (rb = 10, the common
// production size, pads to 12); 0 selects the per-mu path.  The FP32 force kernel reads the
// shared rows: without shared rows of that length it takes the per-mu path.
template <typename Real>
int uniform_nb(const EdgeLaunch<Real>& a, bool force)
{
  if (!a.rt.uniform) return 0;
  const int nb = a.rt.rb <= 8 ? 8 : (a.rt.rb <= 12 ? 12 : 16);
  if (force && sizeof(Real) == sizeof(float) &&
      !(a.rt.shared && a.rt.sigma && a.rt.shared_nb == nb))
    return 0;
  return nb;
}

#define SUS2SH_FOR_EACH_K(X) X(1) X(2) X(3) X(4) X(5) X(6)

}  // namespace

template <typename Real, int LL>
bool launch_basics_l(const EdgeLaunch<Real>& a, Real* basics, hipStream_t s)
{
  if (!edge_shape_valid(a.threads, a.npl)) return false;
  const int per_block = a.threads / a.npl;
  const int blocks = (a.N + per_block - 1) / per_block;
  const int T = a.threads;
  const int nb = a.gate_m ? 0 : uniform_nb(a, false);  // the gated kernel uses the per-mu path
  const size_t lds = table_bytes(a, nb, false);
#define SUS2SH_BASICS_CASE(KK)                                                    \
  if (a.K == KK) {                                                                \
    if (a.gate_m)                                                                 \
      k_basics<Real, LL, KK, 0, true><<<blocks, T, lds, s>>>(a, basics);        \
    else if (nb == 8)                                                             \
      k_basics<Real, LL, KK, 8, false><<<blocks, T, lds, s>>>(a, basics);       \
    else if (nb == 12)                                                            \
      k_basics<Real, LL, KK, 12, false><<<blocks, T, lds, s>>>(a, basics);      \
    else if (nb == 16)                                                            \
      k_basics<Real, LL, KK, 16, false><<<blocks, T, lds, s>>>(a, basics);      \
    else                                                                          \
      k_basics<Real, LL, KK, 0, false><<<blocks, T, lds, s>>>(a, basics);       \
    return true;                                                                  \
  }
  SUS2SH_FOR_EACH_K(SUS2SH_BASICS_CASE)
#undef SUS2SH_BASICS_CASE
  return false;
}

template <typename Real, int LL>
bool launch_forces_l(const EdgeLaunch<Real>& a, const Real* grads, double* f_s, double* v_s,
                     double* pg_rad, const int* win, int win_cap, hipStream_t s)
{
  if (!edge_shape_valid(a.threads, a.npl)) return false;
  const int per_block = a.threads / a.npl;
  const int blocks = (a.N + per_block - 1) / per_block;
  const int T = a.threads;
  // windows serve workgroups that lie inside one window tile (per_block divides kWinTile)
  // the parameter gradient, the gated kernel and deterministic summation use the per-mu path
  const bool fix = a.fix_flag != nullptr;
  const int nb = (pg_rad || a.gate_m || fix) ? 0 : uniform_nb(a, true);
  const bool use_win = win && !pg_rad && !a.gate_m && kWinTile % per_block == 0;
  size_t lds = table_bytes(a, nb, true);
  if (pg_rad) {
    const int Mr = a.K * (a.L + 1);
    lds = ((lds + 15) & ~size_t(15)) +
          sizeof(double) * a.rt.S * a.rt.S * (Mr * a.rt.rb + 2 * Mr) * a.pg_copies;
  }
  if (use_win)
    lds = ((lds + 15) & ~size_t(15)) + 16 * ((sizeof(int) * (2 * kWinMax + 1) + 15) / 16) +
          3 * sizeof(double) * static_cast<size_t>(win_cap);
#define SUS2SH_FORCES_LAUNCH(LL, KK, PGV, NBV, WINV, GV, FV)                                   \
  {                                                                                            \
    auto kern = &k_forces<Real, LL, KK, PGV, NBV, WINV, GV, FV>;                               \
    if (lds > 65536) allow_large_lds<&k_forces<Real, LL, KK, PGV, NBV, WINV, GV, FV>>();       \
    kern<<<blocks, T, lds, s>>>(a, grads, f_s, v_s, pg_rad, win);                            \
  }
#define SUS2SH_FORCES_CASE(KK)                                                               \
  if (a.K == KK) {                                                                           \
    if (pg_rad) {                                                                            \
      SUS2SH_FORCES_LAUNCH(LL, KK, true, 0, false, false, false)                             \
    } else if (a.gate_m) {                                                                   \
      SUS2SH_FORCES_LAUNCH(LL, KK, false, 0, false, true, false)                             \
    } else if (fix) {                                                                        \
      if (use_win) SUS2SH_FORCES_LAUNCH(LL, KK, false, 0, true, false, true)                 \
      else SUS2SH_FORCES_LAUNCH(LL, KK, false, 0, false, false, true)                        \
    } else if (use_win) {                                                                    \
      if (nb == 8) SUS2SH_FORCES_LAUNCH(LL, KK, false, 8, true, false, false)                \
      else if (nb == 12) SUS2SH_FORCES_LAUNCH(LL, KK, false, 12, true, false, false)         \
      else if (nb == 16) SUS2SH_FORCES_LAUNCH(LL, KK, false, 16, true, false, false)         \
      else SUS2SH_FORCES_LAUNCH(LL, KK, false, 0, true, false, false)                        \
    } else {                                                                                 \
      if (nb == 8) SUS2SH_FORCES_LAUNCH(LL, KK, false, 8, false, false, false)               \
      else if (nb == 12) SUS2SH_FORCES_LAUNCH(LL, KK, false, 12, false, false, false)        \
      else if (nb == 16) SUS2SH_FORCES_LAUNCH(LL, KK, false, 16, false, false, false)        \
      else SUS2SH_FORCES_LAUNCH(LL, KK, false, 0, false, false, false)                       \
    }                                                                                        \
    return true;                                                                             \
  }
  SUS2SH_FOR_EACH_K(SUS2SH_FORCES_CASE)
#undef SUS2SH_FORCES_CASE
#undef SUS2SH_FORCES_LAUNCH
  return false;
}

template bool launch_basics_l<SUS2SH_EDGE_REAL, SUS2SH_EDGE_L>(const EdgeLaunch<SUS2SH_EDGE_REAL>&,
                                                              SUS2SH_EDGE_REAL*, hipStream_t);
template bool launch_forces_l<SUS2SH_EDGE_REAL, SUS2SH_EDGE_L>(
    const EdgeLaunch<SUS2SH_EDGE_REAL>&, const SUS2SH_EDGE_REAL*, double*, double*, double*,
    const int*, int, hipStream_t);

#else  // dispatch on l_max

template <typename Real>
bool launch_basics(const EdgeLaunch<Real>& a, Real* basics, hipStream_t s)
{
  if (a.N <= 0) return true;
  switch (a.L) {
    case 0: return launch_basics_l<Real, 0>(a, basics, s);
    case 1: return launch_basics_l<Real, 1>(a, basics, s);
    case 2: return launch_basics_l<Real, 2>(a, basics, s);
    case 3: return launch_basics_l<Real, 3>(a, basics, s);
    case 4: return launch_basics_l<Real, 4>(a, basics, s);
    case 5: return launch_basics_l<Real, 5>(a, basics, s);
    case 6: return launch_basics_l<Real, 6>(a, basics, s);
    default: return false;
  }
}

template <typename Real>
bool launch_forces(const EdgeLaunch<Real>& a, const Real* grads, double* f_s, double* v_s,
                   double* pg_rad, const int* win, int win_cap, hipStream_t s)
{
  if (a.N <= 0) return true;
#define SUS2SH_FORCES_L(LL) \
  case LL: return launch_forces_l<Real, LL>(a, grads, f_s, v_s, pg_rad, win, win_cap, s);
  switch (a.L) {
    SUS2SH_FORCES_L(0)
    SUS2SH_FORCES_L(1)
    SUS2SH_FORCES_L(2)
    SUS2SH_FORCES_L(3)
    SUS2SH_FORCES_L(4)
    SUS2SH_FORCES_L(5)
    SUS2SH_FORCES_L(6)
    default: return false;
  }
#undef SUS2SH_FORCES_L
}

template bool launch_basics<float>(const EdgeLaunch<float>&, float*, hipStream_t);
template bool launch_basics<double>(const EdgeLaunch<double>&, double*, hipStream_t);
template bool launch_forces<float>(const EdgeLaunch<float>&, const float*, double*, double*,
                                   double*, const int*, int, hipStream_t);
template bool launch_forces<double>(const EdgeLaunch<double>&, const double*, double*, double*,
                                    double*, const int*, int, hipStream_t);

#endif

}  // namespace sus2sh
