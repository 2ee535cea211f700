// Device-side building blocks shared by the SUS2-SH kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sus2sh {

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin (workgroup b runs on XCD
// b mod 8), and each XCD has its own L2.  xcd_block() renumbers the workgroups so that XCD x
// runs one contiguous range of logical blocks: atoms are in cell order, so the neighbours a
// workgroup reads (positions, adjoints, force accumulators) then stay in the L2 of the XCD
// that works on their region.  A bijection of [0, gridDim.x) for any grid size.
__device__ __forceinline__ int xcd_block()
{
  const int nb = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
  const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  return x * q + (x < r ? x : r) + k;
}

// One neighbour edge, stored k-major (edge (k, p) at k * N + p) so that a wavefront whose
// lanes are 64 consecutive atoms reads 64 consecutive edges.  j packs the neighbour's
// cell-sorted index (low 27 bits) and its species (bits 27-30), so one 16-byte load gives
// everything an edge kernel needs.
constexpr int kEdgeTypeShift = 27;
constexpr int kEdgeIndexMask = (1 << kEdgeTypeShift) - 1;
constexpr int kMaxEdgeAtoms = 1 << kEdgeTypeShift;

// Deterministic force summation: force contributions are rounded to 64-bit fixed point
// (quantum 2^-34 eV/A) and summed with integer atomics, which are associative, so the totals
// do not depend on the order in which threads add them.  Bounds that keep every per-atom
// total inside the signed 64-bit range (2^29 eV/A): one pair contribution below 2^17 eV/A,
// one lane's own sum below 2^26 eV/A, at most kFixMaxNeighbors neighbours per atom -- twice
// (many-body and ZBL terms) 2^26 + 1023 * 2^17 < 2^28.  A contribution outside its bound
// sets *flag and adds nothing.
constexpr double kFixScale = 17179869184.0;        // 2^34
constexpr double kFixInvScale = 1.0 / 17179869184.0;
constexpr double kFixPairMax = 131072.0;           // 2^17 eV/A
constexpr double kFixSumMax = 67108864.0;          // 2^26 eV/A
constexpr int kFixMaxNeighbors = 1023;

__device__ __forceinline__ unsigned long long fix_encode(double v, double limit, int* flag)
{
  if (!(fabs(v) < limit)) {  // also NaN
    atomicOr(flag, 1);
    return 0ull;
  }
  return static_cast<unsigned long long>(__double2ll_rn(v * kFixScale));
}

template <typename Real>
struct alignas(16) Edge {
  Real dx, dy, dz;  // r_q(image) - r_p
  int j;            // neighbour index | species << kEdgeTypeShift
};
template <>
struct alignas(16) Edge<double> {
  double dx, dy, dz;
  int j;
  int pad;
};

template <typename Real>
struct Term {
  uint32_t idx;
  Real c;
};
template <>
struct alignas(16) Term<double> {
  uint32_t idx;
  uint32_t pad;
  double c;
};

template <typename T, int N>
using vec_t = T __attribute__((ext_vector_type(N)));

// Radial tables for pair p = t_i S + t_j (built from the model on the host).
template <typename Real>
struct RadialTable {
  const Real* coeff;  // [S*S][Mr][rb]  scaling * w_ti * w_tj * radial_coeffs[mu][n]
  const Real* scal;   // [S*S][Mr][2]   (a, s)
  int S;
  int rb;
  Real rc;
  int kind;     // 0: xi = tanh(z), 1: xi = z / sqrt(1 + z^2); z = a (r - s) / 2
  int uniform;  // 1 when (a, s) does not depend on mu for every pair
  // uniform models: coeff[pair][mu][n] = sigma[pair] shared[mu][n], the shared matrix with
  // rows zero-padded to the kernel's row length NB ([Mr][NB]; null: use coeff)
  const Real* shared = nullptr;
  const Real* sigma = nullptr;  // [S*S]
  int shared_nb = 0;            // row length of `shared`
};

struct CellGrid {
  double h[9];
  double hi[9];
  int pbc[3];
  int n[3];  // cells per direction
  int R[3];  // cell search half-range per direction
};

template <typename Real>
__device__ __forceinline__ Real dev_sqrt(Real x)
{
  return sqrt(x);
}

// 1/sqrt(x) and tanh(z).  FP32 uses the hardware reciprocal square root and exponential
// (tanh z = 1 - 2 / (exp(2z) + 1), absolute error ~1e-7 on the |z| < 10 range of the radial
// map); FP64 uses the correctly rounded library functions.
template <typename Real>
__device__ __forceinline__ Real dev_rsqrt(Real x)
{
  return Real(1) / sqrt(x);
}
template <>
__device__ __forceinline__ float dev_rsqrt<float>(float x)
{
  return __builtin_amdgcn_rsqf(x);
}
template <typename Real>
__device__ __forceinline__ Real dev_tanh(Real z)
{
  return tanh(z);
}
template <>
__device__ __forceinline__ float dev_tanh<float>(float z)
{
  const float t = __builtin_amdgcn_exp2f(z * 2.8853900817779268f);  // exp(2z) = 2^(2 z log2 e)
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(t + 1.0f);
}

// tanh(z) and its derivative 1 - tanh(z)^2 = 4 q (1 - q), q = 1 / (exp(2|z|) + 1) <= 1/2: the
// derivative keeps its relative precision where tanh saturates (1 - t*t would cancel).
template <typename Real>
__device__ __forceinline__ void dev_tanh_sech2(Real z, Real& t, Real& sech2)
{
  const Real az = z < Real(0) ? -z : z;
  Real q;
  if constexpr (sizeof(Real) == 4)
    q = __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(az * 2.8853900817779268f) + 1.0f);
  else
    q = Real(1) / (exp(Real(2) * az) + Real(1));
  const Real ta = Real(1) - Real(2) * q;
  t = z < Real(0) ? -ta : ta;
  sech2 = Real(4) * q * (Real(1) - q);
}

// xi(r) and dxi/dr for one (a, s)
template <typename Real>
__device__ __forceinline__ void radial_map(int kind, Real a, Real s, Real r, Real& xi, Real& dxi)
{
  const Real z = Real(0.5) * a * (r - s);
  if (kind == 1) {
    const Real q = dev_rsqrt<Real>(Real(1) + z * z);
    xi = z * q;
    dxi = Real(0.5) * a * q * q * q;
  } else {
    xi = dev_tanh<Real>(z);
    dxi = Real(0.5) * a * (Real(1) - xi * xi);
  }
}

// NB values from a 16-byte aligned LDS row, as NB / (16 / sizeof(Real)) vector loads.
template <typename Real, int NB>
__device__ __forceinline__ void load_row16(const Real* __restrict__ row, Real (&out)[NB])
{
  if constexpr (sizeof(Real) == 4) {
    static_assert(NB % 4 == 0, "rows of 4-byte values come in 16-byte vectors");
    const float4* v = reinterpret_cast<const float4*>(row);
#pragma unroll
    for (int k = 0; k < NB / 4; ++k) {
      const float4 q = v[k];
      out[4 * k] = q.x;
      out[4 * k + 1] = q.y;
      out[4 * k + 2] = q.z;
      out[4 * k + 3] = q.w;
    }
  } else {
    static_assert(NB % 2 == 0, "rows of 8-byte values come in 16-byte vectors");
    const double2* v = reinterpret_cast<const double2*>(row);
#pragma unroll
    for (int k = 0; k < NB / 2; ++k) {
      const double2 q = v[k];
      out[2 * k] = q.x;
      out[2 * k + 1] = q.y;
    }
  }
}

// R_mu(r) (and R'_mu(r)) for all MR radial functions of one pair: Chebyshev recurrence on
// xi with the (r - rc)^2 envelope (src/force/sus2_sh.cu:3171-3220).  `coeff` and `scal`
// point at the pair's block of the tables (usually in LDS).
// NB > 0: every mu shares (a, s) ("uniform"); coefficient rows are NB long, zero beyond rb,
//         so T_n(xi), T'_n(xi) are computed once per edge and the sums need no bound test.
// NB = 0: (a, s) per mu; coefficient rows are rb long.
template <typename Real, int MR, bool DERIV, int NB>
__device__ __forceinline__ void eval_radial(const Real* __restrict__ coeff,
                                            const Real* __restrict__ scal, int rb, Real rc,
                                            int kind, Real r, Real (&R)[MR], Real (&dR)[MR])
{
  const Real dr = r - rc;
  const Real env = dr * dr;
  const Real denv = Real(2) * dr;
  Real xi = 0, dxi = 0;
  if constexpr (NB > 0) {
    radial_map<Real>(kind, scal[0], scal[1], r, xi, dxi);
    Real T[NB], dT[NB];
    T[0] = Real(1);
    dT[0] = Real(0);
    T[1] = xi;
    dT[1] = Real(1);
#pragma unroll
    for (int n = 2; n < NB; ++n) {
      T[n] = Real(2) * xi * T[n - 1] - T[n - 2];
      if (DERIV) dT[n] = Real(2) * (T[n - 1] + xi * dT[n - 1]) - dT[n - 2];
    }
#pragma unroll
    for (int mu = 0; mu < MR; ++mu) {
      // the row is 16-byte aligned (see the table stride): read it as 16-byte vectors
      Real c[NB];
      load_row16<Real, NB>(coeff + mu * NB, c);
      Real acc = 0, accd = 0;
#pragma unroll
      for (int n = 0; n < NB; ++n) {
        const Real cn = c[n];
        acc += cn * T[n];
        if (DERIV) accd += cn * dT[n];
      }
      R[mu] = env * acc;
      if (DERIV) dR[mu] = denv * acc + env * dxi * accd;
    }
    return;
  }
  if (kind == 2) {
    // RBLaguerre_log1p (src/force/sus2_sh.cu:3299-3354): phi_n = (r - rc)^2 e^(-u/2) L_n(u),
    // u = a log(1 + r / rho) with (a, rho) = the pair's (scal, shift) and rho >= 1e-8;
    // Laguerre recurrence (n + 1) L_{n+1} = (2n + 1 - u) L_n - n L_{n-1}, L'_{n+1} = L'_n - L_n
#pragma unroll
    for (int mu = 0; mu < MR; ++mu) {
      const Real a = scal[2 * mu];
      const Real rho = scal[2 * mu + 1] > Real(1e-8) ? scal[2 * mu + 1] : Real(1e-8);
      const Real u = a * log1p(r / rho);
      const Real du = a / (rho + r);
      const Real ex = exp(Real(-0.5) * u);
      const Real* c = coeff + mu * rb;
      Real l0 = Real(1), l1 = Real(1) - u, d1 = Real(-1);
      Real acc = c[0], accd = Real(0);
      if (rb > 1) {
        acc += c[1] * l1;
        accd += c[1] * d1;
      }
      for (int n = 1; n + 1 < rb; ++n) {
        const Real l2 = ((Real(2 * n + 1) - u) * l1 - Real(n) * l0) / Real(n + 1);
        const Real d2 = d1 - l1;
        acc += c[n + 1] * l2;
        if (DERIV) accd += c[n + 1] * d2;
        l0 = l1;
        l1 = l2;
        d1 = d2;
      }
      R[mu] = env * ex * acc;
      if (DERIV) dR[mu] = ex * (denv * acc + env * du * (accd - Real(0.5) * acc));
    }
    return;
  }
#pragma unroll
  for (int mu = 0; mu < MR; ++mu) {
    radial_map<Real>(kind, scal[2 * mu], scal[2 * mu + 1], r, xi, dxi);
    const Real* c = coeff + mu * rb;
    Real t0 = Real(1), t1 = xi, d0 = Real(0), d1 = Real(1);
    Real acc = c[0];
    Real accd = Real(0);
    if (rb > 1) {
      acc += c[1] * xi;
      accd = c[1];
    }
    for (int n = 2; n < rb; ++n) {
      const Real t2 = Real(2) * xi * t1 - t0;
      const Real d2 = Real(2) * (t1 + xi * d1) - d0;
      acc += c[n] * t2;
      if (DERIV) accd += c[n] * d2;
      t0 = t1;
      t1 = t2;
      d0 = d1;
      d1 = d2;
    }
    R[mu] = env * acc;
    if (DERIV) dR[mu] = denv * acc + env * dxi * accd;
  }
}

// The uniform case of eval_radial with the coefficient rows shared by all pairs: every lane
// reads the same row (`c`, constant address space: scalar loads, SGPR operands) and the pair
// enters through sigma_p and (a, s):
//   R_mu = sigma_p (r - rc)^2 sum_n c[mu][n] T_n(xi_p(r))   (src/force/sus2_sh.cu:3171-3220).
template <typename Real, int MR, bool DERIV, int NB>
__device__ __forceinline__ void eval_radial_shared(const __attribute__((address_space(4))) Real* c,
                                                   Real sigma, Real a, Real s, Real rc, int kind,
                                                   Real r, Real (&R)[MR], Real (&dR)[MR])
{
  const Real dr = r - rc;
  const Real env = sigma * dr * dr;
  const Real denv = Real(2) * sigma * dr;
  Real xi = 0, dxi = 0;
  radial_map<Real>(kind, a, s, r, xi, dxi);
  Real T[NB], dT[NB];
  T[0] = Real(1);
  dT[0] = Real(0);
  T[1] = xi;
  dT[1] = Real(1);
#pragma unroll
  for (int n = 2; n < NB; ++n) {
    T[n] = Real(2) * xi * T[n - 1] - T[n - 2];
    if (DERIV) dT[n] = Real(2) * (T[n - 1] + xi * dT[n - 1]) - dT[n - 2];
  }
#pragma unroll
  for (int mu = 0; mu < MR; ++mu) {
    Real acc = 0, accd = 0;
#pragma unroll
    for (int n = 0; n < NB; ++n) {
      const Real cn = c[mu * NB + n];
      acc += cn * T[n];
      if (DERIV) accd += cn * dT[n];
    }
    R[mu] = env * acc;
    if (DERIV) dR[mu] = denv * acc + env * dxi * accd;
  }
}

constexpr __host__ __device__ int sh_index(int l, int m) { return l * l + m + l; }

// Racah-normalised real regular solid harmonics S_lm at a unit vector u (values only):
//   S_{l+1,l+1}  = c_l (x S_ll - y S_l,-l),  S_{l+1,-l-1} = c_l (y S_ll + x S_l,-l)
//   S_{l+1,m}    = ((2l+1) z S_lm - sqrt((l+m)(l-m)) r^2 S_{l-1,m}) / sqrt((l+m+1)(l-m+1))
// with c_l = sqrt(2^{delta_l0} (2l+1)/(2l+2)); Y_lm = sqrt((2l+1)/4pi) S_lm(u).  This is the
// same basis (no Condon-Shortley phase) as the explicit table of src/force/sus2_sh.cu:3058-3168.
template <typename Real, int L>
__device__ __forceinline__ void solid_values(Real x, Real y, Real z, Real (&S)[(L + 1) * (L + 1)])
{
  S[0] = Real(1);
  if constexpr (L >= 1) {
    S[sh_index(1, -1)] = y;
    S[sh_index(1, 0)] = z;
    S[sh_index(1, 1)] = x;
  }
#pragma unroll
  for (int l = 1; l < L; ++l) {
    const Real c = Real(__builtin_sqrt((2.0 * l + 1.0) / (2.0 * l + 2.0)));
    const Real sll = S[sh_index(l, l)], slm = S[sh_index(l, -l)];
    S[sh_index(l + 1, l + 1)] = c * (x * sll - y * slm);
    S[sh_index(l + 1, -l - 1)] = c * (y * sll + x * slm);
#pragma unroll
    for (int m = -l; m <= l; ++m) {
      const double den = __builtin_sqrt(double((l + m + 1) * (l - m + 1)));
      const Real A = Real((2.0 * l + 1.0) / den);
      Real v = A * z * S[sh_index(l, m)];
      if (m > -l && m < l) {
        const Real Bc = Real(__builtin_sqrt(double((l + m) * (l - m))) / den);
        v -= Bc * S[sh_index(l - 1, m)];  // r^2 = 1 at the unit vector
      }
      S[sh_index(l + 1, m)] = v;
    }
  }
}

// Values and polynomial gradients dS_lm/d(x,y,z) at the unit vector u.
template <typename Real, int L>
__device__ __forceinline__ void solid_values_grads(Real x, Real y, Real z,
                                                   Real (&S)[(L + 1) * (L + 1)],
                                                   Real (&Gx)[(L + 1) * (L + 1)],
                                                   Real (&Gy)[(L + 1) * (L + 1)],
                                                   Real (&Gz)[(L + 1) * (L + 1)])
{
  S[0] = Real(1);
  Gx[0] = Gy[0] = Gz[0] = Real(0);
  if constexpr (L >= 1) {
    S[sh_index(1, -1)] = y;
    S[sh_index(1, 0)] = z;
    S[sh_index(1, 1)] = x;
    Gx[sh_index(1, -1)] = 0;
    Gy[sh_index(1, -1)] = 1;
    Gz[sh_index(1, -1)] = 0;
    Gx[sh_index(1, 0)] = 0;
    Gy[sh_index(1, 0)] = 0;
    Gz[sh_index(1, 0)] = 1;
    Gx[sh_index(1, 1)] = 1;
    Gy[sh_index(1, 1)] = 0;
    Gz[sh_index(1, 1)] = 0;
  }
#pragma unroll
  for (int l = 1; l < L; ++l) {
    const Real c = Real(__builtin_sqrt((2.0 * l + 1.0) / (2.0 * l + 2.0)));
    const int ip = sh_index(l, l), im = sh_index(l, -l);
    const Real sp = S[ip], sm = S[im];
    const int op = sh_index(l + 1, l + 1), om = sh_index(l + 1, -l - 1);
    S[op] = c * (x * sp - y * sm);
    Gx[op] = c * (sp + x * Gx[ip] - y * Gx[im]);
    Gy[op] = c * (x * Gy[ip] - sm - y * Gy[im]);
    Gz[op] = c * (x * Gz[ip] - y * Gz[im]);
    S[om] = c * (y * sp + x * sm);
    Gx[om] = c * (y * Gx[ip] + sm + x * Gx[im]);
    Gy[om] = c * (sp + y * Gy[ip] + x * Gy[im]);
    Gz[om] = c * (y * Gz[ip] + x * Gz[im]);
#pragma unroll
    for (int m = -l; m <= l; ++m) {
      const double den = __builtin_sqrt(double((l + m + 1) * (l - m + 1)));
      const Real A = Real((2.0 * l + 1.0) / den);
      const int i0 = sh_index(l, m), o = sh_index(l + 1, m);
      Real v = A * z * S[i0];
      Real gx = A * z * Gx[i0];
      Real gy = A * z * Gy[i0];
      Real gz = A * (S[i0] + z * Gz[i0]);
      if (m > -l && m < l) {
        const Real Bc = Real(__builtin_sqrt(double((l + m) * (l - m))) / den);
        const int i1 = sh_index(l - 1, m);
        const Real s1 = S[i1];
        v -= Bc * s1;
        gx -= Bc * (Real(2) * x * s1 + Gx[i1]);
        gy -= Bc * (Real(2) * y * s1 + Gy[i1]);
        gz -= Bc * (Real(2) * z * s1 + Gz[i1]);
      }
      S[o] = v;
      Gx[o] = gx;
      Gy[o] = gy;
      Gz[o] = gz;
    }
  }
}

// w grad S_lm added to (gx, gy, gz) from the values of degree l - 1: the gradient of a
// Racah-normalised real regular solid harmonic is an exact combination of at most two
// harmonics of the next lower degree (with A = sqrt((l+|m|)(l+|m|-1)) / 2,
// B = sqrt((l-|m|)(l-|m|-1)) / 2):
//   dz S_lm = sqrt((l+m)(l-m)) S_{l-1,m}
//   m = 0:  dx S_l0 = -sqrt(l(l-1)/2) S_{l-1,1},  dy S_l0 = -sqrt(l(l-1)/2) S_{l-1,-1}
//   m > 0:  dx S_lm = A S_{l-1,m-1} - B S_{l-1,m+1},  dy S_lm = -A S_{l-1,-(m-1)} - B S_{l-1,-(m+1)}
//   m < 0:  dx S_lm = A S_{l-1,m+1} - B S_{l-1,m-1},  dy S_lm =  A S_{l-1,-m-1} + B S_{l-1,-m+1}
// where a term whose harmonic does not exist (|m'| > l - 1) is absent, and a term on
// S_{l-1,0} reached from |m| = 1 carries A sqrt(2) in dx (m = 1) or dy (m = -1) only.
// Checked against central differences of solid_values for l <= 6.
template <typename Real, int L>
__device__ __forceinline__ void add_solid_grad(int l, int m, Real w,
                                               const Real (&S)[(L + 1) * (L + 1)], Real& gx,
                                               Real& gy, Real& gz)
{
  if (l == 0) return;
  const int mu = m < 0 ? -m : m;
  if (mu <= l - 1) gz += w * Real(__builtin_sqrt(double((l + m) * (l - m)))) * S[sh_index(l - 1, m)];
  if (m == 0) {
    if (l >= 2) {
      const Real c = Real(-__builtin_sqrt(double(l * (l - 1)) / 2.0));
      gx += w * c * S[sh_index(l - 1, 1)];
      gy += w * c * S[sh_index(l - 1, -1)];
    }
    return;
  }
  const double A = 0.5 * __builtin_sqrt(double((l + mu) * (l + mu - 1)));
  const double B = l - mu - 1 >= 0 ? 0.5 * __builtin_sqrt(double((l - mu) * (l - mu - 1))) : 0.0;
  if (m > 0) {
    if (mu == 1) {
      gx += w * Real(A * 1.4142135623730951) * S[sh_index(l - 1, 0)];
    } else {
      gx += w * Real(A) * S[sh_index(l - 1, mu - 1)];
      gy -= w * Real(A) * S[sh_index(l - 1, -(mu - 1))];
    }
    if (mu + 1 <= l - 1) {
      gx -= w * Real(B) * S[sh_index(l - 1, mu + 1)];
      gy -= w * Real(B) * S[sh_index(l - 1, -(mu + 1))];
    }
  } else {
    if (mu == 1) {
      gy += w * Real(A * 1.4142135623730951) * S[sh_index(l - 1, 0)];
    } else {
      gx += w * Real(A) * S[sh_index(l - 1, -(mu - 1))];
      gy += w * Real(A) * S[sh_index(l - 1, mu - 1)];
    }
    if (mu + 1 <= l - 1) {
      gx -= w * Real(B) * S[sh_index(l - 1, -(mu + 1))];
      gy += w * Real(B) * S[sh_index(l - 1, mu + 1)];
    }
  }
}

template <int L>
__device__ __forceinline__ constexpr double ylm_norm()
{
  return __builtin_sqrt((2.0 * L + 1.0) / (4.0 * 3.14159265358979323846264338327950288));
}

__device__ __forceinline__ double norm_of_l(int l)
{
  return __builtin_sqrt((2.0 * l + 1.0) / (4.0 * 3.14159265358979323846264338327950288));
}

// Full (k, l, m) grid index in the standard basic order: l = L..0, k = K-1..0, m = -l..l
// (src/force/sus2_sh.cu:1451-1465, 2519-2551).
template <int L, int K>
__host__ __device__ constexpr int full_index(int l, int k, int m)
{
  int off = 0;
  for (int lp = L; lp > l; --lp) off += K * (2 * lp + 1);
  return off + (K - 1 - k) * (2 * l + 1) + (m + l);
}

}  // namespace sus2sh
