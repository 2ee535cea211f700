// ZBL short-range pair term over the neighbour list, for models with zbl_enabled
// (reference kernel: gpu_sh_apply_zbl, src/force/sus2_sh.cu:7522-7627; pair function
// src/force/sus2_zbl_common.cuh:171-225).  One lane per cell-sorted atom i; every edge
// i -> j inside the pair's outer cutoff adds phi(r) / 2 to E_i, and its gradient
// g = (phi'(r) / 2) d / r is applied as F_i += g, F_j -= g, W_i -= d (x) g -- the same edge
// convention as the many-body force kernel.  Pairs this close are rare in MD, so the kernel
// is a cheap pass over the list; its arithmetic is FP64 in both precisions.
#include "kernels.hpp"

namespace sus2sh {

namespace {

template <typename Real>
__global__ __launch_bounds__(256) void k_zbl(int N, const int* __restrict__ count,
                                             const Edge<Real>* __restrict__ nl,
                                             const int* __restrict__ type_s, ZblTable t,
                                             double* __restrict__ e_s, double* __restrict__ f_s,
                                             double* __restrict__ v_s, int* __restrict__ fix_flag)
{
  unsigned long long* f_q = reinterpret_cast<unsigned long long*>(f_s);  // fix_flag != null
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int ti = type_s[p];
  const int zi = t.z[ti];
  const int n = count[p];
  double e = 0.0, fs[3] = {0.0, 0.0, 0.0}, w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool any = false;
  for (int c = 0; c < n; ++c) {
    const Edge<Real> ed = nl[static_cast<size_t>(c) * N + p];
    const int tj = ed.j >> kEdgeTypeShift;
    const int pair = ti * t.S + tj;
    const double outer = t.outer[pair];
    const double d[3] = {static_cast<double>(ed.dx), static_cast<double>(ed.dy),
                         static_cast<double>(ed.dz)};
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 <= 0.0 || r2 >= outer * outer) continue;
    const double r = sqrt(r2);
    const int zj = t.z[tj];
    const double a_inv = 2.134563 * (pow(static_cast<double>(zi), 0.23) +
                                     pow(static_cast<double>(zj), 0.23));
    const double amp[4] = {0.18175, 0.50986, 0.28022, 0.02817};
    const double dec[4] = {3.1998, 0.94229, 0.4029, 0.20162};
    double phi = 0.0, dphi = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double x = amp[k] * exp(-dec[k] * a_inv * r);
      phi += x;
      dphi -= dec[k] * a_inv * x;
    }
    const double q = 14.3996454784255 * zi * zj;
    const double base = q * phi / r, dbase = q * (dphi / r - phi / r2);
    double sw = 1.0, dsw = 0.0;
    const double inner = t.inner[pair];
    if (r > inner) {
      const double k = 3.141592653589793 / (outer - inner);
      sw = 0.5 + 0.5 * cos(k * (r - inner));
      dsw = -0.5 * k * sin(k * (r - inner));
    }
    const double pe = sw * base, dedr = sw * dbase + dsw * base;
    const double s = 0.5 * dedr / r;
    const double g[3] = {s * d[0], s * d[1], s * d[2]};
    e += 0.5 * pe;
    const int j = ed.j & kEdgeIndexMask;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      fs[a] += g[a];
      if (fix_flag)
        atomicAdd(f_q + static_cast<size_t>(a) * N + j, fix_encode(-g[a], kFixPairMax, fix_flag));
      else
        unsafeAtomicAdd(f_s + static_cast<size_t>(a) * N + j, -g[a]);
    }
    // xx yy zz xy xz yz yx zx zy
    w[0] -= d[0] * g[0];
    w[1] -= d[1] * g[1];
    w[2] -= d[2] * g[2];
    w[3] -= d[0] * g[1];
    w[4] -= d[0] * g[2];
    w[5] -= d[1] * g[2];
    w[6] -= d[1] * g[0];
    w[7] -= d[2] * g[0];
    w[8] -= d[2] * g[1];
    any = true;
  }
  if (!any) return;
  e_s[p] += e;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (fix_flag)
      atomicAdd(f_q + static_cast<size_t>(a) * N + p, fix_encode(fs[a], kFixSumMax, fix_flag));
    else
      unsafeAtomicAdd(f_s + static_cast<size_t>(a) * N + p, fs[a]);
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) v_s[static_cast<size_t>(k) * N + p] += w[k];
}

}  // namespace

template <typename Real>
void launch_zbl(int N, const int* count, const Edge<Real>* nl, const int* type_s,
                const ZblTable& t, double* e_s, double* f_s, double* v_s, int* fix_flag,
                hipStream_t s)
{
  if (N <= 0) return;
  k_zbl<Real><<<(N + 255) / 256, 256, 0, s>>>(N, count, nl, type_s, t, e_s, f_s, v_s, fix_flag);
}

template void launch_zbl<float>(int, const int*, const Edge<float>*, const int*, const ZblTable&,
                                double*, double*, double*, int*, hipStream_t);
template void launch_zbl<double>(int, const int*, const Edge<double>*, const int*,
                                 const ZblTable&, double*, double*, double*, int*, hipStream_t);

}  // namespace sus2sh
