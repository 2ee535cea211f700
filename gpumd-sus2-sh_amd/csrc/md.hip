// Velocity-Verlet (NVE) driver kernels used for the energy-drift configurations.
// Units: positions A, velocities A/fs, masses amu, energies eV, forces eV/A; an acceleration
// F/m in eV/(A amu) is 9.648533212e-3 A/fs^2.  GPUMD's own integrator is not part of the
// reference overlay; this is the minimal NVE step used around the potential.
#include "kernels.hpp"

namespace sus2sh {

namespace {

constexpr double kAccel = 9.648533212331e-3;  // (eV/A)/amu -> A/fs^2
constexpr double kKinetic = 1.0 / kAccel;      // amu A^2/fs^2 -> eV

__global__ void k_half_kick(int N, double half_dt, const int* __restrict__ type,
                            const double* __restrict__ mass, const double* __restrict__ f,
                            double* __restrict__ v)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double s = half_dt * kAccel / mass[type[i]];
  v[i] += s * f[i];
  v[N + i] += s * f[N + i];
  v[2 * N + i] += s * f[2 * N + i];
}

__global__ void k_drift(int N, double dt, const double* __restrict__ v, double* __restrict__ x)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  x[i] += dt * v[i];
  x[N + i] += dt * v[N + i];
  x[2 * N + i] += dt * v[2 * N + i];
}

// out[0] += sum potential, out[1] += sum kinetic (one atomic pair per workgroup); with a
// device step counter the pair is out[2 * step], out[2 * step + 1]
__global__ void k_energies(int N, const int* __restrict__ type, const double* __restrict__ mass,
                           const double* __restrict__ pot, const double* __restrict__ v,
                           double* __restrict__ out, const int* __restrict__ step)
{
  __shared__ double sp[4], sk[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double p = 0.0, k = 0.0;
  if (i < N) {
    p = pot[i];
    const double vv = v[i] * v[i] + v[N + i] * v[N + i] + v[2 * N + i] * v[2 * N + i];
    k = 0.5 * mass[type[i]] * vv * kKinetic;
  }
  for (int off = 32; off > 0; off >>= 1) {
    p += __shfl_xor(p, off, 64);
    k += __shfl_xor(k, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sp[threadIdx.x >> 6] = p;
    sk[threadIdx.x >> 6] = k;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (step) out += 2 * static_cast<size_t>(*step);
    double a = 0.0, b = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x) / 64; ++w) {
      a += sp[w];
      b += sk[w];
    }
    unsafeAtomicAdd(out, a);
    unsafeAtomicAdd(out + 1, b);
  }
}

// NPT: out[0..2] += sum m v_a^2 (eV, per Cartesian axis), out[3] += sum potential,
// out[4..6] += sum of the diagonal per-atom virials xx yy zz
__global__ void k_npt_reduce(int N, const int* __restrict__ type, const double* __restrict__ mass,
                             const double* __restrict__ pot, const double* __restrict__ v,
                             const double* __restrict__ vir, double* __restrict__ out)
{
  __shared__ double part[4][7];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double q[7] = {0, 0, 0, 0, 0, 0, 0};
  if (i < N) {
    const double m = mass[type[i]] * kKinetic;
    for (int a = 0; a < 3; ++a) q[a] = m * v[static_cast<size_t>(a) * N + i] * v[static_cast<size_t>(a) * N + i];
    q[3] = pot[i];
    for (int a = 0; a < 3; ++a) q[4 + a] = vir[static_cast<size_t>(a) * N + i];
  }
  for (int k = 0; k < 7; ++k)
    for (int off = 32; off > 0; off >>= 1) q[k] += __shfl_xor(q[k], off, 64);
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 7; ++k) part[threadIdx.x >> 6][k] = q[k];
  __syncthreads();
  if (threadIdx.x < 7) {
    double s = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x) / 64; ++w) s += part[w][threadIdx.x];
    unsafeAtomicAdd(out + threadIdx.x, s);
  }
}

// x_a *= s_a (one factor per Cartesian axis)
__global__ void k_scale3(int N, double s0, double s1, double s2, double* __restrict__ x)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  x[i] *= s0;
  x[N + i] *= s1;
  x[2 * N + i] *= s2;
}

__global__ void k_step_advance(int* step)
{
  if (threadIdx.x == 0) *step += 1;
}

inline int nblocks(int n) { return (n + 255) / 256; }

}  // namespace

void launch_half_kick(int N, double half_dt, const int* type, const double* mass, const double* f,
                      double* v, hipStream_t s)
{
  k_half_kick<<<nblocks(N), 256, 0, s>>>(N, half_dt, type, mass, f, v);
}

void launch_drift(int N, double dt, const double* v, double* x, hipStream_t s)
{
  k_drift<<<nblocks(N), 256, 0, s>>>(N, dt, v, x);
}

void launch_energies(int N, const int* type, const double* mass, const double* pot,
                     const double* v, double* out, const int* step, hipStream_t s)
{
  k_energies<<<nblocks(N), 256, 0, s>>>(N, type, mass, pot, v, out, step);
}

void launch_step_advance(int* step, hipStream_t s) { k_step_advance<<<1, 64, 0, s>>>(step); }

void launch_npt_reduce(int N, const int* type, const double* mass, const double* pot,
                       const double* v, const double* vir, double* out, hipStream_t s)
{
  k_npt_reduce<<<nblocks(N), 256, 0, s>>>(N, type, mass, pot, v, vir, out);
}

void launch_scale3(int N, double s0, double s1, double s2, double* x, hipStream_t s)
{
  k_scale3<<<nblocks(N), 256, 0, s>>>(N, s0, s1, s2, x);
}

}  // namespace sus2sh
