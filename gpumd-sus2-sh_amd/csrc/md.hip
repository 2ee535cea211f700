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

// Reductions for the MD drivers in two passes with a fixed order of additions (a butterfly
// within each wave, the waves in index order, the blocks in index order), so that energy
// traces -- and the NPT barostat driven by them -- are bit-identical between runs.
// Pass 1: block b writes its NV sums to partial[b * NV + k].
template <int NV>
__device__ __forceinline__ void block_partials(double (&q)[NV], double* __restrict__ partial)
{
  __shared__ double part[4][NV];
#pragma unroll
  for (int k = 0; k < NV; ++k)
    for (int off = 32; off > 0; off >>= 1) q[k] += __shfl_xor(q[k], off, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) part[threadIdx.x >> 6][k] = q[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    double sum = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x) / 64; ++w) sum += part[w][threadIdx.x];
    partial[static_cast<size_t>(blockIdx.x) * NV + threadIdx.x] = sum;
  }
}

// (potential, kinetic) of each atom
__global__ void k_energy_partials(int N, const int* __restrict__ type,
                                  const double* __restrict__ mass, const double* __restrict__ pot,
                                  const double* __restrict__ v, double* __restrict__ partial)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double q[2] = {0.0, 0.0};
  if (i < N) {
    q[0] = pot[i];
    const double vv = v[i] * v[i] + v[N + i] * v[N + i] + v[2 * N + i] * v[2 * N + i];
    q[1] = 0.5 * mass[type[i]] * vv * kKinetic;
  }
  block_partials<2>(q, partial);
}

// NPT: (m v_x^2, m v_y^2, m v_z^2 in eV, potential, virial xx, yy, zz) of each atom
__global__ void k_npt_partials(int N, const int* __restrict__ type, const double* __restrict__ mass,
                               const double* __restrict__ pot, const double* __restrict__ v,
                               const double* __restrict__ vir, double* __restrict__ partial)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double q[7] = {0, 0, 0, 0, 0, 0, 0};
  if (i < N) {
    const double m = mass[type[i]] * kKinetic;
    for (int a = 0; a < 3; ++a) q[a] = m * v[static_cast<size_t>(a) * N + i] * v[static_cast<size_t>(a) * N + i];
    q[3] = pot[i];
    for (int a = 0; a < 3; ++a) q[4 + a] = vir[static_cast<size_t>(a) * N + i];
  }
  block_partials<7>(q, partial);
}

// Pass 2 (one block of 256 threads): out[k] = sum over blocks of partial[b * NV + k]; with a
// device step counter the NV values go to out + NV * *step.
template <int NV>
__global__ __launch_bounds__(256) void k_finish(int nblocks, const double* __restrict__ partial,
                                                double* __restrict__ out,
                                                const int* __restrict__ step)
{
  double q[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) q[k] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < NV; ++k) q[k] += partial[static_cast<size_t>(b) * NV + k];
  __shared__ double part[4][NV];
#pragma unroll
  for (int k = 0; k < NV; ++k)
    for (int off = 32; off > 0; off >>= 1) q[k] += __shfl_xor(q[k], off, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) part[threadIdx.x >> 6][k] = q[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    double sum = 0.0;
    for (int w = 0; w < 4; ++w) sum += part[w][threadIdx.x];
    out[(step ? static_cast<size_t>(*step) * NV : 0) + threadIdx.x] = sum;
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

size_t md_partials_size(int N) { return 8 * static_cast<size_t>(nblocks(N)); }

void launch_energies(int N, const int* type, const double* mass, const double* pot,
                     const double* v, double* out, const int* step, double* partial,
                     hipStream_t s)
{
  k_energy_partials<<<nblocks(N), 256, 0, s>>>(N, type, mass, pot, v, partial);
  k_finish<2><<<1, 256, 0, s>>>(nblocks(N), partial, out, step);
}

void launch_step_advance(int* step, hipStream_t s) { k_step_advance<<<1, 64, 0, s>>>(step); }

void launch_npt_reduce(int N, const int* type, const double* mass, const double* pot,
                       const double* v, const double* vir, double* out, double* partial,
                       hipStream_t s)
{
  k_npt_partials<<<nblocks(N), 256, 0, s>>>(N, type, mass, pot, v, vir, partial);
  k_finish<7><<<1, 256, 0, s>>>(nblocks(N), partial, out, nullptr);
}

void launch_scale3(int N, double s0, double s1, double s2, double* x, hipStream_t s)
{
  k_scale3<<<nblocks(N), 256, 0, s>>>(N, s0, s1, s2, x);
}

}  // namespace sus2sh
