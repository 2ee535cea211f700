// Issue-rate calibration for the roofline notes: sustained cycles per wave64 instruction of
// a few VALU kinds and of LDS reads on one MI355X, at 16 waves per SIMD with eight
// independent chains per lane (no dependency stalls).  Prints, per kind, the kernel time and
// (SIMD cycles at 2.4 GHz) / (wave-instructions per SIMD).
//
//   make -C gpumd-sus2-sh_amd valu_rate && gpumd-sus2-sh_amd/build_obj/valu_rate
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int kBlocks = 4096, kThreads = 256, kIters = 2048, kChains = 8;

template <int KIND>
__global__ __launch_bounds__(kThreads) void k_rate(float* out, double* outd, int iters)
{
  __shared__ float lds[kThreads * kChains];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  float a[kChains], b[kChains];
  double d[kChains];
  uint32_t u[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) {
    a[i] = 1.0f + 1e-7f * static_cast<float>(t + i);
    b[i] = 0.999f - 1e-8f * static_cast<float>(t ^ i);
    d[i] = 1.0 + 1e-9 * (t + i);
    u[i] = static_cast<uint32_t>(t) * 7u + i;
    lds[threadIdx.x + i * kThreads] = a[i];
  }
  __syncthreads();
  const float x = 0.999999f, y = 1e-7f;
  const double xd = 0.9999999, yd = 1e-9;
  for (int k = 0; k < iters; ++k) {
#pragma unroll
    for (int i = 0; i < kChains; ++i) {
      if (KIND == 0) a[i] = fmaf(a[i], x, y);           // v_fmamk_f32 (literal operand)
      if (KIND == 1) a[i] = fmaf(b[i], a[i], a[i]);      // v_fmac_f32, VGPR operands
      if (KIND == 2) a[i] = a[i] * x;                    // v_mul_f32
      if (KIND == 3) a[i] = a[i] + b[i];                 // v_add_f32
      if (KIND == 4) d[i] = fma(d[i], xd, yd);           // v_fma_f64
      if (KIND == 5) u[i] = u[i] + 0x9e3779b9u;          // v_add_u32
      if (KIND == 6)                                     // ds_read_b32, lane-strided rows
        a[i] += lds[threadIdx.x + ((i + k) & (kChains - 1)) * kThreads];
    }
  }
  float s = 0.0f;
  double sd = 0.0;
#pragma unroll
  for (int i = 0; i < kChains; ++i) {
    s += a[i] + static_cast<float>(u[i]);
    sd += d[i];
  }
  out[t] = s;
  outd[t] = sd;
}

template <int KIND>
float time_kind(float* o, double* od)
{
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1.0f;
  k_rate<KIND><<<kBlocks, kThreads>>>(o, od, kIters);
  (void)hipEventRecord(e0);
  for (int r = 0; r < 5; ++r) k_rate<KIND><<<kBlocks, kThreads>>>(o, od, kIters);
  (void)hipEventRecord(e1);
  (void)hipEventSynchronize(e1);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / 5;
}

}  // namespace

int main()
{
  float* o = nullptr;
  double* od = nullptr;
  if (hipMalloc(&o, sizeof(float) * kBlocks * kThreads) != hipSuccess ||
      hipMalloc(&od, sizeof(double) * kBlocks * kThreads) != hipSuccess) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  int simds = 0;
  (void)hipDeviceGetAttribute(&simds, hipDeviceAttributeMultiprocessorCount, 0);
  simds *= 4;
  const char* names[] = {"v_fmamk_f32 (literal)", "v_fmac_f32", "v_mul_f32", "v_add_f32",
                         "v_fma_f64", "v_add_u32", "ds_read_b32 + v_add_f32"};
  const float ms[] = {time_kind<0>(o, od), time_kind<1>(o, od), time_kind<2>(o, od),
                      time_kind<3>(o, od), time_kind<4>(o, od), time_kind<5>(o, od),
                      time_kind<6>(o, od)};
  const double per_simd = 4.0 * kBlocks * static_cast<double>(kChains) * kIters / simds;
  std::printf("# %d SIMDs, %d waves, %d loop operations per wave\n", simds, 4 * kBlocks,
              kChains * kIters);
  for (int k = 0; k < 7; ++k)
    std::printf("%-26s %8.3f ms %7.3f cycles per wave64 operation at 2.4 GHz\n", names[k], ms[k],
                ms[k] * 1e-3 * 2.4e9 / per_simd);
  const bool ok = hipDeviceSynchronize() == hipSuccess;
  (void)hipFree(o);
  (void)hipFree(od);
  return ok ? 0 : 1;
}
