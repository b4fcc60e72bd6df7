// Throughput microbenchmarks for the instruction classes the form-factor kernels issue: FP32 FMA,
// FP64 FMA, the hardware sine (v_sin_f32), exp2 (v_exp_f32), the FP32 and fp16 matrix cores
// (v_mfma_f32_32x32x2_f32, v_mfma_f32_32x32x16_f16), and whether VALU work overlaps the matrix core. Each kernel runs independent dependency chains so that issue rate, not
// latency, bounds it. Prints one JSON line of rates in operations per second for the whole device.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kChains = 16;

__global__ void k_fma32(float* out, int iters, float x, float y) {
  float a[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) a[c] = threadIdx.x * 1e-3f + c;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) a[c] = __builtin_fmaf(a[c], x, y);
  }
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += a[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_fma64(double* out, int iters, double x, double y) {
  double a[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) a[c] = threadIdx.x * 1e-3 + c;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) a[c] = __builtin_fma(a[c], x, y);
  }
  double s = 0.;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += a[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_sin(float* out, int iters) {
  float a[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) a[c] = threadIdx.x * 1e-3f + 0.01f * c;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) a[c] = __builtin_amdgcn_sinf(a[c]);
  }
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += a[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_exp2(float* out, int iters) {
  float a[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) a[c] = -(threadIdx.x * 1e-3f + 0.01f * c);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) a[c] = -__builtin_amdgcn_exp2f(a[c]);
  }
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += a[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// four independent 32x32 accumulators per wave
__global__ void k_mfma32(float* out, int iters) {
  f32x16 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[c][r] = 0.f;
  float a = threadIdx.x * 1e-3f, b = 1.0f - threadIdx.x * 1e-4f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[c], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 16; ++r) s += acc[c][r];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// the fp16 matrix core (v_mfma_f32_32x32x16_f16), four independent accumulators per wave
__global__ void k_mfma16(float* out, int iters) {
  f32x16 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[c][r] = 0.f;
  f16x8 a, b;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    a[k] = (_Float16)(threadIdx.x * 1e-3f + k);
    b[k] = (_Float16)(1.0f - threadIdx.x * 1e-4f);
  }
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[c], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 16; ++r) s += acc[c][r];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Whether VALU work overlaps the matrix core: per iteration NM fp16 MFMAs (4 accumulators) and
// NV independent FP32 FMAs (16 chains), in one instruction stream; NM = 0 or NV = 0 times each part
// alone.
template <int NM, int NV>
__global__ void k_mix(float* out, int iters, float x, float y) {
  f32x16 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[c][r] = 0.f;
  f16x8 a, b;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    a[k] = (_Float16)(threadIdx.x * 1e-3f + k);
    b[k] = (_Float16)(1.0f - threadIdx.x * 1e-4f);
  }
  float v[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) v[c] = threadIdx.x * 1e-3f + c;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int m = 0; m < NM; ++m) acc[m & 3] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[m & 3], 0, 0, 0);
#pragma unroll
    for (int n = 0; n < NV; ++n) v[n % kChains] = __builtin_fmaf(v[n % kChains], x, y);
  }
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 16; ++r) s += acc[c][r];
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += v[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <typename F>
static double time_ms(F launch) {
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  launch();
  (void)hipDeviceSynchronize();
  (void)hipEventRecord(e0, 0);
  launch();
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

int main(int argc, char** argv) {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    std::fprintf(stderr, "no device\n");
    return 1;
  }
  const int cus = prop.multiProcessorCount;
  const int blocks = cus * 16, threads = 256;  // 16 waves per CU
  const int iters = argc > 1 ? std::atoi(argv[1]) : 4096;
  const double lanes = (double)blocks * threads;
  float* out = nullptr;
  double* out64 = nullptr;
  if (hipMalloc(&out, sizeof(float) * blocks * threads) != hipSuccess ||
      hipMalloc(&out64, sizeof(double) * blocks * threads) != hipSuccess) {
    std::fprintf(stderr, "alloc failed\n");
    return 1;
  }
  double t;
  t = time_ms([&] { k_fma32<<<blocks, threads>>>(out, iters, 0.999f, 1e-3f); });
  const double fma32 = lanes * iters * kChains / (t * 1e-3);
  t = time_ms([&] { k_fma64<<<blocks, threads>>>(out64, iters, 0.999, 1e-3); });
  const double fma64 = lanes * iters * kChains / (t * 1e-3);
  t = time_ms([&] { k_sin<<<blocks, threads>>>(out, iters); });
  const double sinr = lanes * iters * kChains / (t * 1e-3);
  t = time_ms([&] { k_exp2<<<blocks, threads>>>(out, iters); });
  const double expr = lanes * iters * kChains / (t * 1e-3);
  const int mblocks = cus * 8;  // 2 waves per SIMD
  t = time_ms([&] { k_mfma32<<<mblocks, threads>>>(out, iters); });
  const double mfma_flops = (double)mblocks * (threads / 64) * iters * 4 * (32.0 * 32 * 2 * 2) / (t * 1e-3);
  t = time_ms([&] { k_mfma16<<<mblocks, threads>>>(out, iters); });
  const double mfma16_flops = (double)mblocks * (threads / 64) * iters * 4 * (32.0 * 32 * 16 * 2) / (t * 1e-3);
  // overlap test at 1 and 2 waves per SIMD: 8 MFMAs (256 matrix-core cycles at full rate) and 64
  // FMAs (about the same issue time) per iteration, each alone and together
  const int mi = iters / 4;
  double mix[2][3];
  for (int w = 0; w < 2; ++w) {
    const int nb = cus * (w + 1);
    mix[w][0] = time_ms([&] { k_mix<8, 0><<<nb, threads>>>(out, mi, 0.999f, 1e-3f); });
    mix[w][1] = time_ms([&] { k_mix<0, 64><<<nb, threads>>>(out, mi, 0.999f, 1e-3f); });
    mix[w][2] = time_ms([&] { k_mix<8, 64><<<nb, threads>>>(out, mi, 0.999f, 1e-3f); });
  }
  if (hipDeviceSynchronize() != hipSuccess) {
    std::fprintf(stderr, "kernel failed\n");
    return 1;
  }
  std::printf(
      "{\"device\": \"%s\", \"cus\": %d, \"clock_mhz\": %d, \"fp32_fma_per_s\": %.4e, \"fp64_fma_per_s\": %.4e, "
      "\"v_sin_f32_per_s\": %.4e, \"v_exp_f32_per_s\": %.4e, \"mfma_f32_32x32x2_flop_per_s\": %.4e, "
      "\"mfma_f32_32x32x16_f16_flop_per_s\": %.4e, \"mix_1wave_ms\": {\"mfma\": %.3f, \"valu\": %.3f, \"both\": %.3f}, "
      "\"mix_2wave_ms\": {\"mfma\": %.3f, \"valu\": %.3f, \"both\": %.3f}}\n",
      prop.gcnArchName, cus, prop.clockRate / 1000, fma32, fma64, sinr, expr, mfma_flops, mfma16_flops, mix[0][0],
      mix[0][1], mix[0][2], mix[1][0], mix[1][1], mix[1][2]);
  (void)hipFree(out);
  (void)hipFree(out64);
  return 0;
}
