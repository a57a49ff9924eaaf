// VALU throughput of the instruction kinds the direct-sum kernels are made of, on the device it
// runs on: every lane runs 8 independent chains of one operation for a fixed number of steps
// (the kernels hold 8 cells or K targets per lane the same way), and every result is stored.
// Prints, per kind, the time of 5 launches and the rate in instructions and flop per second.
// Then the cost of the smallest FP32 matrix instruction, v_mfma_f32_4x4x1_16b_f32 (16 blocks of
// 4x4 outer products: 256 FMAs per wave), alone and between FP32 FMA streams, in cycles per
// wave-step of one SIMD: whether the matrix core could take the direct sum's a += w d updates
// off the vector ALU.
#include <hip/hip_runtime.h>
#include <stdio.h>

typedef float v4f __attribute__((ext_vector_type(4)));

#define CHK(x)                                                                     \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);        \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

enum { FMA64, MUL64, ADD64, FMA32, MUL32, ADD32, CVT, RSQ32, RSQ64, NOPS };
static const char* kName[NOPS] = {"v_fma_f64", "v_mul_f64", "v_add_f64", "v_fma_f32", "v_mul_f32",
                                  "v_add_f32", "v_cvt_f32_f64 + v_cvt_f64_f32", "v_rsq_f32",
                                  "v_rsq_f64"};
static const double kFlop[NOPS] = {2, 1, 1, 2, 1, 1, 0, 1, 1};
static const int kInstr[NOPS] = {1, 1, 1, 1, 1, 1, 2, 1, 1};

template <int OP>
__global__ __launch_bounds__(256) void k_chain(double* out, int steps, double seed) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double a[8];
  float f[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    a[i] = seed + t * 1e-9 + i;
    f[i] = (float)a[i];
  }
  for (int s = 0; s < steps; ++s) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (OP == FMA64) a[i] = fma(a[i], 0.999999, 1e-7);
      if (OP == MUL64) a[i] = a[i] * 0.999999;
      if (OP == ADD64) a[i] = a[i] + 1e-7;
      if (OP == FMA32) f[i] = fmaf(f[i], 0.999999f, 1e-7f);
      if (OP == MUL32) f[i] = f[i] * 0.999999f;
      if (OP == ADD32) f[i] = f[i] + 1e-7f;
      if (OP == CVT) a[i] = (double)(float)a[i];
      if (OP == RSQ32) f[i] = __builtin_amdgcn_rsqf(f[i]);
      if (OP == RSQ64) a[i] = __builtin_amdgcn_rsq(a[i]);
    }
  }
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 8; ++i) s += a[i] + (double)f[i];
  out[t] = s;
}

template <int OP>
static int run(double* d, int nblk, int steps, int ncu) {
  hipEvent_t e0, e1;
  CHK(hipEventCreate(&e0));
  CHK(hipEventCreate(&e1));
  for (int w = 0; w < 3; ++w) k_chain<OP><<<nblk, 256>>>(d, steps, 1.0);
  CHK(hipGetLastError());
  CHK(hipDeviceSynchronize());
  CHK(hipEventRecord(e0));
  for (int r = 0; r < 5; ++r) k_chain<OP><<<nblk, 256>>>(d, steps, 1.0);
  CHK(hipEventRecord(e1));
  CHK(hipEventSynchronize(e1));
  float ms = 0.f;
  CHK(hipEventElapsedTime(&ms, e0, e1));
  const double lane_ops = 5.0 * nblk * 256.0 * steps * 8.0;            // chain steps, all lanes
  const double instr = lane_ops * kInstr[OP] / 64.0;                    // wave instructions
  const double per_simd_cycle_ghz = instr / (4.0 * ncu) / (ms * 1e-3) / 1e9;
  printf("%-32s %9.3f ms  %8.2f G wave-instr/s per SIMD  %8.2f TFLOP/s\n", kName[OP], ms,
         per_simd_cycle_ghz, lane_ops * kFlop[OP] / (ms * 1e-3) / 1e12);
  CHK(hipEventDestroy(e0));
  CHK(hipEventDestroy(e1));
  return 0;
}

// NF FP32 FMAs (8 chains) then NM 4x4x1 matrix instructions (8 accumulators) per step.
template <int NF, int NM>
__global__ __launch_bounds__(256) void k_mfma_mix(float* out, int steps, float seed) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const float b = 1e-3f * (float)(threadIdx.x & 63);
  float f[8];
  v4f acc[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    f[i] = seed + t * 1e-9f + i;
    acc[i] = (v4f){0.f, 0.f, 0.f, 0.f};
  }
  for (int s = 0; s < steps; ++s) {
#pragma unroll
    for (int r = 0; r < NF / 8; ++r)
#pragma unroll
      for (int i = 0; i < 8; ++i) f[i] = fmaf(f[i], 0.999999f, 1e-7f);
#pragma unroll
    for (int i = 0; i < NM; ++i) acc[i] = __builtin_amdgcn_mfma_f32_4x4x1f32(f[i], b, acc[i], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) s += f[i] + acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[t] = s;
}

template <int NF, int NM>
static int run_mfma(float* d, int nblk, int steps, int ncu) {
  hipEvent_t e0, e1;
  CHK(hipEventCreate(&e0));
  CHK(hipEventCreate(&e1));
  for (int w = 0; w < 3; ++w) k_mfma_mix<NF, NM><<<nblk, 256>>>(d, steps, 1.0f);
  CHK(hipGetLastError());
  CHK(hipDeviceSynchronize());
  CHK(hipEventRecord(e0));
  for (int r = 0; r < 5; ++r) k_mfma_mix<NF, NM><<<nblk, 256>>>(d, steps, 1.0f);
  CHK(hipEventRecord(e1));
  CHK(hipEventSynchronize(e1));
  float ms = 0.f;
  CHK(hipEventElapsedTime(&ms, e0, e1));
  const double wave_steps_per_simd = 5.0 * nblk * 4.0 * steps / (4.0 * ncu);
  const double ns = ms * 1e6 / wave_steps_per_simd;
  printf("%2d v_fma_f32 + %d v_mfma_f32_4x4x1_16b_f32 per step %9.3f ms  %7.2f ns = %6.1f cycles at 2.4 GHz\n",
         NF, NM, ms, ns, ns * 2.4);
  CHK(hipEventDestroy(e0));
  CHK(hipEventDestroy(e1));
  return 0;
}

int main() {
  int dev = 0, ncu = 0, clk = 0;
  CHK(hipGetDevice(&dev));
  CHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  CHK(hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, dev));
  const int nblk = ncu * 16, steps = 4096;
  printf("%d CUs, reported peak clock %.3f GHz; %d workgroups of 256 lanes, 8 chains x %d steps per lane\n",
         ncu, clk * 1e-6, nblk, steps);
  double* d = nullptr;
  CHK(hipMalloc(&d, sizeof(double) * (size_t)nblk * 256));
  if (run<FMA64>(d, nblk, steps, ncu) || run<MUL64>(d, nblk, steps, ncu) || run<ADD64>(d, nblk, steps, ncu) ||
      run<FMA32>(d, nblk, steps, ncu) || run<MUL32>(d, nblk, steps, ncu) || run<ADD32>(d, nblk, steps, ncu) ||
      run<CVT>(d, nblk, steps, ncu) || run<RSQ32>(d, nblk, steps, ncu) || run<RSQ64>(d, nblk, steps, ncu))
    return 1;
  printf("per wave-step of one SIMD (8 chains / accumulators per lane, %d steps):\n", steps / 2);
  float* f = reinterpret_cast<float*>(d);
  if (run_mfma<8, 0>(f, nblk, steps / 2, ncu) || run_mfma<0, 8>(f, nblk, steps / 2, ncu) ||
      run_mfma<8, 8>(f, nblk, steps / 2, ncu) || run_mfma<48, 0>(f, nblk, steps / 2, ncu) ||
      run_mfma<48, 8>(f, nblk, steps / 2, ncu))
    return 1;
  CHK(hipFree(d));
  return 0;
}
