// VALU throughput of the instruction kinds the direct-sum kernels are made of, on the device it
// runs on: every lane runs 8 independent chains of one operation for a fixed number of steps
// (the kernels hold 8 cells or K targets per lane the same way), and every result is stored.
// Prints, per kind, the time of 5 launches and the rate in instructions and flop per second.
#include <hip/hip_runtime.h>
#include <stdio.h>

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
  CHK(hipFree(d));
  return 0;
}
