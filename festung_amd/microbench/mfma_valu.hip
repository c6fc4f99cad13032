// Issue model of a gfx950 SIMD for FP64 matrix work: can VALU instructions
// run while v_mfma_f64_16x16x4f64 instructions execute?
//
// Each wave runs n iterations of (a) four independent MFMA accumulation
// chains (operands in registers), (b) 16 independent VALU chains of one
// instruction kind (FP64 FMA, FP32 FMA or 32-bit integer multiply-add), or
// (c) both, interleaved one MFMA to four VALU instructions by
// sched_group_barrier.  If the two kinds of work overlapped, (c) would take
// max((a), (b)); it takes (a) + (b).  1024 waves put one wave on each SIMD,
// 2048 two.
//
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -mllvm -amdgpu-mfma-vgpr-form=1 festung_amd/microbench/mfma_valu.hip -o mfma_valu
//   ./mfma_valu
#include <hip/hip_runtime.h>

#include <cstdio>

typedef double dbl4 __attribute__((ext_vector_type(4)));

#define CK(x)                                                              \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                            \
    }                                                                      \
  } while (0)

enum Kind { MFMA = 1, F64 = 2, F32 = 4, INT = 8 };

template <int KIND>
__global__ __launch_bounds__(256) void issue(double* out, int n) {
  const double x = threadIdx.x * 1e-3, y = 1.0 + threadIdx.x * 1e-4;
  dbl4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
  double v[16];
  float w[16];
  unsigned u[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    v[i] = i * 1e-3 + x;
    w[i] = float(v[i]);
    u[i] = threadIdx.x * 7u + i;
  }
  for (int it = 0; it < n; ++it) {
    if (KIND & MFMA) {
      c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(y, x, c1, 0, 0, 0);
      c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(x, x, c2, 0, 0, 0);
      c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(y, y, c3, 0, 0, 0);
    }
    if (KIND & F64) {
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = __builtin_fma(v[i], y, x);
    }
    if (KIND & F32) {
#pragma unroll
      for (int i = 0; i < 16; ++i) w[i] = __builtin_fmaf(w[i], 1.0001f, 0.5f);
    }
    if (KIND & INT) {
#pragma unroll
      for (int i = 0; i < 16; ++i) u[i] = u[i] * 3u + (u[i] >> 3);
    }
    if ((KIND & MFMA) && KIND != MFMA) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
        __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);  // four VALU
      }
    }
  }
  const dbl4 r = c0 + c1 + c2 + c3;
  double s = r[0] + r[1] + r[2] + r[3];
#pragma unroll
  for (int i = 0; i < 16; ++i) s += v[i] + w[i] + double(u[i]);
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

template <int KIND>
int run(const char* name, double* o, int waves, int n, hipEvent_t e0, hipEvent_t e1) {
  const int blocks = waves / 4;
  issue<KIND><<<blocks, 256>>>(o, n);
  CK(hipGetLastError());
  CK(hipDeviceSynchronize());
  const int R = 10;
  CK(hipEventRecord(e0));
  for (int r = 0; r < R; ++r) issue<KIND><<<blocks, 256>>>(o, n);
  CK(hipEventRecord(e1));
  CK(hipEventSynchronize(e1));
  float ms;
  CK(hipEventElapsedTime(&ms, e0, e1));
  const double us = 1000.0 * ms / R;
  const double mf = (KIND & MFMA) ? 4.0 * n * waves * 2048.0 : 0.0;
  printf("%5d waves  %-12s %8.2f us   MFMA %5.1f TFLOP/s\n", waves, name, us, mf / us * 1e-6);
  return 0;
}

int main() {
  double* o = nullptr;
  CK(hipMalloc(&o, size_t(2048) * 64 * sizeof(double)));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  for (int waves : {1024, 2048}) {
    const int n = 200;
    if (run<MFMA>("mfma", o, waves, n, e0, e1) || run<F64>("f64", o, waves, n, e0, e1) ||
        run<MFMA | F64>("mfma+f64", o, waves, n, e0, e1) || run<F32>("f32", o, waves, n, e0, e1) ||
        run<MFMA | F32>("mfma+f32", o, waves, n, e0, e1) || run<INT>("int", o, waves, n, e0, e1) ||
        run<MFMA | INT>("mfma+int", o, waves, n, e0, e1))
      return 1;
  }
  CK(hipFree(o));
  printf("done\n");
  return 0;
}
