// Attainable HBM bandwidth of an MI355X for the access shapes of the DG
// step kernels: what a kernel that only moves the stage kernel's bytes can
// reach, as the ceiling its measured GB/s is compared with.
//
//   read       one pass over n doubles, summed per thread (8 or 16 B loads)
//   copy       y = x over n doubles (8 or 16 B per lane)
//   soa R/W    one thread per element k < K reads R fields x[f K + k] and
//              writes W fields y[f K + k] (y = sum of the reads): the
//              structure-of-arrays streams of the stage (R = 33, W = 6 at
//              p = 2 with the limiter), dc/dt limiter (R = 24, W = 6) and
//              solution limiter (R = 10, W = 6) kernels, without arithmetic
//              or neighbour gathers
//
// Every kernel is timed over 20 launches with HIP events after 3 warm-up
// launches; buffers larger than the 256 MB Infinity Cache are used, and the
// SoA shapes are also run at K = 524 288 (the p = 2 benchmark mesh, whose
// working set fits in that cache).
//
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         festung_amd/microbench/hbm_stream.hip -o hbm_stream
//   ./hbm_stream
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#define CK(x)                                                              \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                            \
    }                                                                      \
  } while (0)

typedef double dbl2 __attribute__((ext_vector_type(2)));

template <int V>
__global__ __launch_bounds__(256) void read_kernel(const double* __restrict__ x, size_t n, double* __restrict__ out) {
  // n is a multiple of V * gridDim.x * blockDim.x (checked on the host)
  const size_t stride = size_t(gridDim.x) * blockDim.x * V;
  double s = 0.0;
  for (size_t i = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) * V; i < n; i += stride) {
    if constexpr (V == 2) {
      const dbl2 v = *reinterpret_cast<const dbl2*>(x + i);
      s += v.x + v.y;
    } else {
      s += x[i];
    }
  }
  if (s == 12345.678) out[0] = s;  // keeps the loads; never true for the fill value
}

template <int V>
__global__ __launch_bounds__(256) void copy_kernel(const double* __restrict__ x, double* __restrict__ y, size_t n) {
  const size_t stride = size_t(gridDim.x) * blockDim.x * V;
  for (size_t i = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) * V; i < n; i += stride) {
    if constexpr (V == 2)
      *reinterpret_cast<dbl2*>(y + i) = *reinterpret_cast<const dbl2*>(x + i);
    else
      y[i] = x[i];
  }
}

template <int R, int W>
__global__ __launch_bounds__(256) void soa_kernel(const double* __restrict__ x, double* __restrict__ y, int K) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  double s = 0.0;
#pragma unroll
  for (int f = 0; f < R; ++f) s += x[size_t(f) * K + k];
#pragma unroll
  for (int f = 0; f < W; ++f) y[size_t(f) * K + k] = s + f;
}

struct Timer {
  hipEvent_t a, b;
  Timer() {
    hipEventCreate(&a);
    hipEventCreate(&b);
  }
  ~Timer() {
    hipEventDestroy(a);
    hipEventDestroy(b);
  }
};

template <typename F>
static double time_ms(F&& launch, int reps = 20) {
  for (int i = 0; i < 3; ++i) launch();
  Timer t;
  hipEventRecord(t.a, 0);
  for (int i = 0; i < reps; ++i) launch();
  hipEventRecord(t.b, 0);
  hipEventSynchronize(t.b);
  float ms = 0.f;
  hipEventElapsedTime(&ms, t.a, t.b);
  return ms / reps;
}

template <int R, int W>
static int run_soa(const double* x, double* y, int K, size_t cap) {
  if (size_t(R) * K > cap || size_t(W) * K > cap) {
    printf("soa %d/%d K=%d: buffer too small\n", R, W, K);
    return 1;
  }
  const int grid = (K + 255) / 256;
  const double ms = time_ms([&] { soa_kernel<R, W><<<grid, 256>>>(x, y, K); });
  CK(hipGetLastError());
  const double bytes = 8.0 * (R + W) * K;
  printf("soa R=%2d W=%d K=%8d: %8.2f us  %7.1f GB/s\n", R, W, K, 1e3 * ms, bytes / (ms * 1e-3) / 1e9);
  return 0;
}

int main() {
  const size_t n = size_t(1) << 27;  // 128 Mi doubles = 1 GiB per buffer
  double *x, *y, *out;
  CK(hipMalloc(&x, n * 8));
  CK(hipMalloc(&y, n * 8));
  CK(hipMalloc(&out, 8));
  {
    std::vector<double> h(1 << 20, 1.0);
    for (size_t o = 0; o < n; o += h.size()) CK(hipMemcpy(x + o, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  }
  CK(hipMemset(y, 0, n * 8));
  hipDeviceProp_t prop;
  CK(hipGetDeviceProperties(&prop, 0));
  printf("%s, %d CUs\n", prop.gcnArchName, prop.multiProcessorCount);
  for (int wgPerCU : {4, 8, 16, 32}) {
    const int grid = wgPerCU * prop.multiProcessorCount;
    for (int V : {1, 2}) {
      if (n % (size_t(V) * grid * 256)) {
        printf("n not a multiple of the launch\n");
        return 1;
      }
      const double mr = time_ms([&] {
        if (V == 2)
          read_kernel<2><<<grid, 256>>>(x, n, out);
        else
          read_kernel<1><<<grid, 256>>>(x, n, out);
      });
      CK(hipGetLastError());
      const double mc = time_ms([&] {
        if (V == 2)
          copy_kernel<2><<<grid, 256>>>(x, y, n);
        else
          copy_kernel<1><<<grid, 256>>>(x, y, n);
      });
      CK(hipGetLastError());
      printf("%2d WG/CU, %2d B loads: read %7.1f GB/s, copy %7.1f GB/s\n", wgPerCU, 8 * V, n * 8.0 / (mr * 1e-3) / 1e9,
             2.0 * n * 8.0 / (mc * 1e-3) / 1e9);
    }
  }
  for (int K : {524288, 1048352, 4000000}) {
    if (run_soa<33, 6>(x, y, K, n) || run_soa<24, 6>(x, y, K, n) || run_soa<10, 6>(x, y, K, n)) return 1;
  }
  CK(hipDeviceSynchronize());
  CK(hipFree(x));
  CK(hipFree(y));
  CK(hipFree(out));
  return 0;
}
