// lds_atomics.hip -- throughput of LDS float atomics on gfx950 for the lane
// -> address patterns that matter to the spreader (standalone program, not
// part of the library).  Each of 4096 workgroups (256 threads) performs 2048
// atomic adds per thread into a 4096-element LDS array; the rate is
// reported in lane-atomics per CU per clock (256 CUs, 2.4 GHz).
//
// Build and run:
//   hipcc --offload-arch=gfx950 -O3 -munsafe-fp-atomics \
//     -Xclang -target-feature -Xclang -packed-fp32-ops \
//     -o /tmp/lds_atomics ducc_amd/csrc/probes/lds_atomics.hip
//   /tmp/lds_atomics
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                     \
  do {                                                               \
    hipError_t e_ = (x);                                             \
    if (e_ != hipSuccess) {                                          \
      printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      exit(1);                                                       \
    }                                                                \
  } while (0)

constexpr int ITER = 2048, N = 4096, NB = 4096;

// MODE: 0 linear (lane + 64 k), 1 lanes distinct mod 32 per half-wave,
// 2 adjacent lane pairs share a class mod 32, 3 lanes distinct mod 16 per
// quarter-wave, 4 lanes distinct mod 64, 5 stride two cells, 6 random
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_probe(T *out) {
  __shared__ T s[N];
  for (int i = threadIdx.x; i < N; i += 256) s[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  uint32_t h = threadIdx.x * 2654435761u + blockIdx.x * 40503u + 1u;
  for (int i = 0; i < ITER; ++i) {
    h ^= h << 13;
    h ^= h >> 17;
    h ^= h << 5;
    const int r = (int)(h >> 8) & 63;
    int a;
    if (MODE == 0) a = lane + 64 * (i & 63);
    else if (MODE == 1) a = (lane & 31) + 32 * (r * 2 + (lane >> 5));
    else if (MODE == 2) a = (lane >> 1) + 64 * r + 32 * (lane & 1);
    else if (MODE == 3) a = (lane & 15) + 16 * (r * 4 + (lane >> 4));
    else if (MODE == 4) a = lane + 64 * r;
    else if (MODE == 5) a = ((2 * lane) & 63) | (64 * r);
    else a = (int)(h & (N - 1));
    atomicAdd(&s[a & (N - 1)], (T)1);
  }
  __syncthreads();
  T acc = 0;
  for (int i = threadIdx.x; i < N; i += 256) acc += s[i];
  out[blockIdx.x * 256 + threadIdx.x] = acc;
}

template <typename T, int MODE>
void run(const char *name, T *out) {
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  hipLaunchKernelGGL((k_probe<T, MODE>), dim3(NB), dim3(256), 0, 0, out);
  CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(a));
  for (int r = 0; r < 5; ++r)
    hipLaunchKernelGGL((k_probe<T, MODE>), dim3(NB), dim3(256), 0, 0, out);
  CHECK(hipEventRecord(b));
  CHECK(hipEventSynchronize(b));
  float ms;
  CHECK(hipEventElapsedTime(&ms, a, b));
  const double n = 5.0 * NB * 256.0 * ITER;
  printf("%-12s %-44s %.3e atomics/s  %5.2f lane-atomics/CU/clk\n",
         sizeof(T) == 8 ? "ds_add_f64" : "ds_add_f32", name, n / (ms * 1e-3),
         n / (ms * 1e-3) / 256 / 2.4e9);
  CHECK(hipEventDestroy(a));
  CHECK(hipEventDestroy(b));
}

int main() {
  double *o64;
  float *o32;
  CHECK(hipMalloc(&o64, sizeof(double) * NB * 256));
  CHECK(hipMalloc(&o32, sizeof(float) * NB * 256));
  run<double, 0>("linear (lane + 64k)", o64);
  run<double, 1>("distinct mod 32 per half-wave", o64);
  run<double, 2>("lane pairs share a class mod 32", o64);
  run<double, 3>("distinct mod 16 per quarter-wave", o64);
  run<double, 4>("distinct mod 64", o64);
  run<double, 5>("stride 2 cells", o64);
  run<double, 6>("random", o64);
  run<float, 0>("linear (lane + 64k)", o32);
  run<float, 6>("random", o32);
  CHECK(hipFree(o64));
  CHECK(hipFree(o32));
  return 0;
}
