// probe_atomics.hip -- rate of 64-bit device atomics at random addresses
// against plain loads at the same addresses (8 MB array, n operations, one
// per thread).  Not part of the library: it measures the cost that the push
// levels' count additions and frontier claims pay per edge.
//   make -C turbobfs_amd/csrc probe && turbobfs_amd/probe_atomics
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#define CK(x)                                                                      \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);  \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

__global__ void k_add_ret(unsigned long long* a, const int* idx, int n, unsigned long long* sink) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long o = atomicAdd(&a[idx[i]], 1ull);
  if (o == 0xdeadbeefull) sink[0] = o;  // keeps the result live
}

__global__ void k_add(unsigned long long* a, const int* idx, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicAdd(&a[idx[i]], 1ull);
}

__global__ void k_load(const unsigned long long* a, const int* idx, int n, unsigned long long* sink) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long o = a[idx[i]];
  if (o == 0xdeadbeefull) sink[0] = o;
}

int main() {
  const int m = 1 << 20;      // words in the target array (8 MB)
  const int nmax = 1 << 24;   // largest operation count
  unsigned long long *a = nullptr, *sink = nullptr;
  int* idx = nullptr;
  CK(hipMalloc(&a, (size_t)m * 8));
  CK(hipMalloc(&sink, 8));
  CK(hipMemset(a, 0, (size_t)m * 8));
  std::vector<int> h(nmax);
  std::mt19937 rng(1);
  for (int i = 0; i < nmax; ++i) h[i] = (int)(rng() % m);
  CK(hipMalloc(&idx, (size_t)nmax * 4));
  CK(hipMemcpy(idx, h.data(), (size_t)nmax * 4, hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const char* names[3] = {"atomicAdd, result used", "atomicAdd, result unused", "plain load"};
  for (int n : {1 << 16, 1 << 19, 1 << 22, 1 << 24}) {
    for (int kind = 0; kind < 3; ++kind) {
      float best = 1e9f;
      for (int rep = 0; rep < 4; ++rep) {
        CK(hipEventRecord(e0));
        const dim3 grid((n + 255) / 256), block(256);
        if (kind == 0) hipLaunchKernelGGL(k_add_ret, grid, block, 0, 0, a, idx, n, sink);
        else if (kind == 1) hipLaunchKernelGGL(k_add, grid, block, 0, 0, a, idx, n);
        else hipLaunchKernelGGL(k_load, grid, block, 0, 0, a, idx, n, sink);
        CK(hipGetLastError());
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) best = ms;
      }
      std::printf("n = %8d  %-26s %9.1f us  %6.1f G/s\n", n, names[kind], best * 1e3,
                  n / (best * 1e-3) / 1e9);
    }
  }
  CK(hipFree(idx));
  CK(hipFree(sink));
  CK(hipFree(a));
  return 0;
}
