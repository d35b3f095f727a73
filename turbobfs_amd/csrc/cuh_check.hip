// cuh_check.hip -- launches the two bfsFunctionsKernel overloads through the
// headers include/bfsgputdcsc.cuh / bfsgputdcooc.cuh, from a separate program
// linked against libturbobfs.so, on a fixed 8-element input; prints the
// outputs as one line each (checked by tests/test_gpu_cuh.py).
#include <hip/hip_runtime.h>

#include <cstdio>

#include "bfsgputdcooc.cuh"
#include "bfsgputdcsc.cuh"

#define CK(x)                                                                          \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                     \
      return 1;                                                                        \
    }                                                                                  \
  } while (0)

int main() {
  const int n = 8;
  const int ft_h[n] = {0, 2, 0, 1, 5, 0, 3, 0};
  const float sg_h[n] = {1, 0, 0, 2, 0, 1, 0, 0};
  int *f, *ft, *S, *c;
  float* sg;
  CK(hipMalloc(&f, n * sizeof(int)));
  CK(hipMalloc(&ft, n * sizeof(int)));
  CK(hipMalloc(&S, n * sizeof(int)));
  CK(hipMalloc(&c, sizeof(int)));
  CK(hipMalloc(&sg, n * sizeof(float)));
  for (int variant = 0; variant < 2; ++variant) {
    int zero[n] = {0};
    CK(hipMemcpy(ft, ft_h, sizeof(ft_h), hipMemcpyHostToDevice));
    CK(hipMemcpy(sg, sg_h, sizeof(sg_h), hipMemcpyHostToDevice));
    CK(hipMemcpy(S, zero, sizeof(zero), hipMemcpyHostToDevice));
    CK(hipMemcpy(f, zero, sizeof(zero), hipMemcpyHostToDevice));
    CK(hipMemcpy(c, zero, sizeof(int), hipMemcpyHostToDevice));
    if (variant == 0) {
      void (*k)(int*, int*, float*, int*, int*, int, int) = bfsFunctionsKernel;
      hipLaunchKernelGGL(k, dim3(1), dim3(64), 0, 0, f, ft, sg, S, c, n, 7);
    } else {
      void (*k)(int*, int*, float*, int*, int) = bfsFunctionsKernel;
      hipLaunchKernelGGL(k, dim3(1), dim3(64), 0, 0, f, ft, sg, c, n);
    }
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    int fo[n], So[n], co;
    float so[n];
    CK(hipMemcpy(fo, f, sizeof(fo), hipMemcpyDeviceToHost));
    CK(hipMemcpy(So, S, sizeof(So), hipMemcpyDeviceToHost));
    CK(hipMemcpy(so, sg, sizeof(so), hipMemcpyDeviceToHost));
    CK(hipMemcpy(&co, c, sizeof(int), hipMemcpyDeviceToHost));
    std::printf("%s c=%d f=", variant == 0 ? "csc" : "cooc", co);
    for (int i = 0; i < n; ++i) std::printf("%d%s", fo[i], i + 1 < n ? "," : "");
    std::printf(" sigma=");
    for (int i = 0; i < n; ++i) std::printf("%g%s", so[i], i + 1 < n ? "," : "");
    std::printf(" S=");
    for (int i = 0; i < n; ++i) std::printf("%d%s", So[i], i + 1 < n ? "," : "");
    std::printf("\n");
  }
  return 0;
}
