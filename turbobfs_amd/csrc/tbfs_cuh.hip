// tbfs_cuh.hip -- the reference's two exported level-finishing kernels, under
// their names and signatures (include/bfsgputdcsc.cuh, include/bfsgputdcooc.cuh),
// for HIP code written against those headers.  Elementwise over i < n.
#include <hip/hip_runtime.h>

#include "bfsgputdcooc.cuh"
#include "bfsgputdcsc.cuh"

__global__ void bfsFunctionsKernel(int* f_d, int* ft_d, float* sigma_d, int* S, int* c, int n,
                                   int d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int t = ft_d[i];
  f_d[i] = t > 0 ? t : 0;
  if (t > 0) {
    *c = 1;
    sigma_d[i] += (float)t;
    S[i] = d;
  }
}

__global__ void bfsFunctionsKernel(int* f_d, int* ft_d, float* sigma_d, int* c, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int f = sigma_d[i] < 0.01f ? ft_d[i] : 0;  // unvisited: no paths counted yet
  f_d[i] = f;
  if (f > 0) {
    *c = 1;
    sigma_d[i] += (float)f;
  }
}
