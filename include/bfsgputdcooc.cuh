/*
 * bfsgputdcooc.cuh -- HIP kernel prototype with the name and signature of
 * the reference header TurboBFS_COOC_TD/bfsgputdcooc.cuh:18-19.  Defined in
 * libturbobfs.so (gfx950), with the reference's semantics
 * (TurboBFS_COOC_TD/bfsgputdcooc_sc.cu:175-188): for every i < n,
 *   f[i] = 0; if sigma[i] < 0.01 then f[i] = ft[i]; if f[i] > 0 then c = 1,
 *   sigma[i] += f[i].
 * The library's own COOC BFS does not use it (k_cooc_finish also writes S).
 */
#ifndef BFSGPUTDCOOC_CUH
#define BFSGPUTDCOOC_CUH
#include <hip/hip_runtime.h>

__global__ void bfsFunctionsKernel(int *f_d, int *ft_d, float *sigma_d, int *c, int n);

#endif /* BFSGPUTDCOOC_CUH */
