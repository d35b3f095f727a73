/*
 * bfsgputdcsc.cuh -- HIP kernel prototype with the name and signature of the
 * reference header TurboBFS_CSC_TD/bfsgputdcsc.cuh:18-19, for HIP code that
 * launches it (the reference's _wa driver launches the _sc translation
 * unit's kernel through this header).  The kernel is defined in
 * libturbobfs.so (gfx950) and does what the reference's does
 * (TurboBFS_CSC_TD/bfsgputdcsc_sc.cu:194-208): for every i < n,
 *   f[i] = 0; if ft[i] > 0 then c = 1, f[i] = ft[i], sigma[i] += ft[i], S[i] = d.
 * The library's own BFS does not use it: its level kernels fuse this step.
 */
#ifndef BFSGPUTDCSC_CUH
#define BFSGPUTDCSC_CUH
#include <hip/hip_runtime.h>

__global__ void bfsFunctionsKernel(int *f_d, int *ft_d, float *sigma_d, int *S, int *c,
                                   int n, int d);

#endif /* BFSGPUTDCSC_CUH */
