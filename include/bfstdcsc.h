/*
 * bfstdcsc.h -- drop-in GPU surface of TurboBFS_CSC_TD (top-down BFS over CSC).
 *
 * Same C signatures as the reference header TurboBFS_CSC_TD/bfstdcsc.h; the
 * symbols are exported by libturbobfs.so (gfx950), so the reference's own
 * driver (TurboBFS_CSC_TD/bfstdcsc_main.c) links against it unchanged.  The
 * sequential prototypes of the reference header (bfs_seq_td_csc, sum_vs, ...)
 * belong to the reference's own C files and are not part of this library.
 *
 * Semantics kept from the reference:
 *   - all arrays are caller-owned host memory; the graph is uploaded, the BFS
 *     runs `repetition` times, S_h/sigma_h receive the last result, and all
 *     device memory is released before return;
 *   - S_h[v] = depth, 0 for the root AND for unreached vertices (the
 *     reference memsets S to 0, TurboBFS_CSC_TD/bfsgputdcsc_sc.cu:75);
 *   - sigma_h[v] = number of shortest r->v paths (0 when unreached).
 * Differences, by design:
 *   - counts are computed exactly in 64-bit integers and rounded once to
 *     float; if any count exceeds 2^24 (where float stops being exact) the
 *     call returns TBFS_SIGMA_INEXACT (5) and says so on stderr;
 *   - invalid input (n <= 0, r out of range, malformed CSC, repetition < 1)
 *     returns TBFS_ERR_INVALID (1) instead of running; HIP failures return
 *     TBFS_ERR_HIP (2); the library never calls exit().
 */
#ifndef BFSTDCSC_H
#define BFSTDCSC_H

/* Structure of Arrays for the CSC adjacency matrix (TurboBFS_CSC_TD/bfstdcsc.h:30-34). */
#ifndef TBFS_STRUCT_CSC_DEFINED
#define TBFS_STRUCT_CSC_DEFINED
struct Csc {
  int *IC;
  int *CP;
};
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces bfs_gpu_td_csc_sc, TurboBFS_CSC_TD/bfstdcsc.h:54-55
 * (implementation TurboBFS_CSC_TD/bfsgputdcsc_sc.cu:47-149): one lane per
 * column pull kernel. */
int bfs_gpu_td_csc_sc(int *IC_h, int *CP_h, int *S_h, float *sigma_h, int r,
                      int nz, int n, int repetition);

/* Replaces bfs_gpu_td_csc_wa, TurboBFS_CSC_TD/bfstdcsc.h:65-66
 * (implementation TurboBFS_CSC_TD/bfsgputdcsc_wa.cu:46-150): one wavefront
 * (64 lanes) per column pull kernel. */
int bfs_gpu_td_csc_wa(int *IC_h, int *CP_h, int *S_h, float *sigma_h, int r,
                      int nz, int n, int repetition);

#ifdef __cplusplus
}
#endif
#endif /* BFSTDCSC_H */
