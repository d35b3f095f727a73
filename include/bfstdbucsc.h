/*
 * bfstdbucsc.h -- drop-in GPU surface of TurboBFS_CSC_TDBU (direction-optimizing
 * BFS over CSC).
 *
 * Same C signatures as TurboBFS_CSC_TDBU/bfstdbucsc.h; exported by
 * libturbobfs.so.  The sequential prototypes of the reference header belong
 * to the reference's own C files and are not part of this library.
 *
 * Semantics kept from the reference:
 *   - S_h[v] = depth, S_h[r] = 0, unreached = -1
 *     (TurboBFS_CSC_TDBU/bfsgputdbucsc_sc.cu:84,118);
 *   - levels switch between top-down and bottom-up expansion.
 * Differences, by design:
 *   - sigma_h is the exact number of shortest paths for EVERY reached vertex.
 *     (The reference does not update sigma in bottom-up levels, so its
 *     sigma is 0 for vertices found there; here bottom-up levels sum the
 *     counts of all parents.)  Counts are exact in 64 bits; a count above
 *     2^31-1 is written modulo 2^32 (the reference's int arithmetic) and the
 *     call returns TBFS_SIGMA_INEXACT (5);
 *   - the switch is the two-way frontier-edge heuristic (push while the
 *     frontier's out-edges are few next to the unexplored edges) rather than
 *     the reference's one-way 0.001*n vertex threshold; it changes the work,
 *     never the result; bbreak has no effect because path counts need every
 *     parent of a column;
 *   - TBFS_ERR_INVALID (1) for invalid input, TBFS_ERR_HIP (2) for HIP
 *     failures, never exit().
 */
#ifndef BFSTDBUCSC_H
#define BFSTDBUCSC_H

/* TurboBFS_CSC_TDBU/bfstdbucsc.h:32-36 */
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

/* Replaces bfs_gpu_tdbu_csc_sc, TurboBFS_CSC_TDBU/bfstdbucsc.h:56-57
 * (implementation TurboBFS_CSC_TDBU/bfsgputdbucsc_sc.cu:53-186). */
int bfs_gpu_tdbu_csc_sc(int *I_h, int *CP_h, int *S_h, int *sigma_h, int r,
                        int nz, int n, int repetition, int bbreak);

/* Replaces bfs_gpu_tdbu_csc_wa, TurboBFS_CSC_TDBU/bfstdbucsc.h:67-68
 * (implementation TurboBFS_CSC_TDBU/bfsgputdbucsc_wa.cu:54-191). */
int bfs_gpu_tdbu_csc_wa(int *IC_h, int *CP_h, int *S_h, int *sigma_h, int r,
                        int nz, int n, int repetition, int bbreak);

#ifdef __cplusplus
}
#endif
#endif /* BFSTDBUCSC_H */
