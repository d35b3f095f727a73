/*
 * bfsbucsc.h -- drop-in GPU surface of TurboBFS_CSC_BU (bottom-up BFS, CSC).
 *
 * Same C signatures as TurboBFS_CSC_BU/bfsbucsc.h; exported by libturbobfs.so.
 * The sequential prototypes of the reference header belong to the reference's
 * own C files and are not part of this library.
 *
 * Semantics kept from the reference:
 *   - S_h[v] = depth, S_h[r] = 0, unreached = -1 (TurboBFS_CSC_BU/bfsgpubucsc_sc.cu:84);
 *   - bbreak != 0 stops a column scan at the first parent found; the result
 *     does not depend on it, only the work does;
 *   - no path counts are produced.
 * Errors: TBFS_ERR_INVALID (1) for invalid input, TBFS_ERR_HIP (2) for HIP
 * failures, never exit().
 */
#ifndef BFSBUCSC_H
#define BFSBUCSC_H

/* TurboBFS_CSC_BU/bfsbucsc.h:31-35 */
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

/* Replaces bfs_gpu_bu_csc_sc, TurboBFS_CSC_BU/bfsbucsc.h:55-56
 * (implementation TurboBFS_CSC_BU/bfsgpubucsc_sc.cu:47-114). */
int bfs_gpu_bu_csc_sc(int *I_h, int *CP_h, int *S_h, int r, int nz, int n,
                      int repetition, int bbreak);

/* Replaces bfs_gpu_bu_csc_wa, TurboBFS_CSC_BU/bfsbucsc.h:66-67
 * (implementation TurboBFS_CSC_BU/bfsgpubucsc_wa.cu:46-113). */
int bfs_gpu_bu_csc_wa(int *I_h, int *CP_h, int *S_h, int r, int nz, int n,
                      int repetition, int bbreak);

#ifdef __cplusplus
}
#endif
#endif /* BFSBUCSC_H */
