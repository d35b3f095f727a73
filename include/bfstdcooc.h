/*
 * bfstdcooc.h -- drop-in GPU surface of TurboBFS_COOC_TD (top-down BFS over
 * COO sorted by column).
 *
 * Same C signature as TurboBFS_COOC_TD/bfstdcooc.h:48-49; exported by
 * libturbobfs.so.  The sequential prototypes of the reference header belong
 * to the reference's own C files and are not part of this library.
 *
 * Semantics kept from the reference:
 *   - edge I_h[k] -> J_h[k] for k < nz; J_h is normally non-decreasing
 *     (TurboBFS_COOC_TD/sparseformatransf.c:62-95) but need not be;
 *   - only sigma_h is produced (number of shortest r->v paths, 0 when
 *     unreached); each level is an edge-parallel pass over all nonzeros.
 * Differences, by design: counts are exact in 64 bits and rounded once to
 * float, with TBFS_SIGMA_INEXACT (5) returned when any exceeds 2^24;
 * TBFS_ERR_INVALID (1) / TBFS_ERR_HIP (2) instead of exit().
 */
#ifndef BFSTDCOOC_H
#define BFSTDCOOC_H

/* TurboBFS_COOC_TD/bfstdcooc.h:29-32 */
struct Cooc {
  int *ICOOC;
  int *JCOOC;
};

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces bfs_gpu_td_cooc_sc, TurboBFS_COOC_TD/bfstdcooc.h:48-49
 * (implementation TurboBFS_COOC_TD/bfsgputdcooc_sc.cu:44-140). */
int bfs_gpu_td_cooc_sc(int *I_h, int *J_h, float *sigma_h, int r, int nz,
                       int n, int repetition);

#ifdef __cplusplus
}
#endif
#endif /* BFSTDCOOC_H */
