/*
 * oracle.c -- CPU restatement of TurboBFS's sequential BFS path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py may load it.  The product library never links or calls it.
 *
 * What it restates (reference file:line, paths relative to the reference root):
 *   orc_bfs_seq_td_csc   TurboBFS_CSC_TD/bfstdcsc_seq.c:27-96 (loop order),
 *                        TurboBFS_CSC_TD/bfstdcsc_seq_utils.c:21-103 (sum_vs,
 *                        check_f, S_v, mult_vs, assign_v),
 *                        TurboBFS_CSC_TD/spmvcsc_seq.c:22-41 (pull SpMV over all
 *                        columns).  S conventions: unreached = 0 as in
 *                        TurboBFS_CSC_TD/bfstdcsc_main.c:136 (calloc), or -1 with
 *                        S[r] = 0 as in TurboBFS_CSC_BU/bfstdcsc_seq.c:35-36 and
 *                        TurboBFS_CSC_TDBU/bfstdcsc_seq.c:35-36.
 *   orc_bfs_seq_td_cooc  TurboBFS_COOC_TD/bfstdcooc_seq.c:29-91,
 *                        TurboBFS_COOC_TD/spmvcooc_seq.c:27-36 (f_t is never
 *                        re-zeroed, as in the reference).
 *   orc_bfs_textbook     not a reference function: an O(n+m) queue BFS with
 *                        shortest-path counting, used as an independent second
 *                        oracle and for graphs where the O(L*(n+nz)) restatement
 *                        is too slow.
 *
 * Number type.  The reference holds path counts in float (CSC_TD, CSC_BU,
 * COOC) or int32 (TDBU).  Here every count is a pair (value mod 2^64, sticky
 * overflow flag).  For non-negative addends both halves are independent of
 * the order of additions, so any parallel summation order gives the same
 * pair.  A count is "nonzero" iff value != 0 or the flag is set; this is what
 * the reference's "> 0.1" / "> 0.9" tests mean on integer-valued data.
 *
 * Parity pinning of this file (tests/test_oracle.py): brute force on tiny
 * graphs, closed forms (paths, cycles, stars, complete graphs, corner-rooted
 * grids C(dx+dy,dx) crossing 2^64, hypercubes j!), the known answers for the
 * reference's bundled .mtx files, agreement of the three functions with each
 * other, and agreement with the reference's own float/int seq code compiled
 * from its sources into oracle/_ref where sigma is small enough to be exact.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_ERR_ARG 1
#define ORC_ERR_NOMEM 2

static inline int nz64(uint64_t v, uint8_t o) { return v != 0 || o != 0; }

/* (a, ao) += (b, bo) with carry into the sticky flag */
static inline void add64(uint64_t *a, uint8_t *ao, uint64_t b, uint8_t bo) {
  uint64_t t = *a + b;
  if (t < *a) *ao = 1;
  if (bo) *ao = 1;
  *a = t;
}

/*
 * Restatement of bfs_seq_td_csc (TurboBFS_CSC_TD/bfstdcsc_seq.c:27-96).
 * Returns the number of passes of the level loop (the reference's final d),
 * or -1 on bad arguments / allocation failure.
 */
static int seq_td_csc(const int32_t *IC, const int32_t *CP, int32_t n,
                      int32_t r, int32_t s_unreached, int32_t *S,
                      uint64_t *sigma, uint8_t *ovf, int32_t max_passes);

int orc_bfs_seq_td_csc(const int32_t *IC, const int32_t *CP, int32_t n,
                       int32_t r, int32_t s_unreached, int32_t *S,
                       uint64_t *sigma, uint8_t *ovf) {
  return seq_td_csc(IC, CP, n, r, s_unreached, S, sigma, ovf, 0);
}

/* The same loop stopped after max_passes passes (a timing sample of a BFS
 * too long to run whole on one CPU thread); returns the passes made. */
int orc_bfs_seq_td_csc_passes(const int32_t *IC, const int32_t *CP, int32_t n,
                              int32_t r, int32_t *S, uint64_t *sigma, uint8_t *ovf,
                              int32_t max_passes) {
  if (max_passes <= 0) return -1;
  return seq_td_csc(IC, CP, n, r, -1, S, sigma, ovf, max_passes);
}

static int seq_td_csc(const int32_t *IC, const int32_t *CP, int32_t n,
                      int32_t r, int32_t s_unreached, int32_t *S,
                      uint64_t *sigma, uint8_t *ovf, int32_t max_passes) {
  if (n <= 0 || r < 0 || r >= n) return -1;
  uint64_t *f = (uint64_t *)calloc((size_t)n, sizeof(uint64_t));
  uint64_t *ft = (uint64_t *)calloc((size_t)n, sizeof(uint64_t));
  uint8_t *fo = (uint8_t *)calloc((size_t)n, 1);
  uint8_t *fto = (uint8_t *)calloc((size_t)n, 1);
  if (!f || !ft || !fo || !fto) {
    free(f); free(ft); free(fo); free(fto);
    return -1;
  }
  for (int32_t i = 0; i < n; i++) {
    S[i] = s_unreached;
    sigma[i] = 0;
    ovf[i] = 0;
  }
  S[r] = 0;          /* 0 under both conventions */
  f[r] = 1;          /* bfstdcsc_seq.c:35 */
  int d = 0, c = 1;
  while (c > 0 && (max_passes <= 0 || d < max_passes)) {    /* bfstdcsc_seq.c:48 */
    d++;
    /* sum_vs: sigma += f where f > 0 (bfstdcsc_seq_utils.c:21-31) */
    for (int32_t k = 0; k < n; k++)
      if (nz64(f[k], fo[k])) add64(&sigma[k], &ovf[k], f[k], fo[k]);
    /* spmv_seq_ug_csc: f_t[i] = sum_{k in col i} f[IC[k]] (spmvcsc_seq.c:22-41) */
    for (int32_t i = 0; i < n; i++) {
      uint64_t s = 0;
      uint8_t so = 0;
      for (int32_t k = CP[i]; k < CP[i + 1]; k++) {
        int32_t u = IC[k];
        add64(&s, &so, f[u], fo[u]);
      }
      ft[i] = s;
      fto[i] = so;
    }
    /* assign_v: f = f_t where f_t > 0, else 0 (bfstdcsc_seq_utils.c:90-103) */
    for (int32_t k = 0; k < n; k++) {
      if (nz64(ft[k], fto[k])) { f[k] = ft[k]; fo[k] = fto[k]; }
      else { f[k] = 0; fo[k] = 0; }
    }
    /* mult_vs: f = 0 where sigma > 0, i.e. already visited (:72-83) */
    for (int32_t k = 0; k < n; k++)
      if (nz64(sigma[k], ovf[k])) { f[k] = 0; fo[k] = 0; }
    /* S_v: S = d where f > 0 (:56-66) */
    for (int32_t k = 0; k < n; k++)
      if (nz64(f[k], fo[k])) S[k] = d;
    /* check_f: c = any(f > 0) (:38-49) */
    c = 0;
    for (int32_t k = 0; k < n; k++)
      if (nz64(f[k], fo[k])) { c = 1; break; }
  }
  free(f); free(ft); free(fo); free(fto);
  return d;
}

/*
 * Restatement of bfs_seq_td_cooc (TurboBFS_COOC_TD/bfstdcooc_seq.c:29-91) with
 * spmv_seq_td_cooc (TurboBFS_COOC_TD/spmvcooc_seq.c:27-36): edge I[k] -> J[k],
 * f_t[J[k]] += f[I[k]] over all nonzeros, f_t never cleared.  No S output.
 */
int orc_bfs_seq_td_cooc(const int32_t *I, const int32_t *J, int32_t nz,
                        int32_t n, int32_t r, uint64_t *sigma, uint8_t *ovf) {
  if (n <= 0 || r < 0 || r >= n || nz < 0) return -1;
  uint64_t *f = (uint64_t *)calloc((size_t)n, sizeof(uint64_t));
  uint64_t *ft = (uint64_t *)calloc((size_t)n, sizeof(uint64_t));
  uint8_t *fo = (uint8_t *)calloc((size_t)n, 1);
  uint8_t *fto = (uint8_t *)calloc((size_t)n, 1);
  if (!f || !ft || !fo || !fto) {
    free(f); free(ft); free(fo); free(fto);
    return -1;
  }
  for (int32_t i = 0; i < n; i++) { sigma[i] = 0; ovf[i] = 0; }
  f[r] = 1;
  int d = 0, c = 1;
  while (c > 0) {
    d++;
    for (int32_t k = 0; k < n; k++)
      if (nz64(f[k], fo[k])) add64(&sigma[k], &ovf[k], f[k], fo[k]);
    for (int32_t k = 0; k < nz; k++)
      add64(&ft[J[k]], &fto[J[k]], f[I[k]], fo[I[k]]);
    for (int32_t k = 0; k < n; k++) {
      if (nz64(ft[k], fto[k])) { f[k] = ft[k]; fo[k] = fto[k]; }
      else { f[k] = 0; fo[k] = 0; }
    }
    for (int32_t k = 0; k < n; k++)
      if (nz64(sigma[k], ovf[k])) { f[k] = 0; fo[k] = 0; }
    c = 0;
    for (int32_t k = 0; k < n; k++)
      if (nz64(f[k], fo[k])) { c = 1; break; }
  }
  free(f); free(ft); free(fo); free(fto);
  return d;
}

/*
 * Independent O(n + m) oracle: FIFO BFS over out-edges (the transpose of the
 * CSC in-lists), sigma[v] += sigma[u] for every edge u->v with
 * S[v] == S[u] + 1.  S unreached = -1, S[r] = 0.  Returns the number of
 * distinct depths (1 + eccentricity of r), or -1 on error.
 */
int orc_bfs_textbook(const int32_t *CP, const int32_t *IC, int32_t n, int32_t r,
                     int32_t *S, uint64_t *sigma, uint8_t *ovf) {
  if (n <= 0 || r < 0 || r >= n) return -1;
  int64_t nz = CP[n];
  int64_t *rp = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  int32_t *jr = (int32_t *)malloc((size_t)(nz > 0 ? nz : 1) * sizeof(int32_t));
  int32_t *queue = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  if (!rp || !jr || !queue) { free(rp); free(jr); free(queue); return -1; }
  for (int64_t k = 0; k < nz; k++) rp[IC[k] + 1]++;
  for (int32_t i = 0; i < n; i++) rp[i + 1] += rp[i];
  {
    int64_t *cur = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!cur) { free(rp); free(jr); free(queue); return -1; }
    memcpy(cur, rp, (size_t)n * sizeof(int64_t));
    for (int32_t v = 0; v < n; v++)
      for (int32_t k = CP[v]; k < CP[v + 1]; k++) jr[cur[IC[k]]++] = v;
    free(cur);
  }
  for (int32_t i = 0; i < n; i++) { S[i] = -1; sigma[i] = 0; ovf[i] = 0; }
  S[r] = 0;
  sigma[r] = 1;
  int32_t head = 0, tail = 0, maxd = 0;
  queue[tail++] = r;
  while (head < tail) {
    int32_t u = queue[head++];
    int32_t du = S[u];
    if (du > maxd) maxd = du;
    for (int64_t k = rp[u]; k < rp[u + 1]; k++) {
      int32_t v = jr[k];
      if (S[v] == -1) { S[v] = du + 1; queue[tail++] = v; }
      if (S[v] == du + 1) add64(&sigma[v], &ovf[v], sigma[u], ovf[u]);
    }
  }
  free(rp); free(jr); free(queue);
  return maxd + 1;
}

/* Number of vertices whose exact count exceeds `limit` (flag set counts too). */
int64_t orc_count_exceed(const uint64_t *sigma, const uint8_t *ovf, int32_t n,
                         uint64_t limit) {
  int64_t c = 0;
  for (int32_t i = 0; i < n; i++)
    if (ovf[i] || sigma[i] > limit) c++;
  return c;
}
