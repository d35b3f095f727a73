/*
 * tbfs_exec.c -- command-line driver with the reference's argument lists and
 * printouts, on this library's entry points (paths relative to the reference
 * root).
 *
 * Built once per reference variant (TBFS_EXEC_VARIANT):
 *   0 tbfs_exec_td    file w ug format p r repet seq
 *                     (TurboBFS_CSC_TD/bfstdcsc_main.c:82-90)
 *   1 tbfs_exec_bu    file w ug format p r repet seq bbreak
 *                     (TurboBFS_CSC_BU/bfsbucsc_main.c:84-91)
 *   2 tbfs_exec_tdbu  file w ug format p r repet seq bbreak
 *                     (TurboBFS_CSC_TDBU/bfstdbucsc_main.c:85-92)
 *   3 tbfs_exec_cooc  file w ug p r repet seq
 *                     (TurboBFS_COOC_TD/bfstdcooc_main.c:79-84)
 * so the reference's run scripts (run_bfstdcsc.sh and the others) run with
 * `exec` pointing at the matching binary.  w: the file has a value column
 * (read and ignored); ug: the entries are the lower triangle of an undirected
 * graph (both directions stored); format: 0 one lane per column entry point
 * (_sc), 1 one wavefront per column (_wa); p: print the results (printBFS,
 * TurboBFS_CSC_TD/utils_bfs.c:26-90); r: root; repet: BFS repetitions; seq:
 * also run a sequential BFS on the host and check the GPU results against it
 * (bfs_check / S_check, TurboBFS_CSC_TD/utils_bfs.c:93-145); bbreak: stop a
 * column scan at its first parent.
 *
 * Differences from the reference's mains, all on the error side:
 *   - the argument count is checked before any argument is read (the TD
 *     main reads argv[8] but checks argc < 8; every main reads its
 *     arguments before the check);
 *   - entries need not be sorted by column (they are sorted here);
 *   - the sequential check is an O(n + m) queue BFS with exact 64-bit path
 *     counts, converted to the reference's output type, and the checks are
 *     exact (the reference compares floats within 1e-5 and also calls every
 *     count above n an error);
 *   - every vertex is binned in the degree distribution (the reference's
 *     loop leaves out degree == max when max is a power of two);
 *   - the exit status is the entry point's when it fails (5, a count beyond
 *     the output type's exact range, is reported and is not a failure).
 * The GPU entry points print their own timings with TBFS_VERBOSE=1.
 */
#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bfsbucsc.h"
#include "bfstdbucsc.h"
#include "bfstdcooc.h"
#include "bfstdcsc.h"
#include "tbfs.h"

#ifndef TBFS_EXEC_VARIANT
#define TBFS_EXEC_VARIANT 0
#endif
#define V_TD 0
#define V_BU 1
#define V_TDBU 2
#define V_COOC 3

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void die(const char *msg, const char *arg) {
  fprintf(stderr, "tbfs_exec: %s%s%s\n", msg, arg ? ": " : "", arg ? arg : "");
  exit(1);
}

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) die("out of host memory", NULL);
  return p;
}

/* ---- Matrix Market input (TurboBFS_CSC_TD/mmiof.c:97-218 banner and size
 * line, readMMfilef.c:33-70 entries) ---------------------------------------
 * Entry "i j [value]" (1-based) is the edge i -> j, stored in CSC column j.
 * With ug the entries are mirrored; columns come out with rows ascending
 * (undirected) or in file order (directed), as the reference builds them
 * from a column-sorted file. */
typedef struct {
  int n, nz;
  int *CP, *IC;
} Csc;

static void read_mm(const char *path, int ug, Csc *g) {
  FILE *f = fopen(path, "r");
  if (!f) die("cannot open", path);
  char line[1024];
  if (!fgets(line, sizeof line, f)) die("empty file", path);
  char mtx[64] = "", obj[64] = "", fmt[64] = "", field[64] = "", sym[64] = "";
  if (sscanf(line, "%63s %63s %63s %63s %63s", mtx, obj, fmt, field, sym) < 4 ||
      strcmp(mtx, "%%MatrixMarket") != 0)
    die("Could not process Matrix Market banner", path);
  for (char *c = fmt; *c; ++c) *c = (char)tolower((unsigned char)*c);
  for (char *c = field; *c; ++c) *c = (char)tolower((unsigned char)*c);
  if (strcmp(fmt, "coordinate") != 0) die("only sparse coordinate matrices are supported", path);
  if (strcmp(field, "complex") == 0) die("complex matrices are not supported", path);
  long M = 0, N = 0, nz0 = 0;
  for (;;) {
    if (!fgets(line, sizeof line, f)) die("missing size line", path);
    if (line[0] == '%' || line[0] == '\n') continue;
    if (sscanf(line, "%ld %ld %ld", &M, &N, &nz0) != 3) die("bad size line", path);
    break;
  }
  const long n = M > N ? M : N;
  const long nz = ug ? 2 * nz0 : nz0;
  if (n <= 0 || n > INT32_MAX - 1 || nz > INT32_MAX) die("matrix too large", path);
  int *src = xcalloc((size_t)nz, sizeof(int));
  int *dst = xcalloc((size_t)nz, sizeof(int));
  long k = 0;
  while (k < nz0 && fgets(line, sizeof line, f)) {
    if (line[0] == '%' || line[0] == '\n') continue;
    long i, j;
    if (sscanf(line, "%ld %ld", &i, &j) != 2) die("bad entry line", path);
    if (i < 1 || j < 1 || i > M || j > N) die("entry index out of range", path);
    src[k] = (int)(i - 1);
    dst[k] = (int)(j - 1);
    ++k;
  }
  fclose(f);
  if (k != nz0) die("fewer entries than the size line says", path);
  if (ug)
    for (long e = 0; e < nz0; ++e) {
      src[nz0 + e] = dst[e];
      dst[nz0 + e] = src[e];
    }
  /* stable counting sorts: by row first when undirected (rows ascending in
   * each column), then by column */
  int *cnt = xcalloc((size_t)n + 1, sizeof(int));
  int *tsrc = xcalloc((size_t)nz, sizeof(int)), *tdst = xcalloc((size_t)nz, sizeof(int));
  if (ug) {
    for (long e = 0; e < nz; ++e) cnt[src[e] + 1]++;
    for (long v = 0; v < n; ++v) cnt[v + 1] += cnt[v];
    for (long e = 0; e < nz; ++e) {
      const int p = cnt[src[e]]++;
      tsrc[p] = src[e];
      tdst[p] = dst[e];
    }
    memcpy(src, tsrc, (size_t)nz * sizeof(int));
    memcpy(dst, tdst, (size_t)nz * sizeof(int));
    memset(cnt, 0, ((size_t)n + 1) * sizeof(int));
  }
  g->n = (int)n;
  g->nz = (int)nz;
  g->CP = xcalloc((size_t)n + 1, sizeof(int));
  g->IC = xcalloc((size_t)nz, sizeof(int));
  for (long e = 0; e < nz; ++e) g->CP[dst[e] + 1]++;
  for (long v = 0; v < n; ++v) g->CP[v + 1] += g->CP[v];
  memcpy(cnt, g->CP, (size_t)n * sizeof(int));
  for (long e = 0; e < nz; ++e) g->IC[cnt[dst[e]]++] = src[e];
  free(cnt);
  free(tsrc);
  free(tdst);
  free(src);
  free(dst);
}

/* ---- degree statistics printout (TurboBFS_CSC_TD/degree.c:22-141, printed
 * as bfstdcsc_main.c:173-230 prints them) ---------------------------------- */
static void print_degrees(const char *what, const int *deg, int n) {
  int mx = 0, imx = 0;
  double sum = 0.0, sq = 0.0;
  for (int v = 0; v < n; ++v) {
    if (deg[v] > mx) {
      mx = deg[v];
      imx = v;
    }
    sum += deg[v];
  }
  const double mean = sum / n;
  for (int v = 0; v < n; ++v) sq += (deg[v] - mean) * (deg[v] - mean);
  const double sd = n > 1 ? sqrt(sq / (n - 1)) : 0.0;
  printf("maximum %s = %d of node = %d\n", what, mx, imx);
  if (strcmp(what, "degree") == 0)
    printf("degree statistics:meanDegree=%6.2f,sdDegree = %6.2f,\n", mean, sd);
  else if (strcmp(what, "in_degree") == 0)
    printf("degree statistics:meanInDegree=%6.2f,sdInDegree = %6.2f,\n", mean, sd);
  else
    printf("degree statistics:meanOutDegree=%6.2f,sdOutDegree = %6.2f,\n", mean, sd);
  long bins[33] = {0};
  for (int v = 0; v < n; ++v) {
    int b = 0;
    for (unsigned d = (unsigned)deg[v]; d; d >>= 1) ++b;  /* 0; [1,2) -> 1; [2,4) -> 2 */
    bins[b]++;
  }
  printf("Degree %d: %ld (%1.3f %%)\n", 0, bins[0], 100.0 * bins[0] / n);
  for (int j = 0, i = 1; i <= mx && j < 32; i *= 2, ++j)
    printf("Degree 2^%d: %ld (%1.3f %%))\n", j, bins[j + 1], 100.0 * bins[j + 1] / n);
}

/* ---- sequential check: queue BFS with exact counts -------------------------
 * S[v] = depth (-1 unreached), sig[v] = shortest-path count mod 2^64, ovf[v]
 * set when it is >= 2^64.  In-lists of a column are its in-neighbours, so
 * the out-lists are built first. */
static void seq_bfs(const Csc *g, int r, int *S, uint64_t *sig, uint8_t *ovf) {
  const int n = g->n, nz = g->nz;
  int *RP = xcalloc((size_t)n + 1, sizeof(int)), *JR = xcalloc((size_t)nz, sizeof(int));
  for (int e = 0; e < nz; ++e) RP[g->IC[e] + 1]++;
  for (int v = 0; v < n; ++v) RP[v + 1] += RP[v];
  int *fill = xcalloc((size_t)n, sizeof(int));
  memcpy(fill, RP, (size_t)n * sizeof(int));
  for (int v = 0; v < n; ++v)
    for (int e = g->CP[v]; e < g->CP[v + 1]; ++e) JR[fill[g->IC[e]]++] = v;
  int *q = xcalloc((size_t)n, sizeof(int));
  for (int v = 0; v < n; ++v) {
    S[v] = -1;
    sig[v] = 0;
    ovf[v] = 0;
  }
  int head = 0, tail = 0;
  S[r] = 0;
  sig[r] = 1;
  q[tail++] = r;
  while (head < tail) {
    const int u = q[head++];
    for (int e = RP[u]; e < RP[u + 1]; ++e) {
      const int v = JR[e];
      if (S[v] < 0) {
        S[v] = S[u] + 1;
        q[tail++] = v;
      }
      if (S[v] == S[u] + 1) {
        const uint64_t t = sig[v] + sig[u];
        if (t < sig[v] || ovf[u]) ovf[v] = 1;
        sig[v] = t;
      }
    }
  }
  free(q);
  free(fill);
  free(RP);
  free(JR);
}

/* checkers (TurboBFS_CSC_TD/utils_bfs.c:93-145), exact */
static void S_check(const int *S, const int *S_hgpu, int n) {
  int count = 0;
  for (int i = 0; i < n; ++i) count += S[i] != S_hgpu[i];
  if (!count) printf("values of the GPU computed S vector are correct \n");
  else printf("error in %d values of the GPU computed S_hgpu vector\n", count);
}

static void sigma_check_f(const float *sigma, const float *sigma_hgpu, int n) {
  int count = 0;
  for (int i = 0; i < n; ++i) count += memcmp(&sigma[i], &sigma_hgpu[i], sizeof(float)) != 0;
  if (!count) printf("values of the GPU computed sigma vector are correct \n");
  else printf("error in %d values of the GPU computed sigma_hgpu vector\n", count);
}

static void sigma_check_i(const int *sigma, const int *sigma_hgpu, int n) {
  int count = 0;
  for (int i = 0; i < n; ++i) count += sigma[i] != sigma_hgpu[i];
  if (!count) printf("values of the GPU computed sigma vector are correct \n");
  else printf("error in %d values of the GPU computed sigma_hgpu vector\n", count);
}

/* printBFS (TurboBFS_CSC_TD/utils_bfs.c:26-90 and its per-variant copies):
 * the first and last 100 rows when n > 100. */
static int shown(int i, int n) { return n <= 100 || i < 100 || i >= n - 100; }

static void print_S(const int *S, const int *S_hgpu, int n) {
  printf("\nS   S_hgpu \n");
  for (int i = 0; i < n; ++i)
    if (shown(i, n)) printf("%d,%d,%d\n", i, S[i], S_hgpu[i]);
}

static void print_sigma_f2(const char *title, const float *a, const float *b, int n) {
  printf("\n%s\n", title);
  for (int i = 0; i < n; ++i)
    if (shown(i, n)) printf("%d,%lf,%lf\n", i, (double)a[i], (double)b[i]);
}

static void print_sigma_f1(const char *title, const float *a, int n) {
  printf("\n%s\n", title);
  for (int i = 0; i < n; ++i)
    if (shown(i, n)) printf("%d,%lf\n", i, (double)a[i]);
}

static void print_sigma_i1(const char *title, const int *a, int n) {
  printf("\n%s\n", title);
  for (int i = 0; i < n; ++i)
    if (shown(i, n)) printf("%d,%d\n", i, a[i]);
}

int main(int argc, char *argv[]) {
  static const char *usage[] = {
      "file w ug format p r repet seq", "file w ug format p r repet seq bbreak",
      "file w ug format p r repet seq bbreak", "file w ug p r repet seq"};
  static const int nargs[] = {9, 10, 10, 8};
  const int V = TBFS_EXEC_VARIANT;
  if (argc < nargs[V]) {
    printf("number of arguments not valid, should be %d\n", nargs[V] - 1);
    fprintf(stderr, "Usage: %s %s\n", argv[0], usage[V]);
    return 1;
  }
  const char *path = argv[1];
  const int ug = atoi(argv[3]);
  int a = 4;
  const int format = V == V_COOC ? 0 : atoi(argv[a++]);
  const int p = atoi(argv[a++]);
  const int r = atoi(argv[a++]);
  const int repet = atoi(argv[a++]);
  const int seq = atoi(argv[a++]);
  const int bbreak = (V == V_BU || V == V_TDBU) ? atoi(argv[a++]) : 0;
  /* argv[2] (w): a value column is read and ignored in any case */

  double t0 = now_s();
  Csc g;
  read_mm(path, ug, &g);
  const double scan_t = now_s() - t0;
  const int n = g.n, nz = g.nz;
  printf("\nN = %d,M = %d, nz = %d\n", n, n, nz);
  if (r < 0 || r >= n) {
    fprintf(stderr, "tbfs_exec: root %d out of range [0, %d)\n", r, n);
    return 1;
  }

  t0 = now_s();
  int *indeg = xcalloc((size_t)n, sizeof(int));
  for (int v = 0; v < n; ++v) indeg[v] = g.CP[v + 1] - g.CP[v];
  if (ug) {
    print_degrees("degree", indeg, n);
  } else {
    int *outdeg = xcalloc((size_t)n, sizeof(int));
    for (int e = 0; e < nz; ++e) outdeg[g.IC[e]]++;
    print_degrees("in_degree", indeg, n);
    print_degrees("out_degree", outdeg, n);
    free(outdeg);
  }
  const double degree_t = now_s() - t0;

  int *S = xcalloc((size_t)n, sizeof(int));
  uint64_t *sig = xcalloc((size_t)n, sizeof(uint64_t));
  uint8_t *ovf = xcalloc((size_t)n, 1);
  double seq_t = 0.0;
  if (seq) {
    t0 = now_s();
    seq_bfs(&g, r, S, sig, ovf);
    seq_t = now_s() - t0;
    printf("bfs %s seq %s time = %lfs\n", V == V_BU ? "bu" : "td", V == V_COOC ? "COOC" : "CSC",
           seq_t);
  }
  /* the sequential results in the reference's types and conventions */
  int *S_ref = xcalloc((size_t)n, sizeof(int));
  float *sig_f = xcalloc((size_t)n, sizeof(float));
  int *sig_i = xcalloc((size_t)n, sizeof(int));
  for (int v = 0; v < n; ++v) {
    S_ref[v] = (V == V_TD && S[v] < 0) ? 0 : S[v];
    sig_f[v] = ovf[v] ? INFINITY : (float)sig[v];
    sig_i[v] = (int)(uint32_t)sig[v];
  }

  int *S_hgpu = xcalloc((size_t)n, sizeof(int));
  float *sig_hgpu_f = xcalloc((size_t)n, sizeof(float));
  int *sig_hgpu_i = xcalloc((size_t)n, sizeof(int));
  const char *tag = format ? "wa" : "sc";
  int rc = 0;
  int visited = 0;
  if (V == V_TD) {
    rc = format ? bfs_gpu_td_csc_wa(g.IC, g.CP, S_hgpu, sig_hgpu_f, r, nz, n, repet)
                : bfs_gpu_td_csc_sc(g.IC, g.CP, S_hgpu, sig_hgpu_f, r, nz, n, repet);
    if (seq && (rc == 0 || rc == TBFS_SIGMA_INEXACT)) {
      printf("\ncheck sigma in CSC(%s) format: ", tag);
      sigma_check_f(sig_f, sig_hgpu_f, n);
      printf("\ncheck values of S_hgpu: ");
      S_check(S_ref, S_hgpu, n);
    }
    for (int v = 0; v < n; ++v) visited += sig_hgpu_f[v] != 0.0f;
  } else if (V == V_BU) {
    rc = format ? bfs_gpu_bu_csc_wa(g.IC, g.CP, S_hgpu, r, nz, n, repet, bbreak)
                : bfs_gpu_bu_csc_sc(g.IC, g.CP, S_hgpu, r, nz, n, repet, bbreak);
    if (seq && rc == 0) {
      printf("\ncheck values of S_hgpu: ");
      S_check(S_ref, S_hgpu, n);
    }
    for (int v = 0; v < n; ++v) visited += S_hgpu[v] >= 0;
  } else if (V == V_TDBU) {
    rc = format ? bfs_gpu_tdbu_csc_wa(g.IC, g.CP, S_hgpu, sig_hgpu_i, r, nz, n, repet, bbreak)
                : bfs_gpu_tdbu_csc_sc(g.IC, g.CP, S_hgpu, sig_hgpu_i, r, nz, n, repet, bbreak);
    if (seq && (rc == 0 || rc == TBFS_SIGMA_INEXACT)) {
      printf("\ncheck values of S_hgpu (%s): ", format ? "td-->bu" : "sc");
      S_check(S_ref, S_hgpu, n);
      /* the reference's TDBU stops counting paths after its switch; these
       * counts are complete, so they are checked too */
      printf("\ncheck sigma_hgpu (%s): ", tag);
      sigma_check_i(sig_i, sig_hgpu_i, n);
    }
    for (int v = 0; v < n; ++v) visited += S_hgpu[v] >= 0;
  } else {
    int *J = xcalloc((size_t)nz, sizeof(int));
    for (int v = 0; v < n; ++v)
      for (int e = g.CP[v]; e < g.CP[v + 1]; ++e) J[e] = v;
    rc = bfs_gpu_td_cooc_sc(g.IC, J, sig_hgpu_f, r, nz, n, repet);
    if (seq && (rc == 0 || rc == TBFS_SIGMA_INEXACT)) {
      printf("\ncheck sigma in COOC(sc) format: ");
      sigma_check_f(sig_f, sig_hgpu_f, n);
    }
    for (int v = 0; v < n; ++v) visited += sig_hgpu_f[v] != 0.0f;
    free(J);
  }
  if (rc != 0 && rc != TBFS_SIGMA_INEXACT) {
    fprintf(stderr, "tbfs_exec: the GPU BFS failed: %s (%d)\n", tbfs_status_str(rc), rc);
    return rc;
  }
  if (rc == TBFS_SIGMA_INEXACT)
    printf("\nnote: some path counts exceed the exact range of the output type\n");
  printf("\nvisited vertices == %d\n", visited);

  if (p) {
    if (V == V_TD) {
      print_S(S_ref, S_hgpu, n);
      print_sigma_f2("sigma   sigma_hgpu ", sig_f, sig_hgpu_f, n);
    } else if (V == V_BU) {
      print_S(S_ref, S_hgpu, n);
      print_sigma_f1("sigma_seq  ", sig_f, n);
    } else if (V == V_TDBU) {
      print_S(S_ref, S_hgpu, n);
      print_sigma_i1("sigma  ", sig_i, n);
    } else {
      print_sigma_f2("sigma   sigma_hgpu ", sig_f, sig_hgpu_f, n);
    }
  }
  printf("\nscan=%2.3lfs,degree_t= %2.3lfs,transf_format_t=%2.3lfs,bfs_%s_seq_t=%2.3lfs\n",
         scan_t, degree_t, 0.0, V == V_BU ? "bu" : "td", seq_t);
  free(indeg);
  free(S);
  free(sig);
  free(ovf);
  free(S_ref);
  free(sig_f);
  free(sig_i);
  free(S_hgpu);
  free(sig_hgpu_f);
  free(sig_hgpu_i);
  free(g.CP);
  free(g.IC);
  return 0;
}
