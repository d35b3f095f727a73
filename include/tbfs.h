/*
 * tbfs.h -- extended C API of libturbobfs (MI355X / gfx950).
 *
 * The reference (pcdslab/TurboBFS) has no library API beyond its per-variant
 * GPU entry points (bfstdcsc.h, bfsbucsc.h, bfstdbucsc.h, bfstdcooc.h, also
 * exported by this library).  Those re-upload the graph on every call.  This
 * header adds what a resident-graph caller needs: graphs kept in HBM, a
 * reusable BFS context bound to one HIP stream, a choice of level kernel and
 * level loop, and exact 64-bit path counts with an overflow flag.
 *
 * Conventions
 *   - CSC in-lists: column v of (CP, IC) lists every u with an edge u -> v,
 *     as in the reference (TurboBFS_CSC_TD/bfstdcsc.h:30-34).
 *   - S[v] = depth of v, -1 when v is not reached, S[root] = 0.
 *   - sigma[v] = number of shortest root->v paths modulo 2^64; ovf[v] = 1 when
 *     the exact count is >= 2^64 (sigma then holds it modulo 2^64).
 *   - Every function returns TBFS_OK (0) or a nonzero status; nothing exits
 *     the process.
 */
#ifndef TBFS_H
#define TBFS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TBFS_OK 0
#define TBFS_ERR_INVALID 1     /* bad argument or malformed graph */
#define TBFS_ERR_HIP 2         /* a HIP runtime call failed */
#define TBFS_ERR_NOMEM 3       /* device or host allocation failed */
#define TBFS_ERR_UNSUPPORTED 4 /* mode not available for this graph */
#define TBFS_SIGMA_INEXACT 5   /* (reference entry points only) a path count
                                  exceeds what the reference's output type
                                  holds exactly: float above 2^24, int32 above
                                  2^31-1; the outputs are still written */

/* graph flags */
#define TBFS_GRAPH_SYMMETRIC 1u /* A == A^T: out-lists alias the in-lists */
#define TBFS_GRAPH_DEVICE 2u    /* input arrays are device pointers */

/* level kernel selection (tbfs_opts.mode) */
#define TBFS_MODE_AUTO 0    /* direction-optimizing: push <-> pull */
#define TBFS_MODE_PULL_SC 1 /* pull, one lane per column, every level */
#define TBFS_MODE_PULL_WA 2 /* pull, one wavefront per column, every level */
#define TBFS_MODE_PUSH 3    /* top-down push from a queue, every level */
#define TBFS_MODE_COOC 4    /* edge-parallel over the COOC nonzeros */

/* kernels used by TBFS_MODE_AUTO / TBFS_MODE_PUSH (tbfs_opts.pull) */
#define TBFS_PULL_SC 0 /* pull: one lane per column; push: one wavefront per slice */
#define TBFS_PULL_WA 1 /* pull: one wavefront per column; push: as above */
#define TBFS_PULL_LB 2 /* load-balanced: wavefronts split work by edge count */

/* level loop (tbfs_opts.loop) */
#define TBFS_LOOP_HOST 0       /* host launches each level, reads its counters */
#define TBFS_LOOP_PERSISTENT 1 /* one cooperative launch per BFS, levels
                                  separated by a device-wide barrier; when
                                  the runtime declines the cooperative
                                  launch the BFS runs the chained loop */
#define TBFS_LOOP_CHAINED 2    /* one kernel per level, enqueued back to back;
                                  each decides its direction from the device
                                  counters, the host checks for the end once
                                  per chunk of levels */
#define TBFS_LOOP_GRAPH 3      /* the chained kernels of one BFS (init,
                                  root, a budget of levels) captured once as
                                  a HIP graph and replayed per source */
/* PERSISTENT, CHAINED and GRAPH apply to AUTO / PUSH with TBFS_PULL_LB;
 * other combinations run the host loop. */

typedef struct tbfs_graph tbfs_graph;
typedef struct tbfs_ctx tbfs_ctx;

typedef struct tbfs_opts {
  int32_t mode;   /* TBFS_MODE_* */
  int32_t loop;   /* TBFS_LOOP_* */
  int32_t sigma;  /* 1: compute path counts, 0: depths only */
  int32_t bbreak; /* depths only: stop a column scan at its first parent */
  float alpha;    /* AUTO: switch push->pull when m_frontier*alpha > m_unexplored */
  float beta;     /* AUTO: switch pull->push when n_frontier*beta < n */
  int32_t block;  /* threads per workgroup, multiple of 64 (0: default);
                     the persistent loop takes 256 or 1024 */
  int32_t grid;   /* workgroups (0: default) */
  int32_t pull;   /* TBFS_PULL_*: AUTO's bottom-up kernel */
  int32_t fb_lds; /* chained loop: 0 stages the frontier bitmap in LDS for
                     pull levels when it fits (n <= ~1.1M), -1 never */
  float batch_alpha; /* batches: expand top-down while the frontier's
                        out-degree sum * batch_alpha < nz * words (0: 32) */
  int32_t batch_count_bits; /* batches: 0 = 32-bit counts, again with 64-bit
                               ones if some count reaches 2^32; 64 = 64-bit
                               counts only; 32 = 32-bit only (a count that
                               reaches 2^32 is then flagged as overflowed) */
  int32_t batch_order; /* batches: 0 = run on a copy of the graph with the
                          vertices numbered by decreasing in-degree when the
                          degrees are skewed (sd > 2 x mean), 1 = always,
                          -1 = never; results are in the caller's numbering */
  float stage_ratio;   /* chained loop: pull levels stage the frontier bitmap
                          in LDS when the unexplored in-list entries exceed
                          stage_ratio * n */
  int32_t lb_grid;     /* chained loop: workgroups of the launches that do
                          not stage the bitmap; -1: one grid per level, from
                          that level's work in the context's earlier runs on
                          the graph (one workgroup per CU up to 2^18 list
                          entries, two up to 2^22, else about one wavefront
                          per four frontier words) */
  int32_t finish;      /* chained loop, AUTO mode: once at most this many
                          in-list entries are unexplored, one workgroup runs
                          the remaining levels (0 or -1: never; at most
                          65536 on graphs of more than 65536 vertices) */
  /* alpha, beta, stage_ratio and lb_grid left at 0 take the autotuner's
   * values for the graph's class (tbfs_autotune) when the table holds them,
   * else alpha 14, beta 24, stage_ratio 1 and lb_grid -1. */
} tbfs_opts;

typedef struct tbfs_stats {
  int32_t levels;      /* number of distinct depths (1 + eccentricity of root) */
  int32_t push_levels; /* levels expanded top-down (push) */
  int32_t pull_levels; /* levels expanded bottom-up (pull) or edge-parallel */
  int32_t pad;
  int64_t reached;     /* vertices with S >= 0 */
  int64_t edges;       /* m_s: in-list entries of the reached vertices */
  int64_t ovf64;       /* -1 when some reached count is >= 2^64, else 0 */
  double ms;           /* device time of the BFS, first to last kernel */
  int32_t launches;    /* chained and graph loops: level kernels the BFS used
                          (the one that found the frontier empty included);
                          0 for the other loops and for batches */
  int32_t pad2;
} tbfs_stats;

/* Fill *o with the defaults: AUTO, load-balanced kernels, chained loop,
 * sigma on; alpha, beta, stage_ratio, lb_grid 0 (tuned, else 14, 24, 1 and
 * -1: a grid per level). */
void tbfs_opts_default(tbfs_opts *o);

/* Human-readable name of a status code. */
const char *tbfs_status_str(int status);

/* Upload a CSC graph (host arrays unless TBFS_GRAPH_DEVICE).  The CSC is
 * validated (CP[0] = 0, CP non-decreasing, CP[n] = nz, 0 <= IC < n).  Unless
 * TBFS_GRAPH_SYMMETRIC is set, the out-list transpose is built on the device
 * (stable radix sort, so each out-list is in increasing column order). */
int tbfs_graph_create(const int32_t *CP, const int32_t *IC, int32_t n,
                      int32_t nz, uint32_t flags, tbfs_graph **out);

/* Upload a COOC graph: edge I[k] -> J[k], k < nz (TurboBFS_COOC_TD/bfstdcooc.h:29-32).
 * When J is non-decreasing the CSC column pointers are derived too, so every
 * mode is available; otherwise only TBFS_MODE_COOC is. */
int tbfs_graph_create_cooc(const int32_t *I, const int32_t *J, int32_t n,
                           int32_t nz, uint32_t flags, tbfs_graph **out);

int tbfs_graph_destroy(tbfs_graph *g);

/* n, nz and flags of a resident graph (any pointer may be NULL). */
int tbfs_graph_info(const tbfs_graph *g, int32_t *n, int32_t *nz,
                    uint32_t *flags);

/* Degree statistics of a resident graph, computed on the device: what the
 * reference's drivers print before a run (TurboBFS_CSC_TD/degree.c:38-141:
 * degree_max, degree_sta, degree_dist / degree_dist_ac, called from
 * bfstdcsc_main.c:173-230).  `which`: TBFS_DEGREE_IN (column lengths, the
 * in-neighbours; on a symmetric graph this is the reference's undirected
 * `degree` vector) or TBFS_DEGREE_OUT (out-list lengths).  max / argmax: the
 * largest degree and the first vertex that has it; mean over all n
 * vertices; sd: sample standard deviation (divides by n - 1; 0 when n < 2);
 * bins: bins[0] counts degree 0, bins[k] (k >= 1) degrees in
 * [2^(k-1), 2^k), so bins[1] is degree 1 and bins[2] degrees 2..3, as the
 * reference's accumulated distribution groups them; every vertex is counted
 * (the reference's loop leaves out degree == max when max is a power of two).
 * nbins: highest nonempty bin + 1.  The statistics also key the batch
 * autotuner's table (tbfs_batch_autotune). */
#define TBFS_DEGREE_IN 0
#define TBFS_DEGREE_OUT 1
typedef struct tbfs_degree_stats {
  int32_t max;
  int32_t argmax;
  double mean;
  double sd;
  int64_t sum;
  int64_t bins[33];
  int32_t nbins;
  int32_t reserved[3];
} tbfs_degree_stats;

int tbfs_graph_degree_stats(const tbfs_graph *g, int32_t which,
                            tbfs_degree_stats *out);

/* A BFS context: one HIP stream plus per-BFS state for graphs of up to
 * n_cap vertices.  One context runs one BFS at a time. */
int tbfs_ctx_create(int32_t n_cap, tbfs_ctx **out);
int tbfs_ctx_destroy(tbfs_ctx *c);

/* The context's hipStream_t (as void*), for callers that time or order work. */
void *tbfs_ctx_stream(tbfs_ctx *c);

/* Run one BFS from root on g.  Results stay in the context's device buffers
 * until the next BFS; st (may be NULL) receives the counters. */
int tbfs_bfs(tbfs_ctx *c, const tbfs_graph *g, int32_t root,
             const tbfs_opts *o, tbfs_stats *st);

/* Run `count` BFS back to back (roots[i]), with no host synchronisation in
 * between when the persistent or chained loop applies.  Results of the last BFS stay in
 * the context; st (may be NULL) receives count entries, each with its own
 * device time. */
int tbfs_bfs_many(tbfs_ctx *c, const tbfs_graph *g, const int32_t *roots,
                  int32_t count, const tbfs_opts *o, tbfs_stats *st);

/* One BFS whose depth-0 frontier is every vertex in roots[0..nroots-1]
 * (repeats allowed): S[v] = distance to the nearest root, sigma[v] = number
 * of shortest paths from any root.  Packing many graphs as diagonal blocks of
 * one CSC with one root per block gives each block exactly its own
 * single-source result in one pass.  Runs the chained loop (mode AUTO or
 * PUSH); results stay in the context as for tbfs_bfs. */
int tbfs_bfs_multi(tbfs_ctx *c, const tbfs_graph *g, const int32_t *roots,
                   int32_t nroots, const tbfs_opts *o, tbfs_stats *st);

/* One BFS from root on g that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory during a call
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  The results go to the caller's device buffers S (n int32),
 * sigma (n uint64) and ovf (n uint8), with the conventions of tbfs_bfs; the
 * level loop's working state lives in a fixed device array of the library,
 * TBFS_NOALLOC_SCRATCH_BYTES per GPU, held from the first call on a GPU until
 * the process ends, so graphs of up to TBFS_NOALLOC_MAX_N vertices are
 * taken (larger ones are rejected with TBFS_ERR_INVALID and a message).
 * Calls on one GPU share that array: the library runs them one at a time
 * per GPU, and each returns with its work on the device complete.  opts may
 * be NULL (defaults); modes AUTO and PUSH, chained level loop.  stream is a
 * hipStream_t (NULL: the null stream).  st->ms is the call's host time. */
#define TBFS_NOALLOC_MAX_N (1 << 24)
#define TBFS_NOALLOC_SCRATCH_BYTES (8 * 1024 * 1024 + 64 * 1024)
int tbfs_bfs_noalloc(const tbfs_graph *g, int32_t root, int32_t *S, uint64_t *sigma,
                     uint8_t *ovf, void *stream, const tbfs_opts *opts, tbfs_stats *st);

/* Copy the last BFS's results to host memory (any pointer may be NULL;
 * n must equal that graph's vertex count). */
int tbfs_ctx_read(tbfs_ctx *c, int32_t n, int32_t *S, uint64_t *sigma,
                  uint8_t *ovf);

/* Stream-ordered timing: record marker `slot` (0..15) on the context's stream;
 * tbfs_ctx_elapsed waits for marker b and gives the device time from a to b.
 * (Lets a caller time work on the library's own stream without another
 * runtime in the process.) */
int tbfs_ctx_mark(tbfs_ctx *c, int32_t slot);
int tbfs_ctx_elapsed(tbfs_ctx *c, int32_t a, int32_t b, double *ms);

/* Copy `bytes` (rounded down to 16) between two fresh device buffers with a
 * plain streaming kernel, `reps` times, on the context's stream; *ms gets the
 * average kernel time.  A known-size reference for hardware byte counters
 * (each launch reads `bytes` and writes `bytes`) and for achievable HBM
 * bandwidth. */
int tbfs_calibrate_copy(tbfs_ctx *c, int64_t bytes, int32_t reps, double *ms);

/* Number of visible HIP devices (0 when none; never fails for lack of one). */
int tbfs_device_count(int32_t *count);

/* Properties of device dev: name (NUL-terminated, up to name_len bytes),
 * compute units, HBM bytes, cooperative-launch support (any may be NULL). */
int tbfs_device_info(int32_t dev, char *name, int32_t name_len, int32_t *cus,
                     int64_t *hbm_bytes, int32_t *coop);

/* Device pointers of the context's S / sigma / ovf buffers. */
int tbfs_ctx_device_ptrs(tbfs_ctx *c, int32_t **S, uint64_t **sigma,
                         uint8_t **ovf);

/* ---- batched sources ---------------------------------------------------
 * A batch runs up to 1024 BFS from different roots on one graph at once: the
 * frontiers are bit-packed, 64 sources per 64-bit word, and every level is
 * one pass over the in-lists for all sources.  Each source's S and sigma are
 * exactly those of tbfs_bfs from that root.  A batch object owns its own HIP
 * stream and device state sized for n_cap vertices times b_cap sources
 * rounded up to a power-of-two number of 64-source words (b_cap <= 1024):
 * 828 bytes per vertex per word.  A batch of `count` sources on a graph of n
 * vertices fits when n * words(count) <= n_cap * words(b_cap). */
typedef struct tbfs_batch tbfs_batch;

int tbfs_batch_create(int32_t n_cap, int32_t b_cap, tbfs_batch **out);
int tbfs_batch_destroy(tbfs_batch *b);
void *tbfs_batch_stream(tbfs_batch *b);

/* BFS from roots[0..count-1] (1 <= count <= b_cap; roots may repeat) on the
 * CSC graph g.  opts may be NULL; only opts->grid (workgroups per level
 * kernel, 0: default) and opts->batch_alpha are used.  st (may be NULL)
 * receives one entry for the whole batch: levels = max over sources of their
 * level count, push_levels / pull_levels = levels expanded top-down /
 * bottom-up, pad = count width of the final run (32 or 64), reached and
 * edges summed over sources, ovf64 = -1 if some count reached 2^64 (or 2^32
 * with batch_count_bits = 32), ms = device time of the batch (both runs when
 * the counts had to be widened). */
int tbfs_batch_run(tbfs_batch *b, const tbfs_graph *g, const int32_t *roots,
                   int32_t count, const tbfs_opts *opts, tbfs_stats *st);

/* Autotune a batch for graph g and these roots: times the batch with launch
 * grids {default, 8, 16, 32 workgroups per CU} x push thresholds
 * batch_alpha {4, 32, 128} (best of two runs each; every candidate gives
 * identical results) and returns the fastest in *best (the other fields from
 * base, or the defaults when base is NULL) and its time in *best_ms.  Costs
 * about 25 batch runs; the batch object then holds the last run's results.
 * The winner is also recorded in a process-wide table keyed by the graph's
 * degree statistics (log2 of n, of the maximum and of the mean in-degree)
 * and the batch's number of 64-source words: a later tbfs_batch_run with
 * opts->grid == 0 and opts->batch_alpha == 0 (the defaults), on this graph
 * or any graph with the same key, uses the recorded grid and threshold. */
int tbfs_batch_autotune(tbfs_batch *b, const tbfs_graph *g, const int32_t *roots,
                        int32_t count, const tbfs_opts *base, tbfs_opts *best,
                        double *best_ms);

/* ---- autotuner -----------------------------------------------------------
 * Single-source runs: times `count` BFS from roots on g with the default
 * thresholds and launch shapes, then searches alpha {6,10,14,20,30}, beta
 * {8,16,24,48,96}, stage_ratio {0.25,0.5,1,2,4} and lb_grid {-1 (per
 * level), 1, 2, 4 workgroups per CU, default} one at a time, twice round (best of two runs
 * per candidate).  Every candidate gives the same results, which is checked
 * on each source's level, vertex and edge counts.  *best gets the winner
 * (other fields from base, or the defaults), *best_ms / *default_ms the time
 * per BFS of the winner and of the defaults.  The winner is recorded in the
 * table for g's class (the in-degree statistics key described at
 * tbfs_batch_autotune): later runs on such graphs with those fields at 0 take
 * it.  The context holds the last run's results afterwards. */
int tbfs_autotune(tbfs_ctx *c, const tbfs_graph *g, const int32_t *roots, int32_t count,
                  const tbfs_opts *base, tbfs_opts *best, double *best_ms,
                  double *default_ms);

/* The single-source values the table holds for g's class, written into *o
 * (alpha, beta, stage_ratio, lb_grid; other fields untouched): TBFS_OK, or
 * TBFS_ERR_UNSUPPORTED when no entry applies. */
int tbfs_tuned(const tbfs_graph *g, tbfs_opts *o);

/* The table (single-source and batch entries) as JSON: write it to `path`,
 * merge the entries of `path` into it, empty it, count its entries.  With
 * the environment variable TBFS_TUNE_CACHE set to a path, the table is read
 * from that file on first use and written back after every autotune. */
int tbfs_tune_save(const char *path);
int tbfs_tune_load(const char *path);
int tbfs_tune_clear(void);
int tbfs_tune_count(void);

/* The tuned grid / batch_alpha the table holds for a batch of `count`
 * sources on g, written into *o (other fields untouched).  Returns TBFS_OK
 * when an entry applies, TBFS_ERR_UNSUPPORTED when none does. */
int tbfs_batch_tuned(const tbfs_graph *g, int32_t count, tbfs_opts *o);

/* Empty the autotuner's table (as tbfs_tune_clear). */
int tbfs_batch_tune_clear(void);

/* Results of sources first .. first+count-1 of the last batch, one row of n
 * entries per source (S[i*n + v] is the depth of v from roots[first+i]; any
 * pointer may be NULL; n must equal the graph's vertex count). */
int tbfs_batch_read(tbfs_batch *b, int32_t first, int32_t count, int32_t n,
                    int32_t *S, uint64_t *sigma, uint8_t *ovf);

/* ---- serving -------------------------------------------------------------
 * A server keeps a set of graphs resident and answers single-source queries
 * submitted at any time.  Each of its slots (batch state in device memory
 * from mem_budget, a stream and a worker thread) takes the queries waiting
 * for one graph and runs them as single-source BFS, as one batch, or -- on
 * deep graphs -- as a continuous batch that queries of that graph join at
 * level boundaries while it runs.  Slots run concurrently.  Every query's
 * S / sigma / overflow flags are exactly those of tbfs_bfs from its root,
 * however it ran. */
typedef struct tbfs_server tbfs_server;

typedef struct tbfs_server_opts {
  int64_t mem_budget; /* device bytes for the slots (0: min(free/4, 64 GiB)),
                         split evenly over them; each slot's share covers its
                         batch state (828 bytes per vertex per 64-source word),
                         its staging of finished rows for the host copies (up
                         to 64 rows of 13 bytes per vertex, at most 256 MiB),
                         an id map (4 bytes per vertex) and its single-source
                         context (~14 bytes per vertex
                         plus 16 bytes per 64 list entries), all sized for the
                         largest graph.  Resident graphs and their per-graph
                         tables (slices, degree-ordered copies) are not
                         counted: they belong to the graphs */
  int32_t slots;      /* batches in flight at once (0: 4) */
  int32_t max_batch;  /* sources per batch, <= 1024 (0: 1024) */
  int32_t open_depth; /* graphs on which a BFS has needed more levels than
                         this run as continuous batches: queries that arrive
                         while one runs join it at the next level boundary
                         (0: 64; -1: never) */
  int32_t short_slots; /* slots kept from long jobs (deep graphs, single-
                         source queries over 5 ms) for short queries (0: 1
                         when slots >= 2; -1: none) */
  int32_t reserved[2];
} tbfs_server_opts;

typedef struct tbfs_query_done {
  uint64_t ticket;
  int32_t graph;
  int32_t root;
  int32_t status;    /* TBFS_OK or an error code */
  int32_t batch;     /* sources in the batch that answered it (a continuous
                        batch: those running in it when it finished) */
  double submit_us;  /* times on the server clock (microseconds since */
  double start_us;   /* creation): submitted, its run started, its      */
  double done_us;    /* results were in its host buffers               */
  int32_t level;     /* level of its continuous batch at which it was
                        seeded (0 for every other run) */
  int32_t pad;
} tbfs_query_done;

/* graphs[i] stay owned by the caller and must outlive the server. */
int tbfs_server_create(const tbfs_graph *const *graphs, int32_t ngraphs,
                       const tbfs_server_opts *opts, tbfs_server **out);
/* The mem_budget that gives `slots` slots room for `words` 64-source words
 * of the largest of graphs[0..ngraphs-1] each (what tbfs_server_create would
 * size from it). */
int tbfs_server_budget_for(const tbfs_graph *const *graphs, int32_t ngraphs,
                           int32_t slots, int32_t words, int64_t *bytes);
/* Stops the dispatcher (queries still queued are dropped) and frees slots. */
int tbfs_server_destroy(tbfs_server *s);
/* Queue a BFS from root on graphs[graph].  S (n int32), sigma (n uint64) and
 * ovf (n uint8) are host buffers that receive the result (any may be NULL;
 * pinned memory from tbfs_host_alloc makes the copies asynchronous); they
 * must stay valid until the query is reported done. */
int tbfs_server_submit(tbfs_server *s, int32_t graph, int32_t root, int32_t *S,
                       uint64_t *sigma, uint8_t *ovf, uint64_t *ticket);
/* Up to max completed queries, oldest first; waits up to wait_us for one. */
int tbfs_server_poll(tbfs_server *s, tbfs_query_done *done, int32_t max,
                     int32_t *ndone, int64_t wait_us);
/* Block until every submitted query has completed. */
int tbfs_server_drain(tbfs_server *s);
/* Hold (hold != 0) or release (0) the dispatch of queued queries: while
 * held, queries submitted wait in their queues and running continuous
 * batches admit none, so a group submitted together is taken together (up
 * to each graph's batch cap) once released; jobs already running go on.
 * tbfs_server_drain releases a hold. */
int tbfs_server_hold(tbfs_server *s, int32_t hold);
/* Sources per batch for graphs[graph] (graph >= 0), batches started and
 * queries they held so far, and the server clock. */
int tbfs_server_info(tbfs_server *s, int32_t graph, int32_t *batch_cap,
                     int64_t *batches, int64_t *batched, double *now_us);

/* ---- the reference's entry points -------------------------------------
 * Kernels behind bfs_gpu_td_csc_sc / _wa, bfs_gpu_bu_csc_sc / _wa,
 * bfs_gpu_tdbu_csc_sc / _wa and bfs_gpu_td_cooc_sc (process-wide; every
 * mapping gives the same results):
 *   TBFS_DROPIN_FAST (default): direction-optimizing, load-balanced levels
 *     on the chained loop, for every entry point (a COOC input whose J is
 *     not sorted runs the edge-parallel sweep);
 *   TBFS_DROPIN_REFERENCE: the reference's own mappings on a host-driven
 *     loop -- TD and BU sweep every unreached column every level, one lane
 *     per column (_sc) or one wavefront per column (_wa); TDBU switches
 *     direction (load-balanced levels for _sc, wavefront-per-column pulls
 *     for _wa); COOC runs the edge-parallel sweep.
 * Returns the previous mapping (mapping -1: only query it). */
#define TBFS_DROPIN_FAST 0
#define TBFS_DROPIN_REFERENCE 1
int tbfs_dropin_mapping(int mapping);

/* The seven entry points keep what one call allocates for the next: idle BFS
 * contexts (up to 4) and the device buffers of their graph uploads (up to
 * 4 GiB), so repeated calls pay no device allocation or free.  This releases
 * all of it (calls in progress keep theirs); later calls allocate again. */
int tbfs_dropin_release(void);

/* Pinned host memory (hipHostMalloc) for result buffers. */
void *tbfs_host_alloc(int64_t bytes);
int tbfs_host_free(void *p);

/* Device memory for callers without a HIP allocator of their own (for
 * instance the output buffers of tbfs_bfs_noalloc): allocate, free, copy in
 * any direction (hipMemcpy, synchronous), fill bytes. */
void *tbfs_device_alloc(int64_t bytes);
int tbfs_device_free(void *p);
int tbfs_device_copy(void *dst, const void *src, int64_t bytes);
int tbfs_device_fill(void *dst, int value, int64_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* TBFS_H */
