// tbfs_engine.h -- host-side objects behind the C API (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "tbfs.h"
#include "tbfs_internal.h"

// A graph resident in HBM.
// What the chained loop's launch plan learns about a graph from the BFS run
// on it (by any context): decayed per-level votes of the sources so far
// (each weighs 0.95 of the one after it) and what they imply.
struct PlanHistory {
  std::mutex mu;
  float stage_votes[64] = {};  // sources whose level d staged the bitmap
  float level_votes[64] = {};  // sources whose BFS ran level d at all
  float work_votes[64] = {};   // sum of their level d's work (log2 entries)
  uint64_t stage = ~0ull;      // levels launched on the staging shape
  uint64_t dual = 0;           // levels launched in both shapes
  uint8_t work[64] = {};       // expected log2 work of level d (0: unknown)
  int levels = 0;              // deepest BFS so far
  bool valid = false;          // no BFS learned from yet: every launch stages
};

struct tbfs_graph {
  int32_t n = 0;
  int32_t nz = 0;
  uint32_t flags = 0;
  int32_t* CP = nullptr;  // in-lists (CSC); null for an unsorted COOC graph
  int32_t* IC = nullptr;
  int32_t* RP = nullptr;  // out-lists; alias CP/IC when symmetric
  int32_t* JR = nullptr;
  int32_t* I = nullptr;   // COOC source of each entry (aliases IC)
  int32_t* J = nullptr;   // COOC destination of each entry
  bool J_owned = false;
  bool has_csc = true;
  uint64_t* bigc = nullptr;  // static pull slices of high in-degree columns
  int32_t nbigc = 0;
  uint64_t* bigo = nullptr;  // static push slices of long out-lists (may alias bigc)
  int32_t nbigo = 0;
  int32_t nheavy_in = 0;     // columns cut into bigc slices (in-degree > kBigIn)
  int32_t nheavy_out = 0;    // vertices cut into bigo slices (out-degree > kBigIn)
  uint64_t* skip = nullptr;  // static: in-degree 0 or > kBigIn (one bit per vertex)
  uint64_t* bsl = nullptr;   // batched pull slices of the in-lists (v | slice << 32),
  int32_t nbsl = 0;          // built on first use
  // a batch on this graph needed 64-bit counts: start with them (batches on
  // several threads read and set it)
  std::atomic<bool> wide_counts{false};
  bool deg_in_valid = false; // deg_in computed (first use)
  tbfs_degree_stats deg_in{};
  // batches: a copy numbered by decreasing in-degree (hubs' state rows
  // together), built on first use; bperm maps an id of this graph to its id
  // there; bview_default: -1 not decided, 0 / 1 whether batches use it by
  // default (skewed degrees)
  tbfs_graph* bview = nullptr;
  int32_t* bperm = nullptr;
  int bview_default = -1;
  mutable PlanHistory plan;  // chained launch plan, shared by every context
};

// Per-BFS device state bound to one stream.
struct tbfs_ctx {
  int32_t n_cap = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;           // false: the stream belongs to the creator
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int32_t* S = nullptr;
  uint64_t* sigma = nullptr;
  uint8_t* ovf = nullptr;
  uint64_t* fb[3] = {nullptr, nullptr, nullptr};
  uint64_t* vis = nullptr;          // reached-or-skipped bitmap
  uint64_t* q[2] = {nullptr, nullptr};
  int64_t q_cap = 0;
  tbfs::LevelCtr* ctr = nullptr;    // device, 3 rotating slots (+1 spare)
  tbfs::LevelCtr* ctr_h = nullptr;  // pinned host mirror
  unsigned* bar = nullptr;          // grid barrier words (kBarWords)
  tbfs::CtrSet* cset = nullptr;     // device, 3 striped counter sets
  tbfs::CtrSet* cset_h = nullptr;   // pinned host mirror of one set
  tbfs::RunStats* stats = nullptr;  // device, stats_cap slots
  tbfs::RunStats* stats_h = nullptr;
  int stats_cap = 0;
  std::vector<hipEvent_t> evs;
  hipEvent_t marks[16] = {};
  int device = 0;
  int persist_grid_max = 0;         // 0: no cooperative launch available
  int persist_grid_max1024 = 0;     // the same for 1024-thread workgroups
  tbfs::LoopState* loop = nullptr;  // [2] chained-loop state slots
  uint32_t* fin_list[2] = {nullptr, nullptr};  // the finisher's column lists
  unsigned* fin_ctr = nullptr;      // its list length and workgroup count
  tbfs::LoopState* loop_h = nullptr;
  int chain_k = 8;                  // levels enqueued per host check
  const tbfs_graph* async_g = nullptr;  // graph of the queued async runs
  size_t fbl_limit = 0;             // bytes of LDS for a staged frontier bitmap
  int cus = 0;
  int32_t last_n = 0;
  int32_t* roots_d = nullptr;       // roots of a multi-root BFS
  int32_t roots_cap = 0;
  // graph loop: the captured chunk and what it was captured with
  hipGraphExec_t gexec = nullptr;
  int gK = 0;
  int ggrid = 0;
  size_t gfbl = 0;
  uint64_t gstage = 0;
  uint64_t gdual = 0;
  uint8_t gwork[64] = {};
  int glevels = 0;
  tbfs::PersistParams gparams{};
  tbfs::RunStats* gstats = nullptr;  // the chunk's stats slot (device)
  int32_t* groots_h = nullptr;       // pinned roots, one per queued source
  int groots_cap = 0;
};

namespace tbfs {
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define HIPCHK(call)                                                      \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) return hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

// Device memory.  While a DevReuse object lives on the calling thread,
// allocations are taken from a per-device cache of freed blocks when one of
// a fitting size is there, and blocks handed out under it go back to the
// cache when freed (they must have no work pending on the device by then);
// otherwise both go straight to the runtime.  The reference's entry points
// upload, run and free everything on every call: with the cache, calls of
// the same sizes pay no hipMalloc / hipFree (each of which waits for the
// device).  The cache holds at most kDevReuseCap bytes; it is kept until
// the process ends.
int dev_alloc_bytes(void** p, size_t bytes);
void dev_free_bytes(void* p);
void dev_reuse_release();  // frees every cached block (blocks in use stay)
constexpr size_t kDevReuseCap = size_t(4) << 30;
struct DevReuse {
  DevReuse();
  ~DevReuse();
  DevReuse(const DevReuse&) = delete;
  DevReuse& operator=(const DevReuse&) = delete;
};

template <typename T>
inline int dev_alloc(T** p, size_t count) {
  if (count == 0) count = 1;
  void* q = nullptr;
  const int rc = dev_alloc_bytes(&q, count * (sizeof(T) > 0 ? sizeof(T) : 1));
  *p = static_cast<T*>(q);
  return rc;
}
inline int dev_alloc(void** p, size_t bytes) { return dev_alloc_bytes(p, bytes ? bytes : 1); }

template <typename T>
inline void dev_free(T*& p) {
  if (p) dev_free_bytes((void*)p);
  p = nullptr;
}

inline int grid_for(int64_t items, int block, int cap) {
  int64_t g = (items + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// The autotuner's table (tbfs_tune.hip).  Key: w = 0 for single-source
// runs, else the batch's 64-source words; the graph's in-degree statistics
// as floor(log2) + 1 of n, of the maximum and of twice the mean.
struct TuneKey {
  int32_t w, ln, lmax, lmean;
  bool operator==(const TuneKey& o) const {
    return w == o.w && ln == o.ln && lmax == o.lmax && lmean == o.lmean;
  }
};
struct TuneEntry {
  TuneKey key;
  int32_t grid_per_cu = 0;   // batches: workgroups per CU (0: work-sized)
  float batch_alpha = 0.f;   // batches: push threshold
  float alpha = 0.f;         // single source: push -> pull threshold
  float beta = 0.f;          // single source: pull -> push threshold
  float stage_ratio = 0.f;   // single source: stage the bitmap above ratio * n
  int32_t lb_grid = 0;       // single source: workgroups of non-staging levels
  double ms = 0.0;           // the winner's time and the defaults' (per run)
  double default_ms = 0.0;
};
int tune_key(tbfs_graph* g, int32_t words, hipStream_t s, TuneKey* k);
bool tune_find(const TuneKey& k, TuneEntry* e);
void tune_put(const TuneEntry& e);
bool tune_empty();

// Static kPullChunk-entry slices of every list of ptr[0..n] longer than
// min_deg (entries v | slice << 32, in vertex order), built on stream s
// (returns with the table complete, *out allocated with *count entries).
int build_slices(const int32_t* ptr, int32_t n, int32_t min_deg, hipStream_t s, uint64_t** out,
                 int32_t* count, int32_t* nlists = nullptr);

// The graph a batch runs on: g itself, or its degree-ordered copy (order:
// the batch_order option) with *perm mapping g's ids to the copy's (else
// null); built on first use, on stream s (waits for it).
int graph_batch_view(tbfs_graph* g, int32_t order, hipStream_t s, tbfs_graph** view,
                     const int32_t** perm);
// Degree statistics of g's in-lists (which == TBFS_DEGREE_IN) or out-lists,
// computed on stream s (waits for it); the in-list statistics are kept in g.
int graph_degree_stats(tbfs_graph* g, int32_t which, hipStream_t s, tbfs_degree_stats* out);
int graph_create_csc(const int32_t* CP, const int32_t* IC, int32_t n,
                     int32_t nz, uint32_t flags, tbfs_graph** out);
int graph_create_cooc(const int32_t* I, const int32_t* J, int32_t n,
                      int32_t nz, uint32_t flags, tbfs_graph** out);
void graph_free(tbfs_graph* g);
int ctx_create(int32_t n_cap, tbfs_ctx** out);
// A context whose work goes to an existing stream (not destroyed with it).
int ctx_create_on(int32_t n_cap, hipStream_t stream, tbfs_ctx** out);
void ctx_free(tbfs_ctx* c);
int run_bfs(tbfs_ctx* c, const tbfs_graph* g, int32_t root,
            const tbfs_opts& o, tbfs_stats* st);
int run_bfs_many(tbfs_ctx* c, const tbfs_graph* g, const int32_t* roots,
                 int count, const tbfs_opts& o, tbfs_stats* st);
int run_multi(tbfs_ctx* c, const tbfs_graph* g, const int32_t* roots, int nroots,
              const tbfs_opts& o, tbfs_stats* st);

// asynchronous single-source runs (query server)
void plan_learn(const tbfs_graph* g, const RunStats* st, int count);
int ctx_async_begin(tbfs_ctx* c, int count);
int ctx_async_enqueue(tbfs_ctx* c, const tbfs_graph* g, int32_t root, const tbfs_opts& o,
                      int idx, int levels, int32_t* S, uint64_t* sigma, uint8_t* ovf);
int ctx_async_finish(tbfs_ctx* c, int count);
int ctx_async_poll(tbfs_ctx* c, int count, bool* ready, std::vector<int>* unfinished,
                   int* maxl);

// batched sources (tbfs_batch.hip)
int batch_words_for(int32_t count);
int batch_create(int32_t n_cap, int32_t b_cap, tbfs_batch** out);
int batch_create_words(int64_t cap_words, tbfs_batch** out);
void batch_free(tbfs_batch* b);
int64_t batch_cap_words(const tbfs_batch* b);
hipEvent_t batch_step_event(const tbfs_batch* b);
hipStream_t batch_stream(const tbfs_batch* b);
int batch_start(tbfs_batch* b, const tbfs_graph* g, const int32_t* roots, int32_t count,
                const tbfs_opts* o);
int batch_progress(tbfs_batch* b, bool* finished);
void batch_fill_stats(const tbfs_batch* b, tbfs_stats* st);
int batch_run(tbfs_batch* b, const tbfs_graph* g, const int32_t* roots, int32_t count,
              const tbfs_opts* o, tbfs_stats* st);
// Continuous batches (the query server): a batch of `words` 64-source words
// whose level loop keeps running while sources come and go.  Sources are
// seeded into free bits at the current boundary (the first level not yet
// enqueued, batch_open_depth); batch_open_chunk enqueues `levels` levels
// and a probe of the boundary after them; batch_open_poll returns the
// oldest unread probe once the device has passed it: the boundary depth D,
// per word the bits present in the frontier of depth D, and that
// frontier's word count.  A source seeded at d0 < D whose bit is absent
// from the probe of D is finished: batch_open_extract enqueues its row
// (depths relative to d0) to host destinations and clears its bit for
// reuse.  64-bit counts throughout.
int batch_reserve(tbfs_batch* b, const int32_t* ns, int32_t count);
int batch_open_begin(tbfs_batch* b, const tbfs_graph* g, int32_t words, const tbfs_opts* o);
int batch_open_seed(tbfs_batch* b, const int32_t* bits, const int32_t* roots, int32_t count);
int batch_open_extract(tbfs_batch* b, const int32_t* bits, const int32_t* offs, int32_t count,
                       int32_t* const* S_dst, uint64_t* const* sig_dst, uint8_t* const* ovf_dst);
int batch_open_chunk(tbfs_batch* b, int32_t levels, bool wide);
int batch_open_poll(tbfs_batch* b, bool* ready, int32_t* boundary, uint64_t* alive,
                    int64_t* frontier_words);
int batch_open_outstanding(const tbfs_batch* b);
int32_t batch_open_depth(const tbfs_batch* b);
int batch_open_end(tbfs_batch* b);
int batch_enqueue_rows(tbfs_batch* b, int32_t first, int32_t count, int32_t* const* S_dst,
                       uint64_t* const* sig_dst, uint8_t* const* ovf_dst);
int batch_read(tbfs_batch* b, int32_t first, int32_t count, int32_t n, int32_t* S,
               uint64_t* sigma, uint8_t* ovf);
int ctx_read(tbfs_ctx* c, int32_t n, int32_t* S, uint64_t* sigma,
             uint8_t* ovf);
int ctx_read_converted(tbfs_ctx* c, int32_t n, int32_t s_unreached, int32_t kind,
                       int32_t* S_out, void* val_out, int64_t* inexact);
}  // namespace tbfs
