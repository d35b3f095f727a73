// tbfs_serve.hip -- BFS query serving: resident graphs, continuous batching.
//
// A server holds a set of graphs resident in HBM and a fixed number of batch
// slots (tbfs_batch objects, one HIP stream each) sized from a device-memory
// budget.  Queries (graph, root, host destinations) are queued per graph as
// they arrive.  One dispatcher thread owns every HIP call after creation:
//   - it advances the slots in flight without blocking (event queries): a
//     slot whose BFS finished enqueues the copy of each query's row (S,
//     sigma, overflow flags) to that query's host buffers; when the copy
//     event completes, the queries are reported done;
//   - whenever a slot is idle it admits work: the graph whose oldest waiting
//     query is oldest, with every query waiting for that graph up to the
//     slot's capacity for it (the largest power-of-two number of 64-source
//     words whose state fits the slot's memory, at most max_batch sources);
//     those queries run as one batch, or one after another as single-source
//     BFS on the slot's context when that costs less for so few queries (by
//     the per-graph costs measured on earlier admissions).
// Queries that arrive while slots are busy join the next batch formed for
// their graph; slots run concurrently on their own streams, so batches of
// different graphs (or of one graph) overlap on the device.  A query's
// result does not depend on the batch it ran in: every source of a batch is
// exactly its own single-source BFS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "tbfs.h"
#include "tbfs_engine.h"

using namespace tbfs;

namespace {

// Device bytes per 64-source word of one vertex in a batch slot: frontier x3,
// seen x2, overflow bits (8 bytes each), word lists x3 (4 bytes each), 64
// depths (4) + 64 counts (8).
constexpr int64_t kSlotBytesPerWord = 6 * 8 + 3 * 4 + 64 * (4 + 8);

// Device bytes a slot holds besides its per-word batch state, for graphs of
// up to nmax vertices and nzmax entries: the rows of finished queries staged
// for their copies to the host (up to 64 rows of 13 bytes per vertex, at
// most 256 MiB, at least one row) and the slot's single-source context
// (depths 4, counts 8, flags 1 per vertex; five bitmaps; two work queues of
// n + nz/64 entries of 8 bytes).
int64_t slot_fixed_bytes(int64_t nmax, int64_t nzmax) {
  const int64_t row = 13 * nmax;
  const int64_t stage = row * std::max<int64_t>(1, std::min<int64_t>(64, ((int64_t)256 << 20) / row));
  const int64_t ctx = 13 * nmax + 5 * 8 * ((nmax + 63) / 64) + 2 * 8 * (nmax + nzmax / 64 + 64);
  return stage + ctx + ((int64_t)1 << 20);  // counters, loop state, alignment
}

struct Query {
  uint64_t ticket = 0;
  int32_t graph = 0;
  int32_t root = 0;
  int32_t* S = nullptr;
  uint64_t* sigma = nullptr;
  uint8_t* ovf = nullptr;
  double t_submit = 0;
};

struct Slot {
  tbfs_batch* b = nullptr;
  tbfs_ctx* c = nullptr;  // single-source runs (small batches)
  int state = 0;  // 0 idle, 1 batch BFS running, 2 rows being copied, 3 single runs
  int graph = -1;
  int words = 0;  // 64-source words of the running batch (0: single runs)
  double t_start = 0;
  std::vector<Query> q;
  std::vector<int32_t> roots;  // kept until the batch has finished
  std::vector<int> redo;
};

}  // namespace

struct tbfs_server {
  std::vector<const tbfs_graph*> graphs;
  std::vector<int> cap;  // sources per batch for each graph
  // measured cost per graph (ms, host clock, results included): one
  // single-source query, and one batch of 1, 2, 4, 8, 16 words (0: unknown)
  std::vector<double> t_single;
  std::vector<std::array<double, 5>> t_batch;
  // completed runs of each kind: the first of each pays one-time costs (the
  // graph's batch slice table, a re-run with wider counts, the single-source
  // level budget found by doubling) and is not taken as a cost sample
  std::vector<int> n_single;
  std::vector<std::array<int, 5>> n_batch;
  std::vector<int> budget;  // level kernels per single-source run, per graph
  int32_t nmax = 0;
  std::vector<std::deque<Query>> pending;
  std::deque<tbfs_query_done> completed;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread th;
  std::atomic<bool> stop{false};
  int64_t in_flight = 0;  // submitted, not yet completed
  uint64_t next_ticket = 1;
  int device = 0;
  tbfs_opts bopts{};
  std::chrono::steady_clock::time_point t0;
  // written by the dispatcher, read by tbfs_server_info on caller threads
  std::atomic<uint64_t> batches{0}, batched{0};

  double now_us() const {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0)
        .count();
  }
};

namespace {

int word_index(int words) {
  int k = 0;
  while ((1 << k) < words) ++k;
  return k;
}

void complete(tbfs_server* s, Slot& sl, int status, double t_done) {
  if (status == TBFS_OK && sl.graph >= 0 && !sl.q.empty()) {
    const double ms = (t_done - sl.t_start) * 1e-3;
    const int wi = word_index(sl.words);
    double& t = sl.words ? s->t_batch[sl.graph][wi] : s->t_single[sl.graph];
    int& runs = sl.words ? s->n_batch[sl.graph][wi] : s->n_single[sl.graph];
    const double x = sl.words ? ms : ms / (double)sl.q.size();
    if (runs++ > 0) t = t > 0 ? 0.7 * t + 0.3 * x : x;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  for (const Query& q : sl.q) {
    tbfs_query_done d{};
    d.ticket = q.ticket;
    d.graph = q.graph;
    d.root = q.root;
    d.status = status;
    d.batch = (int32_t)sl.q.size();
    d.submit_us = q.t_submit;
    d.start_us = sl.t_start;
    d.done_us = t_done;
    s->completed.push_back(d);
  }
  s->in_flight -= (int64_t)sl.q.size();
  sl.q.clear();
  sl.state = 0;
  sl.graph = -1;
  sl.words = 0;
  s->cv_done.notify_all();
}

// Advance one slot; true if anything changed.
bool step_slot(tbfs_server* s, Slot& sl) {
  if (sl.state == 1) {
    bool fin = false;
    int rc = batch_progress(sl.b, &fin);
    if (rc) {
      complete(s, sl, rc, s->now_us());
      return true;
    }
    if (!fin) return false;
    {  // a single-source run on this graph needs at most the batch's depth
       // in launches, plus the one that finds the frontier empty
      tbfs_stats bst{};
      batch_fill_stats(sl.b, &bst);
      int& bud = s->budget[sl.graph];
      bud = std::max(bud, std::min(1 << 16, bst.levels + 2));
    }
    const size_t k = sl.q.size();
    std::vector<int32_t*> sd(k);
    std::vector<uint64_t*> gd(k);
    std::vector<uint8_t*> od(k);
    for (size_t i = 0; i < k; ++i) {
      sd[i] = sl.q[i].S;
      gd[i] = sl.q[i].sigma;
      od[i] = sl.q[i].ovf;
    }
    rc = batch_enqueue_rows(sl.b, 0, (int32_t)k, sd.data(), gd.data(), od.data());
    if (rc) {
      complete(s, sl, rc, s->now_us());
      return true;
    }
    sl.state = 2;
    return true;
  }
  if (sl.state == 2) {
    const hipError_t e = hipEventQuery(batch_step_event(sl.b));
    if (e == hipErrorNotReady) return false;
    complete(s, sl, e == hipSuccess ? TBFS_OK : TBFS_ERR_HIP, s->now_us());
    return true;
  }
  if (sl.state == 3) {
    bool ready = false;
    int maxl = 0;
    int rc = ctx_async_poll(sl.c, (int)sl.q.size(), &ready, &sl.redo, &maxl);
    if (rc) {
      complete(s, sl, rc, s->now_us());
      return true;
    }
    if (!ready) return false;
    int& bud = s->budget[sl.graph];
    // the most launches a BFS needed plus one spare (never lowered: sources
    // of one graph differ in depth, and a run-out costs a whole BFS); twice
    // the budget if one ran out
    bud = std::max(bud, std::min(1 << 16, sl.redo.empty() ? maxl + 1 : 2 * bud));
    if (sl.redo.empty()) {
      complete(s, sl, TBFS_OK, s->now_us());
      return true;
    }
    // sources that ran out of levels: again, with the larger budget
    for (int i : sl.redo) {
      const Query& q = sl.q[i];
      rc = ctx_async_enqueue(sl.c, s->graphs[sl.graph], q.root, s->bopts, i, bud, q.S, q.sigma,
                             q.ovf);
      if (rc) break;
    }
    if (!rc) rc = ctx_async_finish(sl.c, (int)sl.q.size());
    if (rc) complete(s, sl, rc, s->now_us());
    return true;
  }
  return false;
}

// Single-source runs or one batch: single runs while k queries cost less
// that way than one batch of their word count, by the costs measured so far
// (a graph's batch cost is measured first, then one single query).  A slot
// takes one single run at a time: more waiting queries of the graph go to
// the other idle slots, where their (latency-bound) level loops overlap.
constexpr size_t kSingleMax = 1;

bool use_single(tbfs_server* s, int g, size_t* k) {
  const int wi = word_index(batch_words_for((int32_t)*k));
  const double tb = s->t_batch[g][wi];
  const double t1 = s->t_single[g];
  if (tb <= 0) return false;
  if (t1 <= 0) {
    *k = 1;
    return true;
  }
  if ((double)*k * t1 >= tb) return false;
  *k = std::min(*k, kSingleMax);
  return true;
}

// Form a batch for an idle slot; true if one was started.
bool admit(tbfs_server* s, Slot& sl) {
  std::vector<Query> take;
  int g = -1;
  bool single = false;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    double oldest = 0;
    for (size_t i = 0; i < s->pending.size(); ++i) {
      if (s->pending[i].empty()) continue;
      const double t = s->pending[i].front().t_submit;
      if (g < 0 || t < oldest) {
        g = (int)i;
        oldest = t;
      }
    }
    if (g < 0) return false;
    std::deque<Query>& pq = s->pending[g];
    size_t k = std::min(pq.size(), (size_t)s->cap[g]);
    single = use_single(s, g, &k);
    take.assign(pq.begin(), pq.begin() + k);
    pq.erase(pq.begin(), pq.begin() + k);
  }
  if (single) {
    sl.q = std::move(take);
    sl.graph = g;
    sl.words = 0;
    sl.t_start = s->now_us();
    int rc = TBFS_OK;
    if (!sl.c) rc = tbfs_ctx_create(s->nmax, &sl.c);
    if (!rc) rc = ctx_async_begin(sl.c, (int)sl.q.size());
    for (size_t i = 0; i < sl.q.size() && !rc; ++i) {
      const Query& q = sl.q[i];
      rc = ctx_async_enqueue(sl.c, s->graphs[g], q.root, s->bopts, (int)i, s->budget[g], q.S,
                             q.sigma, q.ovf);
    }
    if (!rc) rc = ctx_async_finish(sl.c, (int)sl.q.size());
    if (rc) {
      complete(s, sl, rc, s->now_us());
      return true;
    }
    sl.state = 3;
    ++s->batches;
    s->batched += sl.q.size();
    return true;
  }
  std::vector<int32_t>& roots = sl.roots;
  roots.resize(take.size());
  for (size_t i = 0; i < take.size(); ++i) roots[i] = take[i].root;
  sl.q = std::move(take);
  sl.graph = g;
  sl.words = batch_words_for((int32_t)roots.size());
  sl.t_start = s->now_us();
  const int rc = batch_start(sl.b, s->graphs[g], roots.data(), (int32_t)roots.size(), &s->bopts);
  if (rc) {
    complete(s, sl, rc, s->now_us());
    return true;
  }
  sl.state = 1;
  ++s->batches;
  s->batched += roots.size();
  return true;
}

void dispatcher(tbfs_server* s) {
  (void)hipSetDevice(s->device);
  while (!s->stop.load(std::memory_order_relaxed)) {
    bool moved = false;
    bool busy = false;
    for (Slot& sl : s->slots) {
      moved |= step_slot(s, sl);
      if (sl.state == 0) moved |= admit(s, sl);
      busy |= sl.state != 0;
    }
    if (moved) continue;
    if (busy) {
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    } else {
      std::unique_lock<std::mutex> lk(s->mu);
      bool any = false;
      for (const auto& pq : s->pending) any |= !pq.empty();
      if (!any && !s->stop.load())
        s->cv_work.wait_for(lk, std::chrono::milliseconds(2));
    }
  }
}

void server_free(tbfs_server* s) {
  if (!s) return;
  if (s->th.joinable()) {
    s->stop.store(true);
    s->cv_work.notify_all();
    s->th.join();
  }
  for (Slot& sl : s->slots) {
    if (sl.c) tbfs_ctx_destroy(sl.c);  // first: it runs on the batch's stream
    batch_free(sl.b);
  }
  delete s;
}

}  // namespace

extern "C" {

int tbfs_server_create(const tbfs_graph* const* graphs, int32_t ngraphs,
                       const tbfs_server_opts* o, tbfs_server** out) {
  if (!graphs || ngraphs < 1 || !out) return TBFS_ERR_INVALID;
  *out = nullptr;
  tbfs_server_opts so{};
  if (o) so = *o;
  if (so.slots < 0 || so.max_batch < 0 || so.max_batch > 1024 || so.mem_budget < 0)
    return TBFS_ERR_INVALID;
  const int nslots = so.slots ? so.slots : 4;
  const int max_batch = so.max_batch ? so.max_batch : 1024;
  int32_t nmax = 0, nzmax = 0;
  for (int i = 0; i < ngraphs; ++i) {
    if (!graphs[i] || !graphs[i]->CP) return TBFS_ERR_INVALID;
    nmax = std::max(nmax, graphs[i]->n);
    nzmax = std::max(nzmax, graphs[i]->nz);
  }
  int64_t budget = so.mem_budget;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    budget = std::min<int64_t>((int64_t)(fr / 4), (int64_t)64 << 30);
  }
  // what is left of each slot's share after its fixed buffers holds its
  // per-word batch state
  const int64_t per_slot = budget / nslots - slot_fixed_bytes(nmax, nzmax);
  const int64_t cap_words = per_slot > 0 ? per_slot / kSlotBytesPerWord : 0;
  if (cap_words < nmax) return TBFS_ERR_NOMEM;  // not even 64 sources of the largest graph
  tbfs_server* s = new (std::nothrow) tbfs_server();
  if (!s) return TBFS_ERR_NOMEM;
  s->t0 = std::chrono::steady_clock::now();
  s->graphs.assign(graphs, graphs + ngraphs);
  s->pending.resize(ngraphs);
  s->t_single.assign(ngraphs, 0.0);
  s->t_batch.assign(ngraphs, std::array<double, 5>{});
  s->n_single.assign(ngraphs, 0);
  s->n_batch.assign(ngraphs, std::array<int, 5>{});
  s->budget.assign(ngraphs, 8);
  s->nmax = nmax;
  for (int i = 0; i < ngraphs; ++i) {
    int w = 1;
    while (w < 16 && (int64_t)graphs[i]->n * (w * 2) <= cap_words) w *= 2;
    s->cap.push_back(std::min(64 * w, max_batch));
  }
  tbfs_opts_default(&s->bopts);
  int rc = TBFS_OK;
  hipError_t e = hipGetDevice(&s->device);
  if (e != hipSuccess) {
    rc = hip_fail(e, "hipGetDevice", __FILE__, __LINE__);
    server_free(s);
    return rc;
  }
  // a slot's batches and single-source runs share the batch's stream: the
  // slots' streams, created one after another, then sit on different
  // hardware queues (a busy stream sharing a queue with another busy one
  // waits for it)
  s->slots.resize(nslots);
  for (Slot& sl : s->slots) {
    rc = batch_create_words(cap_words, &sl.b);
    if (!rc) rc = ctx_create_on(nmax, batch_stream(sl.b), &sl.c);
    if (rc) {
      server_free(s);
      return rc;
    }
  }
  s->th = std::thread(dispatcher, s);
  *out = s;
  return TBFS_OK;
}

int tbfs_server_budget_for(const tbfs_graph* const* graphs, int32_t ngraphs, int32_t slots,
                           int32_t words, int64_t* bytes) {
  if (!graphs || ngraphs < 1 || slots < 1 || words < 1 || !bytes) return TBFS_ERR_INVALID;
  int64_t nmax = 0, nzmax = 0;
  for (int i = 0; i < ngraphs; ++i) {
    if (!graphs[i]) return TBFS_ERR_INVALID;
    nmax = std::max<int64_t>(nmax, graphs[i]->n);
    nzmax = std::max<int64_t>(nzmax, graphs[i]->nz);
  }
  *bytes = (int64_t)slots * (slot_fixed_bytes(nmax, nzmax) + (int64_t)words * nmax * kSlotBytesPerWord);
  return TBFS_OK;
}

int tbfs_server_destroy(tbfs_server* s) {
  server_free(s);
  return TBFS_OK;
}

int tbfs_server_submit(tbfs_server* s, int32_t graph, int32_t root, int32_t* S,
                       uint64_t* sigma, uint8_t* ovf, uint64_t* ticket) {
  if (!s || graph < 0 || graph >= (int32_t)s->graphs.size()) return TBFS_ERR_INVALID;
  if (root < 0 || root >= s->graphs[graph]->n) return TBFS_ERR_INVALID;
  Query q;
  q.graph = graph;
  q.root = root;
  q.S = S;
  q.sigma = sigma;
  q.ovf = ovf;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    q.ticket = s->next_ticket++;
    q.t_submit = s->now_us();
    s->pending[graph].push_back(q);
    ++s->in_flight;
  }
  s->cv_work.notify_one();
  if (ticket) *ticket = q.ticket;
  return TBFS_OK;
}

int tbfs_server_poll(tbfs_server* s, tbfs_query_done* done, int32_t max, int32_t* ndone,
                     int64_t wait_us) {
  if (!s || !ndone || max < 0 || (max > 0 && !done)) return TBFS_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->completed.empty() && wait_us > 0)
    s->cv_done.wait_for(lk, std::chrono::microseconds(wait_us),
                        [s] { return !s->completed.empty(); });
  int32_t k = 0;
  while (k < max && !s->completed.empty()) {
    done[k++] = s->completed.front();
    s->completed.pop_front();
  }
  *ndone = k;
  return TBFS_OK;
}

int tbfs_server_drain(tbfs_server* s) {
  if (!s) return TBFS_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [s] { return s->in_flight == 0; });
  return TBFS_OK;
}

int tbfs_server_info(tbfs_server* s, int32_t graph, int32_t* batch_cap, int64_t* batches,
                     int64_t* batched, double* now_us) {
  if (!s || graph < -1 || graph >= (int32_t)s->graphs.size()) return TBFS_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (batch_cap) *batch_cap = graph >= 0 ? s->cap[graph] : 0;
  if (batches) *batches = (int64_t)s->batches;
  if (batched) *batched = (int64_t)s->batched;
  if (now_us) *now_us = s->now_us();
  return TBFS_OK;
}

void* tbfs_host_alloc(int64_t bytes) {
  void* p = nullptr;
  if (bytes <= 0) return nullptr;
  if (hipHostMalloc(&p, (size_t)bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}

int tbfs_host_free(void* p) {
  if (p && hipHostFree(p) != hipSuccess) return TBFS_ERR_HIP;
  return TBFS_OK;
}

void* tbfs_device_alloc(int64_t bytes) {
  void* p = nullptr;
  if (bytes <= 0) return nullptr;
  if (hipMalloc(&p, (size_t)bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}

int tbfs_device_free(void* p) {
  if (p && hipFree(p) != hipSuccess) return TBFS_ERR_HIP;
  return TBFS_OK;
}

int tbfs_device_copy(void* dst, const void* src, int64_t bytes) {
  if (bytes < 0 || (bytes > 0 && (!dst || !src))) return TBFS_ERR_INVALID;
  if (bytes == 0) return TBFS_OK;
  HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDefault));
  return TBFS_OK;
}

int tbfs_device_fill(void* dst, int value, int64_t bytes) {
  if (bytes < 0 || (bytes > 0 && !dst)) return TBFS_ERR_INVALID;
  if (bytes == 0) return TBFS_OK;
  HIPCHK(hipMemset(dst, value, (size_t)bytes));
  return TBFS_OK;
}

}  // extern "C"
