// tbfs_serve.hip -- BFS query serving: resident graphs, continuous batching.
//
// A server holds a set of graphs resident in HBM and a fixed number of batch
// slots (a tbfs_batch and a single-source context on one HIP stream each)
// sized from a device-memory budget.  Queries (graph, root, host
// destinations) are queued per graph as they arrive.  Each slot has a worker
// thread of its own, so the host work of one slot (enqueueing thousands of
// level kernels for a deep graph) never delays another.  An idle worker
// takes the graph whose oldest waiting query is oldest, with every query
// waiting for it up to the slot's capacity for that graph (the largest
// power-of-two number of 64-source words whose state fits the slot's memory,
// at most max_batch sources), and runs them
//   - as single-source BFS on the slot's context when that costs less for so
//     few queries than a batch (by the per-graph costs measured so far), or
//   - as one batch (tbfs_batch), its rows copied to the queries' buffers, or
//   - for a deep graph (a BFS on it has needed more than open_depth levels):
//     as a continuous batch, whose level loop keeps running while sources
//     come and go: queries of that graph that arrive while it runs are seeded
//     into its free bits at the next level boundary (depths counted from
//     there), finished sources are read out at chunk boundaries and their
//     bits reused.
// Long jobs (deep graphs, or single-source queries slower than kLongMs) hold
// at most slots - short_slots slots at once, so short queries always find a
// slot.  A query's result does not depend on how it ran: every source of a
// batch is exactly its own single-source BFS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
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
// most 256 MiB, at least one row), the id map of a degree-ordered copy (4
// per vertex) and the slot's single-source context (depths 4, counts 8,
// flags 1 per vertex; five bitmaps; two work queues of n + nz/64 entries of
// 8 bytes).
int64_t slot_fixed_bytes(int64_t nmax, int64_t nzmax) {
  const int64_t stage = std::min<int64_t>(64 * 13 * nmax, std::max<int64_t>((int64_t)256 << 20, 13 * nmax));
  const int64_t ctx = 13 * nmax + 5 * 8 * ((nmax + 63) / 64) + 2 * 8 * (nmax + nzmax / 64 + 64);
  return stage + 4 * nmax + ctx + ((int64_t)1 << 20);  // counters, loop state, alignment
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

// Work a slot takes: single-source runs, one batch, or a continuous batch.
enum JobKind { kJobSingle = 0, kJobBatch = 1, kJobOpen = 2 };

struct Job {
  int kind = kJobSingle;
  int graph = -1;
  int words = 0;        // continuous batches: 64-source words
  bool long_job = false;
  std::vector<Query> q;
};

struct Slot {
  tbfs_batch* b = nullptr;
  tbfs_ctx* c = nullptr;  // single-source runs (small batches)
  std::thread th;         // the slot's worker
  std::vector<int32_t> roots;
  std::vector<int> redo;
};

constexpr int kDefaultOpenDepth = 64;  // deeper graphs: continuous batches
constexpr double kLongMs = 5.0;        // single-source queries slower than this are long
constexpr int kOpenChunk = 16;         // levels per chunk of a continuous batch
constexpr int kOpenAhead = 4;          // chunks enqueued ahead of the device

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
  std::vector<int> budget;   // level kernels per single-source run, per graph
  std::vector<int> depth;    // most levels a BFS on the graph has needed so far
  std::vector<int> open_free;  // free bits of the graph's running continuous batches
  int open_depth = kDefaultOpenDepth;  // 0: no continuous batches
  int long_cap = 1;      // slots that long jobs may hold at once
  int long_running = 0;  // slots holding one
  int32_t nmax = 0;
  std::vector<std::deque<Query>> pending;
  std::deque<tbfs_query_done> completed;
  std::vector<Slot> slots;
  std::mutex mu;  // everything above (and the slots' job hand-over)
  std::condition_variable cv_work, cv_done;
  std::atomic<bool> stop{false};
  bool held = false;      // tbfs_server_hold: queued queries wait (guarded by mu)
  int64_t in_flight = 0;  // submitted, not yet completed
  uint64_t next_ticket = 1;
  int device = 0;
  tbfs_opts bopts{};
  std::chrono::steady_clock::time_point t0;
  // written by the workers, read by tbfs_server_info on caller threads
  std::atomic<uint64_t> batches{0}, batched{0};
  // counts printed at destroy with TBFS_VERBOSE set (under mu)
  uint64_t jobs[3] = {0, 0, 0};
  uint64_t admitted_mid = 0;  // sources seeded into a running continuous batch
  double enqueue_us = 0;      // host time spent enqueueing work

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

bool is_deep(const tbfs_server* s, int g) {
  return s->open_depth > 0 && s->depth[g] > s->open_depth;
}

bool is_long(const tbfs_server* s, int g) {
  return is_deep(s, g) || s->t_single[g] > kLongMs;
}

// Report queries done (status, times on the server clock); the caller holds
// no lock.
void report(tbfs_server* s, const Query* q, size_t k, int status, int32_t batch, double t_start,
            double t_done, const int32_t* level) {
  std::lock_guard<std::mutex> lk(s->mu);
  for (size_t i = 0; i < k; ++i) {
    tbfs_query_done d{};
    d.ticket = q[i].ticket;
    d.graph = q[i].graph;
    d.root = q[i].root;
    d.status = status;
    d.batch = batch;
    d.submit_us = q[i].t_submit;
    d.start_us = t_start;
    d.done_us = t_done;
    d.level = level ? level[i] : 0;
    s->completed.push_back(d);
  }
  s->in_flight -= (int64_t)k;
  s->cv_done.notify_all();
}

// A cost sample (ms) of a completed run; `words` 0 for single-source runs
// (ms per query).
void note_cost(tbfs_server* s, int g, int words, double ms) {
  std::lock_guard<std::mutex> lk(s->mu);
  const int wi = word_index(words);
  double& t = words ? s->t_batch[g][wi] : s->t_single[g];
  int& runs = words ? s->n_batch[g][wi] : s->n_single[g];
  if (runs++ > 0) t = t > 0 ? 0.7 * t + 0.3 * ms : ms;
}

void sleep_poll() { std::this_thread::sleep_for(std::chrono::microseconds(20)); }

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

// Choose work for an idle slot (s->mu held): the graph whose oldest waiting
// query is oldest, skipping graphs whose running continuous batches have
// room for all their waiting queries (those join at the next level
// boundary) and long graphs while long jobs hold s->long_cap slots (the
// remaining slots stay free for short queries).
bool pick(tbfs_server* s, Job* job) {
  int g = -1;
  double oldest = 0;
  for (size_t i = 0; i < s->pending.size(); ++i) {
    const std::deque<Query>& pq = s->pending[i];
    if (pq.empty()) continue;
    if (s->open_free[i] > 0 && (size_t)s->open_free[i] >= pq.size()) continue;
    if (is_long(s, (int)i) && s->long_running >= s->long_cap) continue;
    const double t = pq.front().t_submit;
    if (g < 0 || t < oldest) {
      g = (int)i;
      oldest = t;
    }
  }
  if (g < 0) return false;
  std::deque<Query>& pq = s->pending[g];
  job->graph = g;
  job->long_job = is_long(s, g);
  size_t k = std::min(pq.size(), (size_t)s->cap[g]);
  if (is_deep(s, g)) {
    // the graph's full capacity: every query that arrives while it runs can
    // join (a deep BFS keeps a slot for thousands of levels whatever its
    // width, and the width costs only at the batch's start)
    job->kind = kJobOpen;
    job->words = batch_words_for(s->cap[g]);
  } else {
    job->kind = use_single(s, g, &k) ? kJobSingle : kJobBatch;
    job->words = job->kind == kJobBatch ? batch_words_for((int32_t)k) : 0;
  }
  job->q.assign(pq.begin(), pq.begin() + k);
  pq.erase(pq.begin(), pq.begin() + k);
  if (job->long_job) ++s->long_running;
  ++s->jobs[job->kind];
  ++s->batches;
  s->batched += k;
  return true;
}

// Single-source runs of job.q (one at a time on the slot's context).
void run_single(tbfs_server* s, Slot& sl, Job& job) {
  const int g = job.graph;
  const tbfs_graph* gr = s->graphs[g];
  const double t_start = s->now_us();
  const int n = (int)job.q.size();
  int bud;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    bud = s->budget[g];
  }
  int rc = ctx_async_begin(sl.c, n);
  for (int i = 0; i < n && !rc; ++i) {
    const Query& q = job.q[i];
    rc = ctx_async_enqueue(sl.c, gr, q.root, s->bopts, i, bud, q.S, q.sigma, q.ovf);
  }
  if (!rc) rc = ctx_async_finish(sl.c, n);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->enqueue_us += s->now_us() - t_start;
  }
  while (!rc) {
    bool ready = false;
    int maxl = 0;
    rc = ctx_async_poll(sl.c, n, &ready, &sl.redo, &maxl);
    if (rc) break;
    if (!ready) {
      sleep_poll();
      continue;
    }
    {
      // the most launches a BFS needed plus one spare (never lowered:
      // sources of one graph differ in depth, and a run-out costs a whole
      // BFS); twice the budget if one ran out
      std::lock_guard<std::mutex> lk(s->mu);
      int& b = s->budget[g];
      b = std::max(b, std::min(1 << 16, sl.redo.empty() ? maxl + 1 : 2 * bud));
      bud = b;
      if (sl.redo.empty()) s->depth[g] = std::max(s->depth[g], maxl - 1);
    }
    if (sl.redo.empty()) break;
    // sources that ran out of levels: again, with the larger budget
    for (int i : sl.redo) {
      const Query& q = job.q[i];
      rc = ctx_async_enqueue(sl.c, gr, q.root, s->bopts, i, bud, q.S, q.sigma, q.ovf);
      if (rc) break;
    }
    if (!rc) rc = ctx_async_finish(sl.c, n);
  }
  const double t_done = s->now_us();
  if (!rc) note_cost(s, g, 0, (t_done - t_start) * 1e-3 / n);
  report(s, job.q.data(), job.q.size(), rc, n, t_start, t_done, nullptr);
}

// One batch of job.q; rows copied to the queries' destinations.
void run_batch(tbfs_server* s, Slot& sl, Job& job) {
  const int g = job.graph;
  const double t_start = s->now_us();
  const int32_t k = (int32_t)job.q.size();
  sl.roots.resize(k);
  for (int32_t i = 0; i < k; ++i) sl.roots[i] = job.q[i].root;
  int rc = batch_start(sl.b, s->graphs[g], sl.roots.data(), k, &s->bopts);
  while (!rc) {
    bool fin = false;
    rc = batch_progress(sl.b, &fin);
    if (rc || fin) break;
    sleep_poll();
  }
  if (!rc) {
    // a single-source run on this graph needs at most the batch's depth in
    // launches, plus the one that finds the frontier empty
    tbfs_stats bst{};
    batch_fill_stats(sl.b, &bst);
    std::lock_guard<std::mutex> lk(s->mu);
    int& bud = s->budget[g];
    bud = std::max(bud, std::min(1 << 16, bst.levels + 2));
    s->depth[g] = std::max(s->depth[g], bst.levels);
  }
  if (!rc) {
    std::vector<int32_t*> sd(k);
    std::vector<uint64_t*> gd(k);
    std::vector<uint8_t*> od(k);
    for (int32_t i = 0; i < k; ++i) {
      sd[i] = job.q[i].S;
      gd[i] = job.q[i].sigma;
      od[i] = job.q[i].ovf;
    }
    rc = batch_enqueue_rows(sl.b, 0, k, sd.data(), gd.data(), od.data());
  }
  while (!rc) {
    const hipError_t e = hipEventQuery(batch_step_event(sl.b));
    if (e == hipSuccess) break;
    if (e != hipErrorNotReady) rc = hip_fail(e, "hipEventQuery", __FILE__, __LINE__);
    else sleep_poll();
  }
  const double t_done = s->now_us();
  if (!rc) note_cost(s, g, batch_words_for(k), (t_done - t_start) * 1e-3);
  report(s, job.q.data(), job.q.size(), rc, k, t_start, t_done, nullptr);
}

// A continuous batch: job.q seeded at depth 0, then chunks of levels; at
// each chunk's probe the finished sources are read out (their results reach
// the host before the next probe, when they are reported) and queries of the
// graph that have arrived meanwhile take the free bits, seeded at the
// current boundary.  Ends when no source is running and none is waiting.
void run_open(tbfs_server* s, Slot& sl, Job& job) {
  const int g = job.graph;
  const int32_t nbits = 64 * job.words;
  struct Src {
    Query q;
    int32_t d0 = -1;  // seeding level; -1: bit free
    double t_start = 0;
  };
  std::vector<Src> src((size_t)nbits);
  std::vector<int32_t> free_bits;
  for (int32_t j = nbits - 1; j >= 0; --j) free_bits.push_back(j);
  struct Done {
    std::vector<Query> q;
    std::vector<double> t_start;
    std::vector<int32_t> level;
    int64_t after;  // reported once the probe of chunk `after` has been read
  };
  std::deque<Done> awaiting;
  int running = 0;
  int64_t polled = 0;  // probes read
  int64_t head = 0;    // chunks enqueued
  bool wide = false;
  int rc = batch_open_begin(sl.b, s->graphs[g], job.words, &s->bopts);
  std::vector<int32_t> bits, roots, offs;
  std::vector<int32_t*> sd;
  std::vector<uint64_t*> gd;
  std::vector<uint8_t*> od;
  // seed queries (at the current boundary) into free bits
  auto seed = [&](std::vector<Query>& qs) {
    bits.clear();
    roots.clear();
    const int32_t d0 = batch_open_depth(sl.b);
    const double t = s->now_us();
    for (Query& q : qs) {
      const int32_t j = free_bits.back();
      free_bits.pop_back();
      src[j].q = q;
      src[j].d0 = d0;
      src[j].t_start = t;
      bits.push_back(j);
      roots.push_back(q.root);
    }
    running += (int)qs.size();
    return batch_open_seed(sl.b, bits.data(), roots.data(), (int32_t)bits.size());
  };
  if (!rc) rc = seed(job.q);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->open_free[g] += (int)free_bits.size();
  }
  if (!rc) {
    rc = batch_open_chunk(sl.b, kOpenChunk, wide);
    ++head;
  }
  uint64_t alive[16];
  while (!rc) {
    bool ready = false;
    int32_t D = 0;
    int64_t fw = 0;
    rc = batch_open_poll(sl.b, &ready, &D, alive, &fw);
    if (rc) break;
    if (!ready) {
      if (head - polled < kOpenAhead && running > 0) {
        rc = batch_open_chunk(sl.b, kOpenChunk, wide);
        ++head;
      } else {
        sleep_poll();
      }
      continue;
    }
    ++polled;
    wide = fw > ((int64_t)1 << 22);
    const double t_now = s->now_us();
    // results whose copies were enqueued before this probe's chunk
    while (!awaiting.empty() && awaiting.front().after < polled) {
      Done& dn = awaiting.front();
      for (size_t i = 0; i < dn.q.size(); ++i)
        report(s, &dn.q[i], 1, TBFS_OK, running + (int)dn.q.size(), dn.t_start[i], t_now,
               &dn.level[i]);
      awaiting.pop_front();
    }
    // sources finished by this probe: read out, bits freed
    bits.clear();
    offs.clear();
    Done dn;
    for (int32_t j = 0; j < nbits; ++j) {
      Src& x = src[j];
      if (x.d0 < 0 || x.d0 >= D || ((alive[j >> 6] >> (j & 63)) & 1ull)) continue;
      bits.push_back(j);
      offs.push_back(x.d0);
      dn.q.push_back(x.q);
      dn.t_start.push_back(x.t_start);
      dn.level.push_back(x.d0);
      x.d0 = -1;
    }
    if (!bits.empty()) {
      sd.resize(bits.size());
      gd.resize(bits.size());
      od.resize(bits.size());
      for (size_t i = 0; i < bits.size(); ++i) {
        sd[i] = dn.q[i].S;
        gd[i] = dn.q[i].sigma;
        od[i] = dn.q[i].ovf;
      }
      rc = batch_open_extract(sl.b, bits.data(), offs.data(), (int32_t)bits.size(), sd.data(),
                              gd.data(), od.data());
      if (rc) break;
      running -= (int)bits.size();
      for (int32_t j : bits) free_bits.push_back(j);
      dn.after = head;  // the next chunk's probe follows the copies
      awaiting.push_back(std::move(dn));
    }
    // arrivals: into the free bits, at this boundary
    std::vector<Query> add;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      std::deque<Query>& pq = s->pending[g];
      // at most the graph's batch capacity running at once (max_batch)
      const bool room = batch_open_depth(sl.b) < (1 << 29) - 4 * kOpenChunk &&
                        !s->stop.load(std::memory_order_relaxed) && !s->held;
      const size_t lim = (size_t)std::max(0, s->cap[g] - running);
      const size_t k = room ? std::min({pq.size(), free_bits.size(), lim}) : 0;
      add.assign(pq.begin(), pq.begin() + k);
      pq.erase(pq.begin(), pq.begin() + k);
      s->open_free[g] += (int)bits.size() - (int)k;
      s->admitted_mid += k;
      s->batched += k;
    }
    if (!add.empty()) {
      rc = seed(add);
      if (rc) break;
    }
    if (running == 0 && awaiting.empty() && head == polled) break;
    if (head - polled < kOpenAhead && (running > 0 || !awaiting.empty())) {
      rc = batch_open_chunk(sl.b, running > 0 ? kOpenChunk : 0, wide);
      ++head;
    }
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->open_free[g] -= (int)free_bits.size();
  }
  if (rc) {  // every query still held by the batch fails
    const double t = s->now_us();
    for (Done& dn : awaiting) report(s, dn.q.data(), dn.q.size(), rc, 0, t, t, nullptr);
    for (Src& x : src)
      if (x.d0 >= 0) report(s, &x.q, 1, rc, 0, x.t_start, t, nullptr);
  }
  (void)batch_open_end(sl.b);
}

void worker(tbfs_server* s, Slot* sl) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop.load(std::memory_order_relaxed)) {
    Job job;
    if (s->held || !pick(s, &job)) {
      s->cv_work.wait_for(lk, std::chrono::milliseconds(2));
      continue;
    }
    lk.unlock();
    if (job.kind == kJobSingle) run_single(s, *sl, job);
    else if (job.kind == kJobBatch) run_batch(s, *sl, job);
    else run_open(s, *sl, job);
    lk.lock();
    if (job.long_job) --s->long_running;
    // a slot came free: another worker may now take a long job
    s->cv_work.notify_all();
  }
}

void server_free(tbfs_server* s) {
  if (!s) return;
  s->stop.store(true);
  s->cv_work.notify_all();
  for (Slot& sl : s->slots)
    if (sl.th.joinable()) sl.th.join();
  const char* v = std::getenv("TBFS_VERBOSE");
  if (v && v[0] && v[0] != '0')
    std::fprintf(stderr,
                 "turbobfs: server: %llu single-source jobs, %llu batches, %llu continuous "
                 "batches (%llu sources admitted into running ones); %.1f ms enqueueing\n",
                 (unsigned long long)s->jobs[kJobSingle], (unsigned long long)s->jobs[kJobBatch],
                 (unsigned long long)s->jobs[kJobOpen], (unsigned long long)s->admitted_mid,
                 s->enqueue_us * 1e-3);
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
  if (so.slots < 0 || so.max_batch < 0 || so.max_batch > 1024 || so.mem_budget < 0 ||
      so.open_depth < -1 || so.short_slots < -1)
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
  s->depth.assign(ngraphs, 0);
  s->open_free.assign(ngraphs, 0);
  s->open_depth = so.open_depth == 0 ? kDefaultOpenDepth : std::max(0, so.open_depth);
  {
    const int keep = so.short_slots == 0 ? (nslots >= 2 ? 1 : 0) : std::max(0, so.short_slots);
    s->long_cap = std::max(1, nslots - keep);
  }
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
  std::vector<int32_t> ns(ngraphs);
  for (int i = 0; i < ngraphs; ++i) ns[i] = graphs[i]->n;
  for (Slot& sl : s->slots) {
    // everything a slot's work allocates is allocated here: growing a
    // buffer later would free device memory, which waits for every stream
    rc = batch_create_words(cap_words, &sl.b);
    if (!rc) rc = batch_reserve(sl.b, ns.data(), ngraphs);
    if (!rc) rc = ctx_create_on(nmax, batch_stream(sl.b), &sl.c);
    if (!rc) rc = ctx_async_begin(sl.c, (int)kSingleMax);
    if (rc) {
      server_free(s);
      return rc;
    }
  }
  for (Slot& sl : s->slots) sl.th = std::thread(worker, s, &sl);
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
  if (s->held) {  // held queries would never finish
    s->held = false;
    s->cv_work.notify_all();
  }
  s->cv_done.wait(lk, [s] { return s->in_flight == 0; });
  return TBFS_OK;
}

int tbfs_server_hold(tbfs_server* s, int32_t hold) {
  if (!s) return TBFS_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->held = hold != 0;
  }
  if (!hold) s->cv_work.notify_all();
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
