// tbfs_dropin.hip -- the reference's seven GPU entry points, same C
// signatures, on top of the resident-graph engine.
//
// Reference contract (paths relative to the reference root): caller-owned
// host arrays; the callee uploads the graph, runs the BFS `repetition` times,
// copies the last S / sigma back and frees every device buffer
// (TurboBFS_CSC_TD/bfsgputdcsc_sc.cu:47-149).  Kept here, with exact 64-bit
// counts converted once to the reference's output type, validated input, and
// nonzero status codes instead of exit().
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "bfsbucsc.h"
#include "bfstdbucsc.h"
#include "bfstdcooc.h"
#include "bfstdcsc.h"
#include "tbfs.h"
#include "tbfs_engine.h"

using namespace tbfs;

namespace {

bool verbose() {
  const char* v = std::getenv("TBFS_VERBOSE");
  return v && v[0] && v[0] != '0';
}

// Kernels behind the seven entry points (tbfs_dropin_mapping): the
// load-balanced direction-optimizing levels on the chained loop, or the
// reference's own thread / warp mappings on a host-driven loop.  Results are
// the same either way.
std::atomic<int> g_mapping{TBFS_DROPIN_FAST};

bool fast() { return g_mapping.load(std::memory_order_relaxed) == TBFS_DROPIN_FAST; }

// The fast engine's options for an entry point: direction-optimizing,
// load-balanced levels, chained loop; sigma off for the depth-only BU
// entry points.
tbfs_opts fast_opts(bool sigma) {
  tbfs_opts o;
  tbfs_opts_default(&o);
  o.mode = TBFS_MODE_AUTO;
  o.pull = TBFS_PULL_LB;
  o.loop = TBFS_LOOP_CHAINED;
  o.sigma = sigma ? 1 : 0;
  return o;
}

// Contexts of finished calls, kept for the next ones (a context's stream,
// pinned host words and level-loop state cost ~1 ms to create and free).  A
// call takes the smallest idle context of its device that holds its graph,
// or creates one; at most kPoolMax stay idle.
constexpr size_t kPoolMax = 4;
std::mutex g_pool_mu;
std::vector<tbfs_ctx*> g_pool;

int pool_take(int32_t n, tbfs_ctx** out) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    size_t best = g_pool.size();
    for (size_t i = 0; i < g_pool.size(); ++i)
      if (g_pool[i]->device == dev && g_pool[i]->n_cap >= n &&
          (best == g_pool.size() || g_pool[i]->n_cap < g_pool[best]->n_cap))
        best = i;
    if (best < g_pool.size()) {
      tbfs_ctx* c = g_pool[best];
      g_pool.erase(g_pool.begin() + (std::ptrdiff_t)best);
      // the launch budget was learned on graphs of earlier calls
      c->chain_k = 8;
      *out = c;
      return TBFS_OK;
    }
  }
  return ctx_create(n, out);
}

void pool_give(tbfs_ctx* c) {
  tbfs_ctx* drop = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool.push_back(c);
    if (g_pool.size() > kPoolMax) {
      drop = g_pool.front();
      g_pool.erase(g_pool.begin());
    }
  }
  if (drop) ctx_free(drop);
}

// Conversion of the exact counts at the boundary (k_dropin_convert, on the
// device): float for the TD CSC and COOC entry points (counts above 2^24, or
// >= 2^64, are not exact in float), int32 modulo 2^32 for TDBU (the
// reference's int arithmetic; counts above 2^31 - 1 are not exact).
constexpr int32_t kNoCounts = 0, kFloatCounts = 1, kIntCounts = 2;

// Upload, run `repetition` BFS from r, write S (unreached as s_unreached;
// S_h may be null) and the counts in the reference's type into the caller's
// arrays.  `cooc` selects the COOC upload (A = I, B = J) instead of CSC
// (A = CP, B = IC).
int run_dropin(const char* name, const int* A, const int* B, int nz, int n, int r,
               int repetition, const tbfs_opts& o, bool cooc, int32_t s_unreached,
               int32_t kind, int* S_h, void* val_h) {
  if (n <= 0 || nz < 0 || r < 0 || r >= n || repetition < 1 || !A ||
      (nz > 0 && !B)) {
    std::fprintf(stderr,
                 "turbobfs: %s: invalid arguments (n=%d nz=%d r=%d "
                 "repetition=%d)\n",
                 name, n, nz, r, repetition);
    return TBFS_ERR_INVALID;
  }
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  // every device buffer of the call comes from (and goes back to) the reuse
  // cache: nothing below frees a buffer with work pending on it
  DevReuse reuse;
  tbfs_graph* g = nullptr;
  tbfs_ctx* c = nullptr;
  int rc = cooc ? graph_create_cooc(A, B, n, nz, 0u, &g)
                : graph_create_csc(A, B, n, nz, 0u, &g);
  const auto t1 = clk::now();
  if (rc) {
    std::fprintf(stderr, "turbobfs: %s: graph upload failed: %s\n", name,
                 tbfs_status_str(rc));
    return rc;
  }
  rc = pool_take(n, &c);
  if (rc) {
    graph_free(g);
    return rc;
  }
  // a COOC graph whose columns are not sorted has no CSC view: only the
  // edge-parallel sweep runs on it
  tbfs_opts oo = o;
  if (cooc && !g->CP) {
    oo.mode = TBFS_MODE_COOC;
    oo.loop = TBFS_LOOP_HOST;
  }
  const auto t2 = clk::now();
  tbfs_stats st{};
  double total_ms = 0.0;
  for (int i = 0; i < repetition && !rc; ++i) {
    rc = run_bfs(c, g, r, oo, &st);
    total_ms += st.ms;
  }
  const auto t3 = clk::now();
  int64_t inexact = 0;
  if (!rc) rc = ctx_read_converted(c, n, s_unreached, kind, S_h, val_h, &inexact);
  const auto t4 = clk::now();
  if (rc) {
    ctx_free(c);  // its state is unknown after a failure
  } else {
    pool_give(c);
  }
  graph_free(g);
  const auto t5 = clk::now();
  auto ms = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  if (rc) {
    std::fprintf(stderr, "turbobfs: %s: BFS failed: %s\n", name, tbfs_status_str(rc));
    return rc;
  }
  if (verbose()) {
    std::printf("\n%s::d = %d,r = %d,t_sum=%lfms \n%s::average time BFS = %lfms \n",
                name, st.levels, r, total_ms, name, total_ms / repetition);
    std::printf("%s::host ms: graph upload + checks %.3f, context %.3f, BFS %.3f, "
                "convert + read %.3f, free %.3f\n", name, ms(t0, t1), ms(t1, t2), ms(t2, t3),
                ms(t3, t4), ms(t4, t5));
  }
  if (inexact) {
    if (kind == kFloatCounts)
      std::fprintf(stderr,
                   "turbobfs: %s: %lld path counts exceed 2^24 and are not exact in float\n",
                   name, (long long)inexact);
    else
      std::fprintf(stderr, "turbobfs: %s: %lld path counts exceed 2^31-1; written modulo 2^32\n",
                   name, (long long)inexact);
    return TBFS_SIGMA_INEXACT;
  }
  return TBFS_OK;
}

int td_csc(const char* name, int* IC_h, int* CP_h, int* S_h, float* sigma_h,
           int r, int nz, int n, int repetition, int mode) {
  if (!S_h || !sigma_h) return TBFS_ERR_INVALID;
  tbfs_opts o = fast_opts(true);
  if (!fast()) {
    o.mode = mode;
    o.loop = TBFS_LOOP_HOST;
    o.pull = mode == TBFS_MODE_PULL_WA ? TBFS_PULL_WA : TBFS_PULL_SC;
  }
  // TD CSC: unreached vertices keep S = 0 (the reference clears S_d to 0)
  return run_dropin(name, CP_h, IC_h, nz, n, r, repetition, o, false, 0, kFloatCounts, S_h,
                    sigma_h);
}

int bu_csc(const char* name, int* I_h, int* CP_h, int* S_h, int r, int nz,
           int n, int repetition, int bbreak, int mode) {
  if (!S_h) return TBFS_ERR_INVALID;
  tbfs_opts o = fast_opts(false);
  if (!fast()) {
    o.mode = mode;
    o.loop = TBFS_LOOP_HOST;
    o.pull = mode == TBFS_MODE_PULL_WA ? TBFS_PULL_WA : TBFS_PULL_SC;
  }
  o.bbreak = bbreak ? 1 : 0;
  return run_dropin(name, CP_h, I_h, nz, n, r, repetition, o, false, -1, kNoCounts, S_h, nullptr);
}

int tdbu_csc(const char* name, int* I_h, int* CP_h, int* S_h, int* sigma_h,
             int r, int nz, int n, int repetition, int pull) {
  if (!S_h || !sigma_h) return TBFS_ERR_INVALID;
  tbfs_opts o = fast_opts(true);
  if (!fast()) {
    o.pull = pull;
    o.loop = pull == TBFS_PULL_LB ? TBFS_LOOP_CHAINED : TBFS_LOOP_HOST;
  }
  return run_dropin(name, CP_h, I_h, nz, n, r, repetition, o, false, -1, kIntCounts, S_h,
                    sigma_h);
}

}  // namespace

extern "C" {

int tbfs_dropin_release(void) {
  std::vector<tbfs_ctx*> idle;
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    idle.swap(g_pool);
  }
  for (tbfs_ctx* c : idle) ctx_free(c);
  dev_reuse_release();
  return TBFS_OK;
}

int tbfs_dropin_mapping(int mapping) {
  if (mapping != TBFS_DROPIN_FAST && mapping != TBFS_DROPIN_REFERENCE && mapping != -1)
    return TBFS_ERR_INVALID;
  if (mapping == -1) return g_mapping.load();
  return g_mapping.exchange(mapping);
}

int bfs_gpu_td_csc_sc(int* IC_h, int* CP_h, int* S_h, float* sigma_h, int r,
                      int nz, int n, int repetition) {
  return td_csc("bfsgputdcsc_sc", IC_h, CP_h, S_h, sigma_h, r, nz, n,
                repetition, TBFS_MODE_PULL_SC);
}

int bfs_gpu_td_csc_wa(int* IC_h, int* CP_h, int* S_h, float* sigma_h, int r,
                      int nz, int n, int repetition) {
  return td_csc("bfsgputdcsc_wa", IC_h, CP_h, S_h, sigma_h, r, nz, n,
                repetition, TBFS_MODE_PULL_WA);
}

int bfs_gpu_bu_csc_sc(int* I_h, int* CP_h, int* S_h, int r, int nz, int n,
                      int repetition, int bbreak) {
  return bu_csc("bfsgpubucsc_sc", I_h, CP_h, S_h, r, nz, n, repetition, bbreak,
                TBFS_MODE_PULL_SC);
}

int bfs_gpu_bu_csc_wa(int* I_h, int* CP_h, int* S_h, int r, int nz, int n,
                      int repetition, int bbreak) {
  return bu_csc("bfsgpubucsc_wa", I_h, CP_h, S_h, r, nz, n, repetition, bbreak,
                TBFS_MODE_PULL_WA);
}

int bfs_gpu_tdbu_csc_sc(int* I_h, int* CP_h, int* S_h, int* sigma_h, int r,
                        int nz, int n, int repetition, int bbreak) {
  (void)bbreak;  // counts need every parent: no early exit with sigma on
  return tdbu_csc("bfsgputdbucsc_sc", I_h, CP_h, S_h, sigma_h, r, nz, n,
                  repetition, TBFS_PULL_LB);
}

int bfs_gpu_tdbu_csc_wa(int* IC_h, int* CP_h, int* S_h, int* sigma_h, int r,
                        int nz, int n, int repetition, int bbreak) {
  (void)bbreak;
  return tdbu_csc("bfsgputdbucsc_wa", IC_h, CP_h, S_h, sigma_h, r, nz, n,
                  repetition, TBFS_PULL_WA);
}

int bfs_gpu_td_cooc_sc(int* I_h, int* J_h, float* sigma_h, int r, int nz,
                       int n, int repetition) {
  if (!sigma_h) return TBFS_ERR_INVALID;
  tbfs_opts o = fast_opts(true);
  if (!fast()) {
    o.mode = TBFS_MODE_COOC;
    o.loop = TBFS_LOOP_HOST;
  }
  return run_dropin("bfsgputdcooc_sc", I_h, J_h, nz, n, r, repetition, o, true, -1,
                    kFloatCounts, nullptr, sigma_h);
}

}  // extern "C"
