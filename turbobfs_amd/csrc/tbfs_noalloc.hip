// tbfs_noalloc.hip -- single-source BFS that allocates nothing.
//
// For callers that cannot allocate device memory during a call (code under
// HIP graph capture, tight serving loops, harnesses that forbid it): the
// same chained level loop as tbfs_bfs, with the caller's device buffers as
// outputs and the loop's working state (frontier bitmaps, reached bitmap,
// striped level counters, loop state, stats) in `scratch`, a fixed device
// array of the library that each GPU holds until the process ends.
//
// In host code `scratch` names a host-side stand-in, not a device address:
// its device address is taken once per device (hipGetSymbolAddress, with
// that device current) and only that address is passed to kernels.  Calls
// on one GPU share the array: the library takes them one at a time per GPU
// (a call returns after its work on the device is done), and the init
// kernel rewrites everything the loop reads before it reads it, so nothing
// left by an earlier call on a larger graph is seen.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <mutex>

#include "tbfs.h"
#include "tbfs_engine.h"

using namespace tbfs;

namespace {

constexpr int64_t kMaxN = TBFS_NOALLOC_MAX_N;
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kBitmapBytes = align256((size_t)((kMaxN + 63) / 64) * sizeof(uint64_t));
// layout: fb[0], fb[1], fb[2], vis, the three striped counter sets, the two
// loop-state slots, the stats slot
constexpr size_t kOffVis = 3 * kBitmapBytes;
constexpr size_t kOffCset = 4 * kBitmapBytes;
constexpr size_t kOffLoop = kOffCset + align256(3 * sizeof(CtrSet));
constexpr size_t kOffStats = kOffLoop + align256(2 * sizeof(LoopState));
constexpr size_t kScratchBytes = kOffStats + align256(sizeof(RunStats));
static_assert(kScratchBytes <= (size_t)TBFS_NOALLOC_SCRATCH_BYTES,
              "the public bound covers the scratch array");

constexpr int kMaxDevices = 64;

// Per-device state, set up on the first call on that device.
struct DevState {
  std::once_flag once;
  int rc = TBFS_OK;
  unsigned char* ptr = nullptr;  // device address of `scratch` on this device
  int cus = 0;
  size_t fbl_limit = 0;          // LDS bytes a staged frontier bitmap may take
  std::mutex mu;                 // one call at a time: calls share `scratch`
  int chain_k = 8;               // level launches per host check (learned)
  LoopState loop_h{};            // host copies, read after each chunk
  RunStats stats_h{};
};
DevState g_dev[kMaxDevices];

}  // namespace

// The working state of tbfs_bfs_noalloc (see the layout above).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

int setup(int dev, DevState& ds) {
  std::call_once(ds.once, [&] {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      ds.rc = hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                       "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
      return;
    }
    ds.ptr = static_cast<unsigned char*>(p);
    e = hipDeviceGetAttribute(&ds.cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) {
      ds.rc = hip_fail(e, "hipDeviceGetAttribute", __FILE__, __LINE__);
      return;
    }
    if (kern::chain_fbl_limit(&ds.fbl_limit) != hipSuccess) {
      (void)hipGetLastError();
      ds.fbl_limit = 0;
    }
  });
  return ds.rc;
}


}  // namespace

extern "C" {

int tbfs_bfs_noalloc(const tbfs_graph* g, int32_t root, int32_t* S, uint64_t* sigma,
                     uint8_t* ovf, void* stream, const tbfs_opts* opts, tbfs_stats* st) {
  if (!g || !S || !sigma || !ovf) return TBFS_ERR_INVALID;
  if (!g->CP) return TBFS_ERR_UNSUPPORTED;
  if ((int64_t)g->n > kMaxN) {
    std::fprintf(stderr,
                 "turbobfs: tbfs_bfs_noalloc: the graph has %d vertices; the library's "
                 "scratch array holds the state of at most %lld\n",
                 g->n, (long long)kMaxN);
    return TBFS_ERR_INVALID;
  }
  if (root < 0 || root >= g->n) return TBFS_ERR_INVALID;
  tbfs_opts o;
  if (opts) o = *opts;
  else tbfs_opts_default(&o);
  if (o.mode != TBFS_MODE_AUTO && o.mode != TBFS_MODE_PUSH) return TBFS_ERR_UNSUPPORTED;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return TBFS_ERR_UNSUPPORTED;
  DevState& ds = g_dev[dev];
  int rc = setup(dev, ds);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(ds.mu);

  unsigned char* base = ds.ptr;
  PersistParams p{};
  p.g = GraphView{g->n, g->nz, g->CP, g->IC, g->RP, g->JR};
  p.S = S;
  p.sigma = sigma;
  p.ovf = ovf;
  for (int i = 0; i < 3; ++i) p.fb[i] = reinterpret_cast<uint64_t*>(base + i * kBitmapBytes);
  p.vis = reinterpret_cast<uint64_t*>(base + kOffVis);
  p.skip = g->skip;
  p.cset = reinterpret_cast<CtrSet*>(base + kOffCset);
  p.bar = nullptr;  // the chained loop has no grid barrier
  p.stats = reinterpret_cast<RunStats*>(base + kOffStats);
  p.bigc = g->bigc;
  p.nbigc = g->nbigc;
  p.nheavy = g->nheavy_in;
  p.bigo = g->bigo;
  p.nbigo = g->nbigo;
  p.root = root;
  p.nroots = 0;
  p.roots = nullptr;
  p.mode = o.mode;
  p.sigma_on = o.sigma ? 1 : 0;
  // tunables left at 0 take the defaults: the autotuner's table is keyed by
  // degree statistics, which a graph computes (and allocates for) on first use
  p.alpha = o.alpha > 0 ? o.alpha : 14.0f;
  p.beta = o.beta > 0 ? o.beta : 24.0f;
  p.stage_ratio = o.stage_ratio > 0 ? o.stage_ratio : 1.0f;
  p.loop = reinterpret_cast<LoopState*>(base + kOffLoop);
  p.step0 = (o.mode == TBFS_MODE_AUTO || o.mode == TBFS_MODE_PUSH) ? 1 : 0;

  // launch shapes: pull levels stage the frontier bitmap in LDS when it fits
  const int64_t nwords = ((int64_t)g->n + 63) / 64;
  const size_t bm = (size_t)nwords * sizeof(uint64_t);
  const size_t fbl = (o.fb_lds >= 0 && ds.fbl_limit && bm <= ds.fbl_limit) ? ((bm + 15) & ~(size_t)15)
                                                                           : 0;
  const int grid_lb = o.grid > 0     ? std::min(o.grid, 65536)
                      : o.lb_grid > 0 ? std::min(o.lb_grid, 65536)
                                      : grid_for(nwords * 16, kLbBlock, 2048);
  const int grid_stage = o.grid > 0 ? std::min(o.grid, 65536) : std::max(1, ds.cus);

  hipStream_t s = static_cast<hipStream_t>(stream);
  const auto t0 = std::chrono::steady_clock::now();
  HIPCHK(kern::chain_init(s, grid_for(g->n, 256, 8192), p));
  int32_t d = p.step0;
  int chunk = std::max(ds.chain_k, 2);
  for (;;) {
    for (int i = 0; i < chunk; ++i) {
      const int lv = d + i;
      const int grid = fbl ? grid_stage
                           : (o.grid > 0 || lv < 16 ? grid_lb : std::min(grid_lb, std::max(1, ds.cus)));
      HIPCHK(kern::chain_level(s, grid, p, lv, fbl));
    }
    d += chunk;
    HIPCHK(hipMemcpyAsync(&ds.loop_h, p.loop + (d & 1), sizeof(LoopState), hipMemcpyDeviceToHost,
                          s));
    HIPCHK(hipStreamSynchronize(s));
    if (ds.loop_h.done) break;
    chunk = std::min(chunk + chunk / 2, 256);
  }
  HIPCHK(hipMemcpyAsync(&ds.stats_h, p.stats, sizeof(RunStats), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  // the next call's first chunk: this BFS's launches plus one spare
  ds.chain_k = std::min(4096, std::max(2, ds.stats_h.launches + 1));
  if (st) {
    st->levels = ds.stats_h.levels;
    st->push_levels = ds.stats_h.push_levels;
    st->pull_levels = ds.stats_h.pull_levels;
    st->pad = 0;
    st->reached = (int64_t)ds.stats_h.reached;
    st->edges = (int64_t)ds.stats_h.edges;
    st->ovf64 = ds.stats_h.ovf_any ? -1 : 0;
    st->ms = ms;  // host wall time of the call (no event is created)
    st->launches = ds.stats_h.launches;
    st->pad2 = 0;
  }
  return TBFS_OK;
}

}  // extern "C"
