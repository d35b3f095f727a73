// tbfs_engine.hip -- host side of libturbobfs: resident graphs, BFS contexts,
// the level loop and the extended C API declared in include/tbfs.h.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <iterator>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <new>
#include <thread>
#include <vector>

#include "tbfs.h"
#include "tbfs_engine.h"
#include "tbfs_internal.h"

using namespace tbfs;

// ----------------------------------------------------------------------------
// error plumbing
namespace tbfs {

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  std::fprintf(stderr, "turbobfs: %s failed: %s (%s:%d)\n", what,
               hipGetErrorString(e), file, line);
  return e == hipErrorOutOfMemory ? TBFS_ERR_NOMEM : TBFS_ERR_HIP;
}

// ---- device allocations with a reuse cache (see DevReuse) -------------------
namespace {
thread_local int t_reuse = 0;
struct CachedBlock {
  int dev;
  size_t bytes;
  void* p;
};
std::mutex g_dc_mu;
std::unordered_map<void*, CachedBlock> g_dc_out;  // handed out under DevReuse
std::vector<CachedBlock> g_dc_free;               // freed, kept for reuse (oldest first)
size_t g_dc_free_bytes = 0;

// Release cached blocks (oldest first) until at most `keep` bytes stay
// cached; dev >= 0: only blocks of that device.  Caller does not hold the lock.
void dc_trim(size_t keep, int dev) {
  std::vector<CachedBlock> victims;
  {
    std::lock_guard<std::mutex> lk(g_dc_mu);
    for (size_t i = 0; i < g_dc_free.size() && g_dc_free_bytes > keep;) {
      if (dev >= 0 && g_dc_free[i].dev != dev) {
        ++i;
        continue;
      }
      victims.push_back(g_dc_free[i]);
      g_dc_free_bytes -= g_dc_free[i].bytes;
      g_dc_free.erase(g_dc_free.begin() + (std::ptrdiff_t)i);
    }
  }
  int cur = -1;
  if (!victims.empty()) (void)hipGetDevice(&cur);
  for (const CachedBlock& b : victims) {
    if (b.dev != cur) (void)hipSetDevice(b.dev);
    (void)hipFree(b.p);
    if (b.dev != cur) (void)hipSetDevice(cur);
  }
}
}  // namespace

DevReuse::DevReuse() { ++t_reuse; }
DevReuse::~DevReuse() { --t_reuse; }

void dev_reuse_release() { dc_trim(0, -1); }

int dev_alloc_bytes(void** p, size_t bytes) {
  *p = nullptr;
  bytes = (bytes + 255) & ~(size_t)255;
  if (t_reuse > 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    {
      // the smallest cached block of this device that fits, at most 1/8 over
      std::lock_guard<std::mutex> lk(g_dc_mu);
      size_t best = g_dc_free.size();
      for (size_t i = 0; i < g_dc_free.size(); ++i) {
        const CachedBlock& b = g_dc_free[i];
        if (b.dev == dev && b.bytes >= bytes && b.bytes <= bytes + bytes / 8 &&
            (best == g_dc_free.size() || b.bytes < g_dc_free[best].bytes))
          best = i;
      }
      if (best < g_dc_free.size()) {
        const CachedBlock b = g_dc_free[best];
        g_dc_free.erase(g_dc_free.begin() + (std::ptrdiff_t)best);
        g_dc_free_bytes -= b.bytes;
        g_dc_out[b.p] = b;
        *p = b.p;
        return TBFS_OK;
      }
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {  // cached blocks hold the memory: give them back
      (void)hipGetLastError();
      dc_trim(0, dev);
      e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
      *p = nullptr;
      return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    }
    std::lock_guard<std::mutex> lk(g_dc_mu);
    g_dc_out[*p] = CachedBlock{dev, bytes, *p};
    return TBFS_OK;
  }
  const hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
  }
  return TBFS_OK;
}

void dev_free_bytes(void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_dc_mu);
    auto it = g_dc_out.find(p);
    if (it != g_dc_out.end()) {
      g_dc_free.push_back(it->second);
      g_dc_free_bytes += it->second.bytes;
      g_dc_out.erase(it);
      if (g_dc_free_bytes <= kDevReuseCap) return;
    } else {
      (void)hipFree(p);
      return;
    }
  }
  dc_trim(kDevReuseCap, -1);
}

}  // namespace tbfs

// ----------------------------------------------------------------------------
// graphs
namespace tbfs {

void graph_free(tbfs_graph* g) {
  if (!g) return;
  graph_free(g->bview);
  dev_free(g->bperm);
  if (g->RP != g->CP) dev_free(g->RP);
  if (g->JR != g->IC) dev_free(g->JR);
  if (g->J_owned) dev_free(g->J);
  if (g->I != g->IC) dev_free(g->I);
  if (g->bigo != g->bigc) dev_free(g->bigo);
  dev_free(g->skip);
  dev_free(g->bsl);
  dev_free(g->bigc);
  dev_free(g->CP);
  dev_free(g->IC);
  delete g;
}

// Build RP/JR (out-lists) from CP/IC with a stable radix sort of the entries
// by row: JR's rows come out in increasing column order, deterministically.
int graph_build_transpose(tbfs_graph* g, hipStream_t s) {
  const int32_t n = g->n, nz = g->nz;
  int rc = dev_alloc(&g->RP, (size_t)n + 1);
  if (rc) return rc;
  rc = dev_alloc(&g->JR, (size_t)nz);
  if (rc) return rc;
  int32_t *col = nullptr, *keys_out = nullptr;
  void* tmp = nullptr;
  size_t tmp_sort = 0;
  int bits = 1;
  while (bits < 31 && (1ll << bits) < (int64_t)n) ++bits;
  rc = dev_alloc(&col, (size_t)nz);
  if (!rc) rc = dev_alloc(&keys_out, (size_t)nz);
  if (rc) goto out;
  {
    // sort (row, column) pairs by row: the columns land in JR as the
    // out-lists, the sorted rows give RP
    hipError_t e;
    e = kern::expand_cols(s, g->CP, n, col);
    if (e != hipSuccess) { rc = hip_fail(e, "expand_cols", __FILE__, __LINE__); goto out; }
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_sort, g->IC, keys_out,
                                           col, g->JR, nz, 0, bits, s);
    if (e != hipSuccess) { rc = hip_fail(e, "SortPairs(size)", __FILE__, __LINE__); goto out; }
    rc = dev_alloc((char**)&tmp, std::max<size_t>(tmp_sort, 1));
    if (rc) goto out;
    e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_sort, g->IC, keys_out, col,
                                           g->JR, nz, 0, bits, s);
    if (e != hipSuccess) { rc = hip_fail(e, "SortPairs", __FILE__, __LINE__); goto out; }
    e = kern::offsets_sorted(s, keys_out, nz, n, g->RP);
    if (e != hipSuccess) { rc = hip_fail(e, "offsets_sorted", __FILE__, __LINE__); goto out; }
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) { rc = hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__); goto out; }
  }
out:
  dev_free(col);
  dev_free(keys_out);
  dev_free(tmp);
  return rc;
}

// Derive CP from a non-decreasing column array J (COOC sorted by column).
int graph_cp_from_sorted(tbfs_graph* g, const int32_t* Jdev, hipStream_t s) {
  const int32_t n = g->n, nz = g->nz;
  int rc = dev_alloc(&g->CP, (size_t)n + 1);
  if (rc) return rc;
  hipError_t e = kern::offsets_sorted(s, Jdev, nz, n, g->CP);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "cp_from_sorted", __FILE__, __LINE__);
  return TBFS_OK;
}

// ----------------------------------------------------------------------------
// Copies between pageable host memory and the device through two pinned
// staging buffers: host threads fill (or empty) one buffer while the DMA
// engine drains (or fills) the other.  A plain copy from pageable memory runs
// at ~10 GB/s here; the graph's in-lists (126 MB at R-MAT scale 20) dominate a
// drop-in call.  The buffers are allocated on first use and kept for the
// process; one mutex serialises the copies that use them.
namespace {
constexpr size_t kStageBytes = size_t(16) << 20;
constexpr size_t kStageMin = size_t(8) << 20;  // smaller copies go direct
std::mutex g_stage_mu;
void* g_stage[2] = {nullptr, nullptr};

// src is pinned or device memory: no staging needed.  A failed query of a
// pageable pointer leaves an error the next launch check would report, so it
// is cleared here.
bool is_pinned(const void* src) {
  hipPointerAttribute_t at;
  const hipError_t e = hipPointerGetAttributes(&at, src);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeHost || at.type == hipMemoryTypeDevice;
}

int stage_threads() {
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(8u, hw / 2));
}

void par_memcpy(void* dst, const void* src, size_t bytes, int nth) {
  if (nth <= 1 || bytes < (size_t(2) << 20)) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t per = (bytes / (size_t)nth + 4095) & ~size_t(4095);
  std::vector<std::thread> ts;
  for (int t = 1; t < nth; ++t) {
    const size_t a = (size_t)t * per;
    if (a >= bytes) break;
    const size_t len = std::min(per, bytes - a);
    ts.emplace_back([=] { std::memcpy((char*)dst + a, (const char*)src + a, len); });
  }
  std::memcpy(dst, src, std::min(per, bytes));
  for (auto& th : ts) th.join();
}

// Caller holds g_stage_mu.  False when pinned memory is not available.
bool stage_ready() {
  for (void*& b : g_stage)
    if (!b && hipHostMalloc(&b, kStageBytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      b = nullptr;
      return false;
    }
  return true;
}
}  // namespace

// Host-to-device copy on stream s; returns once the host buffer may change
// (the device copy may still be in flight for the unstaged path, as with
// hipMemcpyAsync from pageable memory).
static hipError_t upload_h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes < kStageMin || is_pinned(src))
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  std::lock_guard<std::mutex> lk(g_stage_mu);
  if (!stage_ready()) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipError_t e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
  const int nth = stage_threads();
  bool used[2] = {false, false};
  int b = 0;
  for (size_t off = 0; e == hipSuccess && off < bytes; off += kStageBytes, b ^= 1) {
    const size_t len = std::min(kStageBytes, bytes - off);
    if (used[b]) e = hipEventSynchronize(ev[b]);
    if (e != hipSuccess) break;
    par_memcpy(g_stage[b], (const char*)src + off, len, nth);
    e = hipMemcpyAsync((char*)dst + off, g_stage[b], len, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(ev[b], s);
    used[b] = true;
  }
  // the buffers are shared: the last copies finish before the lock is let go
  for (int k = 0; k < 2; ++k) {
    if (ev[k]) {
      const hipError_t w = used[k] ? hipEventSynchronize(ev[k]) : hipSuccess;
      if (e == hipSuccess) e = w;
      (void)hipEventDestroy(ev[k]);
    }
  }
  return e;
}

// Device-to-host copy on stream s, after the work already on s; returns with
// the data in dst.
static hipError_t download_d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes < kStageMin || is_pinned(dst)) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
  }
  std::lock_guard<std::mutex> lk(g_stage_mu);
  if (!stage_ready()) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
  }
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipError_t e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
  const int nth = stage_threads();
  const size_t nchunks = (bytes + kStageBytes - 1) / kStageBytes;
  // chunk k goes through buffer k & 1: chunk k + 1's copy is in flight while
  // the host empties chunk k
  auto issue = [&](size_t k) {
    const size_t off = k * kStageBytes;
    const size_t len = std::min(kStageBytes, bytes - off);
    hipError_t r = hipMemcpyAsync(g_stage[k & 1], (const char*)src + off, len,
                                  hipMemcpyDeviceToHost, s);
    return r == hipSuccess ? hipEventRecord(ev[k & 1], s) : r;
  };
  if (e == hipSuccess) e = issue(0);
  for (size_t k = 0; e == hipSuccess && k < nchunks; ++k) {
    e = hipEventSynchronize(ev[k & 1]);
    if (e == hipSuccess && k + 1 < nchunks) e = issue(k + 1);
    if (e != hipSuccess) break;
    const size_t off = k * kStageBytes;
    par_memcpy((char*)dst + off, g_stage[k & 1], std::min(kStageBytes, bytes - off), nth);
  }
  if (e != hipSuccess) (void)hipStreamSynchronize(s);
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  return e;
}

int graph_build_bigc(tbfs_graph* g, hipStream_t s);

// TBFS_VERBOSE=1: host time of each stage of a graph upload on stdout.
namespace {
struct StageClock {
  using clk = std::chrono::steady_clock;
  bool on = std::getenv("TBFS_VERBOSE") != nullptr;
  clk::time_point t = clk::now();
  double lap() {
    const clk::time_point u = clk::now();
    const double ms = std::chrono::duration<double, std::milli>(u - t).count();
    t = u;
    return ms;
  }
};
}  // namespace

int graph_create_csc(const int32_t* CP, const int32_t* IC, int32_t n,
                     int32_t nz, uint32_t flags, tbfs_graph** out) {
  if (!out) return TBFS_ERR_INVALID;
  *out = nullptr;
  if (n <= 0 || nz < 0 || !CP || (nz > 0 && !IC)) return TBFS_ERR_INVALID;
  const bool on_dev = (flags & TBFS_GRAPH_DEVICE) != 0;
  // the full check (monotone CP, IC in range) runs on the device after the
  // upload: an O(n + nz) host pass cost more than the upload itself
  if (!on_dev && (CP[0] != 0 || CP[n] != nz)) return TBFS_ERR_INVALID;
  tbfs_graph* g = new (std::nothrow) tbfs_graph();
  if (!g) return TBFS_ERR_NOMEM;
  g->n = n;
  g->nz = nz;
  g->flags = flags & (TBFS_GRAPH_SYMMETRIC);
  hipStream_t s = nullptr;
  int rc = TBFS_OK;
  StageClock clock;
  double ms_up = 0, ms_check = 0, ms_tr = 0;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) { rc = hip_fail(e, "hipStreamCreate", __FILE__, __LINE__); goto fail; }
  rc = dev_alloc(&g->CP, (size_t)n + 1);
  if (!rc) rc = dev_alloc(&g->IC, (size_t)nz);
  if (rc) goto fail;
  {
    const hipMemcpyKind kind = on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    e = hipMemcpyAsync(g->CP, CP, ((size_t)n + 1) * sizeof(int32_t), kind, s);
    if (e == hipSuccess && nz > 0)
      e = on_dev ? hipMemcpyAsync(g->IC, IC, (size_t)nz * sizeof(int32_t), kind, s)
                 : upload_h2d(g->IC, IC, (size_t)nz * sizeof(int32_t), s);
    if (e != hipSuccess) { rc = hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__); goto fail; }
    if (clock.on) {
      (void)hipStreamSynchronize(s);
      ms_up = clock.lap();
    }
  }
  {
    int32_t* bad = nullptr;
    rc = dev_alloc(&bad, 1);
    if (rc) goto fail;
    int32_t hbad = 0;
    e = hipMemsetAsync(bad, 0, sizeof(int32_t), s);
    if (e == hipSuccess) e = kern::check_csc(s, g->CP, g->IC, n, nz, bad);
    if (e == hipSuccess) e = hipMemcpyAsync(&hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dev_free(bad);
    if (e != hipSuccess) { rc = hip_fail(e, "check_csc", __FILE__, __LINE__); goto fail; }
    if (hbad) { rc = TBFS_ERR_INVALID; goto fail; }
  }
  ms_check = clock.lap();
  if (g->flags & TBFS_GRAPH_SYMMETRIC) {
    g->RP = g->CP;
    g->JR = g->IC;
  } else {
    rc = graph_build_transpose(g, s);
    if (rc) goto fail;
  }
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) { rc = hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__); goto fail; }
  ms_tr = clock.lap();
  rc = graph_build_bigc(g, s);
  if (rc) goto fail;
  if (clock.on)
    std::printf("graph_create_csc ms: alloc + upload %.3f, check %.3f, transpose %.3f, "
                "slices %.3f\n", ms_up, ms_check, ms_tr, clock.lap());
  (void)hipStreamDestroy(s);
  *out = g;
  return TBFS_OK;
fail:
  if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  graph_free(g);
  return rc;
}

int graph_create_cooc(const int32_t* I, const int32_t* J, int32_t n, int32_t nz,
                      uint32_t flags, tbfs_graph** out) {
  if (!out) return TBFS_ERR_INVALID;
  *out = nullptr;
  if (n <= 0 || nz < 0 || (nz > 0 && (!I || !J))) return TBFS_ERR_INVALID;
  if (flags & TBFS_GRAPH_DEVICE) return TBFS_ERR_UNSUPPORTED;
  bool sorted = true;
  for (int32_t k = 0; k < nz; ++k) {
    if ((uint32_t)I[k] >= (uint32_t)n || (uint32_t)J[k] >= (uint32_t)n)
      return TBFS_ERR_INVALID;
    if (k > 0 && J[k] < J[k - 1]) sorted = false;
  }
  tbfs_graph* g = new (std::nothrow) tbfs_graph();
  if (!g) return TBFS_ERR_NOMEM;
  g->n = n;
  g->nz = nz;
  g->flags = flags & TBFS_GRAPH_SYMMETRIC;
  hipStream_t s = nullptr;
  int rc = TBFS_OK;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) { rc = hip_fail(e, "hipStreamCreate", __FILE__, __LINE__); goto fail; }
  // I doubles as IC when J is sorted (CSC of the same matrix)
  rc = dev_alloc(&g->IC, (size_t)nz);
  if (!rc) rc = dev_alloc(&g->J, (size_t)nz);
  if (rc) goto fail;
  g->I = g->IC;
  g->J_owned = true;
  e = upload_h2d(g->IC, I, (size_t)nz * sizeof(int32_t), s);
  if (e == hipSuccess) e = upload_h2d(g->J, J, (size_t)nz * sizeof(int32_t), s);
  if (e != hipSuccess) { rc = hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__); goto fail; }
  if (sorted) {
    rc = graph_cp_from_sorted(g, g->J, s);
    if (rc) goto fail;
    g->has_csc = true;
    if (g->flags & TBFS_GRAPH_SYMMETRIC) {
      g->RP = g->CP;
      g->JR = g->IC;
    } else {
      rc = graph_build_transpose(g, s);
      if (rc) goto fail;
    }
  } else {
    g->has_csc = false;
  }
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) { rc = hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__); goto fail; }
  if (sorted) {
    rc = graph_build_bigc(g, s);
    if (rc) goto fail;
  }
  (void)hipStreamDestroy(s);
  *out = g;
  return TBFS_OK;
fail:
  if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  graph_free(g);
  return rc;
}

// Column-index array for the COOC kernel on a CSC graph (built on first use).
int graph_ensure_cooc(tbfs_graph* g, hipStream_t s) {
  if (g->J) return TBFS_OK;
  int rc = dev_alloc(&g->J, (size_t)g->nz);
  if (rc) return rc;
  g->J_owned = true;
  g->I = g->IC;
  HIPCHK(kern::expand_cols(s, g->CP, g->n, g->J));
  return TBFS_OK;
}

// Static slices (kPullChunk entries) of every list longer than min_deg, from
// the device list offsets ptr: slices per vertex, their exclusive scan, then
// the entries in vertex order (in-lists for the pull levels, out-lists for
// the push levels; min_deg 0 cuts every nonempty list, as batches use).
// Returns with the table complete.
int build_slices(const int32_t* ptr, int32_t n, int32_t min_deg, hipStream_t s, uint64_t** out,
                 int32_t* count, int32_t* nlists) {
  int32_t *cnt = nullptr, *off = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int32_t total = 0, lists = 0;
  hipError_t e = hipSuccess;
  int rc = dev_alloc(&cnt, (size_t)n + 2);  // cnt[n + 1]: lists cut
  if (!rc) rc = dev_alloc(&off, (size_t)n + 1);
  if (!rc) {
    e = hipMemsetAsync(cnt + n + 1, 0, sizeof(int32_t), s);
    if (e == hipSuccess) e = kern::slice_count(s, ptr, n, min_deg, cnt, cnt + n + 1);
    if (e == hipSuccess)
      e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, off, n + 1, s);
    if (e == hipSuccess) rc = dev_alloc((char**)&tmp, std::max<size_t>(tmp_bytes, 1));
  }
  if (!rc && e == hipSuccess) {
    e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, off, n + 1, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(&total, off + n, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(&lists, cnt + n + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (!rc && e == hipSuccess) rc = dev_alloc(out, (size_t)total);
  if (!rc && e == hipSuccess && total > 0) e = kern::slice_fill(s, cnt, off, n, *out);
  if (!rc && e == hipSuccess) e = hipStreamSynchronize(s);
  if (!rc && e != hipSuccess) rc = hip_fail(e, "build_slices", __FILE__, __LINE__);
  if (!rc) {
    *count = total;
    if (nlists) *nlists = lists;
  }
  dev_free(cnt);
  dev_free(off);
  dev_free(tmp);
  return rc;
}

// Slices of the long in-lists (and out-lists) and the bitmap of columns the
// light-column pull skips, built on the device from g->CP / g->RP (complete
// on stream s or earlier).  Returns with every table complete.
int graph_build_bigc(tbfs_graph* g, hipStream_t s) {
  int rc = build_slices(g->CP, g->n, kBigIn, s, &g->bigc, &g->nbigc, &g->nheavy_in);
  if (rc) return rc;
  const int64_t nw = ((int64_t)g->n + 63) / 64;
  rc = dev_alloc(&g->skip, (size_t)nw);
  if (rc) return rc;
  HIPCHK(kern::skip_bits(s, g->CP, g->n, g->skip, nw));
  HIPCHK(hipStreamSynchronize(s));
  if (g->RP == g->CP) {
    g->bigo = g->bigc;
    g->nbigo = g->nbigc;
    g->nheavy_out = g->nheavy_in;
    return TBFS_OK;
  }
  return build_slices(g->RP, g->n, kBigIn, s, &g->bigo, &g->nbigo, &g->nheavy_out);
}

// Guards the in-degree statistics kept in each graph (a server's dispatcher
// and the caller's threads may ask for them at the same time).
static std::mutex g_deg_mu;

int graph_degree_stats(tbfs_graph* g, int32_t which, hipStream_t s, tbfs_degree_stats* out) {
  if (!g || !out || (which != TBFS_DEGREE_IN && which != TBFS_DEGREE_OUT)) return TBFS_ERR_INVALID;
  if (which == TBFS_DEGREE_IN) {
    std::lock_guard<std::mutex> lk(g_deg_mu);
    if (g->deg_in_valid) {
      *out = g->deg_in;
      return TBFS_OK;
    }
  }
  const int32_t* ptr = which == TBFS_DEGREE_IN ? g->CP : g->RP;
  if (!ptr) return TBFS_ERR_UNSUPPORTED;
  DegAcc* acc = nullptr;
  int rc = dev_alloc(&acc, 1);
  if (rc) return rc;
  DegAcc h{};
  hipError_t e = hipMemsetAsync(acc, 0, sizeof(DegAcc), s);
  if (e == hipSuccess) e = kern::degree_stats(s, ptr, g->n, acc);
  if (e == hipSuccess) e = hipMemcpyAsync(&h, acc, sizeof(DegAcc), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dev_free(acc);
  if (e != hipSuccess) return hip_fail(e, "degree_stats", __FILE__, __LINE__);
  tbfs_degree_stats r{};
  const int64_t n = g->n;
  r.max = n > 0 ? (int32_t)(h.maxkey >> 32) : 0;
  r.argmax = n > 0 ? (int32_t)~(uint32_t)h.maxkey : 0;
  r.sum = (int64_t)h.sum;
  r.mean = n > 0 ? (double)h.sum / (double)n : 0.0;
  // sample variance from exact integer sums: (n sum(d^2) - sum(d)^2) / (n (n - 1))
  if (n > 1) {
    const unsigned __int128 sq = ((unsigned __int128)h.sq_hi << 32) + h.sq_lo;
    const unsigned __int128 num = (unsigned __int128)n * sq -
                                  (unsigned __int128)h.sum * (unsigned __int128)h.sum;
    const long double var = (long double)num / ((long double)n * (long double)(n - 1));
    r.sd = (double)std::sqrt(var);
  }
  r.nbins = 0;
  for (int k = 0; k < 33; ++k) {
    r.bins[k] = (int64_t)h.bins[k];
    if (h.bins[k]) r.nbins = k + 1;
  }
  if (which == TBFS_DEGREE_IN) {
    std::lock_guard<std::mutex> lk(g_deg_mu);
    g->deg_in = r;
    g->deg_in_valid = true;
  }
  *out = r;
  return TBFS_OK;
}

// Degree-ordered copy of g: vertices sorted by decreasing in-degree (stable,
// so ties keep their order), every list renamed and placed in its column's
// new slot; the copy is then set up like any device-resident CSC graph.
static int graph_build_view(tbfs_graph* g, hipStream_t s) {
  const int32_t n = g->n, nz = g->nz;
  uint32_t *keys = nullptr, *keys_out = nullptr;
  int32_t *ids = nullptr, *order = nullptr, *deg2 = nullptr, *CP2 = nullptr, *col = nullptr,
          *IC2 = nullptr, *perm = nullptr;
  void* tmp = nullptr;
  size_t tmp_sort = 0, tmp_scan = 0;
  tbfs_graph* v = nullptr;
  int rc = dev_alloc(&keys, (size_t)n);
  if (!rc) rc = dev_alloc(&keys_out, (size_t)n);
  if (!rc) rc = dev_alloc(&ids, (size_t)n);
  if (!rc) rc = dev_alloc(&order, (size_t)n);
  if (!rc) rc = dev_alloc(&deg2, (size_t)n + 1);
  if (!rc) rc = dev_alloc(&CP2, (size_t)n + 1);
  if (!rc) rc = dev_alloc(&col, (size_t)std::max(nz, 1));
  if (!rc) rc = dev_alloc(&IC2, (size_t)std::max(nz, 1));
  if (!rc) rc = dev_alloc(&perm, (size_t)n);
  if (rc) goto out;
  {
    hipError_t e = kern::order_keys(s, g->CP, n, keys, ids);
    if (e == hipSuccess)
      e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_sort, keys, keys_out, ids, order, n, 0,
                                             32, s);
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_scan, deg2, CP2, n + 1, s);
    if (e != hipSuccess) { rc = hip_fail(e, "view sizes", __FILE__, __LINE__); goto out; }
    rc = dev_alloc((char**)&tmp, std::max(tmp_sort, tmp_scan));
    if (rc) goto out;
    e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_sort, keys, keys_out, ids, order, n, 0, 32, s);
    if (e == hipSuccess) e = kern::order_perm(s, order, g->CP, n, perm, deg2);
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_scan, deg2, CP2, n + 1, s);
    if (e == hipSuccess && nz > 0) e = kern::expand_cols(s, g->CP, n, col);
    if (e == hipSuccess && nz > 0) e = kern::order_entries(s, g->IC, col, g->CP, CP2, perm, nz, IC2);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { rc = hip_fail(e, "view build", __FILE__, __LINE__); goto out; }
  }
  rc = graph_create_csc(CP2, IC2, n, nz, (g->flags & TBFS_GRAPH_SYMMETRIC) | TBFS_GRAPH_DEVICE, &v);
  if (rc) goto out;
  g->bview = v;
  g->bperm = perm;
  perm = nullptr;
out:
  dev_free(tmp);
  dev_free(keys);
  dev_free(keys_out);
  dev_free(ids);
  dev_free(order);
  dev_free(deg2);
  dev_free(CP2);
  dev_free(col);
  dev_free(IC2);
  dev_free(perm);
  return rc;
}

static std::mutex g_view_mu;

int graph_batch_view(tbfs_graph* g, int32_t order, hipStream_t s, tbfs_graph** view,
                     const int32_t** perm) {
  *view = g;
  *perm = nullptr;
  if (order < 0 || !g->CP || g->n < 2) return TBFS_OK;
  std::lock_guard<std::mutex> lk(g_view_mu);
  if (order == 0) {
    if (g->bview_default < 0) {
      tbfs_degree_stats ds;
      const int rc = graph_degree_stats(g, TBFS_DEGREE_IN, s, &ds);
      if (rc) return rc;
      g->bview_default = (ds.mean > 0 && ds.sd > 2.0 * ds.mean) ? 1 : 0;
    }
    if (!g->bview_default) return TBFS_OK;
  }
  if (!g->bview) {
    const int rc = graph_build_view(g, s);
    if (rc) return rc;
  }
  *view = g->bview;
  *perm = g->bperm;
  return TBFS_OK;
}

// ----------------------------------------------------------------------------
// contexts
void ctx_free(tbfs_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dev_free(c->S);
  dev_free(c->sigma);
  dev_free(c->ovf);
  for (int i = 0; i < 3; ++i) dev_free(c->fb[i]);
  dev_free(c->vis);
  for (int i = 0; i < 2; ++i) dev_free(c->q[i]);
  dev_free(c->ctr);
  dev_free(c->bar);
  dev_free(c->stats);
  dev_free(c->loop);
  dev_free(c->cset);
  dev_free(c->fin_list[0]);
  dev_free(c->fin_list[1]);
  dev_free(c->fin_ctr);
  dev_free(c->roots_d);
  dev_free(c->gstats);
  if (c->groots_h) (void)hipHostFree(c->groots_h);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->cset_h) (void)hipHostFree(c->cset_h);
  if (c->loop_h) (void)hipHostFree(c->loop_h);
  if (c->ctr_h) (void)hipHostFree(c->ctr_h);
  if (c->stats_h) (void)hipHostFree(c->stats_h);
  for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->marks)
    if (e) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int ctx_create(int32_t n_cap, tbfs_ctx** out) { return ctx_create_on(n_cap, nullptr, out); }

int ctx_create_on(int32_t n_cap, hipStream_t stream, tbfs_ctx** out) {
  if (!out || n_cap <= 0) return TBFS_ERR_INVALID;
  *out = nullptr;
  tbfs_ctx* c = new (std::nothrow) tbfs_ctx();
  if (!c) return TBFS_ERR_NOMEM;
  c->n_cap = n_cap;
  const size_t nw = ((size_t)n_cap + 63) / 64;
  int rc = TBFS_OK;
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) {
    if (stream) {
      c->stream = stream;
      c->own_stream = false;
    } else {
      e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    }
  }
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) { rc = hip_fail(e, "ctx stream/events", __FILE__, __LINE__); goto fail; }
  rc = dev_alloc(&c->S, (size_t)n_cap);
  if (!rc) rc = dev_alloc(&c->sigma, (size_t)n_cap);
  if (!rc) rc = dev_alloc(&c->ovf, (size_t)n_cap);
  for (int i = 0; i < 3 && !rc; ++i) rc = dev_alloc(&c->fb[i], nw);
  if (!rc) rc = dev_alloc(&c->vis, nw);
  if (!rc) rc = dev_alloc(&c->ctr, 4);
  if (!rc) rc = dev_alloc(&c->bar, kBarWords);
  if (!rc) rc = dev_alloc(&c->cset, 3);
  if (!rc) rc = dev_alloc(&c->stats, 1);
  if (!rc) rc = dev_alloc(&c->loop, 2);
  if (!rc) rc = dev_alloc(&c->fin_list[0], (size_t)std::min<uint32_t>(kFinCap, (uint32_t)n_cap));
  if (!rc) rc = dev_alloc(&c->fin_list[1], (size_t)std::min<uint32_t>(kFinCap, (uint32_t)n_cap));
  if (!rc) rc = dev_alloc(&c->fin_ctr, 16);
  if (rc) goto fail;
  e = hipHostMalloc((void**)&c->cset_h, sizeof(CtrSet), hipHostMallocDefault);
  if (e != hipSuccess) { rc = hip_fail(e, "hipHostMalloc", __FILE__, __LINE__); goto fail; }
  e = hipHostMalloc((void**)&c->loop_h, 2 * sizeof(LoopState), hipHostMallocDefault);
  if (e != hipSuccess) { rc = hip_fail(e, "hipHostMalloc", __FILE__, __LINE__); goto fail; }
  e = hipHostMalloc((void**)&c->ctr_h, 4 * sizeof(LevelCtr), hipHostMallocDefault);
  if (e != hipSuccess) { rc = hip_fail(e, "hipHostMalloc", __FILE__, __LINE__); goto fail; }
  c->stats_cap = 1;
  e = hipHostMalloc((void**)&c->stats_h, sizeof(RunStats), hipHostMallocDefault);
  if (e != hipSuccess) { rc = hip_fail(e, "hipHostMalloc", __FILE__, __LINE__); goto fail; }
  if (kern::persist_grid(c->device, 256, &c->persist_grid_max) != hipSuccess) {
    (void)hipGetLastError();
    c->persist_grid_max = 0;  // no cooperative launch: no persistent loop
  }
  if (kern::persist_grid(c->device, 1024, &c->persist_grid_max1024) != hipSuccess) {
    (void)hipGetLastError();
    c->persist_grid_max1024 = 0;
  }
  if (kern::chain_fbl_limit(&c->fbl_limit) != hipSuccess) {
    (void)hipGetLastError();
    c->fbl_limit = 0;
  }
  if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) {
    (void)hipGetLastError();
    c->cus = 256;
  }
  *out = c;
  return TBFS_OK;
fail:
  ctx_free(c);
  return rc;
}

// Queue capacity for graph g: one entry per kPushChunk slice of every
// out-list in a level, bounded by n + nz / kPushChunk.
static int ctx_reserve_queue(tbfs_ctx* c, const tbfs_graph* g) {
  const int64_t need = (int64_t)g->n + (int64_t)g->nz / kPushChunk + 64;
  if (c->q_cap >= need) return TBFS_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < 2; ++i) dev_free(c->q[i]);
  c->q_cap = 0;
  for (int i = 0; i < 2; ++i) {
    int rc = dev_alloc(&c->q[i], (size_t)need);
    if (rc) return rc;
  }
  c->q_cap = need;
  return TBFS_OK;
}

static int ctx_reserve_stats(tbfs_ctx* c, int count) {
  if (c->stats_cap >= count) return TBFS_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  dev_free(c->stats);
  if (c->stats_h) (void)hipHostFree(c->stats_h);
  c->stats_h = nullptr;
  c->stats_cap = 0;
  int rc = dev_alloc(&c->stats, (size_t)count);
  if (rc) return rc;
  HIPCHK(hipHostMalloc((void**)&c->stats_h, (size_t)count * sizeof(RunStats),
                       hipHostMallocDefault));
  c->stats_cap = count;
  return TBFS_OK;
}

static int ctx_reserve_events(tbfs_ctx* c, size_t count) {
  while (c->evs.size() < count) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    c->evs.push_back(e);
  }
  return TBFS_OK;
}

// ----------------------------------------------------------------------------
// argument checks shared by both level loops
// Defaults of the single-source tunables (tbfs_opts fields left at 0 when
// the autotuner's table has no entry for the graph).
constexpr float kDefaultAlpha = 14.0f;
constexpr float kDefaultBeta = 24.0f;
constexpr float kDefaultStageRatio = 1.0f;

// Options with the tunable fields left at 0 filled from the autotuner's
// table entry for g's class, when there is one (the others keep 0: defaults).
static tbfs_opts resolve_opts(const tbfs_graph* g, const tbfs_opts& o, hipStream_t s) {
  tbfs_opts r = o;
  if (r.alpha > 0 && r.beta > 0 && r.stage_ratio > 0 && r.lb_grid != 0) return r;
  if (tune_empty() || !g->CP) return r;
  TuneKey k;
  TuneEntry e;
  if (tune_key(const_cast<tbfs_graph*>(g), 0, s, &k) != TBFS_OK) {
    (void)hipGetLastError();
    return r;
  }
  if (!tune_find(k, &e)) return r;
  if (r.alpha <= 0) r.alpha = e.alpha;
  if (r.beta <= 0) r.beta = e.beta;
  if (r.stage_ratio <= 0) r.stage_ratio = e.stage_ratio;
  if (r.lb_grid == 0) r.lb_grid = e.lb_grid;
  return r;
}

static int check_run(tbfs_ctx* c, const tbfs_graph* g, int32_t root,
                     const tbfs_opts& o) {
  if (!c || !g) return TBFS_ERR_INVALID;
  if (g->n > c->n_cap) return TBFS_ERR_INVALID;
  if (root < 0 || root >= g->n) return TBFS_ERR_INVALID;
  if (o.mode < TBFS_MODE_AUTO || o.mode > TBFS_MODE_COOC) return TBFS_ERR_INVALID;
  if (o.pull < TBFS_PULL_SC || o.pull > TBFS_PULL_LB) return TBFS_ERR_INVALID;
  if (o.loop < TBFS_LOOP_HOST || o.loop > TBFS_LOOP_GRAPH) return TBFS_ERR_INVALID;
  if (o.mode != TBFS_MODE_COOC && !g->CP) return TBFS_ERR_UNSUPPORTED;
  if (o.mode == TBFS_MODE_COOC && !o.sigma) return TBFS_ERR_UNSUPPORTED;
  if (o.block != 0 && (o.block % 64 != 0 || o.block > 1024 || o.block < 64))
    return TBFS_ERR_INVALID;
  if (o.loop == TBFS_LOOP_PERSISTENT && o.block != 0 && o.block != 256 && o.block != 1024)
    return TBFS_ERR_INVALID;
  return TBFS_OK;
}

// Which level loop runs: the device-resident loops (persistent, chained)
// cover the load-balanced AUTO / PUSH path; the COOC edge sweep and the
// one-lane / one-wavefront reference mappings run host-driven.
static int pick_loop(const tbfs_ctx* c, const tbfs_opts& o) {
  if (o.loop == TBFS_LOOP_HOST) return TBFS_LOOP_HOST;
  if (o.mode == TBFS_MODE_COOC || o.mode == TBFS_MODE_PULL_SC ||
      o.mode == TBFS_MODE_PULL_WA || o.pull != TBFS_PULL_LB)
    return TBFS_LOOP_HOST;
  if (o.loop == TBFS_LOOP_PERSISTENT) {
    const int gmax = o.block == 1024 ? c->persist_grid_max1024 : c->persist_grid_max;
    return gmax > 0 ? TBFS_LOOP_PERSISTENT : TBFS_LOOP_CHAINED;
  }
  if (o.loop == TBFS_LOOP_GRAPH) return TBFS_LOOP_GRAPH;
  return TBFS_LOOP_CHAINED;
}

static PersistParams make_params(tbfs_ctx* c, const tbfs_graph* g, int32_t root,
                                 const tbfs_opts& o, RunStats* stats_slot) {
  PersistParams p{};
  p.g = GraphView{g->n, g->nz, g->CP, g->IC, g->RP, g->JR};
  p.nroots = 0;
  p.roots = nullptr;
  p.S = c->S;
  p.sigma = c->sigma;
  p.ovf = c->ovf;
  for (int i = 0; i < 3; ++i) p.fb[i] = c->fb[i];
  p.vis = c->vis;
  p.skip = g->skip;
  p.cset = c->cset;
  p.bar = c->bar;
  p.stats = stats_slot;
  p.bigc = g->bigc;
  p.nbigc = g->nbigc;
  p.nheavy = g->nheavy_in;
  p.bigo = g->bigo;
  p.nbigo = g->nbigo;
  p.root = root;
  p.mode = o.mode;
  p.sigma_on = o.sigma ? 1 : 0;
  p.alpha = o.alpha > 0 ? o.alpha : kDefaultAlpha;
  p.beta = o.beta > 0 ? o.beta : kDefaultBeta;
  p.stage_ratio = o.stage_ratio > 0 ? o.stage_ratio : kDefaultStageRatio;
  p.loop = c->loop;
  // the finisher's lists hold min(n_cap, kFinCap) columns: the unreached
  // columns with in-lists number at most min(n, unexplored entries), so
  // fin_mu is capped at kFinCap only when n_cap exceeds it
  p.fin_list[0] = c->fin_list[0];
  p.fin_list[1] = c->fin_list[1];
  p.fin_ctr = c->fin_ctr;
  p.fin_mu = o.finish <= 0 ? 0u
             : (uint32_t)c->n_cap <= kFinCap ? (uint32_t)o.finish
                                             : std::min<uint32_t>((uint32_t)o.finish, kFinCap);
  // one root, AUTO / PUSH: the init kernel may run depth 0 (chained loop)
  p.step0 = (root >= 0 && (o.mode == TBFS_MODE_AUTO || o.mode == TBFS_MODE_PUSH)) ? 1 : 0;
  return p;
}

// Enqueue one device-resident BFS (no host synchronisation).  *refused is
// set when the runtime declines the cooperative launch (for instance while
// another process on the device holds the cooperative-launch resources):
// nothing was enqueued then, and the caller runs the chained loop instead.
static int enqueue_persistent(tbfs_ctx* c, const tbfs_graph* g, int32_t root,
                              const tbfs_opts& o, RunStats* stats_slot,
                              bool* refused) {
  const PersistParams p = make_params(c, g, root, o, stats_slot);
  const int block = o.block == 1024 ? 1024 : 256;
  int grid = block == 1024 ? c->persist_grid_max1024 : c->persist_grid_max;
  if (o.grid > 0 && o.grid < grid) grid = o.grid;
  HIPCHK(hipMemsetAsync(c->bar, 0, kBarWords * sizeof(unsigned), c->stream));
  const hipError_t e = kern::persist(c->stream, grid, block, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (std::getenv("TBFS_VERBOSE"))
      std::fprintf(stderr, "turbobfs: cooperative launch declined (%s); "
                           "using the chained level loop\n", hipGetErrorString(e));
    *refused = true;
  }
  return TBFS_OK;
}

// LDS bytes of the staged frontier bitmap for the chained levels (0: none).
static size_t chain_fbl(const tbfs_ctx* c, const tbfs_graph* g, const tbfs_opts& o) {
  const size_t bytes = (((size_t)g->n + 63) / 64) * sizeof(uint64_t);
  if (o.fb_lds < 0 || c->fbl_limit == 0 || bytes > c->fbl_limit) return 0;
  return (bytes + 15) & ~(size_t)15;
}

// Levels past kChainWideLevels belong to a long thin tail (high-diameter
// graph): without a staged bitmap they run on one workgroup per CU, where a
// level costs ~10 us instead of ~16 us on the full grid (any grid is correct).
constexpr int kChainWideLevels = 16;

// Shape of one chained level launch: workgroups, dynamic LDS (0: the
// kLbBlock-thread kernel without a staged bitmap) and which levels it runs.
struct LevelShape {
  int grid;
  size_t fbl;
  int32_t want = kLevelAlways;
};

// The launch shapes of a chained BFS on g: every shape gives the same
// results; they differ in speed.  Learned from the context's earlier calls on
// g (decayed votes per level):
//   - staging workgroups (1024 threads, one per CU, the bitmap in LDS) for
//     the levels that staged for most sources; both shapes, each running the
//     level only if the device's staging decision matches, for the levels
//     that staged for a good part of them (an unstaged heavy pull costs ~17
//     us more on R-MAT scale 20, a staged push ~10 us, a launch that
//     returns ~5 us);
//   - 256-thread workgroups for the others, as many as the level's expected
//     work calls for: one per CU up to 2^18 entries, two up to 2^22, else
//     the default grid (a launch's dispatch and drain grow with its grid: a
//     level of a few thousand entries takes ~9 us on 256 workgroups and
//     ~15 us on 1024);
//   - launches past the levels those calls needed (the one that finds the
//     frontier empty, spare launches) on a small grid.
// No history: every launch stages.  o.grid > 0 fixes the grid, o.lb_grid > 0
// the unstaged launches' grid.
struct LevelPlan {
  int grid_stage = 0, grid_lb = 0, grid_spare = 0, cus = 1;
  size_t fbl = 0;
  uint64_t stage = ~0ull;
  uint64_t dual = 0;
  uint8_t work[64] = {};
  int levels = 0;
  bool fixed = false;
  bool lb_fixed = false;

  int lb_grid(int d) const {
    const int w = d < 64 ? work[d] : 0;
    if (!lb_fixed && w > 0) {
      if (w <= 18) return std::min(grid_lb, cus);
      if (w <= 22) return std::min(grid_lb, 2 * cus);
      return grid_lb;
    }
    return d < kChainWideLevels ? grid_lb : std::min(grid_lb, cus);
  }
  // the launches of step d (one or two); returns how many
  int at(int d, LevelShape (&out)[2]) const {
    const uint64_t bit = 1ull << (d < 63 ? d : 63);
    const bool st = fbl && (stage & bit);
    if (fixed) {
      out[0] = {st ? grid_stage : grid_lb, st ? fbl : 0};
      return 1;
    }
    if (fbl && (dual & bit) && !(levels > 0 && d >= levels)) {
      out[0] = {grid_stage, fbl, kLevelIfStaged};
      out[1] = {lb_grid(d), 0, kLevelIfNotStaged};
      return 2;
    }
    if (st) {
      out[0] = {grid_stage, fbl};
      return 1;
    }
    if (levels > 0 && d >= levels) {
      out[0] = {grid_spare, 0};
      return 1;
    }
    out[0] = {lb_grid(d), 0};
    return 1;
  }
};

// Enqueue the launches of chained step d.
static hipError_t chain_step(hipStream_t s, const LevelPlan& pl, const PersistParams& p, int d) {
  LevelShape sh[2];
  const int k = pl.at(d, sh);
  for (int i = 0; i < k; ++i) {
    const hipError_t e = kern::chain_level(s, sh[i].grid, p, d, sh[i].fbl, sh[i].want);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Learn the launch plan of g from finished chained BFS (stats with levels <
// 0 did not finish and are skipped): per level d, decayed votes (each source
// weighs 0.95 of the one after it).  A level stages where more than 0.45 of
// the sources that ran it staged, runs in both shapes where more than 0.27
// did (on R-MAT scale 20 a lone staging launch loses ~10 us per push it
// runs, a lone unstaged one ~17 us per heavy pull, and the second launch of
// a pair costs ~5.5 us); its grid follows the sources' mean work there.
void plan_learn(const tbfs_graph* g, const RunStats* st, int count) {
  PlanHistory& h = g->plan;
  std::lock_guard<std::mutex> lk(h.mu);
  bool any = false;
  for (int i = 0; i < count; ++i) {
    const RunStats& r = st[i];
    if (r.levels < 0) continue;
    any = true;
    const int lv = std::min(64, (int)r.levels);
    for (int d = 0; d < 64; ++d) {
      h.stage_votes[d] *= 0.95f;
      h.level_votes[d] *= 0.95f;
      h.work_votes[d] *= 0.95f;
    }
    for (int d = 0; d < lv; ++d) {
      h.level_votes[d] += 1.f;
      h.work_votes[d] += (float)r.work_lg[d];
      if ((r.stage_mask >> d) & 1ull) h.stage_votes[d] += 1.f;
    }
    h.levels = std::max(h.levels, (int)r.levels);
  }
  if (!any) return;
  uint64_t stage = 0, dual = 0;
  for (int d = 0; d < 64; ++d) {
    const float lvv = h.level_votes[d];
    if (h.stage_votes[d] > 0.45f * lvv) stage |= 1ull << d;
    else if (h.stage_votes[d] > 0.27f * lvv) dual |= 1ull << d;
    const float w = lvv > 0.f ? h.work_votes[d] / lvv : 0.f;
    h.work[d] = (uint8_t)std::min(63.f, std::ceil(w));
  }
  h.stage = stage;
  h.dual = dual;
  h.valid = true;
}

// about four frontier words per wavefront: fewer workgroups to dispatch per
// level and fewer counter atomics, still >> 256 CUs' worth of waves
static int default_lb_grid(const tbfs_graph* g) {
  return grid_for((((int64_t)g->n + 63) / 64) * 16, kLbBlock, 2048);
}

static LevelPlan level_plan(const tbfs_ctx* c, const tbfs_graph* g, const tbfs_opts& o) {
  LevelPlan pl;
  pl.fbl = chain_fbl(c, g, o);
  pl.grid_lb = o.lb_grid > 0 ? std::min(o.lb_grid, 65536) : default_lb_grid(g);
  pl.lb_fixed = o.lb_grid > 0;
  pl.cus = std::max(1, c->cus);
  pl.grid_stage = std::max(1, c->cus);
  pl.grid_spare = std::max(1, c->cus / 8);
  if (o.grid > 0) {
    pl.fixed = true;
    pl.grid_lb = pl.grid_stage = std::min(o.grid, 65536);
  }
  {
    PlanHistory& h = g->plan;
    std::lock_guard<std::mutex> lk(h.mu);
    if (h.valid) {
      pl.stage = h.stage;
      pl.dual = h.dual;
      std::memcpy(pl.work, h.work, sizeof(pl.work));
      pl.levels = h.levels;
    }
  }
  return pl;
}

// Chained loop, enqueue only: init plus the level kernels of steps
// p.step0 .. levels - 1.  The BFS has finished within them iff
// stats_slot->levels >= 0 afterwards.
static int enqueue_chained(tbfs_ctx* c, const tbfs_graph* g, int32_t root,
                           const tbfs_opts& o, RunStats* stats_slot, int levels) {
  const PersistParams p = make_params(c, g, root, o, stats_slot);
  const LevelPlan pl = level_plan(c, g, o);
  hipStream_t s = c->stream;
  HIPCHK(kern::chain_init(s, grid_for(g->n, 256, 8192), p));
  for (int d = p.step0; d < levels; ++d) HIPCHK(chain_step(s, pl, p, d));
  return TBFS_OK;
}

// Chained loop for one BFS: init, then chunks of level kernels (chain_k, then
// growing); the host reads the loop state once per chunk.  Ends with the
// stream drained.
static int run_chained_p(tbfs_ctx* c, const tbfs_graph* g, const PersistParams& p,
                         const tbfs_opts& o) {
  LevelPlan pl = level_plan(c, g, o);
  pl.levels = 0;  // a run the host checks: no launch is taken for a spare
  hipStream_t s = c->stream;
  HIPCHK(kern::chain_init(s, grid_for(g->n, 256, 8192), p));
  if (p.nroots > 0) HIPCHK(kern::chain_seed(s, p));
  // chunks of level kernels: chain_k, then growing by half up to 256, so a
  // BFS far deeper than the budget costs few host checks and at most one
  // chunk of spare launches
  int32_t d = p.step0;
  int chunk = std::max(c->chain_k, 8);
  for (;;) {
    for (int i = 0; i < chunk; ++i) HIPCHK(chain_step(s, pl, p, d + i));
    d += chunk;
    chunk = std::min(chunk + chunk / 2, 256);
    HIPCHK(hipMemcpyAsync(c->loop_h, c->loop + (d & 1), sizeof(LoopState),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (c->loop_h->done) break;
  }
  return TBFS_OK;
}

static int run_chained(tbfs_ctx* c, const tbfs_graph* g, int32_t root,
                       const tbfs_opts& o, RunStats* stats_slot) {
  return run_chained_p(c, g, make_params(c, g, root, o, stats_slot), o);
}

static void fill_stats(tbfs_stats* st, const RunStats& rs, float ms) {
  st->levels = rs.levels;
  st->push_levels = rs.push_levels;
  st->pull_levels = rs.pull_levels;
  st->pad = 0;
  st->launches = rs.launches;
  st->pad2 = 0;
  st->reached = (int64_t)rs.reached;
  st->edges = (int64_t)rs.edges;
  st->ovf64 = rs.ovf_any ? -1 : 0;  // count known after tbfs_ctx_read
  st->ms = (double)ms;
}

// ----------------------------------------------------------------------------
// the BFS: host-driven level loop (one launch per level, counters read back)
static int run_host_loop(tbfs_ctx* c, tbfs_graph* g, int32_t root,
                         const tbfs_opts& o, tbfs_stats* st) {
  if (o.mode == TBFS_MODE_COOC) {
    int rc = graph_ensure_cooc(g, c->stream);
    if (rc) return rc;
  }
  const bool csc_ok = (g->CP != nullptr);
  const int block = o.block > 0 ? o.block : 256;
  const int32_t n = g->n;
  const int32_t nwords = (int32_t)(((int64_t)n + 63) / 64);
  hipStream_t s = c->stream;
  GraphView gv{n, g->nz, g->CP, g->IC, g->RP, g->JR};
  const int fill_grid = grid_for(n, 256, 8192);
  const int lb_grid = grid_for((int64_t)nwords * 16, kLbBlock, 2048);

  // root degrees (the level-0 counters)
  int32_t cp2[2] = {0, 0}, rp2[2] = {0, 0};
  if (csc_ok) {
    HIPCHK(hipMemcpyAsync(cp2, g->CP + root, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(rp2, g->RP + root, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  HIPCHK(hipEventRecord(c->ev0, s));
  HIPCHK(kern::init(s, fill_grid, n, root, rp2[1] - rp2[0], c->S, c->sigma,
                    c->ovf, c->fb[0], c->fb[1], c->fb[2], c->q[0], c->vis));

  int64_t qsize = 1;
  int64_t qe = (rp2[1] - rp2[0] + kPushChunk - 1) / kPushChunk;
  uint64_t edges = (uint64_t)(cp2[1] - cp2[0]);
  uint64_t m_frontier = (uint64_t)(rp2[1] - rp2[0]);
  uint64_t m_unexplored = (uint64_t)g->nz - edges;
  int64_t reached = 1;
  uint32_t ovf_any = 0;
  int cur = 0;  // fb[cur] holds depth d
  int32_t d = 0;
  int push_levels = 0, pull_levels = 0;
  bool pushing = (o.mode == TBFS_MODE_PUSH || o.mode == TBFS_MODE_AUTO);
  const float alpha = o.alpha > 0 ? o.alpha : kDefaultAlpha;
  const float beta = o.beta > 0 ? o.beta : kDefaultBeta;
  const int sig = o.sigma ? 1 : 0;

  for (;;) {
    if (o.mode == TBFS_MODE_AUTO) {
      if (pushing && (double)m_frontier * alpha > (double)m_unexplored) pushing = false;
      else if (!pushing && (double)qsize * beta < (double)n) pushing = true;
    }
    const int nxt = (cur + 1) % 3;
    StateView sv{c->S, c->sigma, c->ovf, c->fb[cur], c->fb[nxt],
                 c->q[d & 1], c->q[(d + 1) & 1], c->ctr, c->cset, c->vis, g->skip};
    HIPCHK(kern::level_prologue(s, grid_for(nwords, 256, 4096), c->fb[nxt],
                                nwords, c->ctr, c->cset, c->fb[cur], c->vis));
    bool striped = false;  // load-balanced kernels count into the stripes
    switch (o.mode) {
      case TBFS_MODE_PULL_SC:
        HIPCHK(kern::pull_sc(s, grid_for(n, block, 65536), block, gv, sv, d,
                             sig, o.bbreak, (int)ovf_any));
        ++pull_levels;
        break;
      case TBFS_MODE_PULL_WA:
        HIPCHK(kern::pull_wa(s, grid_for((int64_t)nwords * 64, block, 65536),
                             block, gv, sv, d, sig, o.bbreak, (int)ovf_any));
        ++pull_levels;
        break;
      case TBFS_MODE_COOC:
        HIPCHK(kern::cooc_scatter(s, grid_for(g->nz, block, 65536), block,
                                  g->I, g->J, g->nz, sv, (int)ovf_any));
        HIPCHK(kern::cooc_finish(s, grid_for(n, block, 65536), block, gv, sv, d));
        ++pull_levels;
        break;
      default:  // PUSH, or AUTO on its current side
        if (pushing) {
          if (o.pull == TBFS_PULL_LB) {
            HIPCHK(kern::push_lb(s, lb_grid, gv, sv, d, g->bigo, g->nbigo,
                                 (int)ovf_any));
            striped = true;
          }
          else
            HIPCHK(kern::push(s, grid_for(qe * 64, block, 65536), block, gv, sv,
                              d, qe, (int)ovf_any));
          ++push_levels;
        } else {
          if (o.pull == TBFS_PULL_LB) {
            HIPCHK(kern::pull_lb(s, lb_grid, gv, sv, d, g->bigc, g->nbigc, sig,
                                 (int)ovf_any));
            striped = true;
          }
          else if (o.pull == TBFS_PULL_WA)
            HIPCHK(kern::pull_wa(s, grid_for((int64_t)nwords * 64, block, 65536),
                                 block, gv, sv, d, sig, o.bbreak, (int)ovf_any));
          else
            HIPCHK(kern::pull_sc(s, grid_for(n, block, 65536), block, gv, sv, d,
                                 sig, o.bbreak, (int)ovf_any));
          ++pull_levels;
        }
        break;
    }
    LevelCtr lc{};
    if (striped) {
      HIPCHK(hipMemcpyAsync(c->cset_h, c->cset, sizeof(CtrSet), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      for (int i = 0; i < kStripes; ++i) {
        lc.q += c->cset_h->s[i].qm >> 32;
        lc.mout += c->cset_h->s[i].qm & 0xffffffffull;
        lc.min += c->cset_h->s[i].min;
        lc.ovf |= c->cset_h->s[i].ovf;
      }
    } else {
      HIPCHK(hipMemcpyAsync(c->ctr_h, c->ctr, sizeof(LevelCtr), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      lc = c->ctr_h[0];
    }
    if (lc.q == 0) break;
    qsize = (int64_t)lc.q;
    qe = (int64_t)lc.qe;
    reached += (int64_t)lc.q;
    edges += lc.min;
    m_frontier = lc.mout;
    m_unexplored = (m_unexplored > lc.min) ? m_unexplored - lc.min : 0;
    ovf_any |= lc.ovf;
    cur = nxt;
    ++d;
  }
  HIPCHK(hipEventRecord(c->ev1, s));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_n = n;
  if (st) {
    RunStats rs{};
    rs.levels = d + 1;
    rs.push_levels = push_levels;
    rs.pull_levels = pull_levels;
    rs.ovf_any = ovf_any;
    rs.reached = (unsigned long long)reached;
    rs.edges = edges;
    fill_stats(st, rs, ms);
  }
  return TBFS_OK;
}

// Graph loop: init + seed + K chained level kernels captured once as a HIP
// graph (the root is read from device memory, the level kernels take their
// depth from the loop state), then replayed per source.  The capture is kept
// until the graph, the options or the level budget change.
static int graph_chunk(tbfs_ctx* c, const tbfs_graph* g, const tbfs_opts& o, int K) {
  PersistParams p = make_params(c, g, -1, o, c->gstats);
  p.nroots = 1;
  p.roots = c->roots_d;
  const LevelPlan pl = level_plan(c, g, o);
  LevelShape sh0[2];
  pl.at(0, sh0);
  const int grid = sh0[0].grid;
  const size_t fbl = pl.fbl;
  if (c->gexec && c->gK == K && c->ggrid == grid && c->gfbl == fbl &&
      c->gstage == pl.stage && c->gdual == pl.dual && c->glevels == pl.levels &&
      std::memcmp(c->gwork, pl.work, sizeof(pl.work)) == 0 &&
      std::memcmp(&p, &c->gparams, sizeof(p)) == 0)
    return TBFS_OK;
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  hipStream_t s = c->stream;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipError_t e = kern::chain_init(s, grid_for(g->n, 256, 8192), p);
  if (e == hipSuccess) e = kern::chain_seed(s, p);
  for (int i = 0; i < K && e == hipSuccess; ++i) e = chain_step(s, pl, p, i);
  hipGraph_t gr = nullptr;
  const hipError_t e2 = hipStreamEndCapture(s, &gr);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (gr) (void)hipGraphDestroy(gr);
    return hip_fail(e != hipSuccess ? e : e2, "level-loop capture", __FILE__, __LINE__);
  }
  e = hipGraphInstantiate(&c->gexec, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  if (e != hipSuccess) {
    c->gexec = nullptr;
    return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
  }
  c->gK = K;
  c->ggrid = grid;
  c->gfbl = fbl;
  c->gstage = pl.stage;
  c->gdual = pl.dual;
  std::memcpy(c->gwork, pl.work, sizeof(pl.work));
  c->glevels = pl.levels;
  c->gparams = p;
  return TBFS_OK;
}

static int ctx_reserve_graph(tbfs_ctx* c, int count) {
  if (!c->gstats) {
    int rc = dev_alloc(&c->gstats, 1);
    if (rc) return rc;
  }
  if (c->roots_cap < 1) {
    int rc = dev_alloc(&c->roots_d, 1);
    if (rc) return rc;
    c->roots_cap = 1;
  }
  if (c->groots_cap < count) {
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->groots_h) (void)hipHostFree(c->groots_h);
    c->groots_h = nullptr;
    c->groots_cap = 0;
    HIPCHK(hipHostMalloc((void**)&c->groots_h, (size_t)count * sizeof(int32_t),
                         hipHostMallocDefault));
    c->groots_cap = count;
  }
  return TBFS_OK;
}

int run_bfs_many(tbfs_ctx* c, const tbfs_graph* gc, const int32_t* roots,
                 int count, const tbfs_opts& o_in, tbfs_stats* st) {
  tbfs_graph* g = const_cast<tbfs_graph*>(gc);
  if (count < 1 || !roots || !c || !g) return TBFS_ERR_INVALID;
  const tbfs_opts o = resolve_opts(g, o_in, c->stream);
  for (int i = 0; i < count; ++i) {
    int rc = check_run(c, g, roots[i], o);
    if (rc) return rc;
  }
  int rc = ctx_reserve_queue(c, g);
  if (rc) return rc;
  int loop = pick_loop(c, o);
  if (loop == TBFS_LOOP_HOST) {
    for (int i = 0; i < count; ++i) {
      rc = run_host_loop(c, g, roots[i], o, st ? &st[i] : nullptr);
      if (rc) return rc;
    }
    return TBFS_OK;
  }
  rc = ctx_reserve_stats(c, count);
  if (!rc) rc = ctx_reserve_events(c, (size_t)count + 1);
  if (rc) return rc;
  hipStream_t s = c->stream;
  // Chained sources are enqueued back to back with a level budget of
  // chain_k kernels each; a source that needs more is found afterwards from
  // its stats slot and run again with host checks.
  const int budget = c->chain_k;
  if (loop == TBFS_LOOP_GRAPH) {
    rc = ctx_reserve_graph(c, count);
    if (!rc) rc = graph_chunk(c, g, o, (budget + 1) & ~1);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(c->evs[0], s));
  for (int i = 0; i < count; ++i) {
    if (loop == TBFS_LOOP_GRAPH) {
      c->groots_h[i] = roots[i];
      HIPCHK(hipMemcpyAsync(c->roots_d, c->groots_h + i, sizeof(int32_t),
                            hipMemcpyHostToDevice, s));
      HIPCHK(hipGraphLaunch(c->gexec, s));
      HIPCHK(hipMemcpyAsync(c->stats + i, c->gstats, sizeof(RunStats),
                            hipMemcpyDeviceToDevice, s));
      HIPCHK(hipEventRecord(c->evs[i + 1], s));
      continue;
    }
    if (loop == TBFS_LOOP_PERSISTENT) {
      bool refused = false;
      rc = enqueue_persistent(c, g, roots[i], o, c->stats + i, &refused);
      if (refused) loop = TBFS_LOOP_CHAINED;  // this source and the rest
    }
    if (loop == TBFS_LOOP_CHAINED) rc = enqueue_chained(c, g, roots[i], o, c->stats + i, budget);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->evs[i + 1], s));
  }
  HIPCHK(hipMemcpyAsync(c->stats_h, c->stats, (size_t)count * sizeof(RunStats),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  std::vector<float> ms((size_t)count, 0.f);
  for (int i = 0; i < count; ++i)
    HIPCHK(hipEventElapsedTime(&ms[i], c->evs[i], c->evs[i + 1]));
  if (loop == TBFS_LOOP_CHAINED || loop == TBFS_LOOP_GRAPH) {
    bool redo_last = false;
    int maxl = 0;
    for (int i = 0; i < count; ++i) {
      if (c->stats_h[i].levels < 0) {
        HIPCHK(hipEventRecord(c->ev0, s));
        rc = run_chained(c, g, roots[i], o, c->stats + i);
        if (rc) return rc;
        HIPCHK(hipEventRecord(c->ev1, s));
        HIPCHK(hipEventSynchronize(c->ev1));
        HIPCHK(hipEventElapsedTime(&ms[i], c->ev0, c->ev1));
        HIPCHK(hipMemcpy(&c->stats_h[i], c->stats + i, sizeof(RunStats), hipMemcpyDeviceToHost));
        redo_last = i != count - 1;
      }
      maxl = std::max(maxl, c->stats_h[i].launches);
    }
    if (redo_last) {  // leave the last source's results in the context
      rc = run_chained(c, g, roots[count - 1], o, c->stats + count - 1);
      if (rc) return rc;
    }
    // the next call's budget: the launches these sources needed plus one
    // spare, so a source needing one launch more than any seen here still
    // finishes without a re-run (a spare launch costs ~4 us, a re-run a BFS)
    c->chain_k = std::min(4096, std::max(2, maxl + 1));
    plan_learn(g, c->stats_h, count);
  }
  c->last_n = g->n;
  if (st) {
    for (int i = 0; i < count; ++i) fill_stats(&st[i], c->stats_h[i], ms[i]);
  }
  return TBFS_OK;
}

// One BFS whose depth 0 is every root in roots[0 .. nroots-1]: S = distance
// to the nearest root, sigma = number of shortest paths from any root.  With
// one root per connected component (block-diagonal packing of many graphs)
// every component gets exactly its single-source result.  Chained loop.
int run_multi(tbfs_ctx* c, const tbfs_graph* g, const int32_t* roots, int nroots,
              const tbfs_opts& o_in, tbfs_stats* st) {
  if (!c || !g || !roots || nroots < 1) return TBFS_ERR_INVALID;
  tbfs_opts o = resolve_opts(g, o_in, c->stream);
  if (o.mode != TBFS_MODE_AUTO && o.mode != TBFS_MODE_PUSH) return TBFS_ERR_UNSUPPORTED;
  o.pull = TBFS_PULL_LB;
  o.loop = TBFS_LOOP_CHAINED;
  for (int i = 0; i < nroots; ++i) {
    const int rc = check_run(c, g, roots[i], o);
    if (rc) return rc;
  }
  int rc = ctx_reserve_stats(c, 1);
  if (rc) return rc;
  if (c->roots_cap < nroots) {
    HIPCHK(hipStreamSynchronize(c->stream));
    // a captured level loop reads its root from roots_d: capture it again
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    dev_free(c->roots_d);
    c->roots_cap = 0;
    rc = dev_alloc(&c->roots_d, (size_t)nroots);
    if (rc) return rc;
    c->roots_cap = nroots;
  }
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(c->roots_d, roots, (size_t)nroots * sizeof(int32_t),
                        hipMemcpyHostToDevice, s));
  PersistParams p = make_params(c, g, -1, o, c->stats);
  p.nroots = nroots;
  p.roots = c->roots_d;
  HIPCHK(hipEventRecord(c->ev0, s));
  rc = run_chained_p(c, g, p, o);
  if (rc) return rc;
  HIPCHK(hipEventRecord(c->ev1, s));
  HIPCHK(hipMemcpyAsync(c->stats_h, c->stats, sizeof(RunStats), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_n = g->n;
  if (st) fill_stats(st, c->stats_h[0], ms);
  return TBFS_OK;
}

// Asynchronous single-source BFS for the query server: each source is a
// chained BFS (level budget chain_k) followed by the copy of its results to
// host buffers, all on the context's stream (`levels` level kernels per
// source); ctx_async_finish queues the stats copy and an event,
// ctx_async_poll reports when that event has passed, which sources ran out
// of levels (the caller enqueues them again with a larger budget) and the
// most level launches a finished source used.
int ctx_async_begin(tbfs_ctx* c, int count) { return ctx_reserve_stats(c, count); }

int ctx_async_enqueue(tbfs_ctx* c, const tbfs_graph* g, int32_t root, const tbfs_opts& o_in,
                      int idx, int levels, int32_t* S, uint64_t* sigma, uint8_t* ovf) {
  if (idx < 0 || idx >= c->stats_cap || levels < 1) return TBFS_ERR_INVALID;
  const tbfs_opts o = resolve_opts(g, o_in, c->stream);
  int rc = check_run(c, g, root, o);
  if (rc) return rc;
  rc = enqueue_chained(c, g, root, o, c->stats + idx, levels);
  if (rc) return rc;
  c->async_g = g;
  hipStream_t s = c->stream;
  const size_t n = (size_t)g->n;
  if (S) HIPCHK(hipMemcpyAsync(S, c->S, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if (sigma) HIPCHK(hipMemcpyAsync(sigma, c->sigma, n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  if (ovf) HIPCHK(hipMemcpyAsync(ovf, c->ovf, n, hipMemcpyDeviceToHost, s));
  c->last_n = g->n;
  return TBFS_OK;
}

int ctx_async_finish(tbfs_ctx* c, int count) {
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(c->stats_h, c->stats, (size_t)count * sizeof(RunStats),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(c->ev1, s));
  return TBFS_OK;
}

int ctx_async_poll(tbfs_ctx* c, int count, bool* ready, std::vector<int>* unfinished,
                   int* maxl) {
  *ready = false;
  const hipError_t e = hipEventQuery(c->ev1);
  if (e == hipErrorNotReady) return TBFS_OK;
  if (e != hipSuccess) return hip_fail(e, "hipEventQuery", __FILE__, __LINE__);
  *ready = true;
  unfinished->clear();
  *maxl = 0;
  for (int i = 0; i < count; ++i) {
    if (c->stats_h[i].levels < 0) unfinished->push_back(i);
    else *maxl = std::max(*maxl, c->stats_h[i].launches);
  }
  if (c->async_g) plan_learn(c->async_g, c->stats_h, count);
  return TBFS_OK;
}

int run_bfs(tbfs_ctx* c, const tbfs_graph* g, int32_t root, const tbfs_opts& o,
            tbfs_stats* st) {
  return run_bfs_many(c, g, &root, 1, o, st);
}

// Single-source autotuner: coordinate search over the switch thresholds,
// the staging ratio and the non-staging grid, timing `count` BFS per
// candidate (device time, best of two runs).  Every candidate must give the
// same per-source level, vertex and edge counts as the defaults.
int autotune_single(tbfs_ctx* c, const tbfs_graph* g, const int32_t* roots, int count,
                    const tbfs_opts* base, tbfs_opts* best, double* best_ms,
                    double* default_ms) {
  if (!c || !g || !roots || count < 1 || !best) return TBFS_ERR_INVALID;
  if (!g->CP) return TBFS_ERR_UNSUPPORTED;
  tbfs_opts o;
  if (base) o = *base;
  else tbfs_opts_default(&o);
  o.mode = TBFS_MODE_AUTO;
  o.pull = TBFS_PULL_LB;
  if (o.loop != TBFS_LOOP_GRAPH) o.loop = TBFS_LOOP_CHAINED;
  o.grid = 0;
  const int cus = std::max(1, c->cus);
  // the defaults, explicit (so no earlier table entry takes part)
  o.alpha = kDefaultAlpha;
  o.beta = kDefaultBeta;
  o.stage_ratio = kDefaultStageRatio;
  o.lb_grid = -1;  // one grid per level, from each level's work
  std::vector<tbfs_stats> st((size_t)count), ref((size_t)count);
  auto eval = [&](const tbfs_opts& t, double* ms, bool is_ref) -> int {
    double best_t = -1;
    for (int rep = 0; rep < 2; ++rep) {
      const int rc = run_bfs_many(c, g, roots, count, t, st.data());
      if (rc) return rc;
      double sum = 0;
      for (int i = 0; i < count; ++i) {
        sum += st[i].ms;
        const tbfs_stats& a = st[i];
        const tbfs_stats& b = is_ref ? a : ref[i];
        if (a.levels != b.levels || a.reached != b.reached || a.edges != b.edges) {
          std::fprintf(stderr, "turbobfs: autotune: a candidate changed the result of source %d\n",
                       roots[i]);
          return TBFS_ERR_HIP;
        }
      }
      if (is_ref) ref = st;
      if (best_t < 0 || sum < best_t) best_t = sum;
    }
    *ms = best_t / count;
    return TBFS_OK;
  };
  double bt = 0;
  int rc = run_bfs_many(c, g, roots, count, o, st.data());  // warm-up: plan history
  if (!rc) rc = eval(o, &bt, true);
  if (rc) return rc;
  const double dt = bt;
  tbfs_opts bo = o;
  const float alphas[] = {6.f, 10.f, 14.f, 20.f, 30.f};
  const float betas[] = {8.f, 16.f, 24.f, 48.f, 96.f};
  const float ratios[] = {0.25f, 0.5f, 1.f, 2.f, 4.f};
  const int grids[] = {-1, cus, 2 * cus, 4 * cus, default_lb_grid(g)};
  for (int round = 0; round < 2; ++round) {
    for (int dim = 0; dim < 4; ++dim) {
      for (int k = 0; k < 5; ++k) {
        tbfs_opts t = bo;
        if (dim == 0) t.alpha = alphas[k];
        else if (dim == 1) t.beta = betas[k];
        else if (dim == 2) t.stage_ratio = ratios[k];
        else t.lb_grid = std::min(grids[k], 65536);
        if (t.alpha == bo.alpha && t.beta == bo.beta && t.stage_ratio == bo.stage_ratio &&
            t.lb_grid == bo.lb_grid)
          continue;
        double ms = 0;
        rc = eval(t, &ms, false);
        if (rc) return rc;
        if (ms < bt * 0.995) {  // a real gain, not noise
          bt = ms;
          bo = t;
        }
      }
    }
  }
  *best = bo;
  if (best_ms) *best_ms = bt;
  if (default_ms) *default_ms = dt;
  TuneEntry e;
  rc = tune_key(const_cast<tbfs_graph*>(g), 0, c->stream, &e.key);
  if (rc) return rc;
  e.alpha = bo.alpha;
  e.beta = bo.beta;
  e.stage_ratio = bo.stage_ratio;
  e.lb_grid = bo.lb_grid;
  e.ms = bt;
  e.default_ms = dt;
  tune_put(e);
  return TBFS_OK;
}

int tuned_single(const tbfs_graph* g, tbfs_opts* o) {
  if (!g || !o || !g->CP) return TBFS_ERR_INVALID;
  if (tune_empty()) return TBFS_ERR_UNSUPPORTED;
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  TuneKey k;
  const int rc = tune_key(const_cast<tbfs_graph*>(g), 0, s, &k);
  (void)hipStreamDestroy(s);
  if (rc) return rc;
  TuneEntry e;
  if (!tune_find(k, &e)) return TBFS_ERR_UNSUPPORTED;
  o->alpha = e.alpha;
  o->beta = e.beta;
  o->stage_ratio = e.stage_ratio;
  o->lb_grid = e.lb_grid;
  return TBFS_OK;
}

// The last BFS's results in the reference entry points' types, converted on
// the device and copied into host arrays: S_out (unreached as s_unreached)
// (unless null) and, with kind 1 / 2, val_out as float / int32 (see
// k_dropin_convert); *inexact gets the number of counts beyond the type's
// exact range.
int ctx_read_converted(tbfs_ctx* c, int32_t n, int32_t s_unreached, int32_t kind,
                       int32_t* S_out, void* val_out, int64_t* inexact) {
  if (!c || n != c->last_n || n <= 0 || (kind != 0 && !val_out)) return TBFS_ERR_INVALID;
  hipStream_t s = c->stream;
  int32_t* tmp = nullptr;  // converted S, then converted counts, then the counter
  int rc = dev_alloc(&tmp, 2 * (size_t)n + 2);
  if (rc) return rc;
  uint32_t* vals = reinterpret_cast<uint32_t*>(tmp + n);
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(tmp + 2 * (size_t)n);
  unsigned long long bad = 0;
  hipError_t e = hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s);
  if (e == hipSuccess)
    e = kern::dropin_convert(s, c->S, c->sigma, c->ovf, n, s_unreached, kind, tmp, vals, cnt);
  if (e == hipSuccess && S_out) e = download_d2h(S_out, tmp, (size_t)n * sizeof(int32_t), s);
  if (e == hipSuccess && kind != 0) e = download_d2h(val_out, vals, (size_t)n * sizeof(uint32_t), s);
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, cnt, sizeof(bad), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dev_free(tmp);
  if (e != hipSuccess) return hip_fail(e, "ctx_read_converted", __FILE__, __LINE__);
  if (inexact) *inexact = (int64_t)bad;
  return TBFS_OK;
}

int ctx_read(tbfs_ctx* c, int32_t n, int32_t* S, uint64_t* sigma, uint8_t* ovf) {
  if (!c || n != c->last_n || n <= 0) return TBFS_ERR_INVALID;
  hipStream_t s = c->stream;
  if (S) HIPCHK(download_d2h(S, c->S, (size_t)n * sizeof(int32_t), s));
  if (sigma) HIPCHK(download_d2h(sigma, c->sigma, (size_t)n * sizeof(uint64_t), s));
  if (ovf) HIPCHK(download_d2h(ovf, c->ovf, (size_t)n, s));
  HIPCHK(hipStreamSynchronize(s));
  return TBFS_OK;
}

}  // namespace tbfs

// ----------------------------------------------------------------------------
// extended C API
extern "C" {

void tbfs_opts_default(tbfs_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->mode = TBFS_MODE_AUTO;
  o->sigma = 1;
  o->bbreak = 0;
  o->pull = TBFS_PULL_LB;
  o->loop = TBFS_LOOP_CHAINED;
  // alpha, beta, stage_ratio, lb_grid: 0, the tuned values or the defaults
}

const char* tbfs_status_str(int s) {
  switch (s) {
    case TBFS_OK: return "ok";
    case TBFS_ERR_INVALID: return "invalid argument";
    case TBFS_ERR_HIP: return "HIP runtime error";
    case TBFS_ERR_NOMEM: return "out of memory";
    case TBFS_ERR_UNSUPPORTED: return "unsupported";
    case TBFS_SIGMA_INEXACT: return "path count exceeds the output type";
    default: return "unknown status";
  }
}

int tbfs_graph_create(const int32_t* CP, const int32_t* IC, int32_t n,
                      int32_t nz, uint32_t flags, tbfs_graph** out) {
  return graph_create_csc(CP, IC, n, nz, flags, out);
}

int tbfs_graph_create_cooc(const int32_t* I, const int32_t* J, int32_t n,
                           int32_t nz, uint32_t flags, tbfs_graph** out) {
  return graph_create_cooc(I, J, n, nz, flags, out);
}

int tbfs_graph_destroy(tbfs_graph* g) {
  if (!g) return TBFS_ERR_INVALID;
  graph_free(g);
  return TBFS_OK;
}

int tbfs_graph_info(const tbfs_graph* g, int32_t* n, int32_t* nz,
                    uint32_t* flags) {
  if (!g) return TBFS_ERR_INVALID;
  if (n) *n = g->n;
  if (nz) *nz = g->nz;
  if (flags) *flags = g->flags;
  return TBFS_OK;
}

int tbfs_graph_degree_stats(const tbfs_graph* g, int32_t which, tbfs_degree_stats* out) {
  if (!g || !out) return TBFS_ERR_INVALID;
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const int rc = graph_degree_stats(const_cast<tbfs_graph*>(g), which, s, out);
  (void)hipStreamDestroy(s);
  return rc;
}

int tbfs_ctx_create(int32_t n_cap, tbfs_ctx** out) {
  return ctx_create(n_cap, out);
}

int tbfs_ctx_destroy(tbfs_ctx* c) {
  if (!c) return TBFS_ERR_INVALID;
  ctx_free(c);
  return TBFS_OK;
}

void* tbfs_ctx_stream(tbfs_ctx* c) { return c ? (void*)c->stream : nullptr; }

int tbfs_bfs(tbfs_ctx* c, const tbfs_graph* g, int32_t root,
             const tbfs_opts* o, tbfs_stats* st) {
  tbfs_opts def;
  tbfs_opts_default(&def);
  return run_bfs(c, g, root, o ? *o : def, st);
}

int tbfs_bfs_many(tbfs_ctx* c, const tbfs_graph* g, const int32_t* roots,
                  int32_t count, const tbfs_opts* o, tbfs_stats* st) {
  tbfs_opts def;
  tbfs_opts_default(&def);
  return run_bfs_many(c, g, roots, count, o ? *o : def, st);
}

int tbfs_autotune(tbfs_ctx* c, const tbfs_graph* g, const int32_t* roots, int32_t count,
                  const tbfs_opts* base, tbfs_opts* best, double* best_ms, double* default_ms) {
  return autotune_single(c, g, roots, count, base, best, best_ms, default_ms);
}

int tbfs_tuned(const tbfs_graph* g, tbfs_opts* o) { return tuned_single(g, o); }

int tbfs_bfs_multi(tbfs_ctx* c, const tbfs_graph* g, const int32_t* roots,
                   int32_t nroots, const tbfs_opts* o, tbfs_stats* st) {
  tbfs_opts def;
  tbfs_opts_default(&def);
  return run_multi(c, g, roots, nroots, o ? *o : def, st);
}

int tbfs_ctx_read(tbfs_ctx* c, int32_t n, int32_t* S, uint64_t* sigma,
                  uint8_t* ovf) {
  return ctx_read(c, n, S, sigma, ovf);
}

int tbfs_ctx_mark(tbfs_ctx* c, int32_t slot) {
  if (!c || slot < 0 || slot >= 16) return TBFS_ERR_INVALID;
  if (!c->marks[slot]) HIPCHK(hipEventCreate(&c->marks[slot]));
  HIPCHK(hipEventRecord(c->marks[slot], c->stream));
  return TBFS_OK;
}

int tbfs_ctx_elapsed(tbfs_ctx* c, int32_t a, int32_t b, double* ms) {
  if (!c || !ms || a < 0 || a >= 16 || b < 0 || b >= 16) return TBFS_ERR_INVALID;
  if (!c->marks[a] || !c->marks[b]) return TBFS_ERR_INVALID;
  HIPCHK(hipEventSynchronize(c->marks[b]));
  float f = 0.f;
  HIPCHK(hipEventElapsedTime(&f, c->marks[a], c->marks[b]));
  *ms = (double)f;
  return TBFS_OK;
}

int tbfs_calibrate_copy(tbfs_ctx* c, int64_t bytes, int32_t reps, double* ms) {
  if (!c || !ms || bytes < 16 || reps < 1) return TBFS_ERR_INVALID;
  const int64_t n16 = bytes / 16;
  void *src = nullptr, *dst = nullptr;
  HIPCHK(hipMalloc(&src, (size_t)n16 * 16));
  hipError_t e = hipMalloc(&dst, (size_t)n16 * 16);
  if (e != hipSuccess) {
    (void)hipFree(src);
    return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
  }
  int rc = TBFS_OK;
  float f = 0.f;
  e = hipMemsetAsync(src, 1, (size_t)n16 * 16, c->stream);
  if (e == hipSuccess) e = kern::copy16(c->stream, src, dst, n16);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
  for (int i = 0; i < reps && e == hipSuccess; ++i) e = kern::copy16(c->stream, src, dst, n16);
  if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
  if (e == hipSuccess) e = hipEventElapsedTime(&f, c->ev0, c->ev1);
  if (e != hipSuccess) rc = hip_fail(e, "calibrate_copy", __FILE__, __LINE__);
  (void)hipFree(src);
  (void)hipFree(dst);
  *ms = (double)f / reps;
  return rc;
}

int tbfs_device_count(int32_t* count) {
  if (!count) return TBFS_ERR_INVALID;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  *count = n;
  return TBFS_OK;
}

int tbfs_device_info(int32_t dev, char* name, int32_t name_len, int32_t* cus,
                     int64_t* hbm_bytes, int32_t* coop) {
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, dev));
  if (name && name_len > 0) {
    std::strncpy(name, p.gcnArchName[0] ? p.gcnArchName : p.name, (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cus) *cus = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
  if (coop) *coop = p.cooperativeLaunch;
  return TBFS_OK;
}

int tbfs_ctx_device_ptrs(tbfs_ctx* c, int32_t** S, uint64_t** sigma,
                         uint8_t** ovf) {
  if (!c) return TBFS_ERR_INVALID;
  if (S) *S = c->S;
  if (sigma) *sigma = c->sigma;
  if (ovf) *ovf = c->ovf;
  return TBFS_OK;
}

}  // extern "C"
