// tbfs_batch.hip -- many sources in one BFS: frontier words of 64 sources.
//
// A batch of up to B = 64*W sources (W = 1, 2, 4, 8 or 16) runs as one
// level-synchronous BFS whose per-vertex state is W 64-bit words; bit j of
// word w belongs to source 64*w + j.  Layout in HBM (batch object):
//   F[3]    uint64[W][n]  frontier words of depth d (current), d+1 (next) and
//                         d-1 (being cleared), rotating with d
//   seen[2] uint64[W][n]  OR of the frontiers of depths < d (slot d & 1); bits
//                         of unused sources (index >= count) start set
//   ovfw    uint64[W][n]  per-source "path count >= 2^bits(CT)" bits
//   lst[3]  uint32[W*n]   word lists of the frontiers (index w * n + v), made
//                         by push levels; a level whose frontier and previous
//                         frontier are both listed touches only those words
// Counts are uint32 first (CT = uint32_t: half the bytes of every gather);
// a batch in which some count reaches 2^32 runs again with CT = uint64_t,
// and later batches on that graph start with uint64_t.
//   S       int32[W][n][64] per-source depth } valid where the pair's seen
//   sigma   CT[W][n][64]    per-source count } bit is set (never-reached
//                                              pairs are not written; reads
//                                              give -1 / 0 for them)
// Word-major throughout: a word's frontier array is n * 8 bytes (it stays in
// L2 while the waves work on that word), and S / sigma keep the 64 sources of
// a word fastest: the 64 lanes of a wavefront are the 64 sources of one word,
// so a vertex's counts for a word are 256 / 512 contiguous bytes and every
// gather of a parent's counts is coalesced.
//
// Every level is a pull over static slices of the in-lists (kBatchSlice
// entries at most).  A wavefront takes one (word w, slice of vertex v) item:
// it loads v's unseen sources of word w, then, 64 in-list entries at a time
// with one lane per entry u, ANDs u's frontier word with them; the nonzero
// results mark shortest-path edges u -> v into depth d+1 for those sources.
// For each such edge (taken in turn, all loads of a group of 16 in flight)
// the lanes switch to one lane per source and add sigma[u][64w + lane] where
// the edge's mask has the lane's bit.  Items run word-major, so the waves in
// flight share one word and the caches hold that word's frontier array and
// the count segments of its hubs.
// Counts start at zero (the batch's init kernel clears them).  A vertex whose
// in-list is one slice is written by that one wavefront with plain stores;
// longer in-lists are split over slices whose partial sums are added with
// atomics.
//
// Small frontiers expand top-down instead (push while the frontier's
// out-degree sum times alpha stays below nz * W): each frontier word pushes
// its bits along the out-lists with atomic ORs (the lane that sets a bit
// first writes its depth) and adds its counts atomically, in one pass.
//
// The loop is chained: one kernel per level, enqueued back to back; each level reads the counters of its
// frontier (written by the previous kernels), chooses its direction, stops
// when the frontier is empty, and clears the buffers of depth d-1.
//
// Semantics per source: exactly the single-source BFS of the reference's
// sequential oracle (TurboBFS_CSC_TD/bfstdcsc_seq.c:27-96): depth and number
// of shortest paths, with exact 64-bit counts and overflow flags.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "tbfs.h"
#include "tbfs_device.h"
#include "tbfs_engine.h"
#include "tbfs_internal.h"

namespace tbfs {
namespace {

using namespace dev;

constexpr int kBatchSlice = 256;  // in-list entries per pull slice
static_assert(kBatchSlice == kPullChunk, "batch slices come from the shared slice builder");
constexpr int kBatchBlock = 256;  // threads per workgroup (4 wavefronts)
constexpr int kBatchMaxW = 16;    // words per vertex: up to 1024 sources
constexpr float kBatchAlpha = 32.0f;       // default push / pull switch
constexpr float kBatchAlphaShort = 4.0f;   // the same for short in-lists
constexpr int kBatchWideLevels = 16;  // levels on the full-occupancy grid
constexpr int kOpenRing = 8;          // chunk probes of a continuous batch in flight
constexpr int kBatchLbDegree = 12;    // mean in-degree below which pulls share passes  // default push / pull switch

// Level counters, striped over kStripes cache lines like the single-source
// ones (a workgroup adds into stripe blockIdx & (kStripes - 1)).
struct alignas(128) BStripe {
  unsigned long long q;      // (v, w) words that became nonzero
  unsigned long long bits;   // (v, source) pairs discovered
  unsigned long long edges;  // sum of in-degrees over those pairs
  unsigned long long mout;   // sum of out-degrees over the new words (push work)
  unsigned int ovf;          // some new count is >= 2^64
  unsigned int tail;         // stripe 0 only: entries appended to the word list
  unsigned long long pad1[11];
};
struct BCtrSet {
  BStripe s[kStripes];
};

// Chained-loop state, two slots alternating by level parity.
struct alignas(64) BLoop {
  int32_t d;        // depth of the current frontier
  int32_t done;     // the frontier of depth d is empty
  uint32_t ovf_any; // some count so far is >= 2^64 (parents' flags matter)
  int32_t pushing;  // level d - 1 expanded top-down
  int32_t push_levels;
  int32_t pull_levels;
  unsigned long long reached;  // (v, source) pairs reached before depth d
  unsigned long long edges;    // their in-degrees
  int32_t list_cur;   // the word list of frontier d is complete (made by a push level)
  int32_t list_prev;  // the same for frontier d - 1
  uint32_t q_prev;    // words in frontier d - 1
  uint32_t pad0;
  unsigned long long pad[2];
};

struct BatchParams {
  int32_t n;
  int32_t nz;
  int32_t nsrc;
  const int32_t* CP;       // in-lists
  const int32_t* IC;
  const int32_t* RP;       // out-lists (alias CP / IC when symmetric)
  const int32_t* JR;
  const uint64_t* slices;  // pull slices: v | slice << 32
  int32_t nslices;
  const uint64_t* bigo;    // out-lists longer than kBigIn, in kPullChunk slices
  int32_t nbigo;
  float alpha;             // push while m_out * alpha < nz * W
  int32_t pull_lb;         // pull with in-list entries of several items per pass
  uint64_t* F[3];
  uint32_t* lst[3];  // word lists of the frontiers (index w * n + v), rotating like F
  uint64_t* seen[2];
  uint64_t* ovfw;
  void* sigma;    // CT[n][B]: uint32_t or uint64_t counts
  int32_t* S;
  BCtrSet* cset;  // [3]
  BLoop* loop;    // [2]
  RunStats* stats;
  const int32_t* roots;
  const int32_t* perm;  // caller's vertex id -> id in the graph run (null: same)
  int32_t open;         // continuous batch: an empty frontier does not end the loop
};

__device__ __forceinline__ uint64_t rl64(uint64_t x, int l) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, l);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t rfl64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Bits of word w that belong to one of the nsrc sources of the batch.
__device__ __forceinline__ uint64_t word_valid(int nsrc, int w) {
  const int k = nsrc - 64 * w;
  return k >= 64 ? ~0ull : (k <= 0 ? 0ull : ((1ull << k) - 1ull));
}

__device__ __forceinline__ uint32_t atomic_add_ct(uint32_t* a, uint32_t x) {
  return atomicAdd((unsigned int*)a, (unsigned int)x);
}
__device__ __forceinline__ uint64_t atomic_add_ct(uint64_t* a, uint64_t x) {
  return atomicAdd((unsigned long long*)a, (unsigned long long)x);
}

template <int W, typename CT>
__global__ __launch_bounds__(256) void k_batch_init(BatchParams p) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  const int64_t nw = (int64_t)p.n * W;
  for (int64_t i = gtid; i < nw; i += gstride) {
    const uint64_t pad = ~word_valid(p.nsrc, (int)(i / p.n));
    p.F[0][i] = 0ull;
    p.F[1][i] = 0ull;
    p.F[2][i] = 0ull;
    p.seen[0][i] = pad;
    p.seen[1][i] = pad;
    p.ovfw[i] = 0ull;
  }
  // every count starts at zero: push levels and split in-lists add into them
  // atomically (pull levels on single-slice vertices store their sums)
  // (16-byte stores: 64 counts per word make the count array a multiple of 16)
  uint4* sg = (uint4*)p.sigma;
  const int64_t n16 = nw * 64 * (int64_t)sizeof(CT) / 16;
  for (int64_t i = gtid; i < n16; i += gstride) sg[i] = make_uint4(0u, 0u, 0u, 0u);
  const int nctr = (int)(3 * sizeof(BCtrSet) / sizeof(unsigned long long));
  unsigned long long* cw = reinterpret_cast<unsigned long long*>(p.cset);
  for (int64_t i = gtid; i < nctr; i += gstride) cw[i] = 0ull;
  if (gtid == 0) {
    BLoop l0{};
    l0.list_cur = 1;   // the seed kernel lists the depth-0 words
    l0.list_prev = 1;  // (no depth -1 words)
    p.loop[0] = l0;
    p.loop[1] = BLoop{};
    RunStats rs{};
    rs.levels = -1;  // set by the level that finds the frontier empty
    *p.stats = rs;
  }
}

// Depth-0 frontier: one bit per source at its root.
template <int W, typename CT>
__global__ __launch_bounds__(256) void k_batch_seed(BatchParams p) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= p.nsrc) return;
  const int r = p.perm ? p.perm[p.roots[j]] : p.roots[j];
  const uint64_t bit = 1ull << (j & 63);
  const unsigned long long old =
      atomicOr((unsigned long long*)&p.F[0][(int64_t)(j >> 6) * p.n + r], (unsigned long long)bit);
  const int64_t ix = ((int64_t)(j >> 6) * p.n + r) * 64 + (j & 63);
  p.S[ix] = 0;
  ((CT*)p.sigma)[ix] = 1;
  BStripe& c = p.cset[0].s[j & (kStripes - 1)];
  if (old == 0) {
    atomicAdd(&c.q, 1ull);
    atomicAdd(&c.mout, (unsigned long long)(p.RP[r + 1] - p.RP[r]));
    p.lst[0][atomicAdd(&p.cset[0].s[0].tail, 1u)] = (uint32_t)((j >> 6) * p.n + r);
  }
  atomicAdd(&c.bits, 1ull);
  atomicAdd(&c.edges, (unsigned long long)(p.CP[r + 1] - p.CP[r]));
}

// Per-lane counters of one wavefront (summed over the wave when flushed).
struct BAcc {
  unsigned long long q = 0, bits = 0, edges = 0, mout = 0;
  unsigned int ovf = 0;
};

__device__ __forceinline__ void flush_acc(BStripe& c, BAcc a) {
  a.q = wave_sum_u64(a.q);
  a.bits = wave_sum_u64(a.bits);
  a.edges = wave_sum_u64(a.edges);
  a.mout = wave_sum_u64(a.mout);
  a.ovf = __ballot(a.ovf != 0) != 0;
  if (lane() == 0) {
    if (a.q) atomicAdd(&c.q, a.q);
    if (a.bits) atomicAdd(&c.bits, a.bits);
    if (a.edges) atomicAdd(&c.edges, a.edges);
    if (a.mout) atomicAdd(&c.mout, a.mout);
    if (a.ovf) atomicOr(&c.ovf, 1u);
  }
}

// Append word index iw (where flagged) to a word list; one atomic per wave.
__device__ __forceinline__ void list_append(uint32_t* lst, unsigned int* tail, bool flag,
                                            uint32_t iw) {
  const uint64_t m = __ballot(flag);
  if (!m) return;
  const int first = __builtin_ctzll(m);
  unsigned int base = 0;
  if (lane() == first) base = atomicAdd(tail, (unsigned int)__popcll(m));
  base = __shfl(base, first);
  if (flag) lst[base + __popcll(m & lanemask_lt())] = iw;
}

// Write the discoveries of one pull item (vertex v, word w): the lanes are
// the word's sources; `hit` marks the sources found at depth d + 1 with
// count `acc`, `ovb` those whose count overflowed.
template <int W, typename CT>
__device__ __forceinline__ void pull_flush(const BatchParams& p, int32_t d, uint64_t* Fn, int w,
                                           int v, int dg, CT acc, bool hit, bool ovb, BAcc& a) {
  const int ln = lane();
  const uint64_t nb = __ballot(hit);
  if (nb == 0) return;
  uint64_t ob = __ballot(ovb);
  const int64_t iw = (int64_t)w * p.n + v;
  const int64_t ix = iw * 64 + ln;
  if (dg <= kBatchSlice) {  // the only slice of v: plain stores
    if (hit) {
      p.S[ix] = d + 1;
      ((CT*)p.sigma)[ix] = acc;
    }
    if (ln == 0) {
      Fn[iw] = nb;
      if (ob) p.ovfw[iw] |= ob;
      a.q += 1;
      a.bits += (unsigned long long)__popcll(nb);
      a.edges += (unsigned long long)__popcll(nb) * (unsigned long long)dg;
      a.mout += (unsigned long long)(p.RP[v + 1] - p.RP[v]);
    }
  } else {  // one of several slices: partial sums added atomically
    bool cy = false;
    if (hit) {
      const CT old = atomic_add_ct((CT*)p.sigma + ix, acc);
      cy = (CT)(old + acc) < old;
      p.S[ix] = d + 1;
    }
    ob |= __ballot(cy);
    if (ln == 0) {
      if (ob) atomicOr((unsigned long long*)&p.ovfw[iw], (unsigned long long)ob);
      const unsigned long long old = atomicOr((unsigned long long*)&Fn[iw], (unsigned long long)nb);
      const uint64_t newly = nb & ~(uint64_t)old;
      a.bits += (unsigned long long)__popcll(newly);
      a.edges += (unsigned long long)__popcll(newly) * (unsigned long long)dg;
      if (old == 0) {
        a.q += 1;
        a.mout += (unsigned long long)(p.RP[v + 1] - p.RP[v]);
      }
    }
  }
  if (ob && ln == 0) a.ovf = 1;
}

struct PullLds {
  int32_t excl[64];  // exclusive prefix of the live items' slice lengths
  int32_t k0[64];    // first in-list entry of each item's slice
  int32_t v[64];
  int32_t deg[64];   // in-degree of the item's vertex
  uint64_t tgt[64];  // its unseen sources
  int32_t po[64];    // parent edges of one pass: item, parent, source mask
  int32_t pu[64];
  uint64_t pm[64];
};

// Pull over (word, slice) items, grouped by word: the waves working on a word
// share its frontier array and the count segments of its hubs in the caches.
// A wavefront takes 64 consecutive items of one word; the in-list entries of
// the live ones (some source of the word still unseen at v) are spread over
// the lanes by a prefix sum of their lengths, 64 entries per pass, so short
// in-lists (grids, road networks) share a pass.  Entry u of item v: u's
// frontier word AND v's unseen sources marks the sources for which u -> v is
// a shortest-path edge into depth d + 1; those parent edges, kept in item
// order, are then taken 16 at a time with the lanes as sources: all 16
// count gathers are issued before any is added, and an item's discoveries
// are written when the next item starts.
template <int W, typename CT>
__device__ __forceinline__ void batch_pull(const BatchParams& p, int32_t d, const uint64_t* Fc,
                                           uint64_t* Fn, const uint64_t* sd, bool pov,
                                           int64_t gtid, int64_t gstride, PullLds& L, BAcc& a) {
  const int ln = lane();
  const int64_t nwaves = gstride >> 6;
  // chunks of 64 items of one word.  With the grid a multiple of 8 the
  // chunks follow the XCDs (workgroup b runs on XCD b % 8): XCD x takes the
  // words w = x (mod 8) one after another, or shares word x % W with the
  // other XCDs of that residue when W < 8, so each XCD's L2 holds the
  // frontier array and hot count segments of its own word only.
  const int64_t cpw = ((int64_t)p.nslices + 63) >> 6;
  const bool xm = (gridDim.x & 7) == 0;
  const int xcd = blockIdx.x & 7;
  const int64_t wpb = blockDim.x >> 6;
  const int nsh = W >= 8 ? 1 : 8 / W;
  const int64_t nloc = !xm ? (int64_t)W * cpw
                           : (W >= 8 ? (int64_t)(W / 8) * cpw
                                     : (cpw + nsh - 1 - xcd / W) / nsh);
  const int64_t l0 = xm ? (int64_t)(blockIdx.x >> 3) * wpb + (threadIdx.x >> 6) : (gtid >> 6);
  const int64_t ls = xm ? (int64_t)(gridDim.x >> 3) * wpb : nwaves;
  constexpr int G = 8;  // fewer in flight than the one-item pull: registers
  for (int64_t li = l0; li < nloc; li += ls) {
    int w;
    int64_t c;
    if (!xm) {
      w = (int)(li / cpw);
      c = li - (int64_t)w * cpw;
    } else if (W >= 8) {
      const int64_t q = li / cpw;
      w = xcd + 8 * (int)q;
      c = li - q * cpw;
    } else {
      w = xcd % W;
      c = li * nsh + xcd / W;
    }
    const int64_t jj = c * 64 + ln;  // slice index within word w
    int v = 0, k0 = 0, len = 0, deg = 0;
    uint64_t tgt = 0;
    if (jj < p.nslices) {
      const uint64_t e = p.slices[jj];
      v = (int)(uint32_t)e;
      const int beg0 = p.CP[v], end0 = p.CP[v + 1];
      k0 = beg0 + (int)(e >> 32) * kBatchSlice;
      deg = end0 - beg0;
      const int64_t iw = (int64_t)w * p.n + v;
      tgt = ~(sd[iw] | Fc[iw]);
      if (tgt) len = min(end0, k0 + kBatchSlice) - k0;
    }
    const uint32_t incl = wave_incl_scan((uint32_t)len);
    const int total = (int)__shfl(incl, 63);
    if (total == 0) continue;
    wave_lds_sync();
    L.excl[ln] = (int)incl - len;
    L.k0[ln] = k0;
    L.v[ln] = v;
    L.deg[ln] = deg;
    L.tgt[ln] = tgt;
    wave_lds_sync();
    const CT* sgw = (const CT*)p.sigma + (int64_t)w * p.n * 64 + ln;
    const uint64_t* Fcw = Fc + (int64_t)w * p.n;
    const uint64_t* ovw = p.ovfw + (int64_t)w * p.n;
    int cur = -1;  // item being accumulated (uniform)
    CT acc = 0;
    bool hit = false, ovb = false;
    for (int e0 = 0; e0 < total; e0 += 64) {
      const int e = e0 + ln;
      int o = 0, u = 0;
      uint64_t mm = 0;
      if (e < total) {
        o = owner_of(L.excl, e);
        u = p.IC[L.k0[o] + (e - L.excl[o])];
        mm = Fcw[u] & L.tgt[o];
      }
      const uint64_t bal = __ballot(mm != 0);
      if (bal == 0) continue;
      const int cnt = __popcll(bal);
      wave_lds_sync();
      if (mm) {
        const int pos = __popcll(bal & lanemask_lt());
        L.po[pos] = o;
        L.pu[pos] = u;
        L.pm[pos] = mm;
      }
      wave_lds_sync();
      for (int g0 = 0; g0 < cnt; g0 += G) {
        CT x[G];
        int og[G];
        uint32_t hm = 0, pm = 0;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          x[g] = 0;
          og[g] = -1;
          if (g0 + g < cnt) {
            og[g] = __builtin_amdgcn_readfirstlane(L.po[g0 + g]);
            const int ue = __builtin_amdgcn_readfirstlane(L.pu[g0 + g]);
            const uint64_t me = rfl64(L.pm[g0 + g]);
            // the edge's mask is the lane mask of its load (no per-lane bit test)
            if (__builtin_amdgcn_inverse_ballot_w64(me)) {
              x[g] = sgw[(int64_t)ue * 64];
              hm |= 1u << g;
              if (pov && ((ovw[ue] >> ln) & 1ull)) pm |= 1u << g;
            }
          }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (og[g] < 0) break;
          if (og[g] != cur) {
            if (cur >= 0)
              pull_flush<W, CT>(p, d, Fn, w, __builtin_amdgcn_readfirstlane(L.v[cur]),
                                __builtin_amdgcn_readfirstlane(L.deg[cur]), acc, hit, ovb, a);
            cur = og[g];
            acc = 0;
            hit = false;
            ovb = false;
          }
          const CT y = acc + x[g];
          if (y < acc) ovb = true;
          acc = y;
          hit |= (hm >> g) & 1u;
          ovb |= (pm >> g) & 1u;
        }
      }
    }
    if (cur >= 0)
      pull_flush<W, CT>(p, d, Fn, w, __builtin_amdgcn_readfirstlane(L.v[cur]),
                        __builtin_amdgcn_readfirstlane(L.deg[cur]), acc, hit, ovb, a);
  }
}

struct PushLds {
  int32_t excl[64];
  int32_t beg[64];
  int32_t u[64];
  int32_t w[64];
  uint64_t f[64];
};

// Pull, one item at a time (graphs with long in-lists): a wavefront tests 64
// consecutive items of one word at once (one per lane), then expands the
// live ones in turn, each with its own pass over its in-list slice: one lane
// per entry for the parent test, then one lane per source for the counts,
// 16 gathers in flight.  Chunks are scheduled as in batch_pull.
template <int W, typename CT>
__device__ __forceinline__ void batch_pull_items(const BatchParams& p, int32_t d, const uint64_t* Fc,
                                           uint64_t* Fn, const uint64_t* sd, bool pov,
                                           int64_t gtid, int64_t gstride, BAcc& a) {
  const int ln = lane();
  const int64_t nwaves = gstride >> 6;
  // chunks of 64 items of one word.  With the grid a multiple of 8 the
  // chunks follow the XCDs (workgroup b runs on XCD b % 8): XCD x takes the
  // words w = x (mod 8) one after another, or shares word x % W with the
  // other XCDs of that residue when W < 8, so each XCD's L2 holds the
  // frontier array and hot count segments of its own word only.
  const int64_t cpw = ((int64_t)p.nslices + 63) >> 6;
  const bool xm = (gridDim.x & 7) == 0;
  const int xcd = blockIdx.x & 7;
  const int64_t wpb = blockDim.x >> 6;
  const int nsh = W >= 8 ? 1 : 8 / W;
  const int64_t nloc = !xm ? (int64_t)W * cpw
                           : (W >= 8 ? (int64_t)(W / 8) * cpw
                                     : (cpw + nsh - 1 - xcd / W) / nsh);
  const int64_t l0 = xm ? (int64_t)(blockIdx.x >> 3) * wpb + (threadIdx.x >> 6) : (gtid >> 6);
  const int64_t ls = xm ? (int64_t)(gridDim.x >> 3) * wpb : nwaves;
  for (int64_t li = l0; li < nloc; li += ls) {
    int w;
    int64_t c;
    if (!xm) {
      w = (int)(li / cpw);
      c = li - (int64_t)w * cpw;
    } else if (W >= 8) {
      const int64_t q = li / cpw;
      w = xcd + 8 * (int)q;
      c = li - q * cpw;
    } else {
      w = xcd % W;
      c = li * nsh + xcd / W;
    }
    const int64_t jj = c * 64 + ln;  // slice index within word w
    int v = 0, k0 = 0, k1 = 0, deg = 0;
    uint64_t tgt = 0;
    if (jj < p.nslices) {
      const uint64_t e = p.slices[jj];
      v = (int)(uint32_t)e;
      const int beg0 = p.CP[v], end0 = p.CP[v + 1];
      k0 = beg0 + (int)(e >> 32) * kBatchSlice;
      k1 = min(end0, k0 + kBatchSlice);
      deg = end0 - beg0;
      const int64_t iw = (int64_t)w * p.n + v;
      tgt = ~(sd[iw] | Fc[iw]);
    }
    uint64_t live = __ballot(tgt != 0);
    while (live) {
      const int src = __builtin_ctzll(live);
      live &= live - 1;
      const int wi = w;
      const int vi = __builtin_amdgcn_readlane(v, src);
      const int ka = __builtin_amdgcn_readlane(k0, src);
      const int kb = __builtin_amdgcn_readlane(k1, src);
      const int dg = __builtin_amdgcn_readlane(deg, src);
      const uint64_t un = rl64(tgt, src);
      const CT* sgw = (const CT*)p.sigma + (int64_t)wi * p.n * 64;  // uniform
      const uint64_t* Fcw = Fc + (int64_t)wi * p.n;
      const uint64_t* ovw = p.ovfw + (int64_t)wi * p.n;
      // 32-bit counts are summed in 64 bits (at most kBatchSlice terms below
      // 2^32: no wrap) and checked once; 64-bit counts check each carry
      uint64_t acc = 0;
      uint64_t hm = 0, om = 0;  // uniform: sources with a parent / an overflowed parent
      bool ovb = false;
      for (int c0 = ka; c0 < kb; c0 += 64) {
        const int k = c0 + ln;
        int u = 0;
        uint64_t mm = 0;
        if (k < kb) {
          u = p.IC[k];
          mm = Fcw[u] & un;  // sources for which u -> v is a parent edge
        }
        uint64_t bal = __ballot(mm != 0);
        // parent edges in groups of G, their vertex and mask read into
        // scalar registers from the lane that found them: all count loads of
        // a group are issued before any is added.  An edge's mask is the
        // lane mask of its load and of its add (inverse ballot: no per-lane
        // bit test), so sparse masks touch one line of the segment
        constexpr int G = 16;
        while (bal) {
          CT x[G];
          uint64_t me[G];
#pragma unroll
          for (int g = 0; g < G; ++g) {
            me[g] = 0;
            if (bal) {  // wave-uniform
              const int pe = __builtin_ctzll(bal);
              bal &= bal - 1;
              const int ue = __builtin_amdgcn_readlane(u, pe);
              me[g] = rl64(mm, pe);
              if (__builtin_amdgcn_inverse_ballot_w64(me[g])) x[g] = sgw[(int64_t)ue * 64 + ln];
              if (pov) om |= me[g] & ovw[ue];
            }
          }
#pragma unroll
          for (int g = 0; g < G; ++g) {
            if (__builtin_amdgcn_inverse_ballot_w64(me[g])) {
              if constexpr (sizeof(CT) == 8) {
                const uint64_t y = acc + (uint64_t)x[g];
                if (y < acc) ovb = true;
                acc = y;
              } else {
                acc += (uint64_t)x[g];
              }
            }
            hm |= me[g];
          }
        }
      }
      const bool hit = (hm >> ln) & 1ull;
      if ((om >> ln) & 1ull) ovb = true;
      if constexpr (sizeof(CT) == 4) {
        if (acc >> 32) ovb = true;
      }
      pull_flush<W, CT>(p, d, Fn, wi, vi, dg, (CT)acc, hit, ovb, a);
    }
  }
}

// Top-down edge u -> v for word w, f = frontier word of u: the sources of f
// unvisited at v (shortest-path edges u -> v into depth d + 1) join the next
// frontier (atomic OR; the lane that sets a bit first writes its depth) and
// get u's count added atomically (counts start at zero), with carry and
// inherited overflow flags.  Returns true when the OR made the word nonzero
// (the word joins the next frontier's word list).
template <int W, typename CT>
__device__ __forceinline__ bool push_edge(const BatchParams& p, const uint64_t* Fc, uint64_t* Fn,
                                          const uint64_t* sd, int32_t d, int u, int v, int w,
                                          uint64_t f, bool pov, BAcc& a) {
  const int64_t iv = (int64_t)w * p.n + v;
  uint64_t m = f & ~(sd[iv] | Fc[iv]);
  if (!m) return false;
  const int64_t iu = (int64_t)w * p.n + u;
  const unsigned long long old = atomicOr((unsigned long long*)&Fn[iv], (unsigned long long)m);
  const uint64_t po = pov ? p.ovfw[iu] : 0ull;
  uint64_t newly = m & ~(uint64_t)old;
  uint64_t ob = 0;
  while (m) {
    const int sb = __builtin_ctzll(m);
    m &= m - 1;
    const CT x = ((const CT*)p.sigma)[iu * 64 + sb];
    const CT prev = atomic_add_ct((CT*)p.sigma + iv * 64 + sb, x);
    if ((CT)(prev + x) < prev || ((po >> sb) & 1ull)) ob |= 1ull << sb;
  }
  if (ob) {
    atomicOr((unsigned long long*)&p.ovfw[iv], (unsigned long long)ob);
    a.ovf = 1;
  }
  if (!newly) return false;
  const unsigned long long c = (unsigned long long)__popcll(newly);
  a.bits += c;
  a.edges += c * (unsigned long long)(p.CP[v + 1] - p.CP[v]);
  if (old == 0) {
    a.q += 1;
    a.mout += (unsigned long long)(p.RP[v + 1] - p.RP[v]);
  }
  while (newly) {
    const int sb = __builtin_ctzll(newly);
    newly &= newly - 1;
    p.S[iv * 64 + sb] = d + 1;
  }
  return old == 0;
}

// Top-down pass over the frontier words (u, w): u's out-list edges.  The
// words come from the frontier's word list when it is complete (`nlist` >= 0
// entries; sparse frontiers cost in proportion to their size), else from a
// scan of every word.  Lists up to kBigIn entries are split over the lanes of
// a wavefront by edge count (prefix scan + owner search); longer lists go by
// their static kPullChunk slices, one (slice, word) item per wavefront.  The
// words made nonzero are listed for the next level (`nlst`, `ntail`).
template <int W, typename CT>
__device__ __forceinline__ void batch_push(const BatchParams& p, const uint64_t* Fc,
                                           uint64_t* Fn, const uint64_t* sd, int32_t d, bool pov,
                                           const uint32_t* lst, int64_t nlist, uint32_t* nlst,
                                           unsigned int* ntail, int64_t gtid, int64_t gstride,
                                           PushLds& L, BAcc& a) {
  const int ln = lane();
  const int64_t nwaves = gstride >> 6;
  const int64_t nW = (int64_t)p.n * W;
  const int64_t span = nlist >= 0 ? nlist : nW;
  for (int64_t base = (gtid >> 6) * 64; base < span; base += nwaves * 64) {
    const int64_t k = base + ln;
    int64_t i = 0;
    uint64_t f = 0;
    int od = 0, rb = 0;
    if (k < span) {
      i = nlist >= 0 ? (int64_t)lst[k] : k;
      f = Fc[i];
      if (f) {
        const int u = (int)(i % p.n);
        rb = p.RP[u];
        od = p.RP[u + 1] - rb;
        if (od > kBigIn) od = 0;  // served by the static slices below
      }
    }
    const uint32_t incl = wave_incl_scan((uint32_t)od);
    const int total = (int)__shfl(incl, 63);
    if (total == 0) continue;
    wave_lds_sync();
    L.excl[ln] = (int)incl - od;
    L.beg[ln] = rb;
    L.u[ln] = (int)(i % p.n);
    L.w[ln] = (int)(i / p.n);
    L.f[ln] = f;
    wave_lds_sync();
    for (int e0 = 0; e0 < total; e0 += 64) {
      const int e = e0 + ln;
      bool nw = false;
      int v = 0, o = 0;
      if (e < total) {
        o = owner_of(L.excl, e);
        v = p.JR[L.beg[o] + (e - L.excl[o])];
        nw = push_edge<W, CT>(p, Fc, Fn, sd, d, L.u[o], v, L.w[o], L.f[o], pov, a);
      }
      list_append(nlst, ntail, nw, (uint32_t)((int64_t)L.w[o] * p.n + v));
    }
  }
  const int64_t hubs = (int64_t)p.nbigo * W;
  for (int64_t t = gtid >> 6; t < hubs; t += nwaves) {
    const int w = (int)(t / p.nbigo);
    const uint64_t e = p.bigo[t - (int64_t)w * p.nbigo];
    const int u = (int)(uint32_t)e;
    const uint64_t f = Fc[(int64_t)w * p.n + u];
    if (!f) continue;
    const int kb = p.RP[u] + (int)(e >> 32) * kPullChunk;
    const int ke = min(p.RP[u + 1], kb + kPullChunk);
    for (int k0 = kb; k0 < ke; k0 += 64) {
      const int k = k0 + ln;
      bool nw = false;
      int v = 0;
      if (k < ke) {
        v = p.JR[k];
        nw = push_edge<W, CT>(p, Fc, Fn, sd, d, u, v, w, f, pov, a);
      }
      list_append(nlst, ntail, nw, (uint32_t)((int64_t)w * p.n + v));
    }
  }
}

// Level d: expand the frontier of depth d into depth d + 1.  LB selects the
// pull that shares passes between items (short in-lists).
template <int W, typename CT, bool LB>
__global__ __launch_bounds__(kBatchBlock) void k_batch_level(BatchParams p, int32_t d) {
  __shared__ union {
    PushLds push;
    typename std::conditional<LB, PullLds, PushLds>::type pull;
  } lds[kBatchBlock / 64];
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;

  BLoop st = p.loop[d & 1];
  if (st.done) {
    if (gtid == 0) p.loop[(d + 1) & 1] = st;
    return;
  }
  unsigned long long q = 0, bits = 0, edges = 0, mout = 0;
  unsigned int ovf = 0;
  {
    const BCtrSet& cs = p.cset[d % 3];
#pragma unroll
    for (int k = 0; k < kStripes; ++k) {
      q += cs.s[k].q;
      bits += cs.s[k].bits;
      edges += cs.s[k].edges;
      mout += cs.s[k].mout;
      ovf |= cs.s[k].ovf;
    }
  }
  st.reached += bits;
  st.edges += edges;
  st.ovf_any |= ovf;
  if (q == 0 && !p.open) {
    st.done = 1;
    if (gtid == 0) {
      p.loop[(d + 1) & 1] = st;
      RunStats rs{};
      rs.levels = d;
      rs.push_levels = st.push_levels;
      rs.pull_levels = st.pull_levels;
      rs.ovf_any = st.ovf_any;
      rs.reached = st.reached;
      rs.edges = st.edges;
      *p.stats = rs;
    }
    return;
  }
  const bool push = (double)mout * (double)p.alpha < (double)p.nz * (double)W;
  if (gtid == 0) {
    BLoop nx = st;
    nx.d = d + 1;
    nx.pushing = push ? 1 : 0;
    nx.push_levels += push ? 1 : 0;
    nx.pull_levels += push ? 0 : 1;
    nx.list_cur = push ? 1 : 0;  // a push level lists the words it makes nonzero
    nx.list_prev = st.list_cur;
    nx.q_prev = (uint32_t)q;
    p.loop[(d + 1) & 1] = nx;
  }

  const uint64_t* Fc = p.F[d % 3];
  uint64_t* Fn = p.F[(d + 1) % 3];
  uint64_t* Fp = p.F[(d + 2) % 3];
  const uint64_t* sd = p.seen[d & 1];
  uint64_t* sn = p.seen[(d + 1) & 1];

  // housekeeping for level d + 1: seen through depth d, clear depth d - 1
  // (slot (d+1)&1 holds the OR through depth d-2: unchanged where neither
  // frontier d-1 nor frontier d has a bit); from the word lists when both are
  // complete, else over every word
  if (st.list_cur && st.list_prev) {
    // from the grid's last threads: the push below deals the same lists
    // from its first, so on small frontiers the two chains of dependent
    // requests run side by side on different CUs (they touch disjoint
    // words: seen slot d + 1 and frontier d - 1 here, seen slot d and
    // frontiers d, d + 1 there)
    const uint32_t* lp = p.lst[(d + 2) % 3];
    const uint32_t* lc = p.lst[d % 3];
    const int64_t rt = gstride - 1 - gtid;
    for (int64_t k = rt; k < (int64_t)st.q_prev; k += gstride) {
      const uint32_t i = lp[k];
      sn[i] = sd[i] | Fc[i];
      Fp[i] = 0ull;
    }
    for (int64_t k = rt; k < (int64_t)q; k += gstride) {
      const uint32_t i = lc[k];
      sn[i] = sd[i] | Fc[i];
    }
  } else {
    const int64_t nw = (int64_t)p.n * W;
    for (int64_t i = gtid; i < nw; i += gstride) {
      const uint64_t fc = Fc[i], fp = Fp[i];
      if (fc | fp) {
        sn[i] = sd[i] | fc;
        if (fp) Fp[i] = 0ull;
      }
    }
  }
  if (gtid < kStripes) {
    BStripe& c = p.cset[(d + 2) % 3].s[gtid];
    c.q = 0;
    c.bits = 0;
    c.edges = 0;
    c.mout = 0;
    c.ovf = 0;
    c.tail = 0;
  }

  BAcc a;
  if (push)
    batch_push<W, CT>(p, Fc, Fn, sd, d, st.ovf_any != 0, p.lst[d % 3],
                             st.list_cur ? (int64_t)q : -1, p.lst[(d + 1) % 3],
                             &p.cset[(d + 1) % 3].s[0].tail, gtid, gstride,
                             lds[threadIdx.x >> 6].push, a);
  else if constexpr (LB)
    batch_pull<W, CT>(p, d, Fc, Fn, sd, st.ovf_any != 0, gtid, gstride,
                      lds[threadIdx.x >> 6].pull, a);
  else
    batch_pull_items<W, CT>(p, d, Fc, Fn, sd, st.ovf_any != 0, gtid, gstride, a);
  flush_acc(p.cset[(d + 1) % 3].s[blockIdx.x & (kStripes - 1)], a);
}

// Per-source rows out of the [n][B] layout: sources first .. first+count-1
// (count <= 64) into So[count][n], sgo[count][n], ovo[count][n].  Pairs whose
// bit is clear in `seen` (never reached) read S = -1, sigma = 0: the level
// kernels never write them.  perm (may be null): the caller's vertex v is
// vertex perm[v] of the graph the batch ran on.
template <typename CT>
__global__ __launch_bounds__(256) void k_batch_extract(
    const int32_t* __restrict__ S, const CT* __restrict__ sigma,
    const uint64_t* __restrict__ ovfw, const uint64_t* __restrict__ seen, int32_t n,
    int32_t first, int32_t count, int32_t* So, uint64_t* sgo,
    uint8_t* ovo, const int32_t* __restrict__ perm) {
  __shared__ int32_t ts[64][65];
  __shared__ uint64_t tg[64][65];
  __shared__ uint8_t to[64][65];
  const int64_t v0 = (int64_t)blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4) {  // r: vertex row of the tile, tx: source
    const int64_t v = v0 + r;
    if (v < n && tx < count) {
      const int j = first + tx;
      const int64_t iw = (int64_t)(j >> 6) * n + (perm ? perm[v] : v);
      const bool reached = (seen[iw] >> (j & 63)) & 1ull;
      ts[r][tx] = reached ? S[iw * 64 + (j & 63)] : -1;
      tg[r][tx] = reached ? (uint64_t)sigma[iw * 64 + (j & 63)] : 0ull;
      to[r][tx] = reached ? (uint8_t)((ovfw[iw] >> (j & 63)) & 1ull) : (uint8_t)0;
    }
  }
  __syncthreads();
  for (int r = ty; r < count; r += 4) {  // r: source, tx: vertex
    const int64_t v = v0 + tx;
    if (v < n) {
      So[(int64_t)r * n + v] = ts[tx][r];
      sgo[(int64_t)r * n + v] = tg[tx][r];
      ovo[(int64_t)r * n + v] = to[tx][r];
    }
  }
}

// ---- continuous batches ------------------------------------------------------
// A continuous batch keeps running levels while sources come and go: a new
// source takes a free bit j at a level boundary d0 (its root enters the
// frontier of depth d0 with depth d0 and count 1), so the level kernels write
// d0 + its depth into S; a source whose bit is absent from a whole frontier
// is finished, its row is read out with d0 subtracted and its column cleared
// (seen bits, overflow bits, counts) so that the bit can take the next
// source.  Counts are 64-bit throughout.  Up to 64 sources per launch.
struct OpenSeeds {
  int32_t bit[64];
  int32_t root[64];
};
struct OpenRows {
  int32_t bit[64];
  int32_t off[64];      // level at which the source was seeded
  uint64_t wmask[16];   // per word: the bits of this launch
};
// What the host reads at a chunk boundary d: the sources present in the
// frontier of depth d (per word) and that frontier's word count.
struct alignas(256) OpenProbe {
  uint64_t alive[16];
  unsigned long long q;
  int32_t d;
  int32_t pad;
};

// Seed count sources at boundary d (after level d - 1, before level d).
__global__ __launch_bounds__(64) void k_batch_open_seed(BatchParams p, int32_t d, OpenSeeds sd,
                                                        int32_t count) {
  const int t = threadIdx.x;
  if (t >= count) return;
  const int j = sd.bit[t];
  const int r = p.perm ? p.perm[sd.root[t]] : sd.root[t];
  const int64_t iw = (int64_t)(j >> 6) * p.n + r;
  const unsigned long long old =
      atomicOr((unsigned long long*)&p.F[d % 3][iw], (unsigned long long)(1ull << (j & 63)));
  p.S[iw * 64 + (j & 63)] = d;
  ((uint64_t*)p.sigma)[iw * 64 + (j & 63)] = 1ull;
  BStripe& c = p.cset[d % 3].s[t & (kStripes - 1)];
  if (old == 0) {
    atomicAdd(&c.q, 1ull);
    atomicAdd(&c.mout, (unsigned long long)(p.RP[r + 1] - p.RP[r]));
    p.lst[d % 3][atomicAdd(&p.cset[d % 3].s[0].tail, 1u)] = (uint32_t)iw;
  }
  atomicAdd(&c.bits, 1ull);
  atomicAdd(&c.edges, (unsigned long long)(p.CP[r + 1] - p.CP[r]));
}

// The probe of boundary d into out (cleared beforehand).  The frontier's
// words come from its word list when that is complete, else from a scan of
// every word.
template <int W>
__global__ __launch_bounds__(256) void k_batch_open_probe(BatchParams p, int32_t d,
                                                          OpenProbe* out) {
  __shared__ unsigned long long acc[W];
  if (threadIdx.x < W) acc[threadIdx.x] = 0ull;
  __syncthreads();
  const BLoop st = p.loop[d & 1];
  unsigned long long q = 0;
#pragma unroll
  for (int k = 0; k < kStripes; ++k) q += p.cset[d % 3].s[k].q;
  const uint64_t* F = p.F[d % 3];
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  const bool listed = st.list_cur != 0;
  const int64_t span = listed ? (int64_t)q : (int64_t)p.n * W;
  const uint32_t* lst = p.lst[d % 3];
  for (int64_t k = gtid; k < span; k += gstride) {
    const int64_t i = listed ? (int64_t)lst[k] : k;
    const uint64_t f = F[i];
    if (f) atomicOr(&acc[i / p.n], (unsigned long long)f);
  }
  __syncthreads();
  if (threadIdx.x < W && acc[threadIdx.x])
    atomicOr((unsigned long long*)&out->alive[threadIdx.x], acc[threadIdx.x]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    out->q = q;
    out->d = d;
  }
}

// The finished probe `in` to host memory (`host`: pinned, coherent, mapped;
// the host reads it once an event after this kernel has passed, with no
// copy engine in between), and the next probe slot cleared.
__global__ __launch_bounds__(64) void k_batch_open_publish(const OpenProbe* in, OpenProbe* host,
                                                           OpenProbe* next) {
  const int t = threadIdx.x;
  if (t < 16) {
    host->alive[t] = in->alive[t];
    next->alive[t] = 0ull;
  }
  if (t == 0) {
    host->q = in->q;
    host->d = in->d;
    host->pad = 0;
    next->q = 0ull;
  }
}

// Rows of finished sources (bits rows.bit[0..count)) out of the [n][64W]
// layout, depth minus the seeding level, then their columns cleared: seen
// bits in both slots, overflow bits, counts.  `seen` holds the OR of every
// frontier so far.  A 64-vertex tile per workgroup; perm as k_batch_extract.
__global__ __launch_bounds__(256) void k_batch_open_extract(
    int32_t* __restrict__ S, uint64_t* __restrict__ sigma, uint64_t* __restrict__ ovfw,
    uint64_t* __restrict__ seen, uint64_t* __restrict__ seen2, int32_t n, int32_t W, OpenRows rows,
    int32_t count, int32_t* So, uint64_t* sgo, uint8_t* ovo, const int32_t* __restrict__ perm) {
  __shared__ int32_t ts[64][65];
  __shared__ uint64_t tg[64][65];
  __shared__ uint8_t to[64][65];
  const int64_t v0 = (int64_t)blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4) {  // r: vertex row of the tile, tx: source
    const int64_t v = v0 + r;
    if (v < n && tx < count) {
      const int j = rows.bit[tx];
      const int64_t iw = (int64_t)(j >> 6) * n + (perm ? perm[v] : v);
      const int64_t ix = iw * 64 + (j & 63);
      const bool reached = (seen[iw] >> (j & 63)) & 1ull;
      ts[r][tx] = reached ? S[ix] - rows.off[tx] : -1;
      tg[r][tx] = reached ? sigma[ix] : 0ull;
      to[r][tx] = reached ? (uint8_t)((ovfw[iw] >> (j & 63)) & 1ull) : (uint8_t)0;
      if (reached) sigma[ix] = 0ull;
    }
  }
  __syncthreads();
  for (int r = ty; r < count; r += 4) {  // r: source, tx: vertex
    const int64_t v = v0 + tx;
    if (v < n) {
      So[(int64_t)r * n + v] = ts[tx][r];
      sgo[(int64_t)r * n + v] = tg[tx][r];
      ovo[(int64_t)r * n + v] = to[tx][r];
    }
  }
  // the columns' bits (every read of this tile is done)
  for (int k = threadIdx.x; k < 64 * W; k += blockDim.x) {
    const int w = k >> 6;
    const int64_t v = v0 + (k & 63);
    const uint64_t m = rows.wmask[w];
    if (m && v < n) {
      const int64_t iw = (int64_t)w * n + (perm ? perm[v] : v);
      seen[iw] &= ~m;
      seen2[iw] &= ~m;
      ovfw[iw] &= ~m;
    }
  }
}

}  // namespace
}  // namespace tbfs

// ----------------------------------------------------------------------------
// host side
using namespace tbfs;

struct tbfs_batch {
  int64_t cap_words = 0;  // n * W of the largest batch shape the buffers hold
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr;    // start of the current batch
  hipEvent_t ev_step = nullptr;  // end of the work enqueued so far
  uint64_t* F[3] = {nullptr, nullptr, nullptr};
  uint32_t* lst[3] = {nullptr, nullptr, nullptr};
  uint64_t* seen[2] = {nullptr, nullptr};
  uint64_t* ovfw = nullptr;
  uint64_t* sigma = nullptr;  // viewed as uint32_t counts by 32-bit runs
  int32_t* S = nullptr;
  BCtrSet* cset = nullptr;
  BLoop* loop = nullptr;
  RunStats* stats = nullptr;
  RunStats* stats_h = nullptr;
  int32_t* roots = nullptr;   // kBatchMaxW * 64 entries
  int32_t* xS = nullptr;      // extraction staging: xrows rows of xn entries
  uint64_t* xsig = nullptr;
  uint8_t* xovf = nullptr;
  int64_t xcap = 0;           // entries of the staging buffers
  int32_t* xperm = nullptr;   // the last batch's id map (a copy: the graph may go)
  int64_t xperm_cap = 0;
  bool use_perm = false;
  int cus = 0;
  int chain_k = 8;
  // the batch in flight (or the last one)
  const tbfs_graph* g = nullptr;  // used only while the batch runs
  int32_t n = 0;                   // vertices of its graph (rows stay readable)
  tbfs_opts o{};
  int32_t count = 0;
  int W = 0;
  int ct = 4;        // bytes per count of the current run
  int grid = 0;
  int d_next = 0;    // first level not yet enqueued
  bool active = false;
  bool done = false;
  RunStats rs{};
  float ms = 0.f;
  const uint64_t* seen_final = nullptr;  // reached bits of the finished batch
  // continuous batch: probes of the chunk boundaries (device, pinned host
  // copies, their events), chunks enqueued and chunks whose probe was read
  bool open = false;
  OpenProbe* probe_d = nullptr;
  OpenProbe* probe_h = nullptr;   // pinned, coherent host memory
  OpenProbe* probe_hd = nullptr;  // its device address
  hipEvent_t probe_ev[kOpenRing] = {};
  int64_t chunk_head = 0;
  int64_t chunk_tail = 0;
};

namespace tbfs {

int batch_words_for(int32_t count) {
  int w = 1;
  while (64 * w < count) w <<= 1;
  return w;
}

void batch_free(tbfs_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (int i = 0; i < 3; ++i) dev_free(b->F[i]);
  for (int i = 0; i < 3; ++i) dev_free(b->lst[i]);
  for (int i = 0; i < 2; ++i) dev_free(b->seen[i]);
  dev_free(b->ovfw);
  dev_free(b->sigma);
  dev_free(b->S);
  dev_free(b->cset);
  dev_free(b->loop);
  dev_free(b->stats);
  dev_free(b->roots);
  dev_free(b->xS);
  dev_free(b->xsig);
  dev_free(b->xovf);
  dev_free(b->xperm);
  dev_free(b->probe_d);
  if (b->probe_h) (void)hipHostFree(b->probe_h);
  for (int i = 0; i < kOpenRing; ++i)
    if (b->probe_ev[i]) (void)hipEventDestroy(b->probe_ev[i]);
  if (b->stats_h) (void)hipHostFree(b->stats_h);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev_step) (void)hipEventDestroy(b->ev_step);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int batch_create_words(int64_t cap_words, tbfs_batch** out) {
  // word lists hold 32-bit word indices
  if (!out || cap_words <= 0 || cap_words > (int64_t)UINT32_MAX) return TBFS_ERR_INVALID;
  *out = nullptr;
  tbfs_batch* b = new (std::nothrow) tbfs_batch();
  if (!b) return TBFS_ERR_NOMEM;
  b->cap_words = cap_words;
  const size_t nw = (size_t)b->cap_words;
  const size_t nb = nw * 64;
  int dev = 0;
  int rc = TBFS_OK;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_step, hipEventDefault);
  if (e != hipSuccess) {
    rc = hip_fail(e, "batch stream/events", __FILE__, __LINE__);
    batch_free(b);
    return rc;
  }
  for (int i = 0; i < 3 && !rc; ++i) rc = dev_alloc(&b->F[i], nw);
  for (int i = 0; i < 3 && !rc; ++i) rc = dev_alloc(&b->lst[i], nw);
  for (int i = 0; i < 2 && !rc; ++i) rc = dev_alloc(&b->seen[i], nw);
  if (!rc) rc = dev_alloc(&b->ovfw, nw);
  if (!rc) rc = dev_alloc(&b->sigma, nb);
  if (!rc) rc = dev_alloc(&b->S, nb);
  if (!rc) rc = dev_alloc(&b->cset, 3);
  if (!rc) rc = dev_alloc(&b->loop, 2);
  if (!rc) rc = dev_alloc(&b->stats, 1);
  if (!rc) rc = dev_alloc(&b->roots, (size_t)64 * kBatchMaxW);
  if (!rc) {
    e = hipHostMalloc((void**)&b->stats_h, sizeof(RunStats), hipHostMallocDefault);
    if (e != hipSuccess) rc = hip_fail(e, "hipHostMalloc", __FILE__, __LINE__);
  }
  if (rc) {
    batch_free(b);
    return rc;
  }
  *out = b;
  return TBFS_OK;
}

int batch_create(int32_t n_cap, int32_t b_cap, tbfs_batch** out) {
  if (!out || n_cap <= 0 || b_cap <= 0 || b_cap > 64 * kBatchMaxW) return TBFS_ERR_INVALID;
  return batch_create_words((int64_t)n_cap * batch_words_for(b_cap), out);
}

int64_t batch_cap_words(const tbfs_batch* b) { return b->cap_words; }
hipEvent_t batch_step_event(const tbfs_batch* b) { return b->ev_step; }
hipStream_t batch_stream(const tbfs_batch* b) { return b->stream; }

namespace {

// Pull slices of every in-list (built on first use on stream s, kept with
// the graph).  A server's dispatcher and a caller's thread may start batches
// on the same graph: the lock makes one of them build the table.
std::mutex g_bsl_mu;

int graph_ensure_batch_slices(tbfs_graph* g, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_bsl_mu);
  if (g->bsl) return TBFS_OK;
  uint64_t* d = nullptr;
  int32_t count = 0;
  int rc = build_slices(g->CP, g->n, 0, s, &d, &count);
  if (rc) return rc;
  g->nbsl = count;
  g->bsl = d;
  return TBFS_OK;
}

BatchParams make_bparams(const tbfs_batch* b) {
  const tbfs_graph* g = b->g;
  BatchParams p{};
  p.n = g->n;
  p.nz = g->nz;
  p.nsrc = b->count;
  p.CP = g->CP;
  p.IC = g->IC;
  p.RP = g->RP;
  p.JR = g->JR;
  p.slices = g->bsl;
  p.nslices = g->nbsl;
  p.bigo = g->bigo;
  p.nbigo = g->nbigo;
  // short in-lists share a pass; long ones pay off one item at a time
  p.pull_lb = (int64_t)g->nz < (int64_t)kBatchLbDegree * g->n ? 1 : 0;
  // a pull costs more per entry on short in-lists: push longer there
  p.alpha = b->o.batch_alpha > 0 ? b->o.batch_alpha
                                 : (p.pull_lb ? kBatchAlphaShort : kBatchAlpha);
  for (int i = 0; i < 3; ++i) p.F[i] = b->F[i];
  for (int i = 0; i < 3; ++i) p.lst[i] = b->lst[i];
  for (int i = 0; i < 2; ++i) p.seen[i] = b->seen[i];
  p.ovfw = b->ovfw;
  p.sigma = b->sigma;
  p.S = b->S;
  p.cset = b->cset;
  p.loop = b->loop;
  p.stats = b->stats;
  p.roots = b->roots;
  p.perm = b->use_perm ? b->xperm : nullptr;
  p.open = b->open ? 1 : 0;
  return p;
}

// Enqueue (init + seed, when d0 == 0) and levels d0 .. d0+levels-1 of the
// current batch, then the stats copy and the step event.
// The full-occupancy level grid: about two 64-item chunks per wavefront of
// the pull (more workgroups than fit at once: those that finish early make
// room for the rest), between one and 32 workgroups per CU, a multiple of 8
// (XCD mapping).
void set_grid(tbfs_batch* b, const BatchParams& p) {
  if (b->grid) return;
  const int64_t chunks = (int64_t)b->W * (((int64_t)p.nslices + 63) >> 6);
  const int64_t cus = std::max(1, b->cus);
  int64_t gw = (chunks + 2 * (kBatchBlock / 64) - 1) / (2 * (kBatchBlock / 64));
  gw = std::min(std::max(gw, cus), cus * 32);
  gw = (gw + 7) & ~(int64_t)7;
  b->grid = b->o.grid > 0 ? std::min(b->o.grid, 65536) : (int)gw;
}

template <int W, typename CT>
int enqueue_w(tbfs_batch* b, int d0, int levels) {
  const BatchParams p = make_bparams(b);
  hipStream_t s = b->stream;
  set_grid(b, p);
  if (d0 == 0) {
    const int igrid = grid_for((int64_t)p.n * W, 256, std::max(1, b->cus) * 8);
    hipLaunchKernelGGL((k_batch_init<W, CT>), dim3(igrid), dim3(256), 0, s, p);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL((k_batch_seed<W, CT>), dim3((p.nsrc + 255) / 256), dim3(256), 0, s, p);
    HIPCHK(hipGetLastError());
  }
  // A BFS still going after kBatchWideLevels levels is in a long thin tail
  // (high-diameter graph): its levels run on one workgroup per CU, where a
  // launch costs far less than a full-occupancy grid (any grid is correct).
  const int narrow = std::max(1, b->cus);
  for (int d = d0; d < d0 + levels; ++d) {
    const int gd = d < kBatchWideLevels ? b->grid : std::min(b->grid, narrow);
    if (p.pull_lb)
      hipLaunchKernelGGL((k_batch_level<W, CT, true>), dim3(gd), dim3(kBatchBlock), 0, s, p, d);
    else
      hipLaunchKernelGGL((k_batch_level<W, CT, false>), dim3(gd), dim3(kBatchBlock), 0, s, p, d);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpyAsync(b->stats_h, b->stats, sizeof(RunStats), hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(b->ev_step, s));
  return TBFS_OK;
}

template <typename CT>
int enqueue_ct(tbfs_batch* b, int d0, int levels) {
  switch (b->W) {
    case 1: return enqueue_w<1, CT>(b, d0, levels);
    case 2: return enqueue_w<2, CT>(b, d0, levels);
    case 4: return enqueue_w<4, CT>(b, d0, levels);
    case 8: return enqueue_w<8, CT>(b, d0, levels);
    default: return enqueue_w<16, CT>(b, d0, levels);
  }
}

int enqueue(tbfs_batch* b, int d0, int levels) {
  const int rc = b->ct == 4 ? enqueue_ct<uint32_t>(b, d0, levels)
                            : enqueue_ct<uint64_t>(b, d0, levels);
  if (!rc) b->d_next = d0 + levels;
  return rc;
}

}  // namespace

namespace {
int ensure_perm(tbfs_batch* b, int32_t n);
int ensure_probes(tbfs_batch* b);
int stage_rows(int32_t n);
int alloc_staging(tbfs_batch* b, int64_t need);
}  // namespace

// Start a batch without waiting: roots are copied, the first chain of levels
// is enqueued on the batch's stream.  32-bit counts first (half the bytes of
// every count gather); a batch in which some count reaches 2^32 is run again
// with 64-bit counts by batch_progress.
int batch_start(tbfs_batch* b, const tbfs_graph* gc, const int32_t* roots, int32_t count,
                const tbfs_opts* o_in) {
  if (!b || !gc || !roots || count < 1 || count > 64 * kBatchMaxW) return TBFS_ERR_INVALID;
  if (b->active) return TBFS_ERR_INVALID;
  tbfs_graph* g = const_cast<tbfs_graph*>(gc);
  if (!g->CP) return TBFS_ERR_UNSUPPORTED;  // needs in-lists (CSC)
  const int W = batch_words_for(count);
  if ((int64_t)g->n * W > b->cap_words) return TBFS_ERR_INVALID;  // too big for this batch object
  for (int32_t i = 0; i < count; ++i)
    if (roots[i] < 0 || roots[i] >= g->n) return TBFS_ERR_INVALID;
  tbfs_opts o;
  if (o_in) o = *o_in; else tbfs_opts_default(&o);
  if (o.batch_count_bits != 0 && o.batch_count_bits != 32 && o.batch_count_bits != 64)
    return TBFS_ERR_INVALID;
  if (o.batch_order < -1 || o.batch_order > 1) return TBFS_ERR_INVALID;
  int rc = TBFS_OK;
  if (o.grid == 0 && o.batch_alpha == 0 && !tune_empty()) {  // defaults: the tuned values
    TuneKey k;
    TuneEntry e;
    rc = tune_key(g, W, b->stream, &k);
    if (rc) return rc;
    if (tune_find(k, &e)) {
      o.grid = e.grid_per_cu * std::max(1, b->cus);
      o.batch_alpha = e.batch_alpha;
    }
  }
  // the graph the batch runs on: g, or its degree-ordered copy (roots and
  // result rows are mapped through a copy of its id map)
  const int32_t* pm = nullptr;
  {
    tbfs_graph* gv = nullptr;
    rc = graph_batch_view(g, o.batch_order, b->stream, &gv, &pm);
    if (rc) return rc;
    b->use_perm = pm != nullptr;
    if (pm) {
      rc = ensure_perm(b, g->n);
      if (rc) return rc;
    }
    g = gv;
  }
  rc = graph_ensure_batch_slices(g, b->stream);
  if (rc) return rc;
  b->g = g;
  b->n = g->n;
  b->o = o;
  b->count = count;
  b->W = W;
  b->ct = (o.batch_count_bits == 64 || (o.batch_count_bits == 0 && g->wide_counts)) ? 8 : 4;
  b->grid = 0;
  b->done = false;
  b->seen_final = nullptr;
  HIPCHK(hipMemcpyAsync(b->roots, roots, (size_t)count * sizeof(int32_t),
                        hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  if (pm)  // inside the timed span: it is part of running on the copy
    HIPCHK(hipMemcpyAsync(b->xperm, pm, (size_t)g->n * sizeof(int32_t),
                          hipMemcpyDeviceToDevice, b->stream));
  rc = enqueue(b, 0, b->chain_k);
  if (rc) return rc;
  b->active = true;
  return TBFS_OK;
}

// Advance the batch in flight without blocking: *finished = true once its
// results are complete on the device (b->rs, b->ms filled).
int batch_progress(tbfs_batch* b, bool* finished) {
  *finished = false;
  if (b->open) return TBFS_ERR_INVALID;  // continuous batches report by probes
  if (!b->active) {
    *finished = b->done;
    return b->done ? TBFS_OK : TBFS_ERR_INVALID;
  }
  const hipError_t q = hipEventQuery(b->ev_step);
  if (q == hipErrorNotReady) return TBFS_OK;
  if (q != hipSuccess) {
    b->active = false;
    return hip_fail(q, "hipEventQuery", __FILE__, __LINE__);
  }
  const RunStats rs = *b->stats_h;
  if (rs.levels < 0) {  // needs more levels than enqueued
    if (b->d_next > b->n + 1) {  // a BFS has at most n levels: the state is corrupt
      std::fprintf(stderr, "turbobfs: batched BFS did not finish within %d levels\n", b->d_next);
      b->active = false;
      return TBFS_ERR_HIP;
    }
    const int rc = enqueue(b, b->d_next, b->chain_k);
    if (rc) b->active = false;
    return rc;
  }
  if (b->ct == 4 && rs.ovf_any && b->o.batch_count_bits != 32) {  // widen and run again
    const_cast<tbfs_graph*>(b->g)->wide_counts = true;  // start wide on this graph next time
    b->ct = 8;
    b->grid = 0;
    const int rc = enqueue(b, 0, b->chain_k);
    if (rc) b->active = false;
    return rc;
  }
  HIPCHK(hipEventElapsedTime(&b->ms, b->ev0, b->ev_step));
  b->rs = rs;
  // one spare level beyond the deepest batch so far
  b->chain_k = std::max(b->chain_k, std::min(4096, rs.levels + 2));
  b->seen_final = b->seen[rs.levels & 1];  // OR of every frontier
  b->active = false;
  b->done = true;
  *finished = true;
  return TBFS_OK;
}

void batch_fill_stats(const tbfs_batch* b, tbfs_stats* st) {
  const RunStats& rs = b->rs;
  st->levels = rs.levels;
  st->push_levels = rs.push_levels;
  st->pull_levels = rs.pull_levels;
  st->pad = b->ct * 8;  // count width of the final run
  st->reached = (int64_t)rs.reached;
  st->edges = (int64_t)rs.edges;
  st->ovf64 = rs.ovf_any ? -1 : 0;
  st->ms = (double)b->ms;
  st->launches = 0;
  st->pad2 = 0;
}

int batch_run(tbfs_batch* b, const tbfs_graph* g, const int32_t* roots, int32_t count,
              const tbfs_opts* o, tbfs_stats* st) {
  int rc = batch_start(b, g, roots, count, o);
  if (rc) return rc;
  for (;;) {
    bool fin = false;
    rc = batch_progress(b, &fin);
    if (rc) return rc;
    if (fin) break;
    HIPCHK(hipEventSynchronize(b->ev_step));
  }
  if (st) batch_fill_stats(b, st);
  return TBFS_OK;
}

// Enqueue the rows of sources first .. first+count-1 of the finished batch
// into per-source host destinations (row i -> S_dst[i], sig_dst[i],
// ovf_dst[i]; any may be null), through a device staging buffer of bounded
// size; ends with the step event recorded (nothing waits here).
namespace {

// Device staging for rows on their way to the host: up to 64 rows of n
// entries, at most 256 MiB (at least one row).
int stage_rows(int32_t n) {
  const int64_t stage_bytes = (int64_t)256 << 20;
  return (int)std::max<int64_t>(1, std::min<int64_t>(64, stage_bytes / ((int64_t)n * 13)));
}

int alloc_staging(tbfs_batch* b, int64_t need) {
  if (b->xcap >= need) return TBFS_OK;
  HIPCHK(hipStreamSynchronize(b->stream));
  dev_free(b->xS);
  dev_free(b->xsig);
  dev_free(b->xovf);
  b->xcap = 0;
  int rc = dev_alloc(&b->xS, (size_t)need);
  if (!rc) rc = dev_alloc(&b->xsig, (size_t)need);
  if (!rc) rc = dev_alloc(&b->xovf, (size_t)need);
  if (rc) return rc;
  b->xcap = need;
  return TBFS_OK;
}

// *rows = the rows of n entries the staging holds (grown when needed).
int ensure_staging(tbfs_batch* b, int32_t n, int* rows) {
  *rows = stage_rows(n);
  return alloc_staging(b, (int64_t)*rows * n);
}

int ensure_probes(tbfs_batch* b) {
  if (b->probe_d) return TBFS_OK;
  int rc = dev_alloc(&b->probe_d, kOpenRing);
  if (rc) return rc;
  HIPCHK(hipHostMalloc((void**)&b->probe_h, kOpenRing * sizeof(OpenProbe),
                       hipHostMallocMapped | hipHostMallocCoherent));
  std::memset(b->probe_h, 0, kOpenRing * sizeof(OpenProbe));
  HIPCHK(hipHostGetDevicePointer((void**)&b->probe_hd, b->probe_h, 0));
  for (int i = 0; i < kOpenRing; ++i)
    HIPCHK(hipEventCreateWithFlags(&b->probe_ev[i], hipEventDisableTiming));
  return TBFS_OK;
}

int ensure_perm(tbfs_batch* b, int32_t n) {
  if (b->xperm_cap >= n) return TBFS_OK;
  HIPCHK(hipStreamSynchronize(b->stream));
  dev_free(b->xperm);
  b->xperm_cap = 0;
  const int rc = dev_alloc(&b->xperm, (size_t)n);
  if (rc) return rc;
  b->xperm_cap = n;
  return TBFS_OK;
}


// Copies of staged rows 0 .. c-1 to the destinations of rows f .. f+c-1.
int copy_rows(tbfs_batch* b, int32_t n, int32_t f, int32_t c, int32_t* const* S_dst,
              uint64_t* const* sig_dst, uint8_t* const* ovf_dst) {
  hipStream_t s = b->stream;
  for (int32_t r = 0; r < c; ++r) {
    const size_t off = (size_t)r * n;
    if (S_dst && S_dst[f + r])
      HIPCHK(hipMemcpyAsync(S_dst[f + r], b->xS + off, (size_t)n * sizeof(int32_t),
                            hipMemcpyDeviceToHost, s));
    if (sig_dst && sig_dst[f + r])
      HIPCHK(hipMemcpyAsync(sig_dst[f + r], b->xsig + off, (size_t)n * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, s));
    if (ovf_dst && ovf_dst[f + r])
      HIPCHK(hipMemcpyAsync(ovf_dst[f + r], b->xovf + off, (size_t)n, hipMemcpyDeviceToHost, s));
  }
  return TBFS_OK;
}

}  // namespace

// Every buffer a batch on graphs of up to nmax vertices may grow into (row
// staging for the largest of graphs' n, the id map, the continuous batch's
// probes), allocated now: growing one later frees device memory, which
// waits for every stream on the device.
int batch_reserve(tbfs_batch* b, const int32_t* ns, int32_t count) {
  if (!b || !ns || count < 1) return TBFS_ERR_INVALID;
  int64_t stage = 0;
  int32_t nmax = 0;
  for (int32_t i = 0; i < count; ++i) {
    if (ns[i] < 1) return TBFS_ERR_INVALID;
    stage = std::max(stage, (int64_t)stage_rows(ns[i]) * ns[i]);
    nmax = std::max(nmax, ns[i]);
  }
  int rc = alloc_staging(b, stage);
  if (!rc) rc = ensure_perm(b, nmax);
  if (!rc) rc = ensure_probes(b);
  return rc;
}

int batch_enqueue_rows(tbfs_batch* b, int32_t first, int32_t count, int32_t* const* S_dst,
                       uint64_t* const* sig_dst, uint8_t* const* ovf_dst) {
  if (!b->done || first < 0 || count < 0 || first + count > b->count) return TBFS_ERR_INVALID;
  const int32_t n = b->n;
  int rows = 0;
  int rc = ensure_staging(b, n, &rows);
  if (rc) return rc;
  hipStream_t s = b->stream;
  for (int32_t f = 0; f < count; f += rows) {
    const int32_t c = std::min(rows, count - f);
    if (b->ct == 4)
      hipLaunchKernelGGL(k_batch_extract<uint32_t>, dim3((n + 63) / 64), dim3(256), 0, s, b->S,
                         (const uint32_t*)b->sigma, b->ovfw, b->seen_final, n, first + f, c,
                         b->xS, b->xsig, b->xovf, b->use_perm ? b->xperm : nullptr);
    else
      hipLaunchKernelGGL(k_batch_extract<uint64_t>, dim3((n + 63) / 64), dim3(256), 0, s, b->S,
                         (const uint64_t*)b->sigma, b->ovfw, b->seen_final, n, first + f, c,
                         b->xS, b->xsig, b->xovf, b->use_perm ? b->xperm : nullptr);
    HIPCHK(hipGetLastError());
    rc = copy_rows(b, n, f, c, S_dst, sig_dst, ovf_dst);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(b->ev_step, s));
  return TBFS_OK;
}

int batch_read(tbfs_batch* b, int32_t first, int32_t count, int32_t n, int32_t* S,
               uint64_t* sigma, uint8_t* ovf) {
  if (!b || !b->done || n != b->n || n <= 0 || first < 0 || count < 0 ||
      first + count > b->count)
    return TBFS_ERR_INVALID;
  if (count == 0) return TBFS_OK;
  std::vector<int32_t*> sd((size_t)count);
  std::vector<uint64_t*> gd((size_t)count);
  std::vector<uint8_t*> od((size_t)count);
  for (int32_t i = 0; i < count; ++i) {
    sd[i] = S ? S + (size_t)i * n : nullptr;
    gd[i] = sigma ? sigma + (size_t)i * n : nullptr;
    od[i] = ovf ? ovf + (size_t)i * n : nullptr;
  }
  int rc = batch_enqueue_rows(b, first, count, sd.data(), gd.data(), od.data());
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(b->stream));
  return TBFS_OK;
}

// ---- continuous batches (host side) -------------------------------------------
namespace {

template <int W>
int open_levels_w(tbfs_batch* b, int levels, bool wide) {
  const BatchParams p = make_bparams(b);
  hipStream_t s = b->stream;
  // narrow levels (one workgroup per CU) unless the last probed frontier
  // was large: a continuous batch lives mostly in thin tails
  const int gd = wide ? b->grid : std::min(b->grid, std::max(1, b->cus));
  for (int i = 0; i < levels; ++i) {
    const int d = b->d_next + i;
    if (p.pull_lb)
      hipLaunchKernelGGL((k_batch_level<W, uint64_t, true>), dim3(gd), dim3(kBatchBlock), 0, s,
                         p, d);
    else
      hipLaunchKernelGGL((k_batch_level<W, uint64_t, false>), dim3(gd), dim3(kBatchBlock), 0, s,
                         p, d);
    HIPCHK(hipGetLastError());
  }
  const int d = b->d_next + levels;
  const int slot = (int)(b->chunk_head % kOpenRing);
  const int pgrid = grid_for((int64_t)p.n * W, 256, std::max(1, b->cus) * 2);
  hipLaunchKernelGGL((k_batch_open_probe<W>), dim3(pgrid), dim3(256), 0, s, p, d,
                     b->probe_d + slot);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_batch_open_publish, dim3(1), dim3(64), 0, s, b->probe_d + slot,
                     b->probe_hd + slot, b->probe_d + (slot + 1) % kOpenRing);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b->probe_ev[slot], s));
  ++b->chunk_head;
  b->d_next = d;
  return TBFS_OK;
}

}  // namespace

int batch_open_begin(tbfs_batch* b, const tbfs_graph* gc, int32_t words, const tbfs_opts* o_in) {
  if (!b || !gc || (words != 1 && words != 2 && words != 4 && words != 8 && words != 16))
    return TBFS_ERR_INVALID;
  if (b->active) return TBFS_ERR_INVALID;
  tbfs_graph* g = const_cast<tbfs_graph*>(gc);
  if (!g->CP) return TBFS_ERR_UNSUPPORTED;
  if ((int64_t)g->n * words > b->cap_words) return TBFS_ERR_INVALID;
  tbfs_opts o;
  if (o_in) o = *o_in; else tbfs_opts_default(&o);
  if (o.batch_order < -1 || o.batch_order > 1) return TBFS_ERR_INVALID;
  int rc = TBFS_OK;
  rc = ensure_probes(b);
  if (rc) return rc;
  const int32_t* pm = nullptr;
  {
    tbfs_graph* gv = nullptr;
    rc = graph_batch_view(g, o.batch_order, b->stream, &gv, &pm);
    if (rc) return rc;
    b->use_perm = pm != nullptr;
    if (pm) {
      rc = ensure_perm(b, g->n);
      if (rc) return rc;
    }
    g = gv;
  }
  rc = graph_ensure_batch_slices(g, b->stream);
  if (rc) return rc;
  b->g = g;
  b->n = g->n;
  b->o = o;
  b->count = 64 * words;
  b->W = words;
  b->ct = 8;
  b->grid = 0;
  b->done = false;
  b->seen_final = nullptr;
  b->open = true;
  b->d_next = 0;
  b->chunk_head = 0;
  b->chunk_tail = 0;
  hipStream_t s = b->stream;
  if (pm)
    HIPCHK(hipMemcpyAsync(b->xperm, pm, (size_t)g->n * sizeof(int32_t), hipMemcpyDeviceToDevice,
                          s));
  HIPCHK(hipMemsetAsync(b->probe_d, 0, kOpenRing * sizeof(OpenProbe), s));
  const BatchParams p = make_bparams(b);
  set_grid(b, p);
  const int igrid = grid_for((int64_t)p.n * words, 256, std::max(1, b->cus) * 8);
  switch (words) {
    case 1: hipLaunchKernelGGL((k_batch_init<1, uint64_t>), dim3(igrid), dim3(256), 0, s, p); break;
    case 2: hipLaunchKernelGGL((k_batch_init<2, uint64_t>), dim3(igrid), dim3(256), 0, s, p); break;
    case 4: hipLaunchKernelGGL((k_batch_init<4, uint64_t>), dim3(igrid), dim3(256), 0, s, p); break;
    case 8: hipLaunchKernelGGL((k_batch_init<8, uint64_t>), dim3(igrid), dim3(256), 0, s, p); break;
    default: hipLaunchKernelGGL((k_batch_init<16, uint64_t>), dim3(igrid), dim3(256), 0, s, p);
  }
  HIPCHK(hipGetLastError());
  b->active = true;
  return TBFS_OK;
}

int batch_open_seed(tbfs_batch* b, const int32_t* bits, const int32_t* roots, int32_t count) {
  if (!b || !b->open || !b->active || count < 0 || (count > 0 && (!bits || !roots)))
    return TBFS_ERR_INVALID;
  for (int32_t i = 0; i < count; ++i)
    if (bits[i] < 0 || bits[i] >= b->count || roots[i] < 0 || roots[i] >= b->n)
      return TBFS_ERR_INVALID;
  const BatchParams p = make_bparams(b);
  for (int32_t f = 0; f < count; f += 64) {
    OpenSeeds sd{};
    const int32_t c = std::min<int32_t>(64, count - f);
    for (int32_t i = 0; i < c; ++i) {
      sd.bit[i] = bits[f + i];
      sd.root[i] = roots[f + i];
    }
    hipLaunchKernelGGL(k_batch_open_seed, dim3(1), dim3(64), 0, b->stream, p, b->d_next, sd, c);
    HIPCHK(hipGetLastError());
  }
  return TBFS_OK;
}

int batch_open_extract(tbfs_batch* b, const int32_t* bits, const int32_t* offs, int32_t count,
                       int32_t* const* S_dst, uint64_t* const* sig_dst,
                       uint8_t* const* ovf_dst) {
  if (!b || !b->open || !b->active || count < 0 || (count > 0 && (!bits || !offs)))
    return TBFS_ERR_INVALID;
  for (int32_t i = 0; i < count; ++i)
    if (bits[i] < 0 || bits[i] >= b->count || offs[i] < 0 || offs[i] > b->d_next)
      return TBFS_ERR_INVALID;
  if (count == 0) return TBFS_OK;
  const int32_t n = b->n;
  int rows = 0;
  int rc = ensure_staging(b, n, &rows);
  if (rc) return rc;
  rows = std::min(rows, 64);
  for (int32_t f = 0; f < count; f += rows) {
    OpenRows rw{};
    const int32_t c = std::min(rows, count - f);
    for (int32_t i = 0; i < c; ++i) {
      rw.bit[i] = bits[f + i];
      rw.off[i] = offs[f + i];
      rw.wmask[bits[f + i] >> 6] |= 1ull << (bits[f + i] & 63);
    }
    hipLaunchKernelGGL(k_batch_open_extract, dim3((n + 63) / 64), dim3(256), 0, b->stream, b->S,
                       b->sigma, b->ovfw, b->seen[b->d_next & 1], b->seen[(b->d_next + 1) & 1], n,
                       b->W, rw, c, b->xS, b->xsig, b->xovf, b->use_perm ? b->xperm : nullptr);
    HIPCHK(hipGetLastError());
    rc = copy_rows(b, n, f, c, S_dst, sig_dst, ovf_dst);
    if (rc) return rc;
  }
  return TBFS_OK;
}

int batch_open_chunk(tbfs_batch* b, int32_t levels, bool wide) {
  if (!b || !b->open || !b->active || levels < 0) return TBFS_ERR_INVALID;
  if (b->chunk_head - b->chunk_tail >= kOpenRing) return TBFS_ERR_INVALID;  // probes not read
  if ((int64_t)b->d_next + levels >= ((int64_t)1 << 30)) return TBFS_ERR_INVALID;
  switch (b->W) {
    case 1: return open_levels_w<1>(b, levels, wide);
    case 2: return open_levels_w<2>(b, levels, wide);
    case 4: return open_levels_w<4>(b, levels, wide);
    case 8: return open_levels_w<8>(b, levels, wide);
    default: return open_levels_w<16>(b, levels, wide);
  }
}

int batch_open_poll(tbfs_batch* b, bool* ready, int32_t* boundary, uint64_t* alive,
                    int64_t* frontier_words) {
  *ready = false;
  if (!b || !b->open) return TBFS_ERR_INVALID;
  if (b->chunk_tail == b->chunk_head) return TBFS_OK;
  const int slot = (int)(b->chunk_tail % kOpenRing);
  const hipError_t e = hipEventQuery(b->probe_ev[slot]);
  if (e == hipErrorNotReady) return TBFS_OK;
  if (e != hipSuccess) return hip_fail(e, "hipEventQuery", __FILE__, __LINE__);
  const OpenProbe& pr = b->probe_h[slot];
  *boundary = pr.d;
  for (int w = 0; w < 16; ++w) alive[w] = w < b->W ? pr.alive[w] : 0ull;
  if (frontier_words) *frontier_words = (int64_t)pr.q;
  ++b->chunk_tail;
  *ready = true;
  return TBFS_OK;
}

int batch_open_outstanding(const tbfs_batch* b) {
  return b && b->open ? (int)(b->chunk_head - b->chunk_tail) : 0;
}

int32_t batch_open_depth(const tbfs_batch* b) { return b ? b->d_next : 0; }

int batch_open_end(tbfs_batch* b) {
  if (!b || !b->open) return TBFS_ERR_INVALID;
  HIPCHK(hipStreamSynchronize(b->stream));
  b->open = false;
  b->active = false;
  b->done = false;
  b->chunk_tail = b->chunk_head;
  return TBFS_OK;
}

// Autotuner: the batch's launch grid and push / pull threshold, chosen by
// timing candidates on the caller's graph and roots (every candidate gives
// identical results; only the time differs).
int batch_autotune(tbfs_batch* b, const tbfs_graph* g, const int32_t* roots, int32_t count,
                   const tbfs_opts* base, tbfs_opts* best, double* best_ms) {
  if (!b || !g || !roots || !best) return TBFS_ERR_INVALID;
  tbfs_opts o;
  if (base) o = *base; else tbfs_opts_default(&o);
  const int cus = std::max(1, b->cus);
  const int grids[] = {0, cus * 8, cus * 16, cus * 32};
  const float alphas[] = {4.0f, 32.0f, 128.0f};
  tbfs_stats st{};
  o.grid = 0;
  o.batch_alpha = 0;
  int rc = batch_run(b, g, roots, count, &o, &st);  // warm-up (slices, first launches)
  if (rc) return rc;
  double bt = -1;
  tbfs_opts bo = o;
  for (int gsz : grids) {
    for (float al : alphas) {
      tbfs_opts t = o;
      t.grid = gsz;
      t.batch_alpha = al;
      double ms = -1;
      for (int rep = 0; rep < 2; ++rep) {
        rc = batch_run(b, g, roots, count, &t, &st);
        if (rc) return rc;
        if (ms < 0 || st.ms < ms) ms = st.ms;
      }
      if (bt < 0 || ms < bt) {
        bt = ms;
        bo = t;
      }
    }
  }
  *best = bo;
  if (best_ms) *best_ms = bt;
  TuneEntry e;
  rc = tune_key(const_cast<tbfs_graph*>(g), batch_words_for(count), b->stream, &e.key);
  if (rc) return rc;
  e.grid_per_cu = bo.grid / cus;
  e.batch_alpha = bo.batch_alpha;
  e.ms = bt;
  tune_put(e);
  return TBFS_OK;
}

int batch_tuned(const tbfs_graph* g, int32_t count, tbfs_opts* o) {
  if (!g || !o || count < 1 || count > 64 * kBatchMaxW) return TBFS_ERR_INVALID;
  if (tune_empty()) return TBFS_ERR_UNSUPPORTED;
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  TuneKey k;
  int rc = tune_key(const_cast<tbfs_graph*>(g), batch_words_for(count), s, &k);
  (void)hipStreamDestroy(s);
  if (rc) return rc;
  TuneEntry e;
  if (!tune_find(k, &e)) return TBFS_ERR_UNSUPPORTED;
  int dev = 0, cus = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  o->grid = e.grid_per_cu * std::max(1, cus);
  o->batch_alpha = e.batch_alpha;
  return TBFS_OK;
}

}  // namespace tbfs

extern "C" {

int tbfs_batch_autotune(tbfs_batch* b, const tbfs_graph* g, const int32_t* roots, int32_t count,
                        const tbfs_opts* base, tbfs_opts* best, double* best_ms) {
  return tbfs::batch_autotune(b, g, roots, count, base, best, best_ms);
}

int tbfs_batch_tuned(const tbfs_graph* g, int32_t count, tbfs_opts* o) {
  return tbfs::batch_tuned(g, count, o);
}

int tbfs_batch_tune_clear(void) { return tbfs_tune_clear(); }

int tbfs_batch_create(int32_t n_cap, int32_t b_cap, tbfs_batch** out) {
  return tbfs::batch_create(n_cap, b_cap, out);
}

int tbfs_batch_destroy(tbfs_batch* b) {
  tbfs::batch_free(b);
  return TBFS_OK;
}

void* tbfs_batch_stream(tbfs_batch* b) { return b ? (void*)b->stream : nullptr; }

int tbfs_batch_run(tbfs_batch* b, const tbfs_graph* g, const int32_t* roots, int32_t count,
                   const tbfs_opts* o, tbfs_stats* st) {
  return tbfs::batch_run(b, g, roots, count, o, st);
}

int tbfs_batch_read(tbfs_batch* b, int32_t first, int32_t count, int32_t n, int32_t* S,
                    uint64_t* sigma, uint8_t* ovf) {
  return tbfs::batch_read(b, first, count, n, S, sigma, ovf);
}

}  // extern "C"
