// tbfs_persist.hip -- load-balanced BFS levels and the device-resident level
// loops: one cooperative launch per BFS with a grid barrier between levels
// (persistent), or one launch per level enqueued back to back (chained).
//
// Direction-optimizing BFS with exact path counts (the reference's TD/BU
// switch, TurboBFS_CSC_TDBU/bfsgputdbucsc_sc.cu:108-153, redone for wave64):
//   push level  expands the frontier, read from its bitmap.  A wavefront takes
//               the set bits of 64 frontier words, numbers them across the
//               words, scans their out-degrees and walks the concatenated
//               out-lists 64 edges at a time; out-lists longer than kBigIn are
//               cut into static slices (graph property) spread over all
//               wavefronts.  Counts are 64-bit atomics whose old value exposes
//               the carry; the adder that finds a count at 0 claims the
//               vertex's bit in the next frontier and records its depth.
//   pull level  sweeps the unvisited columns.  A wavefront takes the 64
//               columns of one frontier word, scans their in-degrees and walks
//               the concatenated in-lists; parents' counts are summed per
//               column in LDS.  Columns with in-degree > kBigIn are instead cut
//               into static kPullChunk slices (graph property, computed once)
//               that many wavefronts reduce with global atomics.  No early exit:
//               a column's count needs all of its parents.
// Level counters (vertices found, their out- and in-degree sums) are tallied
// per lane and added to the device counters once per workgroup per level.
// The level loop keeps every decision on the device: each workgroup reads the
// same level counters after the grid barrier and takes the same branch.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tbfs.h"
#include "tbfs_device.h"
#include "tbfs_internal.h"

namespace tbfs {
namespace {

using namespace dev;

constexpr int kUnroll = 4;  // independent edges in flight per lane
constexpr int kPullRegs = 8;  // in-list entries per lane per light-column pass
constexpr int kSliceRegs = kPullChunk / kWave;  // slice entries per lane
constexpr int kSlicesPerPass = 2;  // pull slices taken together by a wavefront

struct WaveLds {
  int32_t excl[64];  // exclusive prefix of per-lane work
  int32_t beg[64];   // first edge index of each lane's slice
  uint64_t x[64];    // push: count of each lane's vertex; pull: per-column sum
                     // of the parents' counts' low 32 bits
  uint64_t xh[64];   // pull: per-column sum of their high 32 bits
};

// Per-lane tallies of the vertices this lane discovered in the current level.
struct LevelAcc {
  uint64_t q = 0, mout = 0, min = 0;
  uint32_t ovf = 0;

  __device__ __forceinline__ void found(const GraphView& g, int v) {
    found_in(g, v, indeg(g, v));
  }
  // v's in-degree already known: on a symmetric graph (out-lists alias the
  // in-lists) nothing more is loaded
  __device__ __forceinline__ void found_in(const GraphView& g, int v, uint32_t di) {
    q += 1;
    min += di;
    mout += (g.RP == g.CP) ? di : outdeg(g, v);
  }
};

// Sum the tallies of the workgroup and add them to its stripe of the level
// counters (two 64-bit atomics).  Every thread of the workgroup must call it.
__device__ __forceinline__ void flush_level(CtrSet* set, LevelAcc& a) {
  __shared__ unsigned long long red[kMaxBlock / 64][3];
  __shared__ unsigned int redo[kMaxBlock / 64];
  const uint64_t q = wave_sum_u64(a.q);
  const uint64_t mo = wave_sum_u64(a.mout);
  const uint64_t mi = wave_sum_u64(a.min);
  const bool ov = __ballot(a.ovf != 0) != 0;
  const int wv = threadIdx.x >> 6;
  if (lane() == 0) {
    red[wv][0] = q;
    red[wv][1] = mo;
    red[wv][2] = mi;
    redo[wv] = ov ? 1u : 0u;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t0 = 0, t1 = 0, t2 = 0;
    unsigned int to = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      t0 += red[i][0];
      t1 += red[i][1];
      t2 += red[i][2];
      to |= redo[i];
    }
    CtrStripe* st = &set->s[blockIdx.x & (kStripes - 1)];
    if (t0) {
      atomicAdd(&st->qm, (t0 << 32) + t1);
      atomicAdd(&st->min, t2);
    }
    if (to) atomicOr(&st->ovf, 1u);
  }
  __syncthreads();
  a = LevelAcc();
}

// This lane's stripe of a striped counter set (lanes 0..kStripes-1 load one
// stripe each).  `agent`: the loads bypass caches that may hold values from
// before the last grid barrier.
struct StripeVals {
  uint64_t qm = 0, mi = 0;
  uint32_t ov = 0;
};
__device__ __forceinline__ StripeVals load_set(const CtrSet* set, bool agent) {
  StripeVals r;
  const int l = lane();
  if (l < kStripes) {
    const CtrStripe* st = &set->s[l];
    if (agent) {
      r.qm = __hip_atomic_load(&st->qm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      r.mi = __hip_atomic_load(&st->min, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      r.ov = __hip_atomic_load(&st->ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      r.qm = st->qm;
      r.mi = st->min;
      r.ov = st->ovf;
    }
  }
  return r;
}

// Totals of the stripes, on every lane of the calling wavefront.
__device__ __forceinline__ LevelCtr sum_set(const StripeVals& r) {
  LevelCtr c;
  c.q = wave_sum_u64(r.qm >> 32);
  c.qe = 0;
  c.mout = wave_sum_u64(r.qm & 0xffffffffull);
  c.min = wave_sum_u64(r.mi);
  c.ovf = __ballot(r.ov != 0) ? 1u : 0u;
  return c;
}

__device__ __forceinline__ LevelCtr read_set(const CtrSet* set, bool agent) {
  return sum_set(load_set(set, agent));
}

// One edge u -> v of a push level, u at depth d with count x (overflow ou).
// "Reached at depth <= d" is read from the reached bitmap or the current
// frontier (both small and cached); otherwise v is reached at depth d+1 and
// the count is added at once.  A vertex not reached before this level holds
// count 0, so only an adder that finds 0 can be the first: those (normally
// one per vertex; more only when a count is 0 mod 2^64) claim v's bit in the
// next frontier, and the one that sets it records the depth.  Device-scope
// atomics go out to memory, so the second one is skipped where it can be.
__device__ __forceinline__ void push_edge(const GraphView& g, const StateView& st,
                                          int dn, int v, uint64_t x, bool ou,
                                          LevelAcc& a) {
  const uint64_t bit = 1ull << (v & 63);
  if ((st.vis[v >> 6] | st.fcur[v >> 6]) & bit) return;
  const unsigned long long old =
      atomicAdd((unsigned long long*)&st.sigma[v], (unsigned long long)x);
  if (old + x < old || ou) {
    st.ovf[v] = 1;
    a.ovf = 1;
  }
  if (old == 0) {
    const unsigned long long oldb =
        atomicOr((unsigned long long*)&st.fnext[v >> 6], (unsigned long long)bit);
    if ((oldb & bit) == 0) {
      st.S[v] = dn;
      a.found(g, v);
    }
  }
}

// ----------------------------------------------------------------------------

// Push level, driven by the frontier bitmap: (a) static slices of the long
// out-lists whose vertex is in the frontier, (b) every other frontier vertex:
// frontier words are dealt to wavefronts round-robin; a wavefront loads 64 of
// its words at once, and for each nonzero one walks the out-lists of its set
// bits load-balanced, kUnroll edges per lane in flight.
// Number of the words wave, wave + nwaves, ... below nwords.
__device__ __forceinline__ int64_t nmine_of(int64_t nwords, int64_t wave, int64_t nwaves) {
  return nwords > wave ? (nwords - wave + nwaves - 1) / nwaves : 0;
}

// Position of the r-th set bit (0-based) of x; x has more than r set bits.
__device__ __forceinline__ int select_bit(uint64_t x, int r) {
  int pos = 0;
  int c = __popc((uint32_t)x);
  if (r >= c) {
    r -= c;
    x >>= 32;
    pos += 32;
  }
  uint32_t w = (uint32_t)x;
  c = __popc(w & 0xffffu);
  if (r >= c) { r -= c; w >>= 16; pos += 16; }
  c = __popc(w & 0xffu);
  if (r >= c) { r -= c; w >>= 8; pos += 8; }
  c = __popc(w & 0xfu);
  if (r >= c) { r -= c; w >>= 4; pos += 4; }
  c = __popc(w & 0x3u);
  if (r >= c) { r -= c; w >>= 2; pos += 2; }
  if (r >= (int)(w & 1u)) pos += 1;
  return pos;
}

__device__ __forceinline__ void push_lb_level(const GraphView& g,
                                              const StateView& st, int dn,
                                              uint32_t ovf_any,
                                              const uint64_t* bigo,
                                              int32_t nbigo, int64_t wave,
                                              int64_t nwaves, WaveLds* L,
                                              LevelAcc& a) {
  // this wave's share: entries wave, wave + nwaves, ...; 64 examined at once.
  // Each lane loads the bounds and count of its own entry; the wavefront then
  // takes the active slices in turn, kSliceRegs out-list entries per lane, all
  // loads of a slice in flight together.
  const int64_t nbo = nbigo > wave ? (nbigo - wave + nwaves - 1) / nwaves : 0;
  for (int64_t b = 0; b < nbo; b += 64) {
    const int64_t i = wave + (b + lane()) * nwaves;
    bool act = false;
    int cb = 0, ce = 0;
    uint64_t x = 0;
    int ou = 0;
    if (b + lane() < nbo) {
      const uint64_t ent = bigo[i];
      const int u = (int)(uint32_t)ent;
      act = fbit(st.fcur, u);
      if (act) {
        cb = g.RP[u] + (int)(ent >> 32) * kPullChunk;
        ce = min(g.RP[u + 1], cb + kPullChunk);
        x = st.sigma[u];
        if (ovf_any) ou = st.ovf[u];
      }
    }
    uint64_t todo = __ballot(act);
    while (todo) {
      const int j = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const int scb = __shfl(cb, j);
      const int sce = __shfl(ce, j);
      const uint64_t sx = __shfl(x, j);
      const bool sou = __shfl(ou, j) != 0;
      int v[kSliceRegs];
#pragma unroll
      for (int t = 0; t < kSliceRegs; ++t) {
        const int k = scb + lane() + 64 * t;
        v[t] = k < sce ? g.JR[k] : -1;
      }
#pragma unroll
      for (int t = 0; t < kSliceRegs; ++t)
        if (v[t] >= 0) push_edge(g, st, dn, v[t], sx, sou, a);
    }
  }
  const int64_t nwords = ((int64_t)g.n + 63) >> 6;
  // this wave's words: wave, wave + nwaves, ...; their bits loaded 64 at once.
  // The frontier vertices of those 64 words are numbered across the words
  // (prefix sum of popcounts) and handed out one per lane, 64 at a time, so a
  // sparse frontier costs one pass per 64 vertices, not one per word.
  const int64_t nm = nmine_of(nwords, wave, nwaves);
  for (int64_t b = 0; b < nm; b += 64) {
    const int64_t wl = wave + (b + lane()) * nwaves;
    const uint64_t bits_l = (b + lane() < nm) ? st.fcur[wl] : 0ull;
    const uint32_t cnt = (uint32_t)__popcll(bits_l);
    const uint32_t wincl = wave_incl_scan(cnt);
    const int nv = (int)__shfl(wincl, 63);
    if (nv == 0) continue;
    for (int k0 = 0; k0 < nv; k0 += 64) {
      // vertex k0 + lane: its word (owner lane) and its bit within the word
      wave_lds_sync();
      L->excl[lane()] = (int)(wincl - cnt);
      L->x[lane()] = bits_l;
      wave_lds_sync();
      int u = -1;
      const int k = k0 + lane();
      if (k < nv) {
        const int j = owner_of(L->excl, k);
        const int bit = select_bit(L->x[j], k - L->excl[j]);
        u = (int)((wave + (b + j) * nwaves) * 64) + bit;
      }
      int beg = 0, deg = 0;
      uint64_t x = 0;
      uint32_t ou = 0;
      if (u >= 0) {
        beg = g.RP[u];
        const int dg = g.RP[u + 1] - beg;
        if (dg <= kBigIn) {
          deg = dg;
          x = st.sigma[u];
          if (ovf_any) ou = st.ovf[u];
        }
      }
      const uint32_t incl = wave_incl_scan((uint32_t)deg);
      const int total = (int)__shfl(incl, 63);
      if (total == 0) continue;
      const uint64_t oumask = __ballot(ou != 0);
      wave_lds_sync();
      L->excl[lane()] = (int)incl - deg;
      L->beg[lane()] = beg;
      L->x[lane()] = x;
      wave_lds_sync();
      for (int e0 = lane(); e0 < total; e0 += 64 * kUnroll) {
        int v[kUnroll], o[kUnroll];
#pragma unroll
        for (int t = 0; t < kUnroll; ++t) {
          const int e = e0 + 64 * t;
          o[t] = -1;
          if (e < total) {
            o[t] = owner_of(L->excl, e);
            v[t] = g.JR[L->beg[o[t]] + (e - L->excl[o[t]])];
          }
        }
#pragma unroll
        for (int t = 0; t < kUnroll; ++t)
          if (o[t] >= 0)
            push_edge(g, st, dn, v[t], L->x[o[t]], (oumask >> o[t]) & 1ull, a);
      }
    }
  }
}

// One pass over in-list entries e0, e0 + 64, ... (R per lane) of the
// columns laid out in L (exclusive prefix excl, first entry beg), every load
// of the pass issued before any is used: owners, then in-list entries, then
// parent bits, then parents' counts.  Lanes past the end read entry 0 and
// are masked out; a count is read for every entry (non-parents read vertex
// 0's, one shared line), so no load sits behind a branch.  Column sums go
// into LDS as 32-bit halves, each added into a 64-bit word: exact for any
// in-degree below 2^32, so no atomic's result is needed to find a carry.
template <int R, bool SIGMA>
__device__ __forceinline__ void pull_light_pass(const GraphView& g, const StateView& st,
                                                WaveLds* L, const uint64_t* fpar,
                                                uint32_t ovf_any, int e0, int total,
                                                uint64_t& fmask, uint64_t& omask) {
  int o[R], u[R];
  bool par[R];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int e = e0 + 64 * t;
    const bool ok = e < total;
    o[t] = owner_of(L->excl, ok ? e : 0);
    const int k = ok ? L->beg[o[t]] + (e - L->excl[o[t]]) : 0;
    u[t] = g.IC[k];
    par[t] = ok;
  }
#pragma unroll
  for (int t = 0; t < R; ++t) par[t] = par[t] && fbit(fpar, u[t]);
  if (SIGMA) {
    uint64_t xs[R];
#pragma unroll
    for (int t = 0; t < R; ++t) xs[t] = st.sigma[par[t] ? u[t] : 0];
    if (ovf_any) {
#pragma unroll
      for (int t = 0; t < R; ++t)
        if (par[t] && st.ovf[u[t]]) omask |= 1ull << o[t];
    }
#pragma unroll
    for (int t = 0; t < R; ++t) {
      if (par[t]) {
        fmask |= 1ull << o[t];
        atomicAdd((unsigned long long*)&L->x[o[t]], (unsigned long long)(xs[t] & 0xffffffffull));
        const unsigned long long h = xs[t] >> 32;
        if (h) atomicAdd((unsigned long long*)&L->xh[o[t]], h);
      }
    }
  } else {
#pragma unroll
    for (int t = 0; t < R; ++t)
      if (par[t]) fmask |= 1ull << o[t];
  }
}

// Pull level: (a) static slices of the high in-degree columns, reduced with
// global atomics; (b) the other unreached columns: words are dealt to
// wavefronts round-robin; a wavefront loads 64 of its words of the
// reached-or-skipped bitmap at once, and for each word with a clear bit walks
// the in-lists of its clear bits load-balanced (kUnroll edges per lane in
// flight), summing parents' counts per column in LDS.
template <bool SIGMA>
__device__ __forceinline__ void pull_lb_level(const GraphView& g,
                                              const StateView& st, int dn,
                                              uint32_t ovf_any,
                                              const uint64_t* bigc,
                                              int32_t nbigc, int64_t wave,
                                              int64_t nwaves, WaveLds* L,
                                              LevelAcc& a,
                                              const uint64_t* fpar) {
  // Slices: each lane loads the bounds of its own entry; the wavefront takes
  // the active slices in turn (kSliceRegs in-list entries per lane, loads in
  // flight together) and the lane owning a slice keeps its sum; the atomics
  // of up to 64 slices then go out together.
  const int64_t nbc = nbigc > wave ? (nbigc - wave + nwaves - 1) / nwaves : 0;
  for (int64_t b = 0; b < nbc; b += 64) {
    const int64_t i = wave + (b + lane()) * nwaves;
    bool act = false;
    int v = 0, cb = 0, ce = 0;
    if (b + lane() < nbc) {
      const uint64_t ent = bigc[i];
      v = (int)(uint32_t)ent;
      const int sv = st.S[v];
      act = (sv == -1 || sv == dn);  // not reached before this level
      if (act) {
        cb = g.CP[v] + (int)(ent >> 32) * kPullChunk;
        ce = min(g.CP[v + 1], cb + kPullChunk);
      }
    }
    uint64_t todo = __ballot(act);
    bool mine_f = false;
    uint64_t mine_sum = 0;
    uint32_t mine_ov = 0;
    // kSlicesPerPass active slices per pass (kSliceRegs entries per lane of
    // each); a slot without a slice repeats the first one's start, empty
    while (todo) {
      int js[kSlicesPerPass], scb[kSlicesPerPass], sce[kSlicesPerPass];
#pragma unroll
      for (int h = 0; h < kSlicesPerPass; ++h) {
        js[h] = todo ? __ffsll((unsigned long long)todo) - 1 : -1;
        todo &= todo - 1;
        const int jj = js[h] >= 0 ? js[h] : js[0];
        scb[h] = __shfl(cb, jj);
        sce[h] = js[h] >= 0 ? __shfl(ce, jj) : scb[h];
      }
      constexpr int R = kSlicesPerPass * kSliceRegs;
      int u[R];
      bool par[R];
#pragma unroll
      for (int h = 0; h < kSlicesPerPass; ++h) {
#pragma unroll
        for (int t = 0; t < kSliceRegs; ++t) {
          const int k = scb[h] + lane() + 64 * t;
          u[h * kSliceRegs + t] = g.IC[k < sce[h] ? k : scb[0]];
          par[h * kSliceRegs + t] = k < sce[h];
        }
      }
#pragma unroll
      for (int t = 0; t < R; ++t) par[t] = par[t] && fbit(fpar, u[t]);
      bool f[kSlicesPerPass];
      uint64_t sum[kSlicesPerPass];
      uint32_t ov[kSlicesPerPass];
#pragma unroll
      for (int h = 0; h < kSlicesPerPass; ++h) {
        f[h] = false;
        sum[h] = 0;
        ov[h] = 0;
      }
      if (SIGMA) {
        uint64_t xs[R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
          const uint64_t xv = st.sigma[par[t] ? u[t] : 0];  // no load behind a branch
          xs[t] = par[t] ? xv : 0ull;
        }
#pragma unroll
        for (int t = 0; t < R; ++t) {
          const int h = t / kSliceRegs;
          f[h] |= par[t];
          add_carry(sum[h], ov[h], xs[t]);
          if (ovf_any && par[t]) ov[h] |= (uint32_t)st.ovf[u[t]];
        }
      } else {
#pragma unroll
        for (int t = 0; t < R; ++t) f[t / kSliceRegs] |= par[t];
      }
#pragma unroll
      for (int h = 0; h < kSlicesPerPass; ++h) {
        if (__ballot(f[h])) {
          if (SIGMA) wave_sum_carry(sum[h], ov[h]);
          if (lane() == js[h]) {
            mine_f = true;
            mine_sum = sum[h];
            mine_ov = ov[h];
          }
        }
      }
    }
    if (mine_f) {
      if (SIGMA) {
        const unsigned long long old =
            atomicAdd((unsigned long long*)&st.sigma[v], (unsigned long long)mine_sum);
        if (old + mine_sum < old) mine_ov = 1;
        if (mine_ov) {
          st.ovf[v] = 1;
          a.ovf = 1;
        }
      }
      const unsigned long long bit = 1ull << (v & 63);
      const unsigned long long old =
          atomicOr((unsigned long long*)&st.fnext[v >> 6], bit);
      if ((old & bit) == 0) {  // first slice of v to find a parent
        st.S[v] = dn;
        a.found(g, v);
      }
    }
  }
  const int64_t nwords = ((int64_t)g.n + 63) >> 6;
  const int64_t nmine = nwords > wave ? (nwords - wave + nwaves - 1) / nwaves : 0;
  for (int64_t b = 0; b < nmine; b += 64) {
    const int64_t wl = wave + (b + lane()) * nwaves;
    // reached (any frontier so far, the current one included) or not a
    // light column: nothing to do for that vertex
    const uint64_t done_l =
        (b + lane() < nmine) ? (st.vis[wl] | st.skip[wl] | st.fcur[wl]) : ~0ull;
    uint64_t todo = __ballot(done_l != ~0ull);
    // column bounds of the next word with work are loaded one word ahead:
    // their latency hides behind the current word's passes
    int nbeg = 0, nend = 0;
    if (todo) {
      const int j = __ffsll((unsigned long long)todo) - 1;
      const int64_t w = wave + (b + j) * nwaves;
      if (!((__shfl(done_l, j) >> lane()) & 1ull)) {
        const int v = (int)(w * 64) + lane();
        nbeg = g.CP[v];
        nend = g.CP[v + 1];
      }
    }
    while (todo) {
      const int j = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const int64_t w = wave + (b + j) * nwaves;
      const int v = (int)(w * 64) + lane();
      const int beg = nbeg;
      const int deg = nend - nbeg;
      nbeg = 0;
      nend = 0;
      if (todo) {
        const int jn = __ffsll((unsigned long long)todo) - 1;
        const int64_t wn = wave + (b + jn) * nwaves;
        if (!((__shfl(done_l, jn) >> lane()) & 1ull)) {
          const int vn = (int)(wn * 64) + lane();
          nbeg = g.CP[vn];
          nend = g.CP[vn + 1];
        }
      }
      const uint32_t incl = wave_incl_scan((uint32_t)deg);
      const int total = (int)__shfl(incl, 63);
      wave_lds_sync();
      L->excl[lane()] = (int)incl - deg;
      L->beg[lane()] = beg;
      if (SIGMA) {
        L->x[lane()] = 0;
        L->xh[lane()] = 0;
      }
      wave_lds_sync();
      uint64_t fmask = 0, omask = 0;
      // one pass of 4 entries per lane for short words, else passes of 8
      if (total <= 64 * 4) {
        pull_light_pass<4, SIGMA>(g, st, L, fpar, ovf_any, lane(), total, fmask, omask);
      } else {
        for (int e0 = lane(); e0 < total; e0 += 64 * kPullRegs)
          pull_light_pass<kPullRegs, SIGMA>(g, st, L, fpar, ovf_any, e0, total, fmask, omask);
      }
      fmask = wave_or_u64(fmask);
      if (fmask == 0) continue;
      if (SIGMA) omask = wave_or_u64(omask);
      wave_lds_sync();
      if ((fmask >> lane()) & 1ull) {
        st.S[v] = dn;
        if (SIGMA) {
          // count = xh * 2^32 + x; at least 2^64 iff xh >= 2^32 or the sum
          // of the low 64 bits carries
          const uint64_t xh = L->xh[lane()];
          const uint64_t hs = xh << 32;
          const uint64_t sum = hs + L->x[lane()];
          st.sigma[v] = sum;
          if ((xh >> 32) != 0 || sum < hs || ((omask >> lane()) & 1ull)) {
            st.ovf[v] = 1;
            a.ovf = 1;
          }
        }
        a.found_in(g, v, (uint32_t)deg);
      }
      if (lane() == 0) atomicOr((unsigned long long*)&st.fnext[w], fmask);
    }
  }
}

// ----------------------------------------------------------------------------
// standalone level kernels (host-driven loop)
__global__ __launch_bounds__(kLbBlock) void k_push_lb(GraphView g, StateView st,
                                                      int32_t d,
                                                      const uint64_t* bigo,
                                                      int32_t nbigo,
                                                      int ovf_any) {
  __shared__ WaveLds lds[kLbBlock / 64];
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  LevelAcc a;
  push_lb_level(g, st, d + 1, (uint32_t)ovf_any, bigo, nbigo, wave, nwaves,
                &lds[threadIdx.x >> 6], a);
  flush_level(st.cset, a);
}

template <bool SIGMA>
__global__ __launch_bounds__(kLbBlock) void k_pull_lb(GraphView g, StateView st,
                                                      int32_t d,
                                                      const uint64_t* bigc,
                                                      int32_t nbigc,
                                                      int ovf_any) {
  __shared__ WaveLds lds[kLbBlock / 64];
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  LevelAcc a;
  pull_lb_level<SIGMA>(g, st, d + 1, (uint32_t)ovf_any, bigc, nbigc, wave,
                       nwaves, &lds[threadIdx.x >> 6], a, st.fcur);
  flush_level(st.cset, a);
}

// ----------------------------------------------------------------------------
// device-resident level loop
// Barrier number k (1, 2, ...) over all workgroups of a cooperative launch,
// in two levels so that no address takes more than ~gridDim/8 atomics: the
// workgroups with blockIdx % 8 == g count into group word g; the last of
// each group counts into the top word; the last group publishes k as the
// generation, which every workgroup polls.  Each wave waits for its own
// memory operations, then one wave per workgroup makes the workgroup's
// writes visible device-wide (release) and, after the barrier, drops stale
// cached lines for the workgroup (acquire).
__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned k) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned ng = gridDim.x < 8 ? gridDim.x : 8u;
    const unsigned g = blockIdx.x & 7u;
    const unsigned gsize = (gridDim.x - g + 7u) / 8u;
    unsigned* top = &bar[0];
    unsigned* gen = &bar[32];
    unsigned* grp = &bar[64 + 32 * g];
    const unsigned old = atomicAdd(grp, 1u);
    bool last = false;
    if (old + 1 == k * gsize) last = (atomicAdd(top, 1u) + 1 == k * ng);
    if (last) {
      __hip_atomic_store(gen, k, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < k)
        __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();
  }
  __syncthreads();
}

__device__ __forceinline__ void init_state(const PersistParams& p, int64_t gtid,
                                           int64_t gstride) {
  const GraphView& g = p.g;
  const int32_t n = g.n;
  const int64_t nwords = ((int64_t)n + 63) >> 6;
  const int r = p.nroots > 0 ? -1 : p.root;  // multi-root: seeded afterwards
  for (int64_t i = gtid; i < n; i += gstride) {
    p.S[i] = (i == r) ? 0 : -1;
    p.sigma[i] = (i == r) ? 1ull : 0ull;
    p.ovf[i] = 0;
  }
  for (int64_t w = gtid; w < nwords; w += gstride) {
    p.fb[0][w] = (w == (r >> 6)) ? (1ull << (r & 63)) : 0ull;
    p.fb[1][w] = 0ull;
    p.fb[2][w] = 0ull;
    p.vis[w] = 0ull;
  }
  if (gtid < 3 * kStripes) {
    CtrStripe* c = &p.cset[gtid / kStripes].s[gtid % kStripes];
    const bool root_slot = gtid == 0 && r >= 0;
    c->qm = root_slot ? ((1ull << 32) + outdeg(g, r)) : 0ull;
    c->min = root_slot ? indeg(g, r) : 0ull;
    c->ovf = 0;
  }
}

__device__ __forceinline__ LoopState first_state(const PersistParams& p) {
  LoopState s{};
  s.d = 0;
  s.pushing = (p.mode == TBFS_MODE_AUTO || p.mode == TBFS_MODE_PUSH) ? 1 : 0;
  s.done = 0;
  s.ovf_any = 0;
  s.mu = (unsigned long long)p.g.nz;
  return s;
}

// Fold the counters of the current frontier (depth s.d) into the loop state,
// choose this level's direction, and report whether the BFS goes on.  Every
// workgroup evaluates it on the same inputs and gets the same answer.
__device__ __forceinline__ bool step_state(const PersistParams& p,
                                           const LevelCtr* c, LoopState& s) {
  const unsigned long long q = c->q;
  if (q == 0) {
    s.done = 1;
    return false;
  }
  const unsigned long long mf = c->mout;
  const unsigned long long mi = c->min;
  s.ovf_any |= c->ovf;
  s.reached += q;
  s.edges += mi;
  s.mu = s.mu > mi ? s.mu - mi : 0ull;
  if (p.mode == TBFS_MODE_AUTO) {
    if (s.pushing && (double)mf * p.alpha > (double)s.mu) s.pushing = 0;
    else if (!s.pushing && (double)q * p.beta < (double)p.g.n) s.pushing = 1;
  }
  if (s.pushing) ++s.push_levels;
  else ++s.pull_levels;
  return true;
}

__device__ __forceinline__ void write_stats(const PersistParams& p,
                                            const LoopState& s, int32_t step) {
  p.stats->levels = s.d;
  p.stats->launches = step + 1;
  p.stats->push_levels = s.push_levels;
  p.stats->pull_levels = s.pull_levels;
  p.stats->ovf_any = s.ovf_any;
  p.stats->reached = s.reached;
  p.stats->edges = s.edges;
}

// Housekeeping for depth d+2 plus the level's work (frontier depth d).
// fb_lds: when not null, pull levels test parents against a copy of the
// frontier bitmap staged in this workgroup's LDS (random bitmap lookups stay
// out of the L2); it must hold ceil(n/64) words.
__device__ __forceinline__ void run_level(const PersistParams& p, int d,
                                          bool pushing, uint32_t ovf_any,
                                          unsigned long long mu,
                                          int64_t gtid, int64_t gstride,
                                          WaveLds* L, LevelAcc& acc,
                                          uint64_t* fb_lds) {
  const GraphView& g = p.g;
  const int64_t nwords = ((int64_t)g.n + 63) >> 6;
  const int64_t wave = gtid >> 6, nwaves = gstride >> 6;
  // buffers of depth d+2 were last used by depth d-1: clear them now; and
  // fold the current frontier into the reached bitmap
  uint64_t* fclr = p.fb[(d + 2) % 3];
  const uint64_t* fcur = p.fb[d % 3];
  for (int64_t w = gtid; w < nwords; w += gstride) {
    fclr[w] = 0ull;
    const uint64_t f = fcur[w];
    if (f) p.vis[w] |= f;
  }
  if (gtid < kStripes) {
    CtrStripe* z = &p.cset[(d + 2) % 3].s[gtid];
    z->qm = 0;
    z->min = 0;
    z->ovf = 0;
  }
  StateView st{p.S, p.sigma, p.ovf, p.fb[d % 3], p.fb[(d + 1) % 3],
               nullptr, nullptr, nullptr, &p.cset[(d + 1) % 3], p.vis, p.skip};
  if (pushing) {
    push_lb_level(g, st, d + 1, ovf_any, p.bigo, p.nbigo, wave, nwaves, L, acc);
  } else {
    const uint64_t* fpar = st.fcur;
    // staging costs every workgroup a read of the whole bitmap: worth it
    // only when the unexplored in-lists (parent tests) outnumber the vertices
    if (fb_lds && mu > (unsigned long long)g.n) {
      const uint4* src = reinterpret_cast<const uint4*>(st.fcur);
      uint4* dst = reinterpret_cast<uint4*>(fb_lds);
      const int64_t n16 = nwords >> 1;
      for (int64_t i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
      if ((nwords & 1) && threadIdx.x == 0) fb_lds[nwords - 1] = st.fcur[nwords - 1];
      __syncthreads();
      fpar = fb_lds;
    }
    if (p.sigma_on)
      pull_lb_level<true>(g, st, d + 1, ovf_any, p.bigc, p.nbigc, wave, nwaves, L, acc, fpar);
    else
      pull_lb_level<false>(g, st, d + 1, ovf_any, p.bigc, p.nbigc, wave, nwaves, L, acc, fpar);
  }
  flush_level(st.cset, acc);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_persist(PersistParams p) {
  __shared__ WaveLds lds[BLOCK / 64];
  WaveLds* L = &lds[threadIdx.x >> 6];
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;

  init_state(p, gtid, gstride);
  unsigned bk = 1;
  grid_barrier(p.bar, bk);

  LoopState s = first_state(p);
  LevelAcc acc;
  for (;;) {
    const LevelCtr c = read_set(&p.cset[s.d % 3], true);
    if (!step_state(p, &c, s)) break;
    run_level(p, s.d, s.pushing != 0, s.ovf_any, s.mu, gtid, gstride, L, acc, nullptr);
    ++s.d;
    grid_barrier(p.bar, ++bk);
  }
  if (gtid == 0) write_stats(p, s, -1);
}

// Chained loop: state of depth 0 and the first loop-state slot.
__global__ __launch_bounds__(256) void k_chain_init(PersistParams p) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  init_state(p, gtid, gstride);
  if (gtid == 0) {
    p.loop[0] = first_state(p);
    p.stats->levels = -1;  // set by the level that finds the frontier empty
  }
}

// Chained loop, several roots: depth 0 is every distinct root (runs after
// k_chain_init, which seeds none when p.nroots > 0).
__global__ __launch_bounds__(256) void k_chain_seed(PersistParams p) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= p.nroots) return;
  const int r = p.roots[j];
  const uint64_t bit = 1ull << (r & 63);
  const unsigned long long old =
      atomicOr((unsigned long long*)&p.fb[0][r >> 6], (unsigned long long)bit);
  if (old & bit) return;  // repeated root
  p.S[r] = 0;
  p.sigma[r] = 1ull;
  CtrStripe* c = &p.cset[0].s[j & (kStripes - 1)];
  atomicAdd(&c->qm, (1ull << 32) + (unsigned long long)outdeg(p.g, r));
  atomicAdd(&c->min, (unsigned long long)indeg(p.g, r));
}

// Chained loop: level d.  Reads slot d & 1 and the counters of depth d (both
// final: written by earlier kernels), writes slot (d + 1) & 1.  FBL: pull
// levels stage the frontier bitmap in dynamic LDS (one workgroup per CU).
// `step` only selects the state slot (its parity equals the depth's): the
// depth itself comes from the state, so one captured chunk of level kernels
// serves every stretch of levels starting at an even depth.
template <int BLOCK, bool FBL>
__global__ __launch_bounds__(BLOCK) void k_chain_level(PersistParams p, int32_t step) {
  __shared__ WaveLds lds[BLOCK / 64];
  extern __shared__ uint64_t fb_dyn[];
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  // launch `step` runs depth `step` (one level per launch): its counters are
  // loaded together with the loop state instead of after it
  const StripeVals early = load_set(&p.cset[step % 3], false);
  LoopState s = p.loop[step & 1];
  if (s.done) {
    if (gtid == 0) p.loop[(step + 1) & 1] = s;
    return;
  }
  const int32_t d = s.d;
  const LevelCtr c = sum_set(d == step ? early : load_set(&p.cset[d % 3], false));
  if (!step_state(p, &c, s)) {
    if (gtid == 0) {
      p.loop[(step + 1) & 1] = s;
      write_stats(p, s, step);
    }
    return;
  }
  LoopState nx = s;
  nx.d = d + 1;
  if (gtid == 0) p.loop[(step + 1) & 1] = nx;
  LevelAcc acc;
  run_level(p, d, s.pushing != 0, s.ovf_any, s.mu, gtid, gstride, &lds[threadIdx.x >> 6], acc,
            FBL ? fb_dyn : nullptr);
}

}  // namespace

// ----------------------------------------------------------------------------
namespace kern {

hipError_t push_lb(hipStream_t s, int grid, const GraphView& g,
                   const StateView& st, int32_t d, const uint64_t* bigo,
                   int32_t nbigo, int ovf_any) {
  hipLaunchKernelGGL(k_push_lb, dim3(grid), dim3(kLbBlock), 0, s, g, st, d,
                     bigo, nbigo, ovf_any);
  return hipGetLastError();
}

hipError_t pull_lb(hipStream_t s, int grid, const GraphView& g,
                   const StateView& st, int32_t d, const uint64_t* bigc,
                   int32_t nbigc, int sigma, int ovf_any) {
  if (sigma)
    hipLaunchKernelGGL(k_pull_lb<true>, dim3(grid), dim3(kLbBlock), 0, s, g, st,
                       d, bigc, nbigc, ovf_any);
  else
    hipLaunchKernelGGL(k_pull_lb<false>, dim3(grid), dim3(kLbBlock), 0, s, g,
                       st, d, bigc, nbigc, ovf_any);
  return hipGetLastError();
}

hipError_t persist_grid(int device, int block, int* grid) {
  int per_cu = 0, cus = 0, coop = 0;
  const void* fn = block == 1024 ? reinterpret_cast<const void*>(k_persist<1024>)
                                 : reinterpret_cast<const void*>(k_persist<256>);
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device);
  if (e != hipSuccess) return e;
  if (!coop || per_cu < 1) return hipErrorNotSupported;
  *grid = per_cu * cus;
  return hipSuccess;
}

hipError_t persist(hipStream_t s, int grid, int block, const PersistParams& p) {
  PersistParams arg = p;
  void* args[] = {&arg};
  const void* fn = block == 1024 ? reinterpret_cast<const void*>(k_persist<1024>)
                                 : reinterpret_cast<const void*>(k_persist<256>);
  return hipLaunchCooperativeKernel(fn, dim3(grid), dim3(block), args, 0, s);
}

hipError_t chain_init(hipStream_t s, int grid, const PersistParams& p) {
  hipLaunchKernelGGL(k_chain_init, dim3(grid), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t chain_seed(hipStream_t s, const PersistParams& p) {
  hipLaunchKernelGGL(k_chain_seed, dim3((p.nroots + 255) / 256), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t chain_level(hipStream_t s, int grid, const PersistParams& p, int32_t d,
                       size_t fbl_bytes) {
  if (fbl_bytes)
    hipLaunchKernelGGL((k_chain_level<1024, true>), dim3(grid), dim3(1024), fbl_bytes, s, p, d);
  else
    hipLaunchKernelGGL((k_chain_level<kLbBlock, false>), dim3(grid), dim3(kLbBlock), 0, s, p, d);
  return hipGetLastError();
}

hipError_t chain_fbl_limit(size_t* bytes) {
  // LDS left for the staged bitmap next to the kernel's static LDS
  hipFuncAttributes fa;
  hipError_t e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(k_chain_level<1024, true>));
  if (e != hipSuccess) return e;
  const size_t lds_cu = 160 * 1024;
  size_t avail = lds_cu > fa.sharedSizeBytes ? lds_cu - fa.sharedSizeBytes : 0;
  avail &= ~(size_t)255;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_chain_level<1024, true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)avail);
  if (e != hipSuccess) return e;
  *bytes = avail;
  return hipSuccess;
}

}  // namespace kern
}  // namespace tbfs
