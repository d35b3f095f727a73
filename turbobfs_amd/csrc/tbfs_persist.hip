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
// in-list entries per lane per light-column pass.  A group's last pass is
// partly empty and its idle slots still issue their (clamped) loads, so
// narrow passes waste fewer requests: on R-MAT 20, 8 / 6 / 4 / 3 / 2 / 1
// entries gave ~99 / 101 / 103 / 104.3 / 104.2 / 99.9 GTEPS.
constexpr int kPullRegs = 2;
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
// Also counted: the vertices with lists cut into static slices (in- and
// out-degree above kBigIn), so that levels can skip the slice tables when no
// such vertex is open (pull) or in the frontier (push).
struct LevelAcc {
  uint64_t q = 0, mout = 0, min = 0;
  uint32_t ovf = 0, hin = 0, hout = 0;

  __device__ __forceinline__ void found(const GraphView& g, int v) {
    found_in(g, v, indeg(g, v));
  }
  // v's in-degree already known: on a symmetric graph (out-lists alias the
  // in-lists) nothing more is loaded
  __device__ __forceinline__ void found_in(const GraphView& g, int v, uint32_t di) {
    const uint32_t dout = (g.RP == g.CP) ? di : outdeg(g, v);
    q += 1;
    min += di;
    mout += dout;
    hin += di > (uint32_t)kBigIn ? 1u : 0u;
    hout += dout > (uint32_t)kBigIn ? 1u : 0u;
  }
};

// The hub counts of one vertex v in CtrStripe::hubs form.
__device__ __forceinline__ unsigned long long hub_bits(const GraphView& g, int v) {
  return ((unsigned long long)(outdeg(g, v) > (uint32_t)kBigIn ? 1u : 0u) << 32) |
         (indeg(g, v) > (uint32_t)kBigIn ? 1u : 0u);
}

// Sum the tallies of the workgroup and add them to its stripe of the level
// counters (two 64-bit atomics).  Every thread of the workgroup must call it.
__device__ __forceinline__ void flush_level(CtrSet* set, LevelAcc& a) {
  __shared__ unsigned long long red[kMaxBlock / 64][4];
  __shared__ unsigned int redo[kMaxBlock / 64];
  const uint64_t q = wave_sum_u64(a.q);
  const uint64_t mo = wave_sum_u64(a.mout);
  const uint64_t mi = wave_sum_u64(a.min);
  const uint64_t hb = wave_sum_u64(((uint64_t)a.hout << 32) | a.hin);
  const bool ov = __ballot(a.ovf != 0) != 0;
  const int wv = threadIdx.x >> 6;
  if (lane() == 0) {
    red[wv][0] = q;
    red[wv][1] = mo;
    red[wv][2] = mi;
    red[wv][3] = hb;
    redo[wv] = ov ? 1u : 0u;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    unsigned int to = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      t0 += red[i][0];
      t1 += red[i][1];
      t2 += red[i][2];
      t3 += red[i][3];
      to |= redo[i];
    }
    CtrStripe* st = &set->s[blockIdx.x & (kStripes - 1)];
    if (t0) {
      atomicAdd(&st->qm, (t0 << 32) + t1);
      atomicAdd(&st->min, t2);
    }
    if (t3) atomicAdd(&st->hubs, t3);
    if (to) atomicOr(&st->ovf, 1u);
  }
  __syncthreads();
  a = LevelAcc();
}

// This lane's stripe of a striped counter set (lanes 0..kStripes-1 load one
// stripe each).  `agent`: the loads bypass caches that may hold values from
// before the last grid barrier.
struct StripeVals {
  uint64_t qm = 0, mi = 0, hb = 0;
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
      r.hb = __hip_atomic_load(&st->hubs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      r.ov = __hip_atomic_load(&st->ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      r.qm = st->qm;
      r.mi = st->min;
      r.hb = st->hubs;
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
  const uint64_t hb = wave_sum_u64(r.hb);
  c.hin = (uint32_t)(hb & 0xffffffffull);
  c.hout = (uint32_t)(hb >> 32);
  c.ovf = __ballot(r.ov != 0) ? 1u : 0u;
  return c;
}

__device__ __forceinline__ LevelCtr read_set(const CtrSet* set, bool agent) {
  return sum_set(load_set(set, agent));
}

// R edges u -> v[t] of a push level (v[t] < 0: no edge), u at depth d with
// count x[t] (overflow ou[t]).  "Reached at depth <= d" is read from the
// reached bitmap or the current frontier (both small and cached); otherwise
// v is reached at depth d+1 and the count is added at once.  A vertex not
// reached before this level holds count 0, so only an adder that finds 0 can
// be the first: those (normally one per vertex; more only when a count is 0
// mod 2^64) claim v's bit in the next frontier, and the one that sets it
// records the depth.  Device-scope atomics are performed at memory, a round
// trip of a microsecond or more, so the R edges go through each step
// together: all bitmap loads, then all count additions (with the targets'
// in-degrees loaded alongside), then all claims -- three round trips for the
// lane's R edges instead of four per edge.  Lanes without an edge read word 0
// of the bitmaps and vertex 0's degree (no load behind a branch) and are
// masked out.
template <int R>
__device__ __forceinline__ void push_edges(const GraphView& g, const StateView& st, int dn,
                                           const int (&v)[R], const uint64_t (&x)[R],
                                           const bool (&ou)[R], LevelAcc& a) {
  uint64_t bw[R];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int w = v[t] < 0 ? 0 : (v[t] >> 6);
    bw[t] = st.vis[w] | st.fcur[w];
  }
  bool live[R];
#pragma unroll
  for (int t = 0; t < R; ++t) live[t] = v[t] >= 0 && !((bw[t] >> (v[t] & 63)) & 1ull);
  unsigned long long old[R];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    old[t] = 1;
    if (live[t])
      old[t] = atomicAdd((unsigned long long*)&st.sigma[v[t]], (unsigned long long)x[t]);
  }
  // in-degrees of every target still open, in flight with the additions
  // (only the claimed ones are used): one round trip less per level
  int32_t c0[R], c1[R];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int vv = live[t] ? v[t] : 0;
    c0[t] = g.CP[vv];
    c1[t] = g.CP[vv + 1];
  }
#pragma unroll
  for (int t = 0; t < R; ++t) {
    if (live[t] && (old[t] + x[t] < old[t] || ou[t])) {
      st.ovf[v[t]] = 1;
      a.ovf = 1;
    }
  }
  bool first[R];
#pragma unroll
  for (int t = 0; t < R; ++t) first[t] = live[t] && old[t] == 0;
  unsigned long long ob[R];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    ob[t] = ~0ull;
    if (first[t])
      ob[t] = atomicOr((unsigned long long*)&st.fnext[v[t] >> 6], 1ull << (v[t] & 63));
  }
#pragma unroll
  for (int t = 0; t < R; ++t) {
    if (!((ob[t] >> (v[t] & 63)) & 1ull)) {
      st.S[v[t]] = dn;
      a.found_in(g, v[t], (uint32_t)(c1[t] - c0[t]));
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

// fclr (device-resident loops; null for the host-driven kernels, whose
// prologue kernel does it): the level's housekeeping rides on the frontier
// word loads -- each word of the frontier bitmap due for clearing is zeroed
// and the current frontier is ORed into the reached bitmap as its words are
// read.  Every word belongs to exactly one wave's share, and readers of the
// reached bitmap in this level test it together with the current frontier,
// so the order of the OR does not matter.
__device__ __forceinline__ void push_lb_level(const GraphView& g,
                                              const StateView& st, int dn,
                                              uint32_t ovf_any,
                                              const uint64_t* bigo,
                                              int32_t nbigo, int64_t wave,
                                              int64_t nwaves, WaveLds* L,
                                              LevelAcc& a, uint64_t* fclr) {
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
      uint64_t xs[kSliceRegs];
      bool os[kSliceRegs];
#pragma unroll
      for (int t = 0; t < kSliceRegs; ++t) {
        const int k = scb + lane() + 64 * t;
        v[t] = k < sce ? g.JR[k] : -1;
        xs[t] = sx;
        os[t] = sou;
      }
      push_edges<kSliceRegs>(g, st, dn, v, xs, os, a);
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
    const bool mine = b + lane() < nm;
    const uint64_t bits_l = mine ? st.fcur[wl] : 0ull;
    if (fclr && mine) {
      const uint64_t vw = st.vis[wl];
      fclr[wl] = 0ull;
      if (bits_l) st.vis[wl] = vw | bits_l;
    }
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
      if (u >= 0) {  // bounds and count of u loaded together
        beg = g.RP[u];
        const int end = g.RP[u + 1];
        const uint64_t xu = st.sigma[u];
        const uint32_t ouu = ovf_any ? st.ovf[u] : 0u;
        if (end - beg <= kBigIn) {  // longer lists go by their static slices
          deg = end - beg;
          x = xu;
          ou = ouu;
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
        int v[kUnroll];
        uint64_t xs[kUnroll];
        bool os[kUnroll];
        // lanes past the end take the last edge (total > 0) and are masked
        // out; the owner searches advance together, no load behind a branch
        int e[kUnroll], o[kUnroll], k[kUnroll];
#pragma unroll
        for (int t = 0; t < kUnroll; ++t) e[t] = min(e0 + 64 * t, total - 1);
        owners_of<kUnroll>(L->excl, e, o);
#pragma unroll
        for (int t = 0; t < kUnroll; ++t) k[t] = L->beg[o[t]] + (e[t] - L->excl[o[t]]);
#pragma unroll
        for (int t = 0; t < kUnroll; ++t) {
          const int vk = g.JR[k[t]];
          v[t] = e0 + 64 * t < total ? vk : -1;
          xs[t] = L->x[o[t]];
          os[t] = (oumask >> o[t]) & 1ull;
        }
        push_edges<kUnroll>(g, st, dn, v, xs, os, a);
      }
    }
  }
}

// Parent test on a frontier bitmap read as 32-bit words (one bit-field
// extract instead of a 64-bit shift).
__device__ __forceinline__ bool fbit32(const uint64_t* fb, int u) {
  return __builtin_amdgcn_ubfe(reinterpret_cast<const uint32_t*>(fb)[u >> 5], (uint32_t)u, 1u) != 0;
}

// A pull group's columns with in-list entries are numbered by rank (lane
// order); the group's in-lists are concatenated in that order, column j's
// entries starting at xj (strictly increasing over the ranks).  The owner of
// group entry e is the last column whose first entry is at most e.
//
// Owner ranks of the entries B + 64 t + lane (t < R, B a multiple of 64):
// for each window of 64 entries, every column starting inside it (at window
// offset 1..63) sends a flag to the lane of its start (ds_permute moves
// register values between lanes without touching LDS memory; lanes nobody
// sends to receive 0).  A lane's owner is then the number of columns starting
// at or before the window's first entry, plus the flags up to its own lane,
// minus one -- no search.
template <int R>
__device__ __forceinline__ void group_owners(int B, int xj, bool hasd, int (&r)[R]) {
  int recv[R];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int p = xj - (B + 64 * t);
    const bool in = hasd && (unsigned)(p - 1) < 63u;
    recv[t] = __builtin_amdgcn_ds_permute(in ? p << 2 : 0, in ? 1 : 0);
  }
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int p = xj - (B + 64 * t);
    const uint64_t m = __ballot(recv[t] != 0) >> 1;  // bit i - 1: a column starts at offset i
    const uint32_t c0 = (uint32_t)__popcll(__ballot(hasd && p <= 0));
    r[t] = (int)(__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                           __builtin_amdgcn_mbcnt_lo((uint32_t)m, c0 - 1u)) &
                 63u);
  }
}

// First half of a pass: owner ranks and in-list entries of the group entries
// B + 64 t + lane.  Entries past the group's end get an in-list index clamped
// into [0, kmax] (kmax: the group's last in-list entry) and are masked out by
// the second half; nothing here sits behind a branch.
template <int R>
__device__ __forceinline__ void group_fetch(const GraphView& g, const WaveLds* L, int B, int xj,
                                            bool hasd, int kmax, int (&r)[R], int (&u)[R]) {
  group_owners<R>(B, xj, hasd, r);
  int k[R];
#pragma unroll
  for (int t = 0; t < R; ++t)
    k[t] = min(max(L->beg[r[t]] + (B + 64 * t + lane() - L->excl[r[t]]), 0), kmax);
#pragma unroll
  for (int t = 0; t < R; ++t) u[t] = g.IC[k[t]];
}

// Passes of R group entries per lane: parent bits, then parents' counts
// (non-parents read vertex 0's, one shared line, so no load sits behind a
// branch), added per column into LDS word x[rank] as (low 32 bits + 2^40):
// a light column has at most kBigIn <= 256 entries, so bits 0..39 hold the
// exact sum of the low halves and bits 40.. the number of parents (found iff
// nonzero, whatever the count).  High halves go to xh[rank] (bits 0..39),
// an overflow flag of a parent to its bit 63.  PIPE: two passes in flight --
// the in-list entries of pass k + 1 are loaded (unconditionally, clamped)
// while the counts of pass k are on their way.  Without PIPE the group must
// fit one pass (total <= 64 R).
template <int R, bool SIGMA, bool PIPE>
__device__ __forceinline__ void pull_passes(const GraphView& g, const StateView& st,
                                            WaveLds* L, const uint64_t* fpar, uint32_t ovf_any,
                                            int total, int xj, bool hasd, int kmax) {
  static_assert(kBigIn <= 256, "x[rank] keeps the parent count above bit 40");
  int r[R], u[R];
  group_fetch<R>(g, L, 0, xj, hasd, kmax, r, u);
  for (int B = 0;; B += 64 * R) {
    bool par[R];
#pragma unroll
    for (int t = 0; t < R; ++t) par[t] = (int)(B + 64 * t + lane() < total) & (int)fbit32(fpar, u[t]);
    uint64_t xs[R];
    if (SIGMA) {
#pragma unroll
      for (int t = 0; t < R; ++t) xs[t] = st.sigma[par[t] ? u[t] : 0];
      if (ovf_any) {
#pragma unroll
        for (int t = 0; t < R; ++t)
          if (par[t] && st.ovf[u[t]])
            atomicOr((unsigned long long*)&L->xh[r[t]], 1ull << 63);
      }
    }
    const bool more = PIPE && B + 64 * R < total;  // wave-uniform
    int rn[R], un[R];
    if (PIPE) group_fetch<R>(g, L, B + 64 * R, xj, hasd, kmax, rn, un);
#pragma unroll
    for (int t = 0; t < R; ++t) {
      if (par[t]) {
        const uint64_t lo = SIGMA ? (xs[t] & 0xffffffffull) : 0ull;
        atomicAdd((unsigned long long*)&L->x[r[t]], (unsigned long long)(lo + (1ull << 40)));
        if (SIGMA) {
          const unsigned long long h = xs[t] >> 32;
          if (h) atomicAdd((unsigned long long*)&L->xh[r[t]], h);
        }
      }
    }
    if (!more) break;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      r[t] = rn[t];
      u[t] = un[t];
    }
  }
}

// One group of up to 64 columns of a pull level, column v with in-list
// [beg, beg + deg) on its lane (deg 0: none; deg <= kBigIn): the in-lists are
// walked in passes of kPullRegs entries per lane (parents' counts summed per
// column in LDS), and the columns that found a parent get their depth, count
// and overflow flag and are tallied.  Returns the wavefront's mask of those
// columns (lane bits).  All lanes must take part.
template <bool SIGMA>
__device__ __forceinline__ uint64_t pull_group(const GraphView& g, const StateView& st,
                                               WaveLds* L, const uint64_t* fpar,
                                               uint32_t ovf_any, int dn, int v, int beg,
                                               int deg, LevelAcc& a) {
  const uint32_t incl = wave_incl_scan((uint32_t)deg);
  const int total = (int)__shfl(incl, 63);
  if (total == 0) return 0;
  const bool hasd = deg > 0;
  const int xj = (int)incl - deg;
  const uint64_t hm = __ballot(hasd);
  const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(hm >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)hm, 0u));
  const int kmax = __shfl(beg + deg - 1, 63 - __clzll((long long)hm));
  wave_lds_sync();
  if (hasd) {
    L->excl[rank] = xj;
    L->beg[rank] = beg;
  }
  L->x[lane()] = 0;
  if (SIGMA) L->xh[lane()] = 0;
  wave_lds_sync();
  // one pass for short groups, else pipelined passes of kPullRegs entries
  if (total <= 64 * kPullRegs) {
    pull_passes<kPullRegs, SIGMA, false>(g, st, L, fpar, ovf_any, total, xj, hasd, kmax);
  } else {
    pull_passes<kPullRegs, SIGMA, true>(g, st, L, fpar, ovf_any, total, xj, hasd, kmax);
  }
  wave_lds_sync();
  const int rk = hasd ? rank : 0;
  const uint64_t xr = L->x[rk];
  const bool f = hasd && (xr >> 40) != 0;
  const uint64_t fmask = __ballot(f);
  if (fmask == 0) return 0;
  if (f) {
    st.S[v] = dn;
    if (SIGMA) {
      // count = xh * 2^32 + (low sums); at least 2^64 iff xh >= 2^32 or
      // the final addition carries
      const uint64_t xhr = L->xh[rk];
      const uint64_t xh = xhr & ~(1ull << 63);
      const uint64_t hs = xh << 32;
      const uint64_t sum = hs + (xr & ((1ull << 40) - 1ull));
      st.sigma[v] = sum;
      if ((xh >> 32) != 0 || sum < hs || (xhr >> 63) != 0) {
        st.ovf[v] = 1;
        a.ovf = 1;
      }
    }
    a.found_in(g, v, (uint32_t)deg);
  }
  return fmask;
}

// The next kSlicesPerPass active slices of a pull level's slice batch (lane
// bits of todo, taken lowest first; js[h] = -1 past the last): their lane
// and entry range [scb, sce).  A slot without a slice repeats the first
// slot's start with an empty range.
__device__ __forceinline__ void pick_slices(uint64_t& todo, int cb, int ce,
                                            int (&js)[kSlicesPerPass],
                                            int (&scb)[kSlicesPerPass],
                                            int (&sce)[kSlicesPerPass]) {
#pragma unroll
  for (int h = 0; h < kSlicesPerPass; ++h) {
    js[h] = todo ? __ffsll((unsigned long long)todo) - 1 : -1;
    todo &= todo - 1;
    const int jj = js[h] >= 0 ? js[h] : (js[0] >= 0 ? js[0] : 0);
    scb[h] = __shfl(cb, jj);
    sce[h] = js[h] >= 0 ? __shfl(ce, jj) : scb[h];
  }
}

// In-list entries of the picked slices, kSliceRegs per lane of each; entries
// past a slice's end read the first slice's first entry (masked later).
__device__ __forceinline__ void load_slices(const GraphView& g, const int (&scb)[kSlicesPerPass],
                                            const int (&sce)[kSlicesPerPass],
                                            int (&u)[kSlicesPerPass * kSliceRegs]) {
#pragma unroll
  for (int h = 0; h < kSlicesPerPass; ++h) {
#pragma unroll
    for (int t = 0; t < kSliceRegs; ++t) {
      const int k = scb[h] + lane() + 64 * t;
      u[h * kSliceRegs + t] = g.IC[k < sce[h] ? k : scb[0]];
    }
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
                                              const uint64_t* fpar, uint64_t* fclr) {
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
    // each); a slot without a slice repeats the first one's start, empty.
    // The next pass's in-list entries are loaded as soon as this pass's
    // count gathers are issued, so the two latencies overlap.
    constexpr int R = kSlicesPerPass * kSliceRegs;
    int js[kSlicesPerPass], scb[kSlicesPerPass], sce[kSlicesPerPass];
    int u[R];
    pick_slices(todo, cb, ce, js, scb, sce);
    if (js[0] >= 0) load_slices(g, scb, sce, u);
    while (js[0] >= 0) {
      bool par[R];
#pragma unroll
      for (int h = 0; h < kSlicesPerPass; ++h)
#pragma unroll
        for (int t = 0; t < kSliceRegs; ++t)
          par[h * kSliceRegs + t] = scb[h] + lane() + 64 * t < sce[h];
#pragma unroll
      for (int t = 0; t < R; ++t) par[t] = (int)par[t] & (int)fbit32(fpar, u[t]);  // u[t] valid: no branch
      uint64_t xs[R];
      if (SIGMA) {
#pragma unroll
        for (int t = 0; t < R; ++t) xs[t] = st.sigma[par[t] ? u[t] : 0];  // no load behind a branch
      }
      int jn[kSlicesPerPass], scbn[kSlicesPerPass], scen[kSlicesPerPass];
      int un[R];
      pick_slices(todo, cb, ce, jn, scbn, scen);
      if (jn[0] >= 0) load_slices(g, scbn, scen, un);
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
#pragma unroll
        for (int t = 0; t < R; ++t) {
          const int h = t / kSliceRegs;
          f[h] |= par[t];
          add_carry(sum[h], ov[h], par[t] ? xs[t] : 0ull);
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
#pragma unroll
      for (int h = 0; h < kSlicesPerPass; ++h) {
        js[h] = jn[h];
        scb[h] = scbn[h];
        sce[h] = scen[h];
      }
#pragma unroll
      for (int t = 0; t < R; ++t) u[t] = un[t];
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
    // light column: nothing to do for that vertex.  The housekeeping of
    // these words rides along (as in push_lb_level).
    uint64_t done_l = ~0ull;
    if (b + lane() < nmine) {
      const uint64_t vw = st.vis[wl], fw = st.fcur[wl];
      done_l = vw | st.skip[wl] | fw;
      if (fclr) {
        fclr[wl] = 0ull;
        if (fw) st.vis[wl] = vw | fw;
      }
    }
    uint64_t todo = __ballot(done_l != ~0ull);
    if (!todo) continue;
    const uint64_t open_l = ~done_l;
    const uint32_t cnt = (uint32_t)__popcll(open_l);
    const uint32_t wincl = wave_incl_scan(cnt);
    const int ncol = (int)__shfl(wincl, 63);
    if (2 * ncol < 64 * __popcll(todo)) {
      // Sparse words (under half their columns open on average): the open
      // columns of the 64 words are numbered across the words (prefix sum
      // of popcounts) and taken 64 at a time, so a few open columns per word
      // cost one pass for many words instead of one per word.  Columns found
      // are gathered per word in LDS: one next-frontier OR per word.
      uint64_t wmask = 0;  // this lane's word: columns found
      for (int k0 = 0; k0 < ncol; k0 += 64) {
        wave_lds_sync();
        L->excl[lane()] = (int)(wincl - cnt);
        L->x[lane()] = open_l;
        wave_lds_sync();
        int v = -1, beg = 0, deg = 0, j = 0;
        const int k = k0 + lane();
        if (k < ncol) {
          j = owner_of(L->excl, k);
          const int bit = select_bit(L->x[j], k - L->excl[j]);
          v = (int)((wave + (b + j) * nwaves) * 64) + bit;
          beg = g.CP[v];
          deg = g.CP[v + 1] - beg;
        }
        const uint64_t fm = pull_group<SIGMA>(g, st, L, fpar, ovf_any, dn, v, beg, deg, a);
        if (fm) {
          wave_lds_sync();
          L->x[lane()] = 0;
          wave_lds_sync();
          if ((fm >> lane()) & 1ull)
            atomicOr((unsigned long long*)&L->x[j], (unsigned long long)(1ull << (v & 63)));
          wave_lds_sync();
          wmask |= L->x[lane()];
        }
      }
      if (wmask) atomicOr((unsigned long long*)&st.fnext[wl], (unsigned long long)wmask);
      continue;
    }
    // column bounds of the next word with work are loaded one word ahead:
    // their latency hides behind the current word's passes
    int nbeg = 0, nend = 0;
    {
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
      const uint64_t fm = pull_group<SIGMA>(g, st, L, fpar, ovf_any, dn, v, beg, deg, a);
      if (fm && lane() == 0) atomicOr((unsigned long long*)&st.fnext[w], (unsigned long long)fm);
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
                &lds[threadIdx.x >> 6], a, nullptr);
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
                       nwaves, &lds[threadIdx.x >> 6], a, st.fcur, nullptr);
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
      // relaxed polls: an acquire load invalidates the caches on every
      // iteration, under the workgroups still running the level; the fence
      // below is the one acquire
      while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k)
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
    c->hubs = root_slot ? hub_bits(g, r) : 0ull;
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
  s.heavy_reached += c->hin;
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
  p.stats->stage_mask = s.stage_mask;
}

// A pull level stages the frontier bitmap (when the launch has the LDS for
// it) if the unexplored in-list entries, one parent test each, outnumber the
// vertices times stage_ratio (1 by default): every workgroup then reads the
// whole bitmap once.
__device__ __forceinline__ bool wants_stage(const PersistParams& p, bool pushing,
                                            unsigned long long mu) {
  return !pushing && (double)mu > (double)p.stage_ratio * (double)p.g.n;
}

// Housekeeping for depth d+2 plus the level's work (frontier depth d).
// fb_lds: when not null, pull levels test parents against a copy of the
// frontier bitmap staged in this workgroup's LDS (random bitmap lookups stay
// out of the L2); it must hold ceil(n/64) words.
// frontier_hubs: some frontier vertex has out-degree > kBigIn (else a push
// skips the out-list slices); heavy_open: some vertex with in-degree > kBigIn
// is not reached yet (else a pull skips the in-list slices).
__device__ __forceinline__ void run_level(const PersistParams& p, int d,
                                          bool pushing, uint32_t ovf_any,
                                          unsigned long long mu, bool frontier_hubs,
                                          bool heavy_open,
                                          int64_t gtid, int64_t gstride,
                                          WaveLds* L, LevelAcc& acc,
                                          uint64_t* fb_lds) {
  const GraphView& g = p.g;
  const int64_t nwords = ((int64_t)g.n + 63) >> 6;
  const int64_t wave = gtid >> 6, nwaves = gstride >> 6;
  // buffers of depth d+2 were last used by depth d-1: clear them now; and
  // fold the current frontier into the reached bitmap
  // (done by the level's own word loop, which reads these words anyway)
  uint64_t* fclr = p.fb[(d + 2) % 3];
  if (gtid < kStripes) {
    CtrStripe* z = &p.cset[(d + 2) % 3].s[gtid];
    z->qm = 0;
    z->min = 0;
    z->hubs = 0;
    z->ovf = 0;
  }
  StateView st{p.S, p.sigma, p.ovf, p.fb[d % 3], p.fb[(d + 1) % 3],
               nullptr, nullptr, nullptr, &p.cset[(d + 1) % 3], p.vis, p.skip};
  if (pushing) {
    push_lb_level(g, st, d + 1, ovf_any, p.bigo, frontier_hubs ? p.nbigo : 0, wave, nwaves,
                  L, acc, fclr);
  } else {
    // A staging launch (fb_lds set) stages every pull level it runs: the
    // bitmap pointer is then the LDS array on every path, so the parent tests
    // compile to LDS reads (a pointer that may be either needs flat loads,
    // which wait for every outstanding memory operation).  The launch plan
    // stages only levels predicted heavy; a light one pays the copy (~3 us).
    const int32_t nbigc = heavy_open ? p.nbigc : 0;
    if (fb_lds) {
      const uint4* src = reinterpret_cast<const uint4*>(st.fcur);
      uint4* dst = reinterpret_cast<uint4*>(fb_lds);
      const int64_t n16 = nwords >> 1;
      for (int64_t i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
      if ((nwords & 1) && threadIdx.x == 0) fb_lds[nwords - 1] = st.fcur[nwords - 1];
      __syncthreads();
      if (p.sigma_on)
        pull_lb_level<true>(g, st, d + 1, ovf_any, p.bigc, nbigc, wave, nwaves, L, acc, fb_lds,
                            fclr);
      else
        pull_lb_level<false>(g, st, d + 1, ovf_any, p.bigc, nbigc, wave, nwaves, L, acc, fb_lds,
                             fclr);
    } else if (p.sigma_on) {
      pull_lb_level<true>(g, st, d + 1, ovf_any, p.bigc, nbigc, wave, nwaves, L, acc, st.fcur,
                          fclr);
    } else {
      pull_lb_level<false>(g, st, d + 1, ovf_any, p.bigc, nbigc, wave, nwaves, L, acc, st.fcur,
                           fclr);
    }
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
    run_level(p, s.d, s.pushing != 0, s.ovf_any, s.mu, c.hout != 0,
              s.heavy_reached < (unsigned long long)p.nheavy, gtid, gstride, L, acc, nullptr);
    ++s.d;
    grid_barrier(p.bar, ++bk);
  }
  if (gtid == 0) write_stats(p, s, -1);
}

// Whether out-list entry t (of the root's, in nb[0 .. kInitPush), -1 past its
// end) is a vertex other than the root that no earlier entry repeats.
__device__ __forceinline__ bool v_first(const int32_t* nb, int t, int r) {
  const int v = nb[t];
  if (v < 0 || v == r) return false;
  for (int j = 0; j < t; ++j)
    if (nb[j] == v) return false;
  return true;
}

// The state after depth 0 of a single-root BFS (the root's own counters
// folded in); reports whether the init kernel runs depth 0 itself: a push
// (as the first level launch would choose), from a root with at most
// kInitPush out-list entries, with no finisher taking over.  Every thread
// computes it from the same inputs and gets the same answer.
__device__ __forceinline__ bool init_runs_depth0(const PersistParams& p, LoopState& s) {
  s = first_state(p);
  if (p.step0 != 1 || p.root < 0 || p.nroots > 0) return false;
  const GraphView& g = p.g;
  const int r = p.root;
  LevelCtr c{};
  c.q = 1;
  c.mout = outdeg(g, r);
  c.min = indeg(g, r);
  const unsigned long long hb = hub_bits(g, r);
  c.hin = (uint32_t)(hb & 0xffffffffull);
  c.hout = (uint32_t)(hb >> 32);
  c.ovf = 0;
  step_state(p, &c, s);
  const bool fin = p.fin_mu && s.mu <= (unsigned long long)p.fin_mu && p.mode == TBFS_MODE_AUTO;
  return s.pushing && !fin && c.mout <= (unsigned long long)kInitPush;
}

// Chained loop: state of depth 0 and the first loop-state slot (p.step0 & 1).
// When init_runs_depth0 holds, the kernel also runs depth 0: every
// workgroup loads the root's out-list (one entry per thread) and, for the
// vertices it initialises, counts their edges from the root in LDS, so each
// vertex is written once, already holding its depth-1 values (S = 1, count =
// number of edges from the root); the depth-1 frontier words come from the
// waves' ballots, and the root joins the reached bitmap.  Workgroup 0 tallies
// the distinct depth-1 vertices into depth 1's counters and writes loop slot
// 1 with depth 1 next.
__global__ __launch_bounds__(256) void k_chain_init(PersistParams p) {
  __shared__ int32_t nb[kInitPush];
  __shared__ uint32_t cnt[256];
  __shared__ unsigned long long red[4][4];
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  LoopState s1;
  const bool fuse = init_runs_depth0(p, s1);
  if (!fuse) {
    init_state(p, gtid, gstride);
  } else {
    const GraphView& g = p.g;
    const int32_t n = g.n;
    const int r = p.root;
    const int rb = g.RP[r];
    const int deg = (int)outdeg(g, r);
    nb[threadIdx.x] = threadIdx.x < deg ? g.JR[rb + threadIdx.x] : -1;
    __syncthreads();
    // vertices: the 256-aligned runs blockIdx * 256 + k * gstride
    for (int64_t base = (int64_t)blockIdx.x * 256; base < n; base += gstride) {
      cnt[threadIdx.x] = 0;
      __syncthreads();
      const int v = nb[threadIdx.x];
      if (v >= 0 && v != r && v >= base && v < base + 256) atomicAdd(&cnt[v - base], 1u);
      __syncthreads();
      const int64_t i = base + threadIdx.x;
      const uint32_t k = cnt[threadIdx.x];
      if (i < n) {
        p.S[i] = i == r ? 0 : (k ? 1 : -1);
        p.sigma[i] = i == r ? 1ull : (unsigned long long)k;
        p.ovf[i] = 0;
      }
      const uint64_t word = __ballot(i < n && k != 0);
      if (lane() == 0 && i < n) p.fb[1][i >> 6] = word;
      __syncthreads();
    }
    const int64_t nwords = ((int64_t)n + 63) >> 6;
    for (int64_t w = gtid; w < nwords; w += gstride) {
      const uint64_t rw = (w == (r >> 6)) ? (1ull << (r & 63)) : 0ull;
      p.fb[0][w] = rw;
      p.fb[2][w] = 0ull;
      p.vis[w] = rw;  // depth 0 is reached
    }
    if (blockIdx.x == 0) {
      // counters: depth 1's in set 1, sets 0 and 2 cleared
      if (threadIdx.x < 3 * kStripes) {
        CtrStripe* c = &p.cset[threadIdx.x / kStripes].s[threadIdx.x % kStripes];
        c->qm = 0;
        c->min = 0;
        c->hubs = 0;
        c->ovf = 0;
      }
      // the distinct depth-1 vertices: first occurrence in the out-list
      bool first = v_first(nb, threadIdx.x, r);
      const int v = first ? nb[threadIdx.x] : 0;
      const uint32_t di = first ? indeg(g, v) : 0u;
      const uint32_t dout = first ? outdeg(g, v) : 0u;
      const uint64_t q = wave_sum_u64(first ? 1u : 0u);
      const uint64_t mo = wave_sum_u64(dout);
      const uint64_t mi = wave_sum_u64(di);
      const uint64_t hb = wave_sum_u64(
          ((uint64_t)(dout > (uint32_t)kBigIn ? 1u : 0u) << 32) | (di > (uint32_t)kBigIn ? 1u : 0u));
      if (lane() == 0) {
        red[threadIdx.x >> 6][0] = q;
        red[threadIdx.x >> 6][1] = mo;
        red[threadIdx.x >> 6][2] = mi;
        red[threadIdx.x >> 6][3] = hb;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        unsigned long long t0 = 0, t1 = 0, t2 = 0, t3 = 0;
        for (int j = 0; j < 4; ++j) {
          t0 += red[j][0];
          t1 += red[j][1];
          t2 += red[j][2];
          t3 += red[j][3];
        }
        CtrStripe* c = &p.cset[1].s[0];
        c->qm = (t0 << 32) + t1;
        c->min = t2;
        c->hubs = t3;
      }
    }
  }
  if (gtid == 0) {
    LoopState s = first_state(p);
    if (fuse) {
      s = s1;
      s.d = 1;
    }
    p.loop[p.step0 & 1] = s;
    p.stats->levels = -1;  // set by the level that finds the frontier empty
    p.stats->fin_level = -1;
    for (int i = 0; i < kWorkLevels; ++i) p.stats->work_lg[i] = 0;
    if (fuse) p.stats->work_lg[0] = (uint8_t)(64 - __clzll((long long)outdeg(p.g, p.root)));
    if (p.fin_ctr) {
      p.fin_ctr[0] = 0;
      p.fin_ctr[1] = 0;
    }
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
  const unsigned long long hb = hub_bits(p.g, r);
  if (hb) atomicAdd(&c->hubs, hb);
}

// ---- the finisher ------------------------------------------------------------
// Once few in-list entries are unexplored (mu <= fin_mu), the rest of the BFS
// costs more in level launches than in work: every workgroup of the level
// lists the unreached columns that have in-lists (grid-stride over the
// bitmap words, appended with one atomic per wavefront), and the last
// workgroup to finish its share runs every remaining level itself, as a pull
// over that list with workgroup barriers between levels.  A column's parents
// at depth k are the in-list entries u with S[u] == k; a column found at
// depth k + 1 gets S, sigma (the parents' counts added with carry into the
// overflow flag) and its flag, the others go on to the next level's list.
// The last workgroup is the only one left, so it needs no grid barrier, and
// it reads and writes S / sigma / ovf with plain stores (one CU: coherent at
// workgroup scope).
constexpr int kFinU = 8;       // columns per thread in flight
constexpr int kFinLight = 16;  // columns with more in-list entries: a wavefront each
constexpr int kFinHeavyCap = 1024;

template <int BLOCK>
struct FinLds {
  int32_t heavy[kFinHeavyCap];  // this pass's columns with long in-lists
  unsigned int nheavy;
  unsigned int nkeep;           // columns kept for the next level
  unsigned int found;           // columns found this level
  unsigned int ovf;
  unsigned long long edges;     // their in-list entries
  unsigned long long hin;       // how many of them have in-degree > kBigIn
  unsigned int last;            // this workgroup finishes the BFS
};

// List the unreached columns (not in vis or the current frontier) that have
// in-lists; skip bits mark in-degree 0 or > kBigIn, so they are checked
// only while some column with in-degree > kBigIn is unreached.
__device__ __forceinline__ void fin_list_columns(const PersistParams& p, int d, bool heavy_open,
                                                 int64_t gtid, int64_t gstride) {
  const GraphView& g = p.g;
  const int64_t nwords = ((int64_t)g.n + 63) >> 6;
  const uint64_t* fcur = p.fb[d % 3];
  uint32_t* out = p.fin_list[0];
  const int ln = lane();
  for (int64_t wb = gtid - ln; wb < nwords; wb += gstride) {
    const int64_t w = wb + ln;
    uint64_t live = 0;
    if (w < nwords) {
      uint64_t m = ~(p.vis[w] | fcur[w]);
      if (w == nwords - 1 && (g.n & 63)) m &= (1ull << (g.n & 63)) - 1ull;
      const uint64_t sk = p.skip[w];
      live = m & ~sk;
      if (heavy_open) {
        uint64_t c = m & sk;
        while (c) {
          const int b = __builtin_ctzll(c);
          c &= c - 1;
          if (indeg(g, (int)(w * 64 + b)) > 0) live |= 1ull << b;
        }
      }
    }
    const uint32_t cnt = (uint32_t)__popcll(live);
    const uint32_t incl = wave_incl_scan(cnt);
    const uint32_t total = __shfl(incl, 63);
    if (total == 0) continue;
    unsigned int base = 0;
    if (ln == 0) base = atomicAdd(&p.fin_ctr[0], total);
    base = __shfl(base, 0);
    uint32_t o = base + incl - cnt;
    while (live) {
      const int b = __builtin_ctzll(live);
      live &= live - 1;
      if (o < kFinCap) out[o] = (uint32_t)(w * 64 + b);
      ++o;
    }
  }
}

// Add x (with its overflow flag xo) into the 64-bit sum acc / flag ao.
__device__ __forceinline__ void fin_add(uint64_t& acc, uint32_t& ao, uint64_t x, uint32_t xo) {
  const uint64_t t = acc + x;
  ao |= xo | (t < acc ? 1u : 0u);
  acc = t;
}

template <int BLOCK>
__device__ __forceinline__ void fin_run(const PersistParams& p, LoopState& s, int32_t step, FinLds<BLOCK>& L) {
  const GraphView& g = p.g;
  const int t = threadIdx.x;
  const int ln = lane();
  const int wave = t >> 6;
  constexpr int kWaves = BLOCK / 64;
  uint32_t* A = p.fin_list[0];
  uint32_t* B = p.fin_list[1];
  uint32_t LA = min(p.fin_ctr[0], kFinCap);
  const int d0 = s.d;
  // level d0 expands as a pull here, whatever step_state chose for it
  if (s.pushing) {
    --s.push_levels;
    ++s.pull_levels;
  }
  for (int k = d0;; ++k) {
    if (t == 0) {
      L.nkeep = 0;
      L.found = 0;
      L.ovf = 0;
      L.edges = 0;
      L.hin = 0;
    }
    __syncthreads();
    unsigned int found = 0, ovfl = 0;
    unsigned long long edges = 0, hin = 0;
    for (uint32_t base = 0; base < LA; base += (uint32_t)BLOCK * kFinU) {
      if (t == 0) L.nheavy = 0;
      __syncthreads();
      int v[kFinU], b[kFinU], e[kFinU];
#pragma unroll
      for (int i = 0; i < kFinU; ++i) {
        const uint32_t j = base + (uint32_t)(i * BLOCK + t);
        v[i] = j < LA ? (int)A[j] : -1;
      }
#pragma unroll
      for (int i = 0; i < kFinU; ++i) {
        b[i] = 0;
        e[i] = 0;
        if (v[i] >= 0) {
          b[i] = g.CP[v[i]];
          e[i] = g.CP[v[i] + 1];
        }
      }
      // long in-lists: a wavefront each, after this pass
#pragma unroll
      for (int i = 0; i < kFinU; ++i) {
        if (v[i] >= 0 && e[i] - b[i] > kFinLight) {
          const unsigned int h = atomicAdd(&L.nheavy, 1u);
          if (h < (unsigned)kFinHeavyCap) {
            L.heavy[h] = v[i];
            v[i] = -2;  // settled by its wavefront
          }
        }
      }
      uint64_t acc[kFinU];
      uint32_t ao[kFinU], hit[kFinU];
#pragma unroll
      for (int i = 0; i < kFinU; ++i) {
        acc[i] = 0;
        ao[i] = 0;
        hit[i] = 0;
      }
      // entry r of every light column at once
      for (int r = 0;; ++r) {
        int u[kFinU];
        bool any = false;
#pragma unroll
        for (int i = 0; i < kFinU; ++i) {
          u[i] = -1;
          if (v[i] >= 0 && b[i] + r < e[i]) {
            u[i] = g.IC[b[i] + r];
            any = true;
          }
        }
        if (!__any(any)) break;  // wave-uniform exit
        int32_t su[kFinU];
#pragma unroll
        for (int i = 0; i < kFinU; ++i) su[i] = u[i] >= 0 ? p.S[u[i]] : -1;
#pragma unroll
        for (int i = 0; i < kFinU; ++i) {
          if (su[i] == k) {
            hit[i] = 1;
            if (p.sigma_on) fin_add(acc[i], ao[i], p.sigma[u[i]], p.ovf[u[i]]);
          }
        }
      }
#pragma unroll
      for (int i = 0; i < kFinU; ++i) {
        if (v[i] < 0) continue;
        if (hit[i]) {
          p.S[v[i]] = k + 1;
          if (p.sigma_on) {
            p.sigma[v[i]] = acc[i];
            p.ovf[v[i]] = (uint8_t)ao[i];
          }
          ++found;
          ovfl |= ao[i];
          edges += (unsigned long long)(e[i] - b[i]);
        } else {
          B[atomicAdd(&L.nkeep, 1u)] = (uint32_t)v[i];
        }
      }
      __syncthreads();
      // the long in-lists: a wavefront per column, lanes over its entries
      const unsigned int nh = min(L.nheavy, (unsigned)kFinHeavyCap);
      for (unsigned int h = wave; h < nh; h += kWaves) {
        const int vv = L.heavy[h];
        const int hb = g.CP[vv], he = g.CP[vv + 1];
        uint64_t lo = 0, hi = 0;
        uint32_t o = 0, any = 0;
        for (int q = hb + ln; q < he; q += 64) {
          const int uu = g.IC[q];
          if (p.S[uu] == k) {
            any = 1;
            if (p.sigma_on) {
              const uint64_t x = p.sigma[uu];
              lo += x & 0xffffffffull;
              hi += x >> 32;
              o |= p.ovf[uu];
            }
          }
        }
        const bool got = __any(any != 0);
        if (got) {
          // sum = hi * 2^32 + lo: exact below 2^64 unless hi >= 2^32 or the
          // final addition carries
          lo = wave_sum_u64(lo);
          hi = wave_sum_u64(hi);
          o = __any(o != 0) ? 1u : 0u;
          const uint64_t hs = hi << 32;
          const uint64_t c = hs + lo;
          o |= (hi >> 32) != 0 || c < hs ? 1u : 0u;
          if (ln == 0) {
            p.S[vv] = k + 1;
            if (p.sigma_on) {
              p.sigma[vv] = c;
              p.ovf[vv] = (uint8_t)o;
            }
            ++found;
            ovfl |= o;
            edges += (unsigned long long)(he - hb);
            if (he - hb > kBigIn) ++hin;
          }
        } else if (ln == 0) {
          B[atomicAdd(&L.nkeep, 1u)] = (uint32_t)vv;
        }
      }
      __syncthreads();
    }
    // this level's totals
    if (found) atomicAdd(&L.found, found);
    if (ovfl) atomicOr(&L.ovf, 1u);
    if (edges) atomicAdd(&L.edges, edges);
    if (hin) atomicAdd(&L.hin, hin);
    __syncthreads();
    const unsigned int q = L.found;
    if (q == 0) {  // the frontier of depth k + 1 is empty
      s.d = k + 1;
      break;
    }
    s.reached += q;
    s.edges += L.edges;
    s.heavy_reached += L.hin;
    s.ovf_any |= L.ovf;
    s.mu = s.mu > L.edges ? s.mu - L.edges : 0ull;
    ++s.pull_levels;
    LA = L.nkeep;
    uint32_t* tmp = A;
    A = B;
    B = tmp;
    __syncthreads();  // L.nkeep read by every thread before the reset
  }
  if (t == 0) {
    s.done = 1;
    p.loop[(step + 1) & 1] = s;
    write_stats(p, s, step);
    p.stats->fin_level = d0;
  }
}

// Chained loop: level d.  Reads slot d & 1 and the counters of depth d (both
// final: written by earlier kernels), writes slot (d + 1) & 1.  FBL: pull
// levels stage the frontier bitmap in dynamic LDS (one workgroup per CU).
// `step` only selects the state slot (its parity equals the depth's): the
// depth itself comes from the state, so one captured chunk of level kernels
// serves every stretch of levels starting at an even depth.  FIN: the
// instantiation that can hand the BFS to the one-workgroup finisher (taken
// only when the option asks for it): compiled into every level kernel, the
// finisher's registers made the staging launch spill 120 bytes per lane and
// left the 256-thread one at 148 VGPRs (three waves per SIMD); without it
// both fit four waves per SIMD.
template <int BLOCK, bool FBL, bool FIN>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
void k_chain_level(PersistParams p, int32_t step, int32_t want) {
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
  const bool heavy_open = s.heavy_reached < (unsigned long long)p.nheavy;
  const bool fin = FIN && p.fin_mu && s.mu <= (unsigned long long)p.fin_mu &&
                   p.mode == TBFS_MODE_AUTO;
  const bool stage = !fin && wants_stage(p, s.pushing != 0, s.mu);
  // one of a pair of launches for this step: the other one runs the level
  if (want == kLevelIfStaged && !stage) return;
  if (want == kLevelIfNotStaged && stage) return;
  if (gtid == 0 && d < kWorkLevels) {
    const unsigned long long w = s.pushing ? c.mout : s.mu;
    p.stats->work_lg[d] = (uint8_t)(64 - __clzll((long long)w));
  }
  if constexpr (FIN) {
    if (fin) {
      // the rest of the BFS in the last workgroup to list its columns
      static_assert(sizeof(FinLds<BLOCK>) <= sizeof(lds), "the finisher's LDS fits the level's");
      FinLds<BLOCK>& F = *reinterpret_cast<FinLds<BLOCK>*>(lds);
      fin_list_columns(p, d, heavy_open, gtid, gstride);
      __syncthreads();
      if (threadIdx.x == 0) {
        __threadfence();
        const unsigned prev = atomicAdd(&p.fin_ctr[1], 1u);
        F.last = prev == gridDim.x - 1 ? 1u : 0u;
        if (F.last) __threadfence();
      }
      __syncthreads();
      if (!F.last) return;
      fin_run<BLOCK>(p, s, step, F);
      return;
    }
  }
  LoopState nx = s;
  nx.d = d + 1;
  // the host picks each launch's shape (staging workgroups or not) from the
  // levels that staged in its last BFS on this graph
  if (stage) nx.stage_mask |= 1ull << (d < 63 ? d : 63);
  if (gtid == 0) p.loop[(step + 1) & 1] = nx;
  LevelAcc acc;
  run_level(p, d, s.pushing != 0, s.ovf_any, s.mu, c.hout != 0, heavy_open, gtid, gstride,
            &lds[threadIdx.x >> 6], acc, FBL ? fb_dyn : nullptr);
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
                       size_t fbl_bytes, int32_t want) {
  // the finisher's instantiations only when the run may hand over to it
  if (fbl_bytes) {
    if (p.fin_mu)
      hipLaunchKernelGGL((k_chain_level<1024, true, true>), dim3(grid), dim3(1024), fbl_bytes, s,
                         p, d, want);
    else
      hipLaunchKernelGGL((k_chain_level<1024, true, false>), dim3(grid), dim3(1024), fbl_bytes, s,
                         p, d, want);
  } else {
    if (p.fin_mu)
      hipLaunchKernelGGL((k_chain_level<kLbBlock, false, true>), dim3(grid), dim3(kLbBlock), 0, s,
                         p, d, want);
    else
      hipLaunchKernelGGL((k_chain_level<kLbBlock, false, false>), dim3(grid), dim3(kLbBlock), 0, s,
                         p, d, want);
  }
  return hipGetLastError();
}

hipError_t chain_fbl_limit(size_t* bytes) {
  // LDS left for the staged bitmap next to the kernels' static LDS (both
  // staging instantiations take the same limit)
  const void* fns[2] = {reinterpret_cast<const void*>(k_chain_level<1024, true, false>),
                        reinterpret_cast<const void*>(k_chain_level<1024, true, true>)};
  const size_t lds_cu = 160 * 1024;
  size_t avail = lds_cu;
  for (const void* fn : fns) {
    hipFuncAttributes fa;
    const hipError_t e = hipFuncGetAttributes(&fa, fn);
    if (e != hipSuccess) return e;
    avail = std::min(avail, lds_cu > fa.sharedSizeBytes ? lds_cu - fa.sharedSizeBytes : 0);
  }
  avail &= ~(size_t)255;
  for (const void* fn : fns) {
    const hipError_t e =
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)avail);
    if (e != hipSuccess) return e;
  }
  *bytes = avail;
  return hipSuccess;
}

}  // namespace kern
}  // namespace tbfs
