// tbfs_device.h -- wavefront (64-lane) helpers shared by the BFS kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tbfs_internal.h"

namespace tbfs {
namespace dev {

__device__ __forceinline__ int lane() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << lane()) - 1ull;
}

__device__ __forceinline__ bool fbit(const uint64_t* __restrict__ fb, int u) {
  return (fb[u >> 6] >> (u & 63)) & 1ull;
}

// s += x with the carry folded into the sticky flag o.
__device__ __forceinline__ void add_carry(uint64_t& s, uint32_t& o, uint64_t x) {
  const uint64_t t = s + x;
  o |= (t < s) ? 1u : 0u;
  s = t;
}

// Butterfly sum of (value, flag) pairs over the 64 lanes; every lane ends with
// the total.  All lanes must be converged.
__device__ __forceinline__ void wave_sum_carry(uint64_t& s, uint32_t& o) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t t = __shfl_xor(s, off);
    const uint32_t to = __shfl_xor(o, off);
    const uint64_t r = s + t;
    o = o | to | ((r < s) ? 1u : 0u);
    s = r;
  }
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

__device__ __forceinline__ uint64_t wave_or_u64(uint64_t x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x |= __shfl_xor(x, off);
  return x;
}

// Inclusive prefix sum over the 64 lanes (all lanes converged).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const int l = lane();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(x, off);
    if (l >= off) x += t;
  }
  return x;
}

// Order this wavefront's LDS accesses before / after this point.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Largest j in [0, 63] with excl[j] <= e (excl[0] == 0 <= e).  With excl an
// exclusive prefix sum of per-lane counts this is the lane owning item e.
__device__ __forceinline__ int owner_of(const int32_t* excl, int e) {
  int lo = 0;
#pragma unroll
  for (int step = 32; step > 0; step >>= 1)
    if (excl[lo + step] <= e) lo += step;
  return lo;
}

// owner_of for R items at once: the R searches advance in step, so each
// step's R LDS reads are issued together and waited for once (six waits per
// call instead of six per item).
template <int R>
__device__ __forceinline__ void owners_of(const int32_t* excl, const int (&e)[R], int (&o)[R]) {
#pragma unroll
  for (int t = 0; t < R; ++t) o[t] = 0;
#pragma unroll
  for (int step = 32; step > 0; step >>= 1) {
    int x[R];
#pragma unroll
    for (int t = 0; t < R; ++t) x[t] = excl[o[t] + step];
#pragma unroll
    for (int t = 0; t < R; ++t) o[t] += x[t] <= e[t] ? step : 0;
  }
}

__device__ __forceinline__ uint32_t outdeg(const GraphView& g, int v) {
  return (uint32_t)(g.RP[v + 1] - g.RP[v]);
}
__device__ __forceinline__ uint32_t indeg(const GraphView& g, int v) {
  return (uint32_t)(g.CP[v + 1] - g.CP[v]);
}

// Publish the vertices this wavefront discovered in level d (found on this
// lane: vertex v with out-degree od, in-degree id, overflow flag ov): queue
// entries for each kPushChunk slice of v's out-list, and the level counters.
// All lanes must be converged.
__device__ __forceinline__ void emit(const StateView& st, int v, bool found,
                                     uint32_t od, uint32_t id, uint32_t ov) {
  const uint64_t m = __ballot(found);
  if (m == 0) return;
  const uint32_t nch = found ? (od + kPushChunk - 1) / kPushChunk : 0u;
  const uint32_t incl = wave_incl_scan(nch);
  const uint32_t total = __shfl(incl, 63);
  unsigned long long base = 0;
  const uint64_t sod = wave_sum_u64(found ? od : 0);
  const uint64_t sid = wave_sum_u64(found ? id : 0);
  const bool anyov = __ballot(found && ov) != 0;
  if (lane() == 0) {
    if (total) base = atomicAdd(&st.ctr->qe, (unsigned long long)total);
    atomicAdd(&st.ctr->q, (unsigned long long)__popcll(m));
    atomicAdd(&st.ctr->mout, (unsigned long long)sod);
    atomicAdd(&st.ctr->min, (unsigned long long)sid);
    if (anyov) atomicOr(&st.ctr->ovf, 1u);
  }
  base = __shfl(base, 0);
  if (found) {
    const unsigned long long first = base + incl - nch;
    for (uint32_t c = 0; c < nch; ++c)
      st.qnext[first + c] = (uint64_t)(uint32_t)v | ((uint64_t)c << 32);
  }
}

}  // namespace dev
}  // namespace tbfs
