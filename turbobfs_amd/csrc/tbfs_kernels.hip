// tbfs_kernels.hip -- gfx950 per-level kernels of the host-driven BFS loop.
//
// All kernels are integer work (no MFMA): depth S, path count sigma
// (uint64 mod 2^64 + sticky overflow flag), frontier bitmaps and queues.
// Wavefronts are 64 lanes; ballots are 64-bit; every per-level counter is
// aggregated per wavefront before its single device-scope atomic.
//
// Correspondence with the reference kernels (paths relative to the reference
// root):
//   pull_sc   spMvUgCscScKernel + bfsFunctionsKernel fused
//             (TurboBFS_CSC_TD/bfsgputdcsc_sc.cu:159-208), and with sigma
//             off, bfsBuScKernel (TurboBFS_CSC_BU/bfsgpubucsc_sc.cu:122-148)
//   pull_wa   spMvUgCscWaKernel (TurboBFS_CSC_TD/bfsgputdcsc_wa.cu:159-219)
//             and bfsBuWaKernel (TurboBFS_CSC_BU/bfsgpubucsc_wa.cu:121-180):
//             one wavefront per column, ballot early exit, butterfly sum
//   cooc_*    spMvUgCoocScKernel + bfsFunctionsKernel
//             (TurboBFS_COOC_TD/bfsgputdcooc_sc.cu:152-188)
//   push      top-down expansion from the frontier queue (the TD half of the
//             direction-optimizing switch, TurboBFS_CSC_TDBU/bfsgputdbucsc_sc.cu:119-153)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tbfs_device.h"
#include "tbfs_internal.h"

namespace tbfs {
namespace {

using namespace dev;

// ----------------------------------------------------------------------------
__global__ void k_init(int32_t n, int32_t root, int32_t root_out_deg,
                       int32_t* S, uint64_t* sigma, uint8_t* ovf,
                       uint64_t* fb0, uint64_t* fb1, uint64_t* fb2,
                       uint64_t* q0, uint64_t* vis) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = tid; i < n; i += stride) {
    S[i] = (i == root) ? 0 : -1;
    sigma[i] = (i == root) ? 1ull : 0ull;
    ovf[i] = 0;
  }
  const int64_t nw = ((int64_t)n + 63) >> 6;
  for (int64_t w = tid; w < nw; w += stride) {
    fb0[w] = (w == (root >> 6)) ? (1ull << (root & 63)) : 0ull;
    fb1[w] = 0ull;
    fb2[w] = 0ull;
    vis[w] = 0ull;
  }
  const int64_t nch = ((int64_t)root_out_deg + kPushChunk - 1) / kPushChunk;
  for (int64_t c = tid; c < nch; c += stride)
    q0[c] = (uint64_t)(uint32_t)root | ((uint64_t)c << 32);
}

__global__ void k_level_prologue(uint64_t* fnext, int32_t nwords,
                                 LevelCtr* ctr, CtrSet* cset,
                                 const uint64_t* fcur, uint64_t* vis) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t w = tid; w < nwords; w += stride) {
    fnext[w] = 0ull;
    const uint64_t f = fcur[w];
    if (f) vis[w] |= f;
  }
  if (tid == 0) {
    ctr->q = 0;
    ctr->qe = 0;
    ctr->mout = 0;
    ctr->min = 0;
    ctr->ovf = 0;
  }
  if (tid < kStripes) {
    cset->s[tid].qm = 0;
    cset->s[tid].min = 0;
    cset->s[tid].hubs = 0;
    cset->s[tid].ovf = 0;
  }
}

// ----------------------------------------------------------------------------
// Pull over unvisited columns, one lane per column.  Lanes of a wavefront own
// 64 consecutive columns, so the ballot of "discovered" is exactly the next
// frontier word for those columns.
template <bool SIGMA>
__global__ __launch_bounds__(256) void k_pull_sc(GraphView g, StateView st,
                                                 int32_t d, int bbreak,
                                                 int ovf_any) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
       base < g.n; base += stride) {
    const int v = (int)base + lane();
    bool found = false;
    uint64_t sum = 0;
    uint32_t ov = 0;
    if (v < g.n && st.S[v] == -1) {
      const int beg = g.CP[v], end = g.CP[v + 1];
      for (int k = beg; k < end; ++k) {
        const int u = g.IC[k];
        if (fbit(st.fcur, u)) {
          found = true;
          if (SIGMA) {
            add_carry(sum, ov, st.sigma[u]);
            if (ovf_any) ov |= st.ovf[u];
          } else if (bbreak) {
            break;
          }
        }
      }
      if (found) {
        st.S[v] = d + 1;
        if (SIGMA) {
          st.sigma[v] = sum;
          if (ov) st.ovf[v] = 1;
        }
      }
    }
    const uint64_t m = __ballot(found);
    if (m) {
      if (lane() == 0) st.fnext[base >> 6] = m;
      emit(st, v, found, found ? outdeg(g, v) : 0u, found ? indeg(g, v) : 0u, ov);
    }
  }
}

// Pull over unvisited columns, one wavefront per column.  A wavefront walks
// the 64 columns of one frontier word in turn; lanes stride the column's
// in-list, the path counts are summed with a 64-lane butterfly, and with
// sigma off and bbreak set the whole wavefront leaves the column at the
// first 64-entry chunk that holds a parent (ballot).
template <bool SIGMA>
__global__ __launch_bounds__(256) void k_pull_wa(GraphView g, StateView st,
                                                 int32_t d, int bbreak,
                                                 int ovf_any) {
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t base = wave * 64; base < g.n; base += nwaves * 64) {
    const int cnt = (int)((g.n - base) < 64 ? (g.n - base) : 64);
    uint64_t word = 0;
    uint32_t myov = 0;
    for (int j = 0; j < cnt; ++j) {
      const int v = (int)base + j;
      if (st.S[v] != -1) continue;
      const int beg = g.CP[v], end = g.CP[v + 1];
      bool f = false;
      uint64_t sum = 0;
      uint32_t ov = 0;
      for (int k0 = beg; k0 < end; k0 += 64) {
        const int k = k0 + lane();
        if (k < end) {
          const int u = g.IC[k];
          if (fbit(st.fcur, u)) {
            f = true;
            if (SIGMA) {
              add_carry(sum, ov, st.sigma[u]);
              if (ovf_any) ov |= st.ovf[u];
            }
          }
        }
        if (!SIGMA && bbreak) {
          if (__ballot(f)) break;
        }
      }
      if (__ballot(f)) {
        if (SIGMA) {
          wave_sum_carry(sum, ov);
          if (lane() == 0) {
            st.sigma[v] = sum;
            if (ov) st.ovf[v] = 1;
          }
        }
        if (lane() == 0) st.S[v] = d + 1;
        word |= 1ull << j;
        if (lane() == j) myov = ov;
      }
    }
    if (word) {
      if (lane() == 0) st.fnext[base >> 6] = word;
      const int v = (int)base + lane();
      const bool found = (word >> lane()) & 1ull;
      emit(st, v, found, found ? outdeg(g, v) : 0u, found ? indeg(g, v) : 0u, myov);
    }
  }
}

// Top-down expansion from the frontier queue: one wavefront per queue entry
// (a slice of at most 64 out-list entries of frontier vertex u), one lane per
// edge.  The first lane to move S[v] from -1 to d+1 enqueues v; every edge
// into a vertex of depth d+1 adds sigma[u] with a 64-bit atomic whose
// returned old value exposes the carry.
__global__ __launch_bounds__(256) void k_push(GraphView g, StateView st,
                                              int32_t d, int64_t qe,
                                              int ovf_any) {
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int dn = d + 1;
  for (int64_t i = wave; i < qe; i += nwaves) {
    const uint64_t ent = st.qcur[i];
    const int u = (int)(uint32_t)ent;
    const int beg = g.RP[u] + (int)(ent >> 32) * kPushChunk;
    const int end = min(g.RP[u + 1], beg + kPushChunk);
    const uint64_t x = st.sigma[u];
    const uint32_t ou = ovf_any ? (uint32_t)st.ovf[u] : 0u;
    const int k = beg + lane();
    bool nw = false;
    int v = 0;
    uint32_t ov = 0;
    if (k < end) {
      v = g.JR[k];
      int s = st.S[v];
      if (s == -1) {
        const int old = atomicCAS(&st.S[v], -1, dn);
        nw = (old == -1);
        s = dn;  // old is -1 or dn: only this level writes S now
      }
      if (s == dn) {
        const unsigned long long old =
            atomicAdd((unsigned long long*)&st.sigma[v], (unsigned long long)x);
        if (old + x < old || ou) {
          st.ovf[v] = 1;
          ov = 1;
        }
      }
      if (nw) atomicOr((unsigned long long*)&st.fnext[v >> 6], 1ull << (v & 63));
    }
    if (__ballot(nw || ov)) {
      emit(st, v, nw, nw ? outdeg(g, v) : 0u, nw ? indeg(g, v) : 0u, ov);
      // an overflow on an edge into a vertex claimed by another lane
      if (__ballot(ov && !nw) && lane() == 0) atomicOr(&st.ctr->ovf, 1u);
    }
  }
}

// COOC, edge-parallel: every nonzero whose source is in the frontier adds the
// source's count into its (unvisited) destination.
__global__ __launch_bounds__(256) void k_cooc_scatter(const int32_t* __restrict__ I,
                                                      const int32_t* __restrict__ J,
                                                      int32_t nz, StateView st,
                                                      int ovf_any) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nz;
       k += stride) {
    const int u = I[k];
    if (!fbit(st.fcur, u)) continue;
    const int v = J[k];
    if (st.S[v] != -1) continue;
    const uint64_t x = st.sigma[u];
    const unsigned long long old =
        atomicAdd((unsigned long long*)&st.sigma[v], (unsigned long long)x);
    if (old + x < old || (ovf_any && st.ovf[u])) st.ovf[v] = 1;
  }
}

// COOC level finish: an unvisited vertex with a nonzero accumulated count was
// reached this level (the reference's bfsFunctionsKernel).
__global__ __launch_bounds__(256) void k_cooc_finish(GraphView g, StateView st,
                                                     int32_t d) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
       base < g.n; base += stride) {
    const int v = (int)base + lane();
    bool found = false;
    uint32_t ov = 0;
    if (v < g.n && st.S[v] == -1) {
      ov = st.ovf[v];
      found = (st.sigma[v] != 0) || ov;
      if (found) st.S[v] = d + 1;
    }
    const uint64_t m = __ballot(found);
    if (m) {
      if (lane() == 0) st.fnext[base >> 6] = m;
      uint32_t od = 0, id = 0;
      if (found && g.RP) od = outdeg(g, v);
      if (found && g.CP) id = indeg(g, v);
      // without out-lists (unsorted COOC) there is no push queue to fill
      emit(st, v, found, g.RP ? od : 0u, id, ov);
    }
  }
}

// ----------------------------------------------------------------------------
// calibration: one 16-byte load and store per thread and iteration
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src,
                                                uint4* __restrict__ dst,
                                                int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += stride)
    dst[i] = src[i];
}

// ----------------------------------------------------------------------------
// setup kernels
__global__ void k_expand_cols(const int32_t* CP, int32_t n, int32_t* col) {
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t v = wave; v < n; v += nwaves) {
    const int beg = CP[v], end = CP[v + 1];
    for (int k = beg + lane(); k < end; k += 64) col[k] = (int32_t)v;
  }
}

// Static slices of the lists longer than kBigIn (entry v | slice << 32, in
// vertex then slice order): slices per vertex (cnt[n] = 0, so an exclusive
// scan of n + 1 counts ends with the total), then the entries at the scanned
// offsets.
__global__ void k_slice_count(const int32_t* ptr, int32_t n, int32_t min_deg, int32_t* cnt,
                              int32_t* nlists) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int32_t c = 0;
  if (v < n) {
    const int64_t deg = (int64_t)ptr[v + 1] - ptr[v];
    if (deg > min_deg) c = (int32_t)((deg + kPullChunk - 1) / kPullChunk);
  }
  if (v <= n) cnt[v] = c;
  // lists cut: one atomic per wavefront
  const uint64_t m = __ballot(c > 0);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(nlists, (int32_t)__popcll(m));
}

__global__ void k_slice_fill(const int32_t* cnt, const int32_t* off, int32_t n,
                             uint64_t* out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int32_t c = cnt[v], o = off[v];
  for (int32_t k = 0; k < c; ++k)
    out[o + k] = (uint64_t)(uint32_t)v | ((uint64_t)(uint32_t)k << 32);
}

// The reference entry points' output types, converted on the device: S with
// unreached vertices as s_unreached (0 for top-down CSC, -1 otherwise); with
// kind 1 the counts as float (INFINITY when >= 2^64), kind 2 as int32 modulo
// 2^32 (the reference's int arithmetic), kind 0 none; *inexact counts the
// vertices whose count is beyond the type's exact range (float: above 2^24,
// int32: above 2^31 - 1, or >= 2^64).  One atomic per wavefront.
__global__ void k_dropin_convert(const int32_t* S, const uint64_t* sigma, const uint8_t* ovf,
                                 int32_t n, int32_t s_unreached, int32_t kind, int32_t* S_out,
                                 uint32_t* val_out, unsigned long long* inexact) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (i < n) {
    const int32_t sv = S[i];
    S_out[i] = sv < 0 ? s_unreached : sv;
    if (kind != 0) {
      const uint64_t x = sigma[i];
      const bool o = ovf[i] != 0;
      if (kind == 1) {
        val_out[i] = __float_as_uint(o ? INFINITY : (float)x);
        bad = o || x > (1ull << 24);
      } else {
        val_out[i] = (uint32_t)x;
        bad = o || x > 0x7fffffffull;
      }
    }
  }
  const uint64_t m = __ballot(bad);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(inexact, (unsigned long long)__popcll(m));
}

// Bitmap of the columns the light-column pull skips (no in-list, or one split
// into slices); bits past n are set.  One wavefront per word.
__global__ void k_skip_bits(const int32_t* ptr, int32_t n, uint64_t* skip, int64_t nw) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w = v >> 6;
  if (w >= nw) return;  // uniform per wavefront: 64 lanes share w
  bool sk = true;
  if (v < n) {
    const int64_t deg = (int64_t)ptr[v + 1] - ptr[v];
    sk = deg == 0 || deg > kBigIn;
  }
  const uint64_t bits = __ballot(sk);
  if ((threadIdx.x & 63) == 0) skip[w] = bits;
}

// Offsets of a non-decreasing key array with keys in [0, n): ptr[r] = the
// first k with keys[k] >= r, for r = 0 .. n (ptr[n] = nz).  One binary search
// per offset and no atomics: counting the keys with atomics serialised on the
// keys of hub lists (1.4 ms for R-MAT scale 20's 31.4M row indices).
__global__ void k_offsets_sorted(const int32_t* keys, int32_t nz, int32_t n, int32_t* ptr) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > n) return;
  int32_t lo = 0, hi = nz;
  while (lo < hi) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if ((int64_t)keys[mid] < r) lo = mid + 1;
    else hi = mid;
  }
  ptr[r] = lo;
}

// Degree statistics (see DegAcc): per workgroup, LDS bins and wavefront
// reductions, then one device atomic per workgroup and field.  Bin of degree
// d: 0 for d == 0, else floor(log2 d) + 1.
__global__ __launch_bounds__(256) void k_degree_stats(const int32_t* ptr, int32_t n,
                                                      DegAcc* acc) {
  __shared__ unsigned long long bins[33];
  __shared__ unsigned long long red[4][4];
  if (threadIdx.x < 33) bins[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long mk = 0, sum = 0, lo = 0, hi = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    const uint32_t d = (uint32_t)(ptr[v + 1] - ptr[v]);
    const unsigned long long key = ((unsigned long long)d << 32) | (uint32_t)~(uint32_t)v;
    mk = key > mk ? key : mk;
    sum += d;
    const unsigned long long sq = (unsigned long long)d * d;
    lo += sq & 0xffffffffull;
    hi += sq >> 32;
    atomicAdd(&bins[d ? 32 - __clz((int)d) : 0], 1ull);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(mk, off);
    mk = o > mk ? o : mk;
    sum += __shfl_xor(sum, off);
    lo += __shfl_xor(lo, off);
    hi += __shfl_xor(hi, off);
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[wv][0] = mk;
    red[wv][1] = sum;
    red[wv][2] = lo;
    red[wv][3] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = 0, t1 = 0, t2 = 0, t3 = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      m = red[i][0] > m ? red[i][0] : m;
      t1 += red[i][1];
      t2 += red[i][2];
      t3 += red[i][3];
    }
    if (m) atomicMax(&acc->maxkey, m);
    if (t1) atomicAdd(&acc->sum, t1);
    if (t2) atomicAdd(&acc->sq_lo, t2);
    if (t3) atomicAdd(&acc->sq_hi, t3);
  }
  if (threadIdx.x < 33 && bins[threadIdx.x]) atomicAdd(&acc->bins[threadIdx.x], bins[threadIdx.x]);
}

// Degree-ordered numbering (batches): sort keys (larger in-degree first,
// ties by id through the stable sort), then the inverse permutation and the
// new column lengths, then every entry to its new place.
__global__ void k_order_keys(const int32_t* CP, int32_t n, uint32_t* keys, int32_t* ids) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    keys[v] = 0x7fffffffu - (uint32_t)(CP[v + 1] - CP[v]);
    ids[v] = (int32_t)v;
  }
}

__global__ void k_order_perm(const int32_t* order, const int32_t* CP, int32_t n,
                             int32_t* perm, int32_t* deg2) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c <= n; c += stride) {
    if (c == n) {
      deg2[n] = 0;
      continue;
    }
    const int32_t o = order[c];
    perm[o] = (int32_t)c;
    deg2[c] = CP[o + 1] - CP[o];
  }
}

__global__ void k_order_entries(const int32_t* IC, const int32_t* col, const int32_t* CP,
                                const int32_t* CP2, const int32_t* perm, int32_t nz,
                                int32_t* IC2) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nz; k += stride) {
    const int32_t o = col[k];
    IC2[CP2[perm[o]] + (k - CP[o])] = perm[IC[k]];
  }
}

// bad[0] |= 1: CP not monotone / CP[0] != 0 / CP[n] != nz; 2: IC out of range
__global__ void k_check_csc(const int32_t* CP, const int32_t* IC, int32_t n,
                            int32_t nz, int32_t* bad) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int b = 0;
  for (int64_t i = tid; i < n; i += stride)
    if (CP[i + 1] < CP[i]) b |= 1;
  if (tid == 0 && (CP[0] != 0 || CP[n] != nz)) b |= 1;
  for (int64_t k = tid; k < nz; k += stride)
    if ((uint32_t)IC[k] >= (uint32_t)n) b |= 2;
  if (b) atomicOr(bad, b);
}

}  // namespace

// ----------------------------------------------------------------------------
namespace kern {

hipError_t init(hipStream_t s, int grid, int32_t n, int32_t root,
                int32_t root_out_deg, int32_t* S, uint64_t* sigma,
                uint8_t* ovf, uint64_t* fb0, uint64_t* fb1, uint64_t* fb2,
                uint64_t* q0, uint64_t* vis) {
  hipLaunchKernelGGL(k_init, dim3(grid), dim3(256), 0, s, n, root,
                     root_out_deg, S, sigma, ovf, fb0, fb1, fb2, q0, vis);
  return hipGetLastError();
}

hipError_t level_prologue(hipStream_t s, int grid, uint64_t* fnext,
                          int32_t nwords, LevelCtr* ctr, CtrSet* cset,
                          const uint64_t* fcur, uint64_t* vis) {
  hipLaunchKernelGGL(k_level_prologue, dim3(grid), dim3(256), 0, s, fnext,
                     nwords, ctr, cset, fcur, vis);
  return hipGetLastError();
}

hipError_t pull_sc(hipStream_t s, int grid, int block, const GraphView& g,
                   const StateView& st, int32_t d, int sigma, int bbreak,
                   int ovf_any) {
  if (sigma)
    hipLaunchKernelGGL(k_pull_sc<true>, dim3(grid), dim3(block), 0, s, g, st,
                       d, bbreak, ovf_any);
  else
    hipLaunchKernelGGL(k_pull_sc<false>, dim3(grid), dim3(block), 0, s, g, st,
                       d, bbreak, ovf_any);
  return hipGetLastError();
}

hipError_t pull_wa(hipStream_t s, int grid, int block, const GraphView& g,
                   const StateView& st, int32_t d, int sigma, int bbreak,
                   int ovf_any) {
  if (sigma)
    hipLaunchKernelGGL(k_pull_wa<true>, dim3(grid), dim3(block), 0, s, g, st,
                       d, bbreak, ovf_any);
  else
    hipLaunchKernelGGL(k_pull_wa<false>, dim3(grid), dim3(block), 0, s, g, st,
                       d, bbreak, ovf_any);
  return hipGetLastError();
}

hipError_t push(hipStream_t s, int grid, int block, const GraphView& g,
                const StateView& st, int32_t d, int64_t qe, int ovf_any) {
  hipLaunchKernelGGL(k_push, dim3(grid), dim3(block), 0, s, g, st, d, qe,
                     ovf_any);
  return hipGetLastError();
}

hipError_t cooc_scatter(hipStream_t s, int grid, int block, const int32_t* I,
                        const int32_t* J, int32_t nz, const StateView& st,
                        int ovf_any) {
  hipLaunchKernelGGL(k_cooc_scatter, dim3(grid), dim3(block), 0, s, I, J, nz,
                     st, ovf_any);
  return hipGetLastError();
}

hipError_t cooc_finish(hipStream_t s, int grid, int block, const GraphView& g,
                       const StateView& st, int32_t d) {
  hipLaunchKernelGGL(k_cooc_finish, dim3(grid), dim3(block), 0, s, g, st, d);
  return hipGetLastError();
}

hipError_t copy16(hipStream_t s, const void* src, void* dst, int64_t n16) {
  int64_t grid = (n16 + 255) / 256;
  if (grid > 8192) grid = 8192;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)grid), dim3(256), 0, s,
                     (const uint4*)src, (uint4*)dst, n16);
  return hipGetLastError();
}

hipError_t expand_cols(hipStream_t s, const int32_t* CP, int32_t n,
                       int32_t* col) {
  int grid = (int)(((int64_t)n + 3) / 4);
  if (grid > 65536) grid = 65536;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_expand_cols, dim3(grid), dim3(256), 0, s, CP, n, col);
  return hipGetLastError();
}

hipError_t slice_count(hipStream_t s, const int32_t* ptr, int32_t n, int32_t min_deg,
                        int32_t* cnt, int32_t* nlists) {
  const int grid = (int)(((int64_t)n + 1 + 255) / 256);
  hipLaunchKernelGGL(k_slice_count, dim3(grid), dim3(256), 0, s, ptr, n, min_deg, cnt, nlists);
  return hipGetLastError();
}

hipError_t slice_fill(hipStream_t s, const int32_t* cnt, const int32_t* off, int32_t n,
                      uint64_t* out) {
  const int grid = (int)(((int64_t)n + 255) / 256);
  hipLaunchKernelGGL(k_slice_fill, dim3(grid), dim3(256), 0, s, cnt, off, n, out);
  return hipGetLastError();
}

hipError_t dropin_convert(hipStream_t s, const int32_t* S, const uint64_t* sigma,
                          const uint8_t* ovf, int32_t n, int32_t s_unreached, int32_t kind,
                          int32_t* S_out, uint32_t* val_out, unsigned long long* inexact) {
  const int grid = (int)(((int64_t)n + 255) / 256);
  hipLaunchKernelGGL(k_dropin_convert, dim3(grid), dim3(256), 0, s, S, sigma, ovf, n, s_unreached,
                     kind, S_out, val_out, inexact);
  return hipGetLastError();
}

hipError_t skip_bits(hipStream_t s, const int32_t* ptr, int32_t n, uint64_t* skip,
                     int64_t nw) {
  const int grid = (int)((nw * 64 + 255) / 256);
  hipLaunchKernelGGL(k_skip_bits, dim3(grid), dim3(256), 0, s, ptr, n, skip, nw);
  return hipGetLastError();
}

hipError_t offsets_sorted(hipStream_t s, const int32_t* keys, int32_t nz, int32_t n,
                          int32_t* ptr) {
  const int grid = (int)(((int64_t)n + 1 + 255) / 256);
  hipLaunchKernelGGL(k_offsets_sorted, dim3(grid), dim3(256), 0, s, keys, nz, n, ptr);
  return hipGetLastError();
}

hipError_t degree_stats(hipStream_t s, const int32_t* ptr, int32_t n, DegAcc* acc) {
  // every workgroup ends with ~37 atomics on the same few words: few
  // workgroups, each looping over many vertices
  int grid = (int)(((int64_t)n + 255) / 256);
  if (grid > 512) grid = 512;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_degree_stats, dim3(grid), dim3(256), 0, s, ptr, n, acc);
  return hipGetLastError();
}

static int fill_grid(int64_t m) {
  int64_t g = (m + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t order_keys(hipStream_t s, const int32_t* CP, int32_t n, uint32_t* keys, int32_t* ids) {
  hipLaunchKernelGGL(k_order_keys, dim3(fill_grid(n)), dim3(256), 0, s, CP, n, keys, ids);
  return hipGetLastError();
}

hipError_t order_perm(hipStream_t s, const int32_t* order, const int32_t* CP, int32_t n,
                      int32_t* perm, int32_t* deg2) {
  hipLaunchKernelGGL(k_order_perm, dim3(fill_grid((int64_t)n + 1)), dim3(256), 0, s, order, CP,
                     n, perm, deg2);
  return hipGetLastError();
}

hipError_t order_entries(hipStream_t s, const int32_t* IC, const int32_t* col, const int32_t* CP,
                         const int32_t* CP2, const int32_t* perm, int32_t nz, int32_t* IC2) {
  hipLaunchKernelGGL(k_order_entries, dim3(fill_grid(nz)), dim3(256), 0, s, IC, col, CP, CP2, perm,
                     nz, IC2);
  return hipGetLastError();
}

hipError_t check_csc(hipStream_t s, const int32_t* CP, const int32_t* IC,
                     int32_t n, int32_t nz, int32_t* bad) {
  int64_t m = (int64_t)(n > nz ? n : nz);
  int grid = (int)((m + 255) / 256);
  if (grid > 65536) grid = 65536;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_check_csc, dim3(grid), dim3(256), 0, s, CP, IC, n, nz,
                     bad);
  return hipGetLastError();
}

}  // namespace kern
}  // namespace tbfs
