// tbfs_internal.h -- types shared by the gfx950 kernels and the host engine.
//
// Data layout in HBM (per resident graph / per BFS context):
//   CP int32[n+1], IC int32[nz]   in-lists (CSC): column v lists the u with u->v
//   RP int32[n+1], JR int32[nz]   out-lists (CSR = transpose); alias CP/IC when
//                                 the graph is symmetric
//   S int32[n]                    depth, -1 = not reached
//   sigma uint64[n], ovf uint8[n] shortest-path count mod 2^64 + sticky
//                                 "count >= 2^64" flag
//   fb[3] uint64[ceil(n/64)]      frontier bitmaps (bit v&63 of word v>>6);
//                                 current / next / being cleared, rotating
//   q[2] int32[n]                 frontier queues (current / next)
//   LevelCtr                      per-level counters written by the kernels
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tbfs {

constexpr int kWave = 64;

// Counters of the vertices discovered in one level.  All fields are updated
// with device-scope atomics, one atomic per wavefront.
struct alignas(64) LevelCtr {
  unsigned long long q;     // vertices discovered (= next queue length)
  unsigned long long mout;  // sum of their out-degrees (push work next level)
  unsigned long long min;   // sum of their in-degrees (leaves the unexplored set)
  unsigned int ovf;         // 1 if any of them has a count >= 2^64
  unsigned int pad0;
  unsigned long long pad1[4];
};

struct GraphView {
  int32_t n;
  int32_t nz;
  const int32_t* CP;  // in-lists
  const int32_t* IC;
  const int32_t* RP;  // out-lists (may alias CP/IC)
  const int32_t* JR;
};

struct StateView {
  int32_t* S;
  uint64_t* sigma;
  uint8_t* ovf;
  const uint64_t* fcur;  // frontier of depth d
  uint64_t* fnext;       // frontier of depth d+1 (must be zero on entry)
  const int32_t* qcur;
  int32_t* qnext;
  LevelCtr* ctr;         // counters of depth d+1
};

// Kernel launch wrappers (defined in tbfs_kernels.hip; every launch is made
// in the translation unit that defines the kernel).
namespace kern {
hipError_t init(hipStream_t s, int grid, int32_t n, int32_t root, int32_t* S,
                uint64_t* sigma, uint8_t* ovf, uint64_t* fb0, uint64_t* fb1,
                uint64_t* fb2, int32_t* q0);
hipError_t level_prologue(hipStream_t s, int grid, uint64_t* fnext,
                          int32_t nwords, LevelCtr* ctr);
hipError_t pull_sc(hipStream_t s, int grid, int block, const GraphView& g,
                   const StateView& st, int32_t d, int sigma, int bbreak,
                   int ovf_any);
hipError_t pull_wa(hipStream_t s, int grid, int block, const GraphView& g,
                   const StateView& st, int32_t d, int sigma, int bbreak,
                   int ovf_any);
hipError_t push(hipStream_t s, int grid, int block, const GraphView& g,
                const StateView& st, int32_t d, int64_t qsize, int ovf_any);
hipError_t cooc_scatter(hipStream_t s, int grid, int block, const int32_t* I,
                        const int32_t* J, int32_t nz, const StateView& st,
                        int ovf_any);
hipError_t cooc_finish(hipStream_t s, int grid, int block, const GraphView& g,
                       const StateView& st, int32_t d);
// setup helpers
hipError_t expand_cols(hipStream_t s, const int32_t* CP, int32_t n,
                       int32_t* col);
hipError_t histogram(hipStream_t s, const int32_t* keys, int32_t nz,
                     int32_t* cnt);
hipError_t check_csc(hipStream_t s, const int32_t* CP, const int32_t* IC,
                     int32_t n, int32_t nz, int32_t* bad);
}  // namespace kern

}  // namespace tbfs
