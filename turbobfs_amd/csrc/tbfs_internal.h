// tbfs_internal.h -- types shared by the gfx950 kernels and the host engine.
//
// Data layout in HBM (per resident graph / per BFS context):
//   CP int32[n+1], IC int32[nz]   in-lists (CSC): column v lists the u with u->v
//   RP int32[n+1], JR int32[nz]   out-lists (CSR = transpose); alias CP/IC when
//                                 the graph is symmetric
//   bigc uint64[nbigc]            static pull work list: one entry per
//                                 kPullChunk-entry slice of every column whose
//                                 in-degree exceeds kBigIn (v | slice << 32)
//   bigo uint64[nbigo]            the same for out-lists longer than kBigIn
//                                 (push work; aliases bigc when symmetric)
//   S int32[n]                    depth, -1 = not reached
//   sigma uint64[n], ovf uint8[n] shortest-path count mod 2^64 + sticky
//                                 "count >= 2^64" flag
//   skip uint64[ceil(n/64)]       static bitmap: in-degree 0, or in-degree
//                                 > kBigIn (pulled through bigc slices)
//   fb[3] uint64[ceil(n/64)]      frontier bitmaps (bit v&63 of word v>>6);
//                                 current / next / being cleared, rotating
//   vis uint64[ceil(n/64)]        every frontier so far (reached vertices);
//                                 vis | skip | current frontier has a clear
//                                 bit exactly for the light pull path's work
//   q[2] uint64[q_cap]            frontier work queues of the per-level
//                                 reference-mapping kernels (current / next):
//                                 one entry per kPushChunk-entry slice of a
//                                 frontier vertex's out-list (v | slice << 32);
//                                 the load-balanced kernels read the frontier
//                                 bitmap instead and need no queue
//   LevelCtr[3]                   per-level counters, rotating
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tbfs {

constexpr int kWave = 64;
constexpr int kPushChunk = 64;    // out-list entries per queue entry
constexpr int kBigIn = 256;      // in-degree above which a column is split
constexpr int kPullChunk = 256;  // list entries per static slice (one pass of a wavefront)
static_assert(kPullChunk % kWave == 0, "slices are whole passes of a wavefront");

// Counters of the vertices discovered in one level.  Every field is updated
// with one device-scope atomic per wavefront.
struct alignas(64) LevelCtr {
  unsigned long long q;     // vertices discovered
  unsigned long long qe;    // queue entries written for them
  unsigned long long mout;  // sum of their out-degrees (push work next level)
  unsigned long long min;   // sum of their in-degrees (leaves the unexplored set)
  unsigned int ovf;         // 1 if any of them has a count >= 2^64
  unsigned int hin;         // how many of them have in-degree > kBigIn
  unsigned int hout;        // how many of them have out-degree > kBigIn
  unsigned int pad0;
  unsigned long long pad1[2];
};

// Level counters of the load-balanced kernels, striped over kStripes cache
// lines so that the per-workgroup atomics do not all hit one address: a
// workgroup adds into stripe (blockIdx & (kStripes-1)); readers sum the
// stripes.  qm packs the vertex count (high 32 bits) with the out-degree sum
// (low 32 bits); both totals stay below 2^31, so one 64-bit atomic carries
// both without interference.
constexpr int kStripes = 16;
struct alignas(128) CtrStripe {
  unsigned long long qm;
  unsigned long long min;
  unsigned long long hubs;  // hout << 32 | hin (see LevelCtr), both < 2^31
  unsigned int ovf;
  unsigned int pad0;
  unsigned long long pad1[12];
};
struct CtrSet {
  CtrStripe s[kStripes];
};

// Result counters of one BFS, written by the device-resident level loop.
struct alignas(64) RunStats {
  int32_t levels;
  int32_t push_levels;
  int32_t pull_levels;
  uint32_t ovf_any;
  unsigned long long reached;
  unsigned long long edges;
  int32_t launches;  // chained loop: level kernels used, the last one included
  int32_t fin_level; // chained loop: depth at which one workgroup finished the
                     // BFS (-1: every level ran on the whole grid)
  unsigned long long stage_mask;  // chained loop: bit min(d, 63) set when level d
                                  // was a pull that stages the frontier bitmap
  unsigned long long pad[2];
  // chained loop: work of level d (d < 64): ceil(log2(1 + entries)) of the
  // frontier's out-list entries (push) or the unexplored in-list entries
  // (pull); 0 for levels not run
  uint8_t work_lg[64];
};
constexpr int kWorkLevels = 64;

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
  const uint64_t* qcur;  // queue of depth d
  uint64_t* qnext;       // queue of depth d+1
  LevelCtr* ctr;         // counters of depth d+1 (must be zero on entry)
  CtrSet* cset;          // the same, striped (load-balanced kernels)
  uint64_t* vis;         // reached bitmap (every frontier so far)
  const uint64_t* skip;  // static skip bitmap of the graph
};

// Loop state carried from one level to the next by the device-resident
// loops (registers in the persistent kernel; two alternating slots in device
// memory for the chained per-level kernels, slot d & 1 read by level d).
struct alignas(64) LoopState {
  int32_t d;            // depth of the current frontier
  int32_t pushing;      // direction used by the previous level
  int32_t done;         // frontier empty: the BFS has finished
  uint32_t ovf_any;     // some count reached so far is >= 2^64
  int32_t push_levels;
  int32_t pull_levels;
  unsigned long long mu;       // in-list entries of the unreached vertices
  unsigned long long reached;  // vertices reached before this level
  unsigned long long edges;    // their in-list entries
  unsigned long long stage_mask;  // levels so far that were bitmap-staging pulls
  unsigned long long heavy_reached;  // reached vertices with in-degree > kBigIn
};

// Everything the device-resident level loops need.
struct PersistParams {
  GraphView g;
  int32_t* S;
  uint64_t* sigma;
  uint8_t* ovf;
  uint64_t* fb[3];
  uint64_t* vis;          // reached bitmap (every frontier so far)
  const uint64_t* skip;   // static: in-degree 0 or split column (graph)
  CtrSet* cset;     // [3] striped level counters
  unsigned* bar;    // kBarWords barrier words (zero at launch)
  RunStats* stats;
  const uint64_t* bigc;
  int32_t nbigc;
  int32_t nheavy;   // columns in bigc (in-degree > kBigIn)
  const uint64_t* bigo;
  int32_t nbigo;
  int32_t root;     // single root; -1 when the roots come from `roots`
  int32_t nroots;   // > 0: multi-root BFS from roots[0 .. nroots-1] (chained loop)
  const int32_t* roots;
  int32_t mode;     // TBFS_MODE_AUTO / _PUSH / _PULL_SC (all pulls are load-balanced)
  int32_t sigma_on;
  float alpha;
  float beta;
  float stage_ratio;  // pull levels stage the bitmap when mu > stage_ratio * n
  LoopState* loop;  // [2] (chained loop only)
  // chained loop: once at most fin_mu in-list entries are unexplored, one
  // workgroup runs the remaining levels over a list of the unreached columns
  // (fin_list[0/1]: kFinCap entries each; fin_ctr: list length, workgroups
  // done; zeroed by the init kernel).  fin_mu 0: never.
  uint32_t* fin_list[2];
  unsigned* fin_ctr;
  uint32_t fin_mu;
  // chained loop: 1 when the init kernel may run depth 0 itself (one root,
  // AUTO / PUSH): the level launches then start at step 1 and the init
  // kernel writes loop slot 1.  0: depth 0 is the first level launch's.
  int32_t step0;
};

// The init kernel runs depth 0 itself (a push from the root) when the root
// has at most this many out-list entries: one per thread of a workgroup.
constexpr int kInitPush = 256;

constexpr uint32_t kFinCap = 1u << 16;  // entries of each finisher list (at least)

// Degree statistics of the lists delimited by ptr[0..n]: acc (zeroed by the
// caller) receives the maximum key (degree << 32 | ~vertex), the degree sum,
// the sum of squares as low / high 32-bit halves, and 33 power-of-two bins.
struct DegAcc {
  unsigned long long maxkey;
  unsigned long long sum;
  unsigned long long sq_lo;
  unsigned long long sq_hi;
  unsigned long long bins[33];
};

// Chained level launches: kLevelAlways runs the level whatever it is; a pair
// of launches for one step, kLevelIfStaged (staging shape) then
// kLevelIfNotStaged, lets the level pick its shape on the device: each runs
// it only when the level's staging decision matches, the other returns at
// once.
constexpr int32_t kLevelAlways = 0;
constexpr int32_t kLevelIfStaged = 1;
constexpr int32_t kLevelIfNotStaged = 2;

// Kernel launch wrappers.  Every launch is made in the translation unit that
// defines the kernel.
namespace kern {
// tbfs_kernels.hip: per-level kernels of the host-driven loop
hipError_t init(hipStream_t s, int grid, int32_t n, int32_t root,
                int32_t root_out_deg, int32_t* S, uint64_t* sigma,
                uint8_t* ovf, uint64_t* fb0, uint64_t* fb1, uint64_t* fb2,
                uint64_t* q0, uint64_t* vis);
hipError_t level_prologue(hipStream_t s, int grid, uint64_t* fnext,
                          int32_t nwords, LevelCtr* ctr, CtrSet* cset,
                          const uint64_t* fcur, uint64_t* vis);
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
// plain streaming copy of n16 16-byte words (calibration)
hipError_t copy16(hipStream_t s, const void* src, void* dst, int64_t n16);
// setup helpers
hipError_t expand_cols(hipStream_t s, const int32_t* CP, int32_t n,
                       int32_t* col);
// slice tables and skip bitmap of a graph, from device list offsets ptr
// (lists longer than min_deg, cut into kPullChunk-entry slices; *nlists, zero
// on entry, receives the number of lists cut)
hipError_t slice_count(hipStream_t s, const int32_t* ptr, int32_t n, int32_t min_deg,
                       int32_t* cnt, int32_t* nlists);
hipError_t slice_fill(hipStream_t s, const int32_t* cnt, const int32_t* off, int32_t n,
                      uint64_t* out);
hipError_t skip_bits(hipStream_t s, const int32_t* ptr, int32_t n, uint64_t* skip,
                     int64_t nw);
// the reference entry points' output types (see k_dropin_convert)
hipError_t dropin_convert(hipStream_t s, const int32_t* S, const uint64_t* sigma,
                          const uint8_t* ovf, int32_t n, int32_t s_unreached, int32_t kind,
                          int32_t* S_out, uint32_t* val_out, unsigned long long* inexact);
// ptr[r] = first k with keys[k] >= r, r = 0 .. n (keys non-decreasing)
hipError_t offsets_sorted(hipStream_t s, const int32_t* keys, int32_t nz, int32_t n,
                          int32_t* ptr);
hipError_t degree_stats(hipStream_t s, const int32_t* ptr, int32_t n, DegAcc* acc);
// Degree-ordered numbering of a CSC graph (see graph_batch_view).
hipError_t order_keys(hipStream_t s, const int32_t* CP, int32_t n, uint32_t* keys, int32_t* ids);
hipError_t order_perm(hipStream_t s, const int32_t* order, const int32_t* CP, int32_t n,
                      int32_t* perm, int32_t* deg2);
hipError_t order_entries(hipStream_t s, const int32_t* IC, const int32_t* col, const int32_t* CP,
                         const int32_t* CP2, const int32_t* perm, int32_t nz, int32_t* IC2);
hipError_t check_csc(hipStream_t s, const int32_t* CP, const int32_t* IC,
                     int32_t n, int32_t nz, int32_t* bad);

// tbfs_persist.hip: load-balanced levels and the device-resident level loop
hipError_t push_lb(hipStream_t s, int grid, const GraphView& g,
                   const StateView& st, int32_t d, const uint64_t* bigo,
                   int32_t nbigo, int ovf_any);
hipError_t pull_lb(hipStream_t s, int grid, const GraphView& g,
                   const StateView& st, int32_t d, const uint64_t* bigc,
                   int32_t nbigc, int sigma, int ovf_any);
// Largest cooperative grid for the persistent kernel with `block` threads
// per workgroup (256 or 1024).
hipError_t persist_grid(int device, int block, int* grid);
hipError_t persist(hipStream_t s, int grid, int block, const PersistParams& p);
// Chained loop: one init kernel, then one kernel per level (level index d),
// each deciding its own direction from the device-side counters.
hipError_t chain_init(hipStream_t s, int grid, const PersistParams& p);
hipError_t chain_seed(hipStream_t s, const PersistParams& p);
// fbl_bytes > 0: 1024-thread workgroups staging the frontier bitmap
// (fbl_bytes of dynamic LDS) for pull levels; 0: kLbBlock-thread workgroups.
// want: which levels the launch runs (kLevelAlways, kLevelIfStaged, kLevelIfNotStaged).
hipError_t chain_level(hipStream_t s, int grid, const PersistParams& p, int32_t d,
                       size_t fbl_bytes, int32_t want = kLevelAlways);
// Largest bitmap (bytes) the LDS-staging variant can take; 0 if unavailable.
hipError_t chain_fbl_limit(size_t* bytes);
}  // namespace kern

constexpr int kBarWords = 512;  // grid barrier: top count, generation, 8 group counts
constexpr int kLbBlock = 256;   // threads per block of the load-balanced kernels
constexpr int kMaxBlock = 1024; // largest workgroup of any level kernel

}  // namespace tbfs
