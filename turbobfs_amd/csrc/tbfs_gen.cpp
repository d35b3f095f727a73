// tbfs_gen.cpp -- deterministic synthetic graphs for the benchmark and test
// workloads (host C++).  Every generator uses only integer arithmetic on a
// splitmix64 stream (probabilities become 32-bit integer thresholds once), so
// the same seed gives the same graph on every machine and thread count.
//
// Output: a CSC with in-lists (column v lists every u with u -> v), rows
// ascending inside each column, no duplicates, no self-loops; all generated
// graphs are undirected, so CSC == CSR.
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "tbfs_gen.h"

struct tbfs_csc_host {
  int32_t n = 0;
  int32_t nz = 0;
  std::vector<int32_t> CP;
  std::vector<int32_t> IC;
};

namespace {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next() {
    s += 0x9e3779b97f4a7c15ull;
    return mix64(s);
  }
  // uniform in [0, bound) by Lemire's multiply-shift with rejection
  uint32_t below(uint32_t bound) {
    uint64_t m = (uint64_t)(uint32_t)(next() >> 32) * bound;
    uint32_t l = (uint32_t)m;
    if (l < bound) {
      const uint32_t t = (uint32_t)(-bound) % bound;
      while (l < t) {
        m = (uint64_t)(uint32_t)(next() >> 32) * bound;
        l = (uint32_t)m;
      }
    }
    return (uint32_t)(m >> 32);
  }
};

// probability -> threshold on a 32-bit draw (draw < thr  <=>  success)
inline uint64_t thr32(double p) {
  if (p <= 0) return 0;
  if (p >= 1) return 1ull << 32;
  return (uint64_t)(p * 4294967296.0 + 0.5);
}

// LSD radix sort of 64-bit keys on the low `bits` bits, 11-bit digits.
void radix_sort(std::vector<uint64_t>& a, int bits) {
  std::vector<uint64_t> b(a.size());
  const int D = 11;
  const size_t R = 1u << D;
  std::vector<size_t> cnt(R);
  for (int sh = 0; sh < bits; sh += D) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (uint64_t x : a) cnt[(x >> sh) & (R - 1)]++;
    size_t acc = 0;
    for (size_t i = 0; i < R; ++i) {
      const size_t c = cnt[i];
      cnt[i] = acc;
      acc += c;
    }
    for (uint64_t x : a) b[cnt[(x >> sh) & (R - 1)]++] = x;
    a.swap(b);
  }
}

int bits_for(uint64_t v) {
  int b = 1;
  while (b < 63 && (1ull << b) <= v) ++b;
  return b;
}

// Undirected edge list (u, v) -> symmetric CSC without self-loops/duplicates.
tbfs_csc_host* build_sym_csc(int32_t n, const std::vector<uint32_t>& eu,
                             const std::vector<uint32_t>& ev) {
  std::vector<uint64_t> keys;
  keys.reserve(eu.size() * 2);
  for (size_t i = 0; i < eu.size(); ++i) {
    const uint64_t u = eu[i], v = ev[i];
    if (u == v) continue;
    keys.push_back((v << 32) | u);  // u -> v lands in column v
    keys.push_back((u << 32) | v);
  }
  radix_sort(keys, 32 + bits_for((uint64_t)n));
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  if (keys.size() > 0x7fffffffull) return nullptr;
  tbfs_csc_host* g = new (std::nothrow) tbfs_csc_host();
  if (!g) return nullptr;
  g->n = n;
  g->nz = (int32_t)keys.size();
  g->CP.assign((size_t)n + 1, 0);
  g->IC.resize(keys.size());
  for (size_t k = 0; k < keys.size(); ++k) {
    g->CP[(keys[k] >> 32) + 1]++;
    g->IC[k] = (int32_t)(keys[k] & 0xffffffffu);
  }
  for (int32_t i = 0; i < n; ++i) g->CP[i + 1] += g->CP[i];
  return g;
}

std::vector<uint32_t> permutation(int32_t n, uint64_t seed) {
  std::vector<uint32_t> p((size_t)n);
  for (int32_t i = 0; i < n; ++i) p[i] = (uint32_t)i;
  SplitMix rng(mix64(seed ^ 0x5eed5eed5eed5eedull));
  for (int32_t i = n - 1; i > 0; --i) {
    const uint32_t j = rng.below((uint32_t)i + 1);
    std::swap(p[i], p[j]);
  }
  return p;
}

}  // namespace

extern "C" {

int tbfs_gen_er(int32_t n, double p, uint64_t seed, tbfs_csc_host** out) {
  if (!out || n <= 1 || !(p >= 0 && p <= 1)) return 1;
  const uint64_t t = thr32(p);
  SplitMix rng(mix64(seed));
  std::vector<uint32_t> eu, ev;
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = i + 1; j < n; ++j)
      if ((rng.next() >> 32) < t) {
        eu.push_back((uint32_t)i);
        ev.push_back((uint32_t)j);
      }
  *out = build_sym_csc(n, eu, ev);
  return *out ? 0 : 3;
}

int tbfs_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                  double c, uint64_t seed, int32_t relabel,
                  tbfs_csc_host** out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1) return 1;
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1)) return 1;
  const int32_t n = (int32_t)(1u << scale);
  const uint64_t m = (uint64_t)edge_factor << scale;
  const uint64_t ta = thr32(a), tb = thr32(a + b), tc = thr32(a + b + c);
  std::vector<uint32_t> eu(m), ev(m);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)m; ++k) {
    SplitMix rng(mix64(seed * 0x9e3779b97f4a7c15ull + (uint64_t)k));
    uint32_t u = 0, v = 0;
    uint64_t bitsrc = 0;
    for (int32_t lvl = 0; lvl < scale; ++lvl) {
      if ((lvl & 1) == 0) bitsrc = rng.next();
      const uint64_t r = (lvl & 1) ? (bitsrc & 0xffffffffull) : (bitsrc >> 32);
      uint32_t bu, bv;
      if (r < ta) { bu = 0; bv = 0; }
      else if (r < tb) { bu = 0; bv = 1; }
      else if (r < tc) { bu = 1; bv = 0; }
      else { bu = 1; bv = 1; }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    eu[k] = u;
    ev[k] = v;
  }
  if (relabel) {
    const std::vector<uint32_t> p = permutation(n, seed);
    for (uint64_t k = 0; k < m; ++k) {
      eu[k] = p[eu[k]];
      ev[k] = p[ev[k]];
    }
  }
  *out = build_sym_csc(n, eu, ev);
  return *out ? 0 : 3;
}

int tbfs_gen_grid(int32_t w, int32_t h, double remove_frac, uint64_t seed,
                  int32_t relabel, tbfs_csc_host** out) {
  if (!out || w < 1 || h < 1 || (int64_t)w * h > 0x7fffffffll) return 1;
  if (!(remove_frac >= 0 && remove_frac <= 1)) return 1;
  const int64_t nh = (int64_t)h * (w - 1), nv = (int64_t)w * (h - 1);
  const int64_t E = nh + nv;
  const int64_t R = (int64_t)std::llround(remove_frac * (double)E);
  // edge e < nh: horizontal (x, y)-(x+1, y), e = y*(w-1) + x
  // edge e >= nh: vertical (x, y)-(x, y+1), e - nh = y*w + x
  std::vector<uint32_t> idx((size_t)E);
  for (int64_t e = 0; e < E; ++e) idx[e] = (uint32_t)e;
  SplitMix rng(mix64(seed ^ 0x6772696467726964ull));
  for (int64_t i = 0; i < R; ++i) {  // partial Fisher-Yates: idx[0, R) removed
    const uint32_t j = (uint32_t)i + rng.below((uint32_t)(E - i));
    std::swap(idx[i], idx[j]);
  }
  std::vector<uint32_t> eu, ev;
  eu.reserve((size_t)(E - R));
  ev.reserve((size_t)(E - R));
  for (int64_t i = R; i < E; ++i) {
    const int64_t e = idx[i];
    uint32_t u, v;
    if (e < nh) {
      const int64_t y = e / (w - 1), x = e % (w - 1);
      u = (uint32_t)(y * w + x);
      v = u + 1;
    } else {
      const int64_t f = e - nh, y = f / w, x = f % w;
      u = (uint32_t)(y * w + x);
      v = u + (uint32_t)w;
    }
    eu.push_back(u);
    ev.push_back(v);
  }
  const int32_t n = w * h;
  if (relabel) {
    const std::vector<uint32_t> p = permutation(n, seed);
    for (size_t k = 0; k < eu.size(); ++k) {
      eu[k] = p[eu[k]];
      ev[k] = p[ev[k]];
    }
  }
  *out = build_sym_csc(n, eu, ev);
  return *out ? 0 : 3;
}

// Watts-Strogatz small world: ring lattice of n vertices, each joined to its
// k/2 nearest neighbours on either side; then, lattice offset by offset, each
// edge (u, u+j) is rewired with probability p to (u, w), w uniform among the
// vertices that are neither u nor already adjacent to u.
int tbfs_gen_ws(int32_t n, int32_t k, double p, uint64_t seed,
                tbfs_csc_host** out) {
  if (!out || n < 3 || k < 2 || k >= n || (k & 1)) return 1;
  if (!(p >= 0 && p <= 1)) return 1;
  const int32_t half = k / 2;
  std::vector<std::vector<uint32_t>> adj((size_t)n);
  for (int32_t u = 0; u < n; ++u) adj[u].reserve((size_t)k + 4);
  auto has = [&](uint32_t u, uint32_t v) {
    for (uint32_t x : adj[u])
      if (x == v) return true;
    return false;
  };
  auto del = [&](uint32_t u, uint32_t v) {
    auto& a = adj[u];
    for (size_t i = 0; i < a.size(); ++i)
      if (a[i] == v) { a[i] = a.back(); a.pop_back(); return; }
  };
  for (int32_t u = 0; u < n; ++u)
    for (int32_t j = 1; j <= half; ++j) {
      const uint32_t v = (uint32_t)((u + j) % n);
      adj[u].push_back(v);
      adj[v].push_back((uint32_t)u);
    }
  const uint64_t t = thr32(p);
  SplitMix rng(mix64(seed ^ 0x7773777377737773ull));
  for (int32_t j = 1; j <= half; ++j)
    for (int32_t u = 0; u < n; ++u) {
      if ((rng.next() >> 32) >= t) continue;
      const uint32_t v = (uint32_t)((u + j) % n);
      if (!has((uint32_t)u, v)) continue;  // already rewired away
      if ((int64_t)adj[u].size() >= (int64_t)n - 1) continue;
      uint32_t w = rng.below((uint32_t)n);
      while (w == (uint32_t)u || has((uint32_t)u, w)) w = rng.below((uint32_t)n);
      del((uint32_t)u, v);
      del(v, (uint32_t)u);
      adj[u].push_back(w);
      adj[w].push_back((uint32_t)u);
    }
  std::vector<uint32_t> eu, ev;
  for (int32_t u = 0; u < n; ++u)
    for (uint32_t v : adj[u])
      if ((uint32_t)u < v) {
        eu.push_back((uint32_t)u);
        ev.push_back(v);
      }
  *out = build_sym_csc(n, eu, ev);
  return *out ? 0 : 3;
}

int tbfs_csc_host_info(const tbfs_csc_host* g, int32_t* n, int32_t* nz) {
  if (!g) return 1;
  if (n) *n = g->n;
  if (nz) *nz = g->nz;
  return 0;
}

int tbfs_csc_host_copy(const tbfs_csc_host* g, int32_t* CP, int32_t* IC) {
  if (!g) return 1;
  if (CP) std::memcpy(CP, g->CP.data(), ((size_t)g->n + 1) * sizeof(int32_t));
  if (IC && g->nz > 0) std::memcpy(IC, g->IC.data(), (size_t)g->nz * sizeof(int32_t));
  return 0;
}

void tbfs_csc_host_free(tbfs_csc_host* g) { delete g; }

}  // extern "C"
