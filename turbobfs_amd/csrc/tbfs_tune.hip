// tbfs_tune.hip -- the autotuner's table: launch shapes and direction
// thresholds that timing chose for a class of graphs.
//
// An entry is keyed by the kind of run (single-source BFS, or a batch of w
// 64-source words) and the graph's in-degree statistics as powers of two
// (floor log2 + 1 of n, of the maximum in-degree and of twice the mean), the
// features the reference's drivers print before a run
// (TurboBFS_CSC_TD/degree.c:38-141).  Runs whose options leave the tuned
// fields at 0 take the entry that matches their graph.  Every tuned value
// changes only how fast a run goes: the results are the same for all.
//
// The table lives in the process; tbfs_tune_save / tbfs_tune_load write and
// read it as JSON, and with TBFS_TUNE_CACHE set to a path the table is read
// from that file on first use and written back after each autotune.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "tbfs.h"
#include "tbfs_engine.h"

namespace tbfs {
namespace {

std::mutex g_mu;
std::vector<TuneEntry> g_table;
bool g_env_loaded = false;

int ilog2p1(double x) { return x < 1.0 ? 0 : (int)std::floor(std::log2(x)) + 1; }

const char* env_path() {
  const char* p = std::getenv("TBFS_TUNE_CACHE");
  return (p && p[0]) ? p : nullptr;
}

// ---- JSON (the format this file writes: one object per entry) -------------
std::string to_json(const std::vector<TuneEntry>& t) {
  std::string s = "{\"version\": 1, \"entries\": [";
  char buf[512];
  for (size_t i = 0; i < t.size(); ++i) {
    const TuneEntry& e = t[i];
    std::snprintf(buf, sizeof buf,
                  "%s\n  {\"words\": %d, \"log2_n\": %d, \"log2_max_in_degree\": %d, "
                  "\"log2_twice_mean_in_degree\": %d, \"grid_per_cu\": %d, \"batch_alpha\": %.9g, "
                  "\"alpha\": %.9g, \"beta\": %.9g, \"stage_ratio\": %.9g, \"lb_grid\": %d, "
                  "\"ms\": %.9g, \"default_ms\": %.9g}",
                  i ? "," : "", e.key.w, e.key.ln, e.key.lmax, e.key.lmean, e.grid_per_cu,
                  (double)e.batch_alpha, (double)e.alpha, (double)e.beta, (double)e.stage_ratio,
                  e.lb_grid, e.ms, e.default_ms);
    s += buf;
  }
  s += "\n]}\n";
  return s;
}

// Value of "key": in the object text [b, e); false when absent.
bool field(const char* b, const char* e, const char* key, double* v) {
  const std::string pat = std::string("\"") + key + "\"";
  const char* p = b;
  for (;;) {
    p = static_cast<const char*>(memmem(p, (size_t)(e - p), pat.data(), pat.size()));
    if (!p) return false;
    p += pat.size();
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    if (p < e && *p == ':') break;
  }
  ++p;
  char* end = nullptr;
  *v = std::strtod(p, &end);
  return end != p && end <= e;
}

int from_json(const std::string& s, std::vector<TuneEntry>* out) {
  const char* ent = std::strstr(s.c_str(), "\"entries\"");
  if (!ent) return TBFS_ERR_INVALID;
  const char* p = std::strchr(ent, '[');
  if (!p) return TBFS_ERR_INVALID;
  const char* stop = s.c_str() + s.size();
  std::vector<TuneEntry> t;
  for (;;) {
    const char* b = std::strchr(p, '{');
    const char* close = std::strchr(p, ']');
    if (!b || (close && close < b)) break;
    const char* e = std::strchr(b, '}');
    if (!e || e > stop) return TBFS_ERR_INVALID;
    double w, ln, lmax, lmean;
    if (!field(b, e, "words", &w) || !field(b, e, "log2_n", &ln) ||
        !field(b, e, "log2_max_in_degree", &lmax) || !field(b, e, "log2_twice_mean_in_degree", &lmean))
      return TBFS_ERR_INVALID;
    TuneEntry x{};
    x.key = TuneKey{(int32_t)w, (int32_t)ln, (int32_t)lmax, (int32_t)lmean};
    double v;
    if (field(b, e, "grid_per_cu", &v)) x.grid_per_cu = (int32_t)v;
    if (field(b, e, "batch_alpha", &v)) x.batch_alpha = (float)v;
    if (field(b, e, "alpha", &v)) x.alpha = (float)v;
    if (field(b, e, "beta", &v)) x.beta = (float)v;
    if (field(b, e, "stage_ratio", &v)) x.stage_ratio = (float)v;
    if (field(b, e, "lb_grid", &v)) x.lb_grid = (int32_t)v;
    if (field(b, e, "ms", &v)) x.ms = v;
    if (field(b, e, "default_ms", &v)) x.default_ms = v;
    if (x.grid_per_cu < 0 || x.batch_alpha < 0 || x.alpha < 0 || x.beta < 0 ||
        x.stage_ratio < 0 || x.lb_grid < -1)
      return TBFS_ERR_INVALID;
    t.push_back(x);
    p = e + 1;
  }
  *out = std::move(t);
  return TBFS_OK;
}

int save_locked(const char* path) {
  FILE* f = std::fopen(path, "w");
  if (!f) return TBFS_ERR_INVALID;
  const std::string s = to_json(g_table);
  const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
  return (std::fclose(f) == 0 && ok) ? TBFS_OK : TBFS_ERR_INVALID;
}

int load_locked(const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) return TBFS_ERR_INVALID;
  std::string s;
  char buf[4096];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, k);
  std::fclose(f);
  std::vector<TuneEntry> t;
  const int rc = from_json(s, &t);
  if (rc) return rc;
  for (const TuneEntry& x : t) {
    bool found = false;
    for (TuneEntry& y : g_table)
      if (y.key == x.key) {
        y = x;
        found = true;
      }
    if (!found) g_table.push_back(x);
  }
  return TBFS_OK;
}

// first use: the cache file named by TBFS_TUNE_CACHE, if any (a missing or
// unreadable file leaves the table as it is)
void env_load_locked() {
  if (g_env_loaded) return;
  g_env_loaded = true;
  if (const char* p = env_path()) (void)load_locked(p);
}

}  // namespace

int tune_key(tbfs_graph* g, int32_t words, hipStream_t s, TuneKey* k) {
  tbfs_degree_stats ds;
  const int rc = graph_degree_stats(g, TBFS_DEGREE_IN, s, &ds);
  if (rc) return rc;
  k->w = words;
  k->ln = ilog2p1((double)g->n);
  k->lmax = ilog2p1((double)ds.max);
  k->lmean = ilog2p1(2.0 * ds.mean);
  return TBFS_OK;
}

bool tune_find(const TuneKey& k, TuneEntry* e) {
  std::lock_guard<std::mutex> lk(g_mu);
  env_load_locked();
  for (const TuneEntry& t : g_table)
    if (t.key == k) {
      *e = t;
      return true;
    }
  return false;
}

void tune_put(const TuneEntry& e) {
  std::lock_guard<std::mutex> lk(g_mu);
  env_load_locked();
  bool found = false;
  for (TuneEntry& t : g_table)
    if (t.key == e.key) {
      t = e;
      found = true;
    }
  if (!found) g_table.push_back(e);
  if (const char* p = env_path()) (void)save_locked(p);
}

bool tune_empty() {
  std::lock_guard<std::mutex> lk(g_mu);
  env_load_locked();
  return g_table.empty();
}

}  // namespace tbfs

extern "C" {

int tbfs_tune_save(const char* path) {
  if (!path) return TBFS_ERR_INVALID;
  std::lock_guard<std::mutex> lk(tbfs::g_mu);
  return tbfs::save_locked(path);
}

int tbfs_tune_load(const char* path) {
  if (!path) return TBFS_ERR_INVALID;
  std::lock_guard<std::mutex> lk(tbfs::g_mu);
  tbfs::g_env_loaded = true;
  return tbfs::load_locked(path);
}

int tbfs_tune_clear(void) {
  std::lock_guard<std::mutex> lk(tbfs::g_mu);
  tbfs::g_env_loaded = true;
  tbfs::g_table.clear();
  return TBFS_OK;
}

int tbfs_tune_count(void) {
  std::lock_guard<std::mutex> lk(tbfs::g_mu);
  tbfs::env_load_locked();
  return (int)tbfs::g_table.size();
}

}  // extern "C"
