// Tuned choices per problem shape.  The host tuner (libnegf_amd/tune.py)
// times the candidates on the device and fills this table, or loads a table
// it wrote before; the launchers consult it whenever no variant is forced
// (negf_gemm_variant / negf_inv_variant = 0) and fall back to their built-in
// thresholds for shapes the table does not hold.  This is the role of the
// reference's timing-based opt_params (src/matinvertndiag_ddrgf_utils.jl:
// 357-453), extended to the GEMM tile and the inverse schedule.
//
// Kinds and keys (the last key is the batch-like one):
//   gemm    m n k problems   -> tile variant (zgemm.hip's variant numbers)
//   inv     n batch          -> inverse schedule variant (zinv.hip's)
//   bndiag  N bmax batch     -> DD-RGF blockSizeD1 per level; [0] = RGF
//   getrs   n nrhs           -> 1: blocked substitution, 2: triangular
//                               inverses and two products (getri: n n)
//   trsm    na nrhs          -> 1: blocked substitution, 2: the triangle's
//                               inverse and one product
// Lookup: the other keys must match; of the entries for them, the one with
// the largest batch key <= the request, else the smallest above it.
#include "../../include/negf_hip.h"
#include "negf_internal.hpp"

#include <atomic>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

namespace negf {

namespace {

constexpr int kMaxKeys = 4, kMaxChoice = 4;

struct Entry {
  int kind;
  int nkey;
  int key[kMaxKeys];
  int nchoice;
  int choice[kMaxChoice];
};

std::mutex g_mu;
std::vector<Entry> g_table;
std::atomic<int> g_count{0};

int nkeys_of(int kind) {
  return kind == 0 ? 4 : (kind == 1 ? 2 : (kind == 2 ? 3 : ((kind == 3 || kind == 4) ? 2 : 0)));
}

int kind_of(const std::string& w) {
  if (w == "gemm") return 0;
  if (w == "inv") return 1;
  if (w == "bndiag") return 2;
  if (w == "getrs") return 3;
  if (w == "trsm") return 4;
  return -1;
}

int set_locked(int kind, const int* key, int nkey, const int* choice, int nchoice) {
  if (nkeys_of(kind) == 0 || nkey != nkeys_of(kind) || nchoice < 1 || nchoice > kMaxChoice)
    return Status::BAD_ARG;
  for (int i = 0; i < nkey; ++i)
    if (key[i] < 0) return Status::BAD_ARG;
  for (auto& e : g_table) {
    if (e.kind == kind && std::memcmp(e.key, key, sizeof(int) * nkey) == 0) {
      e.nchoice = nchoice;
      std::memcpy(e.choice, choice, sizeof(int) * nchoice);
      return Status::OK;
    }
  }
  Entry e{};
  e.kind = kind;
  e.nkey = nkey;
  std::memcpy(e.key, key, sizeof(int) * nkey);
  e.nchoice = nchoice;
  std::memcpy(e.choice, choice, sizeof(int) * nchoice);
  g_table.push_back(e);
  g_count.store((int)g_table.size(), std::memory_order_release);
  return Status::OK;
}

// number of choice values written, 0 when the table holds nothing for the key
int lookup(int kind, const int* key, int nkey, int* choice, int maxchoice) {
  if (g_count.load(std::memory_order_acquire) == 0) return 0;
  std::lock_guard<std::mutex> lk(g_mu);
  const Entry* below = nullptr;
  const Entry* above = nullptr;
  const int last = nkey - 1;
  for (const auto& e : g_table) {
    if (e.kind != kind || e.nkey != nkey) continue;
    if (std::memcmp(e.key, key, sizeof(int) * last) != 0) continue;
    if (e.key[last] <= key[last]) {
      if (!below || e.key[last] > below->key[last]) below = &e;
    } else if (!above || e.key[last] < above->key[last]) {
      above = &e;
    }
  }
  const Entry* e = below ? below : above;
  if (!e) return 0;
  const int n = e->nchoice < maxchoice ? e->nchoice : maxchoice;
  std::memcpy(choice, e->choice, sizeof(int) * n);
  return n;
}

}  // namespace

int tune_gemm_variant(int m, int n, int k, long long problems) {
  const int key[4] = {m, n, k, problems > 0x7fffffff ? 0x7fffffff : (int)problems};
  int v = 0;
  return lookup(0, key, 4, &v, 1) ? v : 0;
}

int tune_inv_variant(int n, int batch) {
  const int key[2] = {n, batch};
  int v = 0;
  return lookup(1, key, 2, &v, 1) ? v : 0;
}

int tune_getrs_route(int n, int nrhs) {
  const int key[2] = {n, nrhs};
  int v = 0;
  return lookup(3, key, 2, &v, 1) ? v : 0;
}

int tune_trsm_route(int na, int nrhs) {
  const int key[2] = {na, nrhs};
  int v = 0;
  return lookup(4, key, 2, &v, 1) ? v : 0;
}

}  // namespace negf

using namespace negf;

extern "C" {

int negf_tune_set(int32_t kind, const int32_t* key, int32_t nkey, const int32_t* choice,
                  int32_t nchoice) {
  if (!key || !choice) return NEGF_BAD_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  return set_locked(kind, key, nkey, choice, nchoice);
}

int negf_tune_lookup(int32_t kind, const int32_t* key, int32_t nkey, int32_t* choice,
                     int32_t maxchoice) {
  if (!key || !choice || maxchoice < 1 || nkey != nkeys_of(kind) || nkey == 0) return NEGF_BAD_ARG;
  return lookup(kind, key, nkey, choice, maxchoice);
}

int negf_tune_clear(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_table.clear();
  g_count.store(0, std::memory_order_release);
  return NEGF_OK;
}

int negf_tune_count(void) { return g_count.load(std::memory_order_acquire); }

int negf_tune_load(const char* path) {
  if (!path) return NEGF_BAD_ARG;
  std::ifstream f(path);
  if (!f) return NEGF_BAD_ARG;
  std::vector<Entry> add;
  std::string line;
  while (std::getline(f, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream in(line);
    std::string word;
    if (!(in >> word)) continue;
    const int kind = kind_of(word);
    if (kind < 0) return NEGF_BAD_ARG;
    Entry e{};
    e.kind = kind;
    e.nkey = nkeys_of(kind);
    for (int i = 0; i < e.nkey; ++i)
      if (!(in >> e.key[i]) || e.key[i] < 0) return NEGF_BAD_ARG;
    int v;
    while (e.nchoice < kMaxChoice && (in >> v)) e.choice[e.nchoice++] = v;
    if (e.nchoice == 0) return NEGF_BAD_ARG;
    if (in >> v) return NEGF_BAD_ARG;   // more choice values than a kind holds
    add.push_back(e);
  }
  std::lock_guard<std::mutex> lk(g_mu);
  for (const auto& e : add) {
    const int rc = set_locked(e.kind, e.key, e.nkey, e.choice, e.nchoice);
    if (rc) return rc;
  }
  return (int)add.size();
}

}  // extern "C"
