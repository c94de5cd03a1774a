// tune.cpp -- autotuner switch and layout cache (see tune.h).
#include "tune.h"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

namespace sssi {
namespace {

using Key = std::tuple<int, int64_t, int64_t, int64_t, int, int>;

std::atomic<int> g_mode{-1};  // -1: not yet read from the environment
std::mutex g_mu;
std::map<Key, std::pair<int, int>> g_cache;
bool g_loaded = false;

Key key_of(const TuneKey &k) { return Key{k.dim, k.nz, k.nx, k.ny, k.N, k.nB}; }

// Lines "dim nz nx ny N nB a b"; unreadable lines are skipped.
void load_file_locked() {
    if (g_loaded) return;
    g_loaded = true;
    const char *path = std::getenv("SSSI_TUNE_CACHE");
    if (!path) return;
    FILE *f = std::fopen(path, "r");
    if (!f) return;
    int dim, N, nB, a, b;
    long long nz, nx, ny;
    while (std::fscanf(f, "%d %lld %lld %lld %d %d %d %d", &dim, &nz, &nx, &ny, &N, &nB, &a,
                       &b) == 8)
        g_cache[Key{dim, nz, nx, ny, N, nB}] = {a, b};
    std::fclose(f);
}

}  // namespace

bool autotune_enabled() {
    int m = g_mode.load();
    if (m < 0) {
        const char *e = std::getenv("SSSI_AUTOTUNE");
        m = (e && std::atoi(e) != 0) ? 1 : 0;
        g_mode.store(m);
    }
    return m == 1;
}

void set_autotune(bool on) { g_mode.store(on ? 1 : 0); }

bool tune_lookup(const TuneKey &k, int &a, int &b) {
    std::lock_guard<std::mutex> lk(g_mu);
    load_file_locked();
    auto it = g_cache.find(key_of(k));
    if (it == g_cache.end()) return false;
    a = it->second.first;
    b = it->second.second;
    return true;
}

void tune_store(const TuneKey &k, int a, int b) {
    std::lock_guard<std::mutex> lk(g_mu);
    load_file_locked();
    g_cache[key_of(k)] = {a, b};
    if (const char *path = std::getenv("SSSI_TUNE_CACHE")) {
        if (FILE *f = std::fopen(path, "a")) {
            std::fprintf(f, "%d %lld %lld %lld %d %d %d %d\n", k.dim, (long long)k.nz,
                         (long long)k.nx, (long long)k.ny, k.N, k.nB, a, b);
            std::fclose(f);
        }
    }
}

}  // namespace sssi
