// proxy.h -- proxy point sets of the H2 build (see proxy.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace tsne {

// Ladder of node half-sides: h_k = kProxyBase * 2^(k / kProxyQ), k = 0..kProxyLevels-1.
// The root box's half-side is rounded up onto the ladder, so every quadtree level
// falls exactly on it (the reference rounds its root radius to minL * 2^k in the same
// way, src/tsne_h2pack.c:679-690).
constexpr double kProxyBase = 1.0 / (1 << 20);
constexpr int kProxyQ = 1;
constexpr int kProxyLevels = 48;  // half-sides 2^-20 .. 2^27
// Outer max-norm radius of the proxy candidates, in units of the half-side.
constexpr double kProxyOuter = 48.0;

struct ProxyTable {
    int D = 0, pmax = 0, levels = 0;
    std::vector<double> pts;  // [levels][pmax][D], relative to the node centre
    std::vector<int> cnt;     // proxies per level
    bool from_disk = false;   // read from the persistent cache instead of generated
};

double proxy_half_side(int k);
// Smallest ladder index whose half-side is >= h.
int proxy_level_for(double h);
// Cached table for (D, tolerance, eta, pmax); thread safe.  Read from the persistent
// cache file when a valid one exists, else generated and written there.
const ProxyTable &proxy_table(int D, double tol, double eta, int pmax);
// Path of the persistent cache file of that table ("" when the cache is off).
std::string proxy_cache_file(int D, double tol, double eta, int pmax);
// FNV-1a of the table's counts and points.
uint64_t proxy_table_digest(const ProxyTable &t);
// The persistent cache directory (created if missing; "" when the cache is off).
std::string cache_dir_path();

}  // namespace tsne
