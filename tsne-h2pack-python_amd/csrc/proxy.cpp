// proxy.cpp -- data-independent proxy point sets for the kernel K(r) = (1 + r^2)^-2.
//
// A node whose points lie in the box [-h, h]^D (about its centre) only interacts through
// its basis with points outside the max-norm ball of radius (1 + 2 eta) h (the
// admissibility rule of the H2 build).  Its skeleton is chosen by an interpolative
// decomposition of K(points, proxies); the proxies must span, up to the tolerance, the
// functions x -> K(x, y) over the box for every y in that far region.  They are
// selected from many candidates y (dense near the inner boundary, log-spaced outwards)
// by column-pivoted Gram-Schmidt on the column-normalised K(box samples, candidates) in
// FP64, stopping when the largest residual column norm falls below tol.
//
// The sets depend only on (D, h, tolerance), not on the data, so they are computed once
// per process for a fixed ladder of half-sides h_k = kProxyBase * 2^(k / kProxyQ) and
// cached.  The reference's proxy points come from H2Pack's generator (cited from
// src/tsne_h2pack.c:979-1060 of xinxing02/tsne-h2pack-python, library absent); this
// ladder plays the role of its box-size-quantised proxy file (src/tsne_h2pack.c:679-690).
//
// Persistent cache.  As the reference keeps its proxy file across runs
// (src/tsne_h2pack.c:1093-1146: TSNE_H2PACK_CACHE_DIR, else $XDG_CACHE_HOME/tsne_h2pack,
// else $HOME/.cache/tsne_h2pack; TSNE_H2PACK_NO_PROXY_CACHE=1 turns it off), a generated
// table is written to that directory and read back by later processes.  The file format
// is this library's own, little-endian:
//   "TSNEPRXY" | u32 version (1) | u32 D | u32 pmax | u32 levels | u32 ladder steps per
//   octave | u32 0 | f64 tolerance | f64 eta | f64 ladder base | f64 outer radius |
//   i32 count[levels] | f64 points[levels][pmax][D] | u64 FNV-1a of all bytes before it.
// A file whose header does not match the requested table exactly, or whose checksum
// fails, is ignored and replaced (written to a temporary name, then renamed).
#include "proxy.h"

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <mutex>
#include <random>
#include <thread>
#include <tuple>
#include <vector>

namespace tsne {

namespace {

inline double kern(const double *a, const double *b, int D)
{
    double r2 = 1.0;
    for (int d = 0; d < D; d++) {
        const double df = a[d] - b[d];
        r2 += df * df;
    }
    return 1.0 / (r2 * r2);
}

// Box samples: a tensor grid over [-h, h]^D.
std::vector<double> box_samples(int D, double h)
{
    const int g = (D == 2) ? 12 : 6;
    std::vector<double> x;
    if (D == 2) {
        for (int i = 0; i < g; i++)
            for (int j = 0; j < g; j++) {
                x.push_back(h * (-1.0 + 2.0 * i / (g - 1)));
                x.push_back(h * (-1.0 + 2.0 * j / (g - 1)));
            }
    } else {
        for (int i = 0; i < g; i++)
            for (int j = 0; j < g; j++)
                for (int k = 0; k < g; k++) {
                    x.push_back(h * (-1.0 + 2.0 * i / (g - 1)));
                    x.push_back(h * (-1.0 + 2.0 * j / (g - 1)));
                    x.push_back(h * (-1.0 + 2.0 * k / (g - 1)));
                }
    }
    return x;
}

// Candidates in the far region between max-norm radii r_in and r_out: a point on the
// unit max-norm sphere (uniform face, uniform position on it) scaled by a radius that is
// log-uniform with density concentrated towards r_in.
std::vector<double> annulus_candidates(int D, double r_in, double r_out, int count, uint64_t seed)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<double> y;
    y.reserve((size_t)count * D);
    const double lr = std::log(r_out / r_in);
    for (int c = 0; c < count; c++) {
        const double u = U(rng);
        const double rho = r_in * std::exp(lr * u * u * u);
        const int face = (int)(U(rng) * 2 * D) % (2 * D);
        double p[3];
        for (int d = 0; d < D; d++) p[d] = -1.0 + 2.0 * U(rng);
        p[face / 2] = (face & 1) ? 1.0 : -1.0;
        for (int d = 0; d < D; d++) y.push_back(rho * p[d]);
    }
    return y;
}

// Column-pivoted modified Gram-Schmidt on M (rows x cols, column-major); returns the
// pivot columns until max residual norm <= tol * max initial norm or pmax columns.
std::vector<int> select_columns(std::vector<double> &M, int rows, int cols, double tol, int pmax)
{
    std::vector<double> nrm(cols);
    for (int c = 0; c < cols; c++) {
        double s = 0.0;
        const double *mc = &M[(size_t)c * rows];
        for (int r = 0; r < rows; r++) s += mc[r] * mc[r];
        nrm[c] = s;
    }
    std::vector<int> piv;
    std::vector<char> used(cols, 0);
    double nrm0 = 0.0;
    std::vector<double> q(rows);
    while ((int)piv.size() < std::min({pmax, rows, cols})) {
        int best = -1;
        double bv = -1.0;
        for (int c = 0; c < cols; c++)
            if (!used[c] && nrm[c] > bv) { bv = nrm[c]; best = c; }
        if (piv.empty()) nrm0 = bv;
        if (best < 0 || !(bv > tol * tol * nrm0) || !(bv > 0.0)) break;
        used[best] = 1;
        piv.push_back(best);
        const double inv = 1.0 / std::sqrt(bv);
        const double *mb = &M[(size_t)best * rows];
        for (int r = 0; r < rows; r++) q[r] = mb[r] * inv;
        for (int c = 0; c < cols; c++) {
            if (used[c]) continue;
            double *mc = &M[(size_t)c * rows];
            double dot = 0.0;
            for (int r = 0; r < rows; r++) dot += q[r] * mc[r];
            double s = 0.0;
            for (int r = 0; r < rows; r++) {
                mc[r] -= dot * q[r];
                s += mc[r] * mc[r];
            }
            nrm[c] = s;
        }
    }
    return piv;
}

std::vector<double> make_set(int D, double h, double tol, double eta, int pmax, uint64_t seed)
{
    const std::vector<double> X = box_samples(D, h);
    const int nx = (int)X.size() / D;
    const double r_in = (1.0 + 2.0 * eta) * h;
    const double r_out = kProxyOuter * h;
    const int ny = (D == 2) ? 1600 : 2400;
    const std::vector<double> Y = annulus_candidates(D, r_in, r_out, ny, seed);
    // columns are normalised: every far direction is resolved to the same relative
    // accuracy, not only the ones closest to the box (which dominate unscaled norms)
    std::vector<double> M((size_t)nx * ny);
    for (int c = 0; c < ny; c++) {
        double s = 0.0;
        for (int r = 0; r < nx; r++) {
            const double v = kern(&X[(size_t)r * D], &Y[(size_t)c * D], D);
            M[(size_t)c * nx + r] = v;
            s += v * v;
        }
        const double inv = s > 0.0 ? 1.0 / std::sqrt(s) : 0.0;
        for (int r = 0; r < nx; r++) M[(size_t)c * nx + r] *= inv;
    }
    const std::vector<int> piv = select_columns(M, nx, ny, tol, pmax);
    std::vector<double> out;
    for (int c : piv)
        for (int d = 0; d < D; d++) out.push_back(Y[(size_t)c * D + d]);
    return out;
}

std::mutex g_mu;
std::map<std::tuple<int, long long, long long, int>, ProxyTable> g_cache;

constexpr char kMagic[8] = {'T', 'S', 'N', 'E', 'P', 'R', 'X', 'Y'};
constexpr uint32_t kFormatVersion = 1;

uint64_t fnv1a(const unsigned char *p, size_t n, uint64_t h = 0xcbf29ce484222325ULL)
{
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}

std::string cache_dir()
{
    const char *off = getenv("TSNE_H2PACK_NO_PROXY_CACHE");
    if (off && atoi(off) != 0) return "";
    const char *env = getenv("TSNE_H2PACK_CACHE_DIR"), *xdg = getenv("XDG_CACHE_HOME"),
               *home = getenv("HOME");
    std::string dir;
    if (env && env[0]) {
        dir = env;
    } else if (xdg && xdg[0]) {
        dir = std::string(xdg) + "/tsne_h2pack";
    } else if (home && home[0]) {
        mkdir((std::string(home) + "/.cache").c_str(), 0700);
        dir = std::string(home) + "/.cache/tsne_h2pack";
    }
    if (!dir.empty()) mkdir(dir.c_str(), 0700);  // an existing directory is fine
    return dir;
}

std::vector<unsigned char> serialise(const ProxyTable &t, double tol, double eta)
{
    std::vector<unsigned char> b;
    auto put = [&](const void *p, size_t n) {
        const unsigned char *c = (const unsigned char *)p;
        b.insert(b.end(), c, c + n);
    };
    const uint32_t hdr[6] = {kFormatVersion, (uint32_t)t.D, (uint32_t)t.pmax, (uint32_t)t.levels,
                             (uint32_t)kProxyQ, 0u};
    const double par[4] = {tol, eta, kProxyBase, kProxyOuter};
    put(kMagic, 8);
    put(hdr, sizeof hdr);
    put(par, sizeof par);
    put(t.cnt.data(), t.cnt.size() * sizeof(int));
    put(t.pts.data(), t.pts.size() * sizeof(double));
    const uint64_t h = fnv1a(b.data(), b.size());
    put(&h, sizeof h);
    return b;
}

// Reads a table for exactly (D, pmax, tol, eta) from `path`; false if absent or invalid.
bool load_table(const std::string &path, int D, int pmax, double tol, double eta, ProxyTable &t)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    std::vector<unsigned char> b;
    unsigned char buf[65536];
    size_t r;
    while ((r = fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + r);
    fclose(f);
    const size_t head = 8 + 6 * 4 + 4 * 8;
    const size_t body = (size_t)kProxyLevels * sizeof(int) + (size_t)kProxyLevels * pmax * D * sizeof(double);
    if (b.size() != head + body + 8 || memcmp(b.data(), kMagic, 8) != 0) return false;
    uint64_t h;
    memcpy(&h, b.data() + head + body, 8);
    if (h != fnv1a(b.data(), head + body)) return false;
    uint32_t hdr[6];
    double par[4];
    memcpy(hdr, b.data() + 8, sizeof hdr);
    memcpy(par, b.data() + 8 + sizeof hdr, sizeof par);
    if (hdr[0] != kFormatVersion || hdr[1] != (uint32_t)D || hdr[2] != (uint32_t)pmax ||
        hdr[3] != (uint32_t)kProxyLevels || hdr[4] != (uint32_t)kProxyQ || par[0] != tol || par[1] != eta ||
        par[2] != kProxyBase || par[3] != kProxyOuter)
        return false;
    t.D = D;
    t.pmax = pmax;
    t.levels = kProxyLevels;
    t.cnt.resize(kProxyLevels);
    t.pts.resize((size_t)kProxyLevels * pmax * D);
    memcpy(t.cnt.data(), b.data() + head, kProxyLevels * sizeof(int));
    memcpy(t.pts.data(), b.data() + head + kProxyLevels * sizeof(int), t.pts.size() * sizeof(double));
    for (int c : t.cnt)
        if (c < 0 || c > pmax) return false;
    return true;
}

void store_table(const std::string &path, const ProxyTable &t, double tol, double eta)
{
    const std::vector<unsigned char> b = serialise(t, tol, eta);
    const std::string tmp = path + "." + std::to_string((long)getpid()) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return;
    const bool ok = fwrite(b.data(), 1, b.size(), f) == b.size();
    if (fclose(f) != 0 || !ok || rename(tmp.c_str(), path.c_str()) != 0) unlink(tmp.c_str());
}

std::string cache_path(int D, double tol, double eta, int pmax)
{
    const std::string dir = cache_dir();
    if (dir.empty()) return "";
    char name[160];
    snprintf(name, sizeof name, "/proxy_amd_v%u_%dd_tol%.6e_eta%.3f_p%d.bin", kFormatVersion, D, tol, eta,
             pmax);
    return dir + name;
}

}  // namespace

std::string cache_dir_path() { return cache_dir(); }

std::string proxy_cache_file(int D, double tol, double eta, int pmax) { return cache_path(D, tol, eta, pmax); }

uint64_t proxy_table_digest(const ProxyTable &t)
{
    uint64_t h = fnv1a((const unsigned char *)t.cnt.data(), t.cnt.size() * sizeof(int));
    return fnv1a((const unsigned char *)t.pts.data(), t.pts.size() * sizeof(double), h);
}

double proxy_half_side(int k) { return kProxyBase * std::exp2((double)k / kProxyQ); }

int proxy_level_for(double h)
{
    if (!(h > kProxyBase)) return 0;
    int k = (int)std::ceil(std::log2(h / kProxyBase) * kProxyQ - 1e-9);
    while (k > 0 && proxy_half_side(k - 1) >= h) k--;
    while (proxy_half_side(k) < h) k++;
    return std::min(k, kProxyLevels - 1);
}

const ProxyTable &proxy_table(int D, double tol, double eta, int pmax)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const auto key = std::make_tuple(D, std::llround(tol * 1e12), std::llround(eta * 1e6), pmax);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) return it->second;
    ProxyTable t;
    const std::string path = cache_path(D, tol, eta, pmax);
    if (!path.empty() && load_table(path, D, pmax, tol, eta, t)) {
        t.from_disk = true;
        return g_cache.emplace(key, std::move(t)).first->second;
    }
    t.D = D;
    t.pmax = pmax;
    t.levels = kProxyLevels;
    t.pts.assign((size_t)kProxyLevels * pmax * D, 0.0);
    t.cnt.assign(kProxyLevels, 0);
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < hw; w++) {
        pool.emplace_back([&, w]() {
            for (int k = (int)w; k < kProxyLevels; k += (int)hw) {
                const std::vector<double> s =
                    make_set(D, proxy_half_side(k), tol, eta, pmax, 0x9e3779b97f4a7c15ULL + (uint64_t)k);
                const int c = (int)s.size() / D;
                t.cnt[k] = c;
                std::copy(s.begin(), s.end(), t.pts.begin() + (size_t)k * pmax * D);
            }
        });
    }
    for (auto &th : pool) th.join();
    if (!path.empty()) store_table(path, t, tol, eta);
    return g_cache.emplace(key, std::move(t)).first->second;
}

}  // namespace tsne
