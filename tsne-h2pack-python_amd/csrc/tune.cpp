// tune.cpp -- autotuner of the launch shapes of a job, and its persistent table.
//
// The tuned choices change how work is launched, never what is computed: the stream the
// attractive pass of a rebuild iteration runs on (overlap), the number of tree levels
// queued between two host checks (level batch), graph replay of force-reusing iterations,
// and the largest internal tree level whose IDs run in one launch of the widest size
// class (the tiled ID gives a node the same bits in every class).  A job's result is the
// same under every choice, bit for bit, and the tuner checks this on every trial.
//
// Table: one entry per (embedding dimension, size bucket round(log2 N)), kept in the cache
// directory of the proxy tables as text, one file per device kind (architecture name and
// CU count):
//     tsne_h2pack tuning v1 <arch> <CUs>
//     <dim> <bucket> <overlap> <level_batch> <graph> <one_class_level> <ms>
// tsne_init_with_P applies the entry of the job's bucket when there is one.  No
// counterpart in the reference (its CPU library has no launch shapes to choose).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <unistd.h>
#include <utility>
#include <vector>

#include "proxy.h"
#include "tune.h"

namespace tsne {

namespace {

constexpr int kTuneVersion = 1;

struct TuneTable {
    std::mutex mu;
    bool loaded = false;
    std::string path;  // "" when the cache is off or no device
    std::map<std::pair<int, int>, tsne_tuning_t> entries;
};

TuneTable &table()
{
    static TuneTable t;
    return t;
}

// Device kind of the current device: "<arch>_<CUs>cu" with the feature suffix of the
// architecture name dropped ("gfx950:sramecc+:xnack-" -> "gfx950"), or "" without one.
std::string device_kind(std::string *arch, int *cus)
{
    int dev = 0;
    hipDeviceProp_t pr;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&pr, dev) != hipSuccess) return "";
    std::string a = pr.gcnArchName;
    const size_t c = a.find(':');
    if (c != std::string::npos) a = a.substr(0, c);
    if (arch) *arch = a;
    if (cus) *cus = pr.multiProcessorCount;
    return a + "_" + std::to_string(pr.multiProcessorCount) + "cu";
}

// Loads the table of the current device kind once (the caller holds the lock).
void load_locked(TuneTable &t)
{
    if (t.loaded) return;
    t.loaded = true;
    std::string arch;
    int cus = 0;
    const std::string kind = device_kind(&arch, &cus);
    const std::string dir = cache_dir_path();
    if (kind.empty() || dir.empty()) return;
    t.path = dir + "/tune_amd_v" + std::to_string(kTuneVersion) + "_" + kind + ".txt";
    FILE *f = fopen(t.path.c_str(), "r");
    if (!f) return;
    char head[256] = {0};
    char a2[128] = {0};
    int ver = 0, cu2 = 0;
    if (!fgets(head, sizeof head, f) ||
        sscanf(head, "tsne_h2pack tuning v%d %127s %d", &ver, a2, &cu2) != 3 ||
        ver != kTuneVersion || arch != a2 || cu2 != cus) {
        fclose(f);
        return;
    }
    char line[256];
    while (fgets(line, sizeof line, f)) {
        int dim, b;
        tsne_tuning_t e{};
        if (sscanf(line, "%d %d %d %d %d %d %lf", &dim, &b, &e.overlap, &e.level_batch, &e.graph,
                   &e.one_class_level, &e.ms) != 7)
            continue;
        if ((dim != 2 && dim != 3) || b < 0 || b > 40 || (e.overlap != 0 && e.overlap != 1) ||
            e.level_batch < 1 || e.level_batch > 8 || (e.graph != 0 && e.graph != 1) ||
            e.one_class_level < 0)
            continue;
        t.entries[{dim, b}] = e;
    }
    fclose(f);
}

// Writes the whole table (the caller holds the lock): to a temporary file, then renamed.
void save_locked(TuneTable &t)
{
    if (t.path.empty()) return;
    std::string arch;
    int cus = 0;
    if (device_kind(&arch, &cus).empty()) return;
    const std::string tmp = t.path + "." + std::to_string((long)getpid()) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) return;
    bool ok = fprintf(f, "tsne_h2pack tuning v%d %s %d\n", kTuneVersion, arch.c_str(), cus) > 0;
    for (const auto &kv : t.entries) {
        const tsne_tuning_t &e = kv.second;
        ok = ok && fprintf(f, "%d %d %d %d %d %d %.4f\n", kv.first.first, kv.first.second, e.overlap,
                           e.level_batch, e.graph, e.one_class_level, e.ms) > 0;
    }
    if (fclose(f) != 0 || !ok || rename(tmp.c_str(), t.path.c_str()) != 0) unlink(tmp.c_str());
}

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

int tune_bucket(int n_point)
{
    return n_point < 1 ? 0 : (int)std::lround(std::log2((double)n_point));
}

tsne_tuning_t tune_defaults(int low_dim, int n_point)
{
    (void)low_dim;
    tsne_tuning_t t{};
    // the attractive pass of rebuild iterations beside the rebuild from 2^18 points on
    t.overlap = n_point >= (1 << 18) ? 1 : 0;
    t.level_batch = 4;
    t.graph = 0;
    t.one_class_level = -1;  // the dimension's default
    t.ms = 0.0;
    return t;
}

bool tune_lookup(int low_dim, int n_point, tsne_tuning_t *out)
{
    TuneTable &t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    load_locked(t);
    auto it = t.entries.find({low_dim, tune_bucket(n_point)});
    if (it == t.entries.end()) return false;
    if (out) *out = it->second;
    return true;
}

void tune_store(int low_dim, int n_point, const tsne_tuning_t &e)
{
    TuneTable &t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    load_locked(t);
    t.entries[{low_dim, tune_bucket(n_point)}] = e;
    save_locked(t);
}

std::string tune_path()
{
    TuneTable &t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    load_locked(t);
    return t.path;
}

}  // namespace tsne

using namespace tsne;

extern "C" {

int tsne_tuning_lookup(int low_dim, int n_point, tsne_tuning_t *out)
{
    if ((low_dim != 2 && low_dim != 3) || n_point < 1 || !out) return -1;
    *out = tune_defaults(low_dim, n_point);
    return tune_lookup(low_dim, n_point, out) ? 1 : 0;
}

int tsne_tuning_path(char *path, int path_len)
{
    if (!path || path_len < 1) return -1;
    const std::string p = tune_path();
    snprintf(path, (size_t)path_len, "%s", p.c_str());
    return p.empty() ? -1 : 0;
}

int tsne_autotune(int low_dim, int n_point, int n_iter, tsne_tuning_t *best)
{
    if ((low_dim != 2 && low_dim != 3) || n_point < 2000 || n_iter < 50) {
        fprintf(stderr, "tsne_autotune: low_dim must be 2 or 3, n_point >= 2000 (the H2 path) "
                        "and n_iter >= 50\n");
        return -1;
    }
    const int n = n_point, D = low_dim;
    // the synthetic job: a seeded 10-cluster Gaussian mixture in 16 dimensions, its exact
    // k = 90 graph and joint P at perplexity 30 (on the device), N(0, 1e-4) start
    constexpr int kFeat = 16, kClusters = 10;
    std::mt19937_64 rng(0x5eed0000ULL + (unsigned long long)n * 4 + (unsigned long long)D);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<double> centres((size_t)kClusters * kFeat), X((size_t)n * kFeat), Y0((size_t)D * n);
    for (double &c : centres) c = 5.0 * nd(rng);
    for (int i = 0; i < n; i++) {
        const int c = (int)(rng() % kClusters);
        for (int f = 0; f < kFeat; f++) X[(size_t)i * kFeat + f] = centres[(size_t)c * kFeat + f] + nd(rng);
    }
    for (double &y : Y0) y = 1e-2 * nd(rng);
    sparse_matrix_t *P = tsne_joint_P_from_points(X.data(), n, kFeat, std::min(90, n - 1), 30.0, nullptr);
    if (!P) {
        fprintf(stderr, "tsne_autotune: the synthetic job's joint P failed\n");
        return -1;
    }
    tsne_params_t prm = tsne_create_default_params();
    prm.n_iter = n_iter;
    prm.verbose = 0;
    prm.h2_tol = D == 2 ? 1e-2 : 3e-2;
    prm.h2_rebuild_freq = 3;
    prm.h2_rebuild_freq_early = 10;
    prm.h2_leaf_size = 100;
    prm.learning_rate = -1.0;
    prm.init_method = 1;
    prm.init_Y = Y0.data();
    prm.enable_profiling = 0;
    prm.profiling_output_file = nullptr;

    std::vector<double> Yref, Y((size_t)D * n);
    bool mismatch = false, failed = false;
    // one timed tsne_run of the synthetic job under cfg (ms), or a negative value
    auto trial = [&](const tsne_tuning_t &cfg) -> double {
        sparse_matrix_t *Pc = sparse_matrix_create(P->n_rows, P->n_cols, P->nnz);
        if (!Pc) return -1.0;
        memcpy(Pc->row_ptr, P->row_ptr, sizeof(int) * ((size_t)P->n_rows + 1));
        memcpy(Pc->col_idx, P->col_idx, sizeof(int) * (size_t)P->nnz);
        memcpy(Pc->values, P->values, sizeof(double) * (size_t)P->nnz);
        tsne_workspace_t *ws = tsne_init_with_P(nullptr, n, 0, D, &prm, Pc);
        if (!ws) {
            sparse_matrix_destroy(Pc);
            return -1.0;
        }
        double ms = -1.0;
        if (tsne_set_tuning(ws, &cfg) == 0 && hipDeviceSynchronize() == hipSuccess) {
            const double t0 = now_ms();
            if (tsne_run(ws) == 0) {
                ms = now_ms() - t0;
                tsne_get_embedding(ws, Y.data());
                if (Yref.empty())
                    Yref = Y;
                else if (memcmp(Yref.data(), Y.data(), Y.size() * sizeof(double)) != 0)
                    mismatch = true;
            }
        }
        tsne_destroy(ws);
        return ms;
    };
    auto timed = [&](const tsne_tuning_t &cfg) -> double {
        double b = 1e300;
        for (int r = 0; r < 2 && !failed; r++) {
            const double ms = trial(cfg);
            if (ms < 0) failed = true;
            b = std::min(b, ms);
        }
        return b;
    };
    // coordinate descent from the defaults, one choice at a time
    tsne_tuning_t cur = tune_defaults(D, n);
    cur.level_batch = 4;
    cur.one_class_level = D == 2 ? 512 : 2048;
    double cur_ms = timed(cur);
    const std::vector<int> one_class = D == 2 ? std::vector<int>{64, 512, 4096}
                                              : std::vector<int>{256, 2048, 16384};
    auto descend = [&](int tsne_tuning_t::*field, const std::vector<int> &values) {
        for (int v : values) {
            if (failed || cur.*field == v) continue;
            tsne_tuning_t c = cur;
            c.*field = v;
            const double ms = timed(c);
            if (!failed && ms < cur_ms) {
                cur = c;
                cur_ms = ms;
            }
        }
    };
    descend(&tsne_tuning_t::one_class_level, one_class);
    descend(&tsne_tuning_t::overlap, {0, 1});
    descend(&tsne_tuning_t::level_batch, {1, 2, 4, 8});
    descend(&tsne_tuning_t::graph, {0, 1});
    sparse_matrix_destroy(P);
    if (failed) {
        fprintf(stderr, "tsne_autotune: a trial run failed\n");
        return -1;
    }
    if (mismatch) {
        fprintf(stderr, "tsne_autotune: launch shapes changed a result (not stored)\n");
        return -1;
    }
    cur.ms = cur_ms;
    tune_store(D, n, cur);
    if (best) *best = cur;
    return 0;
}

}  // extern "C"
