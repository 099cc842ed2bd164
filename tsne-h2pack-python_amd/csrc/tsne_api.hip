// tsne_api.hip -- the C ABI of libtsne_h2pack.so (include/tsne_h2pack.h) and the host
// side of the optimisation loop.  Device work is queued on one HIP stream per
// workspace; the loop synchronises only where the host needs a value (rebuild checks,
// verbose prints, the end of a run).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <mutex>
#include <string>
#include <vector>

#include "api_internal.h"
#include "common.h"
#include "h2.h"
#include "joint_p.h"
#include "kernels.h"
#include "renumber.h"
#include "tiles.h"
#include "tune.h"
#include "../../include/tsne_h2pack.h"

using namespace tsne;

struct tsne_workspace {
    int n = 0, pt_dim = 0, dim = 0;
    int device = 0;             // the HIP device the workspace's memory and streams are on
    tsne_params_t params{};
    std::string profile_file;
    double lr = 200.0;
    sparse_matrix_t *P = nullptr;
    hipStream_t st = nullptr;
    // second stream: the attractive pass of a rebuild iteration runs on it next to the
    // rebuild and the product (same default priority as st)
    hipStream_t st2 = nullptr;
    hipEvent_t ev_y = nullptr, ev_a = nullptr;
    float *d_A = nullptr;       // attractive sums of a split iteration, n x D
    int overlap = 1;
    // launch shapes of the H2 builds (tsne_set_tuning / the tuning table): tree levels per
    // host check (0: default) and the one-launch ID level limit (-1: default)
    int tune_level_batch = 0, tune_one_class = -1;
    int nnz = 0;
    int *d_rowptr = nullptr, *d_col = nullptr;
    float *d_val32 = nullptr;
    double *d_val64 = nullptr;
    double *d_Y64 = nullptr, *d_uY = nullptr, *d_gains = nullptr;
    void *d_Yg[2] = {nullptr, nullptr};
    int cur = 0;
    float *d_s = nullptr, *d_F = nullptr;
    double *d_scal = nullptr;
    double *d_block_sums = nullptr;
    int nblocks_upd = 0;
    double *d_partial = nullptr;
    int nblocks_red = 0;
    double *d_rowkl = nullptr;
    int rep_valid = 0;
    double kl = 0.0;
    int precision = TSNE_H2_PLAIN;
    H2 *h2 = nullptr;
    int h2_ok = 0;  // last rebuild produced a usable H2 structure
    // timings of the last run (ms) and counters
    double t_attr = 0, t_rebuild = 0, t_product = 0, t_other = 0;
    int n_rebuild = 0;
    int timing = 0;             // 1: per-phase events (synchronising), 2: attractive kernel only
    double t_attr_kernel = 0;   // timing 2: summed attractive-kernel time (ms) and launches
    int n_attr_kernel = 0;
    double t_attr_tiled = 0;    // the part of it spent in the tiled kernel, and its launches
    int n_attr_tiled = 0;
    double knn_time = 0, p_time = 0;
    std::vector<double> prof;  // per-iteration phase times when profiling
    // renumbering along the tree order (reference src/tsne_h2pack.c:1706-1791)
    std::vector<int> order;     // current point k is the caller's point order[k] (host)
    int *d_order = nullptr;     // same on the device
    std::vector<int> reorder_iters;
    int next_reorder = 0;
    int n_renumber = 0;
    RenumberCsr P2;             // second CSR buffer set, alive while renumberings remain
    RenumberScratch rs;
    void *d_tmp = nullptr;      // n * 4 doubles: permutation staging
    // row tiles of P for the attractive pass, built after each renumbering (ntiles == 0:
    // the points are still in the caller's order and the per-row kernel runs)
    AttrTiles tiles;
    int tile_rows = 0;          // 0: tiles off
    int attr_after_front = 0;   // rebuild iterations: attractive sums queued after the build's front
    int attr_split_grid = 0;    // workgroups of that split attractive launch (0: one per block)
    // check mode (TSNE_DIRECT_REUSE=1): the direct O(N^2) sum in place of the H2 product, on
    // the H2 path's rebuild cadence and force reuse -- the exact counterpart of an H2 run,
    // differing from it only by the H2 approximation
    int direct_reuse = 0;
    int tile_lds = 0;           // LDS bytes per tile workgroup for the staged records
    int tile_nt = 256;          // threads per tile workgroup
    // graph mode: iterations that reuse the repulsive force (attractive + update + centring,
    // two kernels) are replayed from HIP graphs captured on first use, one per (gather
    // buffer parity, exaggeration, momentum, learning rate); renumbering (which swaps P's
    // buffers) and tile builds drop them
    int graph_mode = 0;
    struct GraphSlot {
        int cur;
        double exag, mom, lr;
        hipGraphExec_t exec;
    };
    std::vector<GraphSlot> graphs;
    int n_graph_launches = 0;
};

namespace {

const size_t kGatherBytes[4] = {0, 0, sizeof(float2), sizeof(float4)};

// Device memory is stream-ordered (hipMallocAsync / hipFreeAsync on the job's stream):
// allocating or releasing one job's buffers never waits for the other jobs on the device.
template <typename T> int dalloc(T **p, size_t count, hipStream_t st)
{
    *p = nullptr;
    if (count == 0) count = 1;
    if (hipMallocAsync((void **)p, count * sizeof(T), st) != hipSuccess) {
        fprintf(stderr, "tsne_h2pack: device allocation of %zu bytes failed\n", count * sizeof(T));
        *p = nullptr;
        return -1;
    }
    return 0;
}

template <typename T> void dfree(T *&p, hipStream_t st)
{
    if (p) (void)hipFreeAsync((void *)p, st);
    p = nullptr;
}

bool use_direct(const tsne_workspace *ws)
{
    // N < 2000 and method='exact' (rebuild frequency > 1e5) take the direct sum, as the
    // reference does (src/tsne_h2pack.c:404-405, :1378); so does the exact-with-reuse
    // check mode (TSNE_DIRECT_REUSE=1)
    return ws->n < 2000 || ws->params.h2_rebuild_freq > 100000 || ws->direct_reuse;
}

int rebuild_iter(const tsne_workspace *ws, int it, int reuse)
{
    // src/tsne_h2pack.c:1765-1776
    int freq = ws->params.h2_rebuild_freq > 0 ? ws->params.h2_rebuild_freq : 10;
    int eff = (it < 20) ? 1 : freq;
    if (it >= 20 && it < ws->params.early_exag_iter && reuse &&
        ws->params.h2_rebuild_freq_early > eff)
        eff = ws->params.h2_rebuild_freq_early;
    return it > 0 && it % eff == 0;
}

// Seeded N(0, 1e-4) Box-Muller from glibc rand(), src/tsne_utils.c:1129-1148 (dim x n).
void init_random(double *Y, int n, int dim, unsigned seed)
{
    srand(seed);
    const int m = n * dim;
    for (int i = 0; i < m; i += 2) {
        double u1 = (double)rand() / RAND_MAX;
        double u2 = (double)rand() / RAND_MAX;
        if (u1 < 1e-10) u1 = 1e-10;
        const double rad = std::sqrt(-2.0 * std::log(u1));
        Y[i] = 1e-4 * rad * std::cos(2.0 * M_PI * u2);
        if (i + 1 < m) Y[i + 1] = 1e-4 * rad * std::sin(2.0 * M_PI * u2);
    }
}

void free_renumber_buffers(tsne_workspace *ws)
{
    dfree(ws->P2.rp, ws->st); dfree(ws->P2.col, ws->st); dfree(ws->P2.v32, ws->st); dfree(ws->P2.v64, ws->st);
    dfree(ws->rs.inv, ws->st); dfree(ws->rs.len, ws->st); dfree(ws->rs.src, ws->st); dfree(ws->rs.src_sorted, ws->st);
    dfree(ws->rs.col_tmp, ws->st);
    if (ws->rs.tmp) { (void)hipFreeAsync(ws->rs.tmp, ws->st); ws->rs.tmp = nullptr; }
    if (ws->d_tmp) { (void)hipFreeAsync(ws->d_tmp, ws->st); ws->d_tmp = nullptr; }
}

void drop_graphs(tsne_workspace *ws)
{
    for (auto &g : ws->graphs) (void)hipGraphExecDestroy(g.exec);
    ws->graphs.clear();
}

void free_device(tsne_workspace *ws)
{
    drop_graphs(ws);
    free_renumber_buffers(ws);
    tiles_free(ws->tiles, ws->st);
    dfree(ws->d_order, ws->st);
    if (ws->h2) { h2_destroy(ws->h2); ws->h2 = nullptr; }
    dfree(ws->d_rowptr, ws->st); dfree(ws->d_col, ws->st); dfree(ws->d_val32, ws->st); dfree(ws->d_val64, ws->st);
    dfree(ws->d_Y64, ws->st); dfree(ws->d_uY, ws->st); dfree(ws->d_gains, ws->st);
    dfree(ws->d_Yg[0], ws->st); dfree(ws->d_Yg[1], ws->st);
    dfree(ws->d_s, ws->st); dfree(ws->d_F, ws->st); dfree(ws->d_scal, ws->st); dfree(ws->d_A, ws->st);
    dfree(ws->d_block_sums, ws->st); dfree(ws->d_partial, ws->st); dfree(ws->d_rowkl, ws->st);
    if (ws->st2) { (void)hipStreamSynchronize(ws->st2); (void)hipStreamDestroy(ws->st2); ws->st2 = nullptr; }
    if (ws->ev_y) { (void)hipEventDestroy(ws->ev_y); ws->ev_y = nullptr; }
    if (ws->ev_a) { (void)hipEventDestroy(ws->ev_a); ws->ev_a = nullptr; }
    if (ws->st) { (void)hipStreamDestroy(ws->st); ws->st = nullptr; }
}

// Upload an embedding given dim x n in the caller's order: Y64 interleaved, gather copy.
int upload_embedding(tsne_workspace *ws, const double *Ydn)
{
    const int n = ws->n, D = ws->dim;
    std::vector<double> Yi((size_t)n * D);
    std::vector<float> g((size_t)n * (D == 2 ? 2 : 4), 0.f);
    const int gs = (D == 2) ? 2 : 4;
    for (int i = 0; i < n; i++) {
        const int o = ws->order.empty() ? i : ws->order[i];  // current point i = caller's o
        for (int d = 0; d < D; d++) {
            Yi[(size_t)i * D + d] = Ydn[(size_t)d * n + o];
            g[(size_t)i * gs + d] = (float)Ydn[(size_t)d * n + o];
        }
    }
    TSNE_HIP_CHECK(hipMemcpyAsync(ws->d_Y64, Yi.data(), Yi.size() * sizeof(double),
                                  hipMemcpyHostToDevice, ws->st));
    TSNE_HIP_CHECK(hipMemcpyAsync(ws->d_Yg[ws->cur], g.data(), g.size() * sizeof(float),
                                  hipMemcpyHostToDevice, ws->st));
    // the pending mean starts at zero: Y64 holds the embedding as given
    std::vector<double> z(SLOT_COUNT, 0.0);
    z[SLOT_Z] = 1.0;
    TSNE_HIP_CHECK(hipMemcpyAsync(ws->d_scal, z.data(), z.size() * sizeof(double),
                                  hipMemcpyHostToDevice, ws->st));
    TSNE_HIP_CHECK(hipStreamSynchronize(ws->st));
    ws->rep_valid = 0;
    return 0;
}

// Repulsive sums s, F for the current gather copy (H2 product or direct sum).
int repulsive_sums(tsne_workspace *ws, int allow_h2)
{
    const void *Yg = ws->d_Yg[ws->cur];
    if (allow_h2 && ws->h2 && ws->h2_ok)
        return h2_product(ws->h2, Yg, ws->d_s, ws->d_F, ws->precision == TSNE_H2_MIXED);
    return launch_repulsive_direct_f32(ws->dim, ws->n, Yg, ws->d_s, ws->d_F, ws->st);
}

// The pending-mean snapshot and Z of the last repulsive sums.
int repulsive_z(tsne_workspace *ws)
{
    if (launch_shift_snapshot(ws->d_scal, ws->st)) return -1;
    return launch_z(ws->dim, ws->n, ws->d_Yg[ws->cur], ws->d_s, ws->d_F, ws->d_scal,
                    ws->d_partial, ws->nblocks_red, ws->st);
}

int repulsive(tsne_workspace *ws, int allow_h2)
{
    const int rc = repulsive_sums(ws, allow_h2);
    return rc ? rc : repulsive_z(ws);
}

// The attractive sums of a rebuild iteration on the second stream (into d_A), after
// what the main stream has queued so far.
int queue_split_attr(void *arg)
{
    tsne_workspace *ws = static_cast<tsne_workspace *>(arg);
    TSNE_HIP_CHECK(hipEventRecord(ws->ev_y, ws->st));
    TSNE_HIP_CHECK(hipStreamWaitEvent(ws->st2, ws->ev_y, 0));
    if (launch_attr_only(ws->dim, ws->n, ws->d_rowptr, ws->d_col, ws->d_val32, ws->d_Yg[ws->cur],
                         ws->d_A, ws->nblocks_upd, ws->attr_split_grid, ws->st2) != 0)
        return -1;
    TSNE_HIP_CHECK(hipEventRecord(ws->ev_a, ws->st2));
    return 0;
}

int rebuild(tsne_workspace *ws)
{
    if (!ws->h2) {
        ws->h2 = h2_create(ws->n, ws->dim, ws->params.h2_leaf_size > 0 ? ws->params.h2_leaf_size : 100,
                           ws->params.h2_tol > 0 ? ws->params.h2_tol : 1e-6, ws->st);
        if (!ws->h2) return -1;
        if (ws->st2) h2_set_aux_stream(ws->h2, ws->st2);
        h2_set_tuning(ws->h2, ws->tune_level_batch, ws->tune_one_class);
    }
    int bad = 0;
    h2_set_id_mixed(ws->h2, ws->precision == TSNE_H2_MIXED);
    h2_set_timing(ws->h2, ws->timing == 1);
    if (h2_build(ws->h2, ws->d_Yg[ws->cur], &bad) != 0) {
        if (bad) fprintf(stderr, "tsne_h2pack: the embedding contains NaN or Inf\n");
        ws->h2_ok = 0;
        return -1;
    }
    ws->h2_ok = h2_usable(ws->h2);
    return 0;
}

// Renumber the points along the tree order of the last H2 build: new point k is the
// current point ord[k].  P, every per-point array and the caller order are permuted
// (reference tsne_renumber_points, src/tsne_reorder.c:112-167).  The caller rebuilds the
// H2 structure right after.
int renumber(tsne_workspace *ws)
{
    const int n = ws->n, D = ws->dim, nnz = ws->nnz;
    const int *ord = h2_order(ws->h2);
    hipStream_t st = ws->st;
    if (!ws->P2.rp) {
        ws->P2.n = n;
        ws->P2.nnz = nnz;
        ws->rs.tmp_bytes = renumber_tmp_bytes(n, nnz);
        const bool ok = dalloc(&ws->P2.rp, (size_t)n + 1, ws->st) == 0 && dalloc(&ws->P2.col, (size_t)nnz + kCsrPad, ws->st) == 0 &&
                        dalloc(&ws->P2.v32, (size_t)nnz + kCsrPad, ws->st) == 0 && dalloc(&ws->P2.v64, nnz, ws->st) == 0 &&
                        dalloc(&ws->rs.inv, n, ws->st) == 0 && dalloc(&ws->rs.len, n, ws->st) == 0 &&
                        dalloc(&ws->rs.src, nnz, ws->st) == 0 && dalloc(&ws->rs.src_sorted, nnz, ws->st) == 0 &&
                        dalloc(&ws->rs.col_tmp, nnz, ws->st) == 0 &&
                        hipMallocAsync(&ws->rs.tmp, ws->rs.tmp_bytes, ws->st) == hipSuccess &&
                        hipMallocAsync(&ws->d_tmp, (size_t)n * 4 * sizeof(double), ws->st) == hipSuccess;
        if (!ok) {
            free_renumber_buffers(ws);
            fprintf(stderr, "WARNING: point renumbering skipped (out of device memory)\n");
            return 1;
        }
    }
    RenumberCsr cur;
    cur.n = n;
    cur.nnz = nnz;
    cur.rp = ws->d_rowptr;
    cur.col = ws->d_col;
    cur.v32 = ws->d_val32;
    cur.v64 = ws->d_val64;
    drop_graphs(ws);  // they hold P's current buffers
    if (renumber_csr(cur, ord, ws->P2, ws->rs, st) != 0) return -1;
    std::swap(ws->d_rowptr, ws->P2.rp);
    std::swap(ws->d_col, ws->P2.col);
    std::swap(ws->d_val32, ws->P2.v32);
    std::swap(ws->d_val64, ws->P2.v64);
    // the attractive pass's row tiles follow the new numbering (a failed build leaves
    // the per-row kernel in use)
    if (ws->tile_rows > 0 &&
        tiles_build(ws->tiles, n, nnz, ws->d_rowptr, ws->d_col, ws->tile_rows, D,
                    tile_ucap(D, ws->tile_lds), st) != 0) {
        (void)hipGetLastError();
        fprintf(stderr, "WARNING: attractive row tiles not built (out of device memory?)\n");
    }
    // per-point state: permute into the staging buffer, copy back
    struct Arr { void *p; int width, bytes; };
    const Arr arrs[] = {{ws->d_Y64, D, 8}, {ws->d_uY, D, 8}, {ws->d_gains, D, 8},
                        {ws->d_Yg[ws->cur], D == 2 ? 2 : 4, 4}, {ws->d_F, D, 4},
                        {ws->d_order, 1, 4}};
    for (const Arr &a : arrs) {
        if (permute_points(n, a.width, ord, a.p, ws->d_tmp, a.bytes, st) != 0) return -1;
        TSNE_HIP_CHECK(hipMemcpyAsync(a.p, ws->d_tmp, (size_t)n * a.width * a.bytes,
                                      hipMemcpyDeviceToDevice, st));
    }
    // host copy of the caller order
    std::vector<int> ho(n);
    TSNE_HIP_CHECK(hipMemcpyAsync(ho.data(), ws->d_order, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    TSNE_HIP_CHECK(hipStreamSynchronize(st));
    ws->order.swap(ho);
    ws->n_renumber++;
    return 0;
}

// Current point index of each caller point.
std::vector<int> inverse_order(const tsne_workspace *ws)
{
    std::vector<int> inv(ws->n);
    for (int k = 0; k < ws->n; k++) inv[ws->order[k]] = k;
    return inv;
}

double read_scalar(tsne_workspace *ws, int slot)
{
    double v = 0.0;
    if (hipMemcpyAsync(&v, ws->d_scal + slot, sizeof(double), hipMemcpyDeviceToHost, ws->st) !=
            hipSuccess ||
        hipStreamSynchronize(ws->st) != hipSuccess)
        return NAN;
    return v;
}

// One force-reusing iteration (attractive + update + centring) from a captured graph.
int graph_step(tsne_workspace *ws, double exag, double mom, double lr)
{
    const int nxt = ws->cur ^ 1;
    tsne_workspace::GraphSlot *slot = nullptr;
    for (auto &g : ws->graphs)
        if (g.cur == ws->cur && g.exag == exag && g.mom == mom && g.lr == lr) slot = &g;
    if (!slot) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (hipStreamBeginCapture(ws->st, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return -1;
        const int lrc = launch_attr_update(ws->dim, ws->n, ws->d_rowptr, ws->d_col, ws->d_val32,
                                           ws->d_Yg[ws->cur], ws->d_F, ws->d_Y64, ws->d_uY,
                                           ws->d_gains, ws->d_Yg[nxt], ws->d_scal, exag, mom, lr,
                                           ws->d_block_sums, ws->nblocks_upd, ws->st);
        const hipError_t ec = hipStreamEndCapture(ws->st, &graph);
        if (lrc != 0 || ec != hipSuccess || !graph ||
            hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            fprintf(stderr, "tsne_h2pack: graph capture of an iteration failed\n");
            return -1;
        }
        (void)hipGraphDestroy(graph);
        if (ws->graphs.size() >= 8) drop_graphs(ws);  // the schedule uses a handful
        ws->graphs.push_back({ws->cur, exag, mom, lr, exec});
        slot = &ws->graphs.back();
    }
    return hipGraphLaunch(slot->exec, ws->st) == hipSuccess ? 0 : -1;
}

// Columns of the per-iteration profile (seconds).
enum { PROF_ATTR = 0, PROF_REBUILD, PROF_PRODUCT, PROF_Z, PROF_UPDATE, PROF_OTHER, kProfCols };

struct EvPair {
    hipEvent_t a = nullptr, b = nullptr;
};

}  // namespace

// ---------------------------------------------------------------------------------
// Scratch of tsne_exact_repulsive_rows_noalloc: one file-scope device array per GPU
// (this code object's copy on each device), so that the entry point allocates nothing.
// Host code reaches its memory only through the address hipGetSymbolAddress returns
// for the device, fetched once per device and kept in g_scratch_ptr; the array's name is
// used for nothing else.  Every call on a device shares that device's array: calls that
// overlap on one GPU race.
// ---------------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = 48u << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

void *g_scratch_ptr[kNoAllocMaxDevices] = {nullptr};
std::mutex g_scratch_mu;

// The device address of `scratch` on `device` (which must be current), or nullptr.
void *scratch_ptr(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "tsne_h2pack: hipGetSymbolAddress(scratch) failed on device %d: %s\n",
                    device, hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// inv[order[k]] = k: current index of every caller point.
__global__ void __launch_bounds__(256) invert_order_kernel(int n, const int *__restrict__ order,
                                                           int *__restrict__ inv)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) inv[order[k]] = k;
}

// Caller rows to current rows.
__global__ void __launch_bounds__(256) map_rows_kernel(int n_rows, const int *__restrict__ rows,
                                                       const int *__restrict__ inv,
                                                       int *__restrict__ out)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n_rows) out[r] = inv[rows[r]];
}

}  // namespace

extern "C" {

tsne_params_t tsne_create_default_params(void)
{
    // src/tsne_h2pack.c:56-106
    tsne_params_t p;
    memset(&p, 0, sizeof(p));
    p.n_iter = 1000;
    p.perplexity = 30.0;
    p.learning_rate = -1.0;
    p.initial_momentum = 0.5;
    p.momentum = 0.8;
    p.momentum_switch_iter = 250;
    p.early_exag = 12.0;
    p.early_exag_iter = 250;
    p.use_vptree = 2;
    p.h2_tol = 1e-3;
    p.h2_rebuild_freq = 10;
    p.h2pack_print_level = 0;
    p.init_method = 0;
    p.init_Y = nullptr;
    p.random_seed = (int)(unsigned)time(nullptr);
    p.verbose = 1;
    p.print_freq = 50;
    p.enable_early_stopping = 0;
    p.convergence_tol = 1e-5;
    p.convergence_patience = 10;
    p.min_iter = 250;
    p.sklearn_compatible = 1;
    p.enable_profiling = 0;
    p.profiling_output_file = nullptr;
    p.num_threads = 1;
    p.h2_leaf_size = 100;
    p.h2_rebuild_freq_early = 10;
    return p;
}

sparse_matrix_t *sparse_matrix_create(int n_rows, int n_cols, int nnz)
{
    if (n_rows < 0 || nnz < 0) return nullptr;
    sparse_matrix_t *m = (sparse_matrix_t *)malloc(sizeof(sparse_matrix_t));
    if (!m) return nullptr;
    m->n_rows = n_rows;
    m->n_cols = n_cols;
    m->nnz = nnz;
    m->row_ptr = (int *)malloc(((size_t)n_rows + 1) * sizeof(int));
    m->col_idx = (int *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    m->values = (double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    if (!m->row_ptr || !m->col_idx || !m->values) {
        sparse_matrix_destroy(m);
        return nullptr;
    }
    return m;
}

void sparse_matrix_destroy(sparse_matrix_t *m)
{
    if (!m) return;
    free(m->row_ptr);
    free(m->col_idx);
    free(m->values);
    free(m);
}

sparse_matrix_t *compute_P_from_knn_distances(const double *knn_distances, const int *knn_indices,
                                              const int *knn_indptr, int n_samples, int nnz,
                                              double perplexity, int num_threads)
{
    (void)num_threads;
    bool args_ok = knn_distances && knn_indices && knn_indptr && n_samples > 0 && nnz >= 0 &&
                   knn_indptr[0] == 0 && knn_indptr[n_samples] == nnz && perplexity > 0.0;
    for (int i = 0; args_ok && i < n_samples; i++) args_ok = knn_indptr[i] <= knn_indptr[i + 1];
    if (!args_ok) {
        fprintf(stderr, "compute_P_from_knn_distances: invalid arguments\n");
        return nullptr;
    }
    sparse_matrix_t *P = sparse_matrix_create(n_samples, n_samples, nnz);
    if (!P) return nullptr;
    memcpy(P->row_ptr, knn_indptr, ((size_t)n_samples + 1) * sizeof(int));
    if (nnz > 0) memcpy(P->col_idx, knn_indices, (size_t)nnz * sizeof(int));
    if (nnz == 0) return P;
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        fprintf(stderr, "compute_P_from_knn_distances: no HIP device available\n");
        sparse_matrix_destroy(P);
        return nullptr;
    }
    double *d_dist = nullptr, *d_p = nullptr;
    int *d_ptr = nullptr;
    int ok = dalloc(&d_dist, nnz, st) == 0 && dalloc(&d_p, nnz, st) == 0 &&
             dalloc(&d_ptr, (size_t)n_samples + 1, st) == 0;
    ok = ok &&
         hipMemcpyAsync(d_dist, knn_distances, (size_t)nnz * sizeof(double),
                        hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(d_ptr, knn_indptr, ((size_t)n_samples + 1) * sizeof(int),
                        hipMemcpyHostToDevice, st) == hipSuccess &&
         launch_perplexity(n_samples, d_dist, d_ptr, perplexity, d_p, st) == 0 &&
         hipMemcpyAsync(P->values, d_p, (size_t)nnz * sizeof(double), hipMemcpyDeviceToHost,
                        st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    dfree(d_dist, st); dfree(d_p, st); dfree(d_ptr, st);
    (void)hipStreamDestroy(st);
    if (!ok) {
        fprintf(stderr, "compute_P_from_knn_distances: device failure\n");
        sparse_matrix_destroy(P);
        return nullptr;
    }
    return P;
}

sparse_matrix_t *tsne_joint_P_from_knn_distances(const double *knn_distances,
                                                 const int *knn_indices, const int *knn_indptr,
                                                 int n_samples, int nnz, double perplexity)
{
    bool args_ok = knn_distances && knn_indices && knn_indptr && n_samples > 0 && nnz >= 0 &&
                   knn_indptr[0] == 0 && knn_indptr[n_samples] == nnz && perplexity > 0.0;
    for (int i = 0; args_ok && i < n_samples; i++) {
        args_ok = knn_indptr[i] <= knn_indptr[i + 1];
        for (int e = knn_indptr[i]; args_ok && e < knn_indptr[i + 1]; e++)
            args_ok = knn_indices[e] >= 0 && knn_indices[e] < n_samples &&
                      (e == knn_indptr[i] || knn_indices[e] > knn_indices[e - 1]);
    }
    if (!args_ok) {
        fprintf(stderr, "tsne_joint_P_from_knn_distances: invalid arguments (rows must have "
                        "increasing columns in [0, n))\n");
        return nullptr;
    }
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        fprintf(stderr, "tsne_joint_P_from_knn_distances: no HIP device available\n");
        return nullptr;
    }
    double *d_dist = nullptr, *d_p = nullptr;
    int *d_ptr = nullptr, *d_col = nullptr;
    JointP J;
    sparse_matrix_t *P = nullptr;
    bool ok = dalloc(&d_dist, nnz, st) == 0 && dalloc(&d_p, nnz, st) == 0 &&
              dalloc(&d_ptr, (size_t)n_samples + 1, st) == 0 && dalloc(&d_col, nnz, st) == 0 &&
              hipMemcpyAsync(d_dist, knn_distances, (size_t)nnz * sizeof(double),
                             hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(d_col, knn_indices, (size_t)nnz * sizeof(int),
                             hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(d_ptr, knn_indptr, ((size_t)n_samples + 1) * sizeof(int),
                             hipMemcpyHostToDevice, st) == hipSuccess &&
              (nnz == 0 || launch_perplexity(n_samples, d_dist, d_ptr, perplexity, d_p, st) == 0) &&
              joint_P_device(n_samples, nnz, d_ptr, d_col, d_p, &J, st) == 0;
    if (ok) {
        P = sparse_matrix_create(n_samples, n_samples, J.nnz);
        ok = P && hipMemcpyAsync(P->row_ptr, J.rp, ((size_t)n_samples + 1) * sizeof(int),
                                 hipMemcpyDeviceToHost, st) == hipSuccess &&
             (J.nnz == 0 ||
              (hipMemcpyAsync(P->col_idx, J.col, (size_t)J.nnz * sizeof(int),
                              hipMemcpyDeviceToHost, st) == hipSuccess &&
               hipMemcpyAsync(P->values, J.val, (size_t)J.nnz * sizeof(double),
                              hipMemcpyDeviceToHost, st) == hipSuccess)) &&
             hipStreamSynchronize(st) == hipSuccess;
    }
    dfree(d_dist, st); dfree(d_p, st); dfree(d_ptr, st); dfree(d_col, st);
    dfree(J.rp, st); dfree(J.col, st); dfree(J.val, st);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (!ok) {
        fprintf(stderr, "tsne_joint_P_from_knn_distances: device failure\n");
        if (P) sparse_matrix_destroy(P);
        return nullptr;
    }
    return P;
}

double tsne_numpy_sum(const double *values, long long n)
{
    return (values && n > 0) ? numpy_sum(values, n) : 0.0;
}

tsne_workspace_t *tsne_init(const double *X, int n_point, int pt_dim, int low_dim,
                            const tsne_params_t *params)
{
    (void)X; (void)n_point; (void)pt_dim; (void)low_dim; (void)params;
    fprintf(stderr,
            "tsne_init: the features path (built-in k-NN search) is not part of this library; "
            "compute k-NN distances, call compute_P_from_knn_distances, symmetrise, and use "
            "tsne_init_with_P\n");
    return nullptr;
}

tsne_workspace_t *tsne_init_with_P(const double *X, int n_point, int pt_dim, int low_dim,
                                   const tsne_params_t *params, sparse_matrix_t *P)
{
    (void)X;
    bool taken = false;
    return tsne::init_with_P(n_point, pt_dim, low_dim, params, P, &taken);
}

}  // extern "C"

namespace tsne {

// Body of tsne_init_with_P.  *taken tells whether P now belongs to the library: it does
// exactly when a workspace is returned.  On every NULL return P is left untouched and
// still belongs to the caller (also when the initial H2 build fails after the workspace
// existed: P is detached before that workspace is destroyed).
tsne_workspace_t *init_with_P(int n_point, int pt_dim, int low_dim, const tsne_params_t *params,
                              sparse_matrix_t *P, bool *taken, const tsne_tuning_t *shapes)
{
    *taken = false;
    if (low_dim != 2 && low_dim != 3) {
        fprintf(stderr, "Error: low_dim must be 2 or 3 (got %d)\n", low_dim);
        return nullptr;
    }
    if (!params || !P || P->n_rows != n_point || P->n_cols != n_point || n_point < 2) {
        fprintf(stderr, "Error: Invalid P matrix (expected %dx%d, got %dx%d)\n", n_point, n_point,
                P ? P->n_rows : 0, P ? P->n_cols : 0);
        return nullptr;
    }
    const int nnz = P->row_ptr[n_point];
    bool csr_ok = nnz >= 0 && P->row_ptr[0] == 0;
    for (int i = 0; csr_ok && i < n_point; i++) csr_ok = P->row_ptr[i] <= P->row_ptr[i + 1];
    for (int e = 0; csr_ok && e < nnz; e++) csr_ok = P->col_idx[e] >= 0 && P->col_idx[e] < n_point;
    if (!csr_ok) {
        fprintf(stderr, "Error: Invalid P matrix (row pointers or column indices out of range)\n");
        return nullptr;
    }
    tsne_workspace *ws = new tsne_workspace();
    ws->n = n_point;
    ws->pt_dim = pt_dim;
    ws->dim = low_dim;
    ws->params = *params;
    if (params->profiling_output_file) ws->profile_file = params->profiling_output_file;
    ws->params.profiling_output_file = nullptr;
    ws->params.init_Y = nullptr;
    ws->lr = params->learning_rate;
    if (ws->lr < 0.0) ws->lr = std::fmax(200.0, n_point / 12.0);  // src/tsne_h2pack.c:1398-1405
    ws->nnz = nnz;
    const char *prec = getenv("TSNE_H2_PRECISION");
    if (prec && strcmp(prec, "mixed") == 0) ws->precision = TSNE_H2_MIXED;

    const int n = n_point, D = low_dim;
    // rows per attractive block (environment TSNE_ATTR_RPB, default 16: one row per
    // 16-lane group; at 1M 2D the fused kernel takes 871 us against 896 us at 32 and 986 us
    // at 128); the block count is a multiple of 8 for the XCD-aware placement
    {
        const char *e = getenv("TSNE_ATTR_RPB");
        int rpb = e ? atoi(e) : 16;
        rpb = std::max(16, std::min(rpb, 4096));
        ws->nblocks_upd = (((n + rpb - 1) / rpb) + 7) & ~7;
    }
    // rows per attractive tile (environment TSNE_ATTR_TILE_ROWS, 0 = per-row kernel only);
    // tiles are at least as large as the per-row blocks, so block_sums holds either
    {
        const char *e = getenv("TSNE_ATTR_TILE_ROWS");
        int tr = e ? atoi(e) : 0;
        const int rpb = (n + ws->nblocks_upd - 1) / ws->nblocks_upd;
        ws->tile_rows = (tr <= 0) ? 0 : std::max(std::max(tr, rpb), 16);
        // LDS for the staged records (TSNE_ATTR_TILE_LDS_KB, 16..kTileLdsMaxKB KiB; the
        // kernel's static LDS takes the rest of the CU's 160 KiB) and the workgroup size
        // (TSNE_ATTR_TILE_NT, 256 or 1024)
        const char *l = getenv("TSNE_ATTR_TILE_LDS_KB");
        ws->tile_lds = 1024 * std::max(16, std::min(l ? atoi(l) : 64, kTileLdsMaxKB));
        const char *t = getenv("TSNE_ATTR_TILE_NT");
        ws->tile_nt = (t && atoi(t) == 1024) ? 1024 : 256;
    }
    ws->nblocks_red = std::max(1, std::min(1024, (n + 1023) / 1024));
    // launch shapes: the device's tuning table entry for this job size, else the built-in
    // defaults (none changes a result); TSNE_GRAPH / TSNE_OVERLAP, when set, win
    tsne_tuning_t tune = tune_defaults(low_dim, n_point);
    (void)tune_lookup(low_dim, n_point, &tune);
    if (shapes) {
        if (shapes->overlap == 0 || shapes->overlap == 1) tune.overlap = shapes->overlap;
        if (shapes->graph == 0 || shapes->graph == 1) tune.graph = shapes->graph;
        if (shapes->level_batch >= 1 && shapes->level_batch <= 8) tune.level_batch = shapes->level_batch;
        if (shapes->one_class_level >= 0) tune.one_class_level = shapes->one_class_level;
    }
    ws->tune_level_batch = tune.level_batch;
    ws->tune_one_class = tune.one_class_level;
    {
        const char *e = getenv("TSNE_GRAPH");  // 1 / 0: graph mode on / off
        ws->graph_mode = e ? (e[0] == '1') : tune.graph;
    }
    {
        // the split attractive pass of rebuild iterations: queued after the build's front
        // (Morton keys and their sort, which stream HBM like it) so that it runs beside the
        // tree levels and the ID, whose launches hold a few workgroups each; and launched
        // on 7/8 of the CUs (one workgroup per CU walking its logical blocks), so those
        // launches find CUs free.  1M 2D whole schedule: 687 -> 750 it/s; 500K 3D:
        // 464 -> 487 it/s; the same bits whatever the grid.  TSNE_ATTR_AFTER_FRONT=0 and
        // TSNE_ATTR_SPLIT_GRID (0: one workgroup per logical block) override.
        const char *e = getenv("TSNE_ATTR_AFTER_FRONT");
        ws->attr_after_front = e ? (e[0] == '1') : 1;
        const char *g = getenv("TSNE_ATTR_SPLIT_GRID");
        int cus = 0, dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 0;
        ws->attr_split_grid = g ? std::max(0, atoi(g)) : ((cus * 7 / 8) & ~7);
    }
    {
        const char *e = getenv("TSNE_DIRECT_REUSE");
        ws->direct_reuse = e ? (e[0] == '1') : 0;
    }
    {
        // the attractive pass of rebuild iterations on a second stream (the tuning's
        // choice; TSNE_OVERLAP=0 / 1 turns it off / on for any size)
        const char *e = getenv("TSNE_OVERLAP");
        ws->overlap = e ? (e[0] != '0') : tune.overlap;
    }
    bool ok = hipGetDevice(&ws->device) == hipSuccess &&
              hipStreamCreateWithFlags(&ws->st, hipStreamNonBlocking) == hipSuccess;
    if (ok && ws->overlap) {
        ok = hipStreamCreateWithFlags(&ws->st2, hipStreamNonBlocking) == hipSuccess;
        ok = ok && hipEventCreateWithFlags(&ws->ev_y, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&ws->ev_a, hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
        fprintf(stderr, "tsne_init_with_P: no HIP device available\n");
        delete ws;
        return nullptr;
    }
    ok = dalloc(&ws->d_rowptr, (size_t)n + 1, ws->st) == 0 && dalloc(&ws->d_col, (size_t)nnz + kCsrPad, ws->st) == 0 &&
         dalloc(&ws->d_val32, (size_t)nnz + kCsrPad, ws->st) == 0 && dalloc(&ws->d_val64, nnz, ws->st) == 0 &&
         dalloc(&ws->d_Y64, (size_t)n * D, ws->st) == 0 && dalloc(&ws->d_uY, (size_t)n * D, ws->st) == 0 &&
         dalloc(&ws->d_gains, (size_t)n * D, ws->st) == 0 &&
         dalloc((char **)&ws->d_Yg[0], (size_t)n * kGatherBytes[D], ws->st) == 0 &&
         dalloc((char **)&ws->d_Yg[1], (size_t)n * kGatherBytes[D], ws->st) == 0 &&
         dalloc(&ws->d_s, n, ws->st) == 0 && dalloc(&ws->d_F, (size_t)n * D, ws->st) == 0 &&
         dalloc(&ws->d_scal, SLOT_COUNT, ws->st) == 0 &&
         dalloc(&ws->d_block_sums, ((size_t)ws->nblocks_upd + kMeanParts) * 3, ws->st) == 0 &&
         dalloc(&ws->d_partial, ws->nblocks_red, ws->st) == 0 && dalloc(&ws->d_rowkl, n, ws->st) == 0 &&
         (!ws->overlap || dalloc(&ws->d_A, (size_t)n * D, ws->st) == 0);
    if (ok) {
        // P uploaded once (FP64 values; the FP32 copy is made on the device), gains = 1
        ok = hipMemcpyAsync(ws->d_rowptr, P->row_ptr, ((size_t)n + 1) * sizeof(int),
                            hipMemcpyHostToDevice, ws->st) == hipSuccess &&
             hipMemcpyAsync(ws->d_col, P->col_idx, (size_t)nnz * sizeof(int),
                            hipMemcpyHostToDevice, ws->st) == hipSuccess &&
             hipMemcpyAsync(ws->d_val64, P->values, (size_t)nnz * sizeof(double),
                            hipMemcpyHostToDevice, ws->st) == hipSuccess &&
             launch_f64_to_f32(ws->d_val64, ws->d_val32, (size_t)nnz, ws->st) == 0 &&
             launch_fill_f64(ws->d_gains, 1.0, (size_t)n * D, ws->st) == 0 &&
             hipMemsetAsync(ws->d_uY, 0, (size_t)n * D * sizeof(double), ws->st) == hipSuccess &&
             hipMemsetAsync(ws->d_F, 0, (size_t)n * D * sizeof(float), ws->st) == hipSuccess &&
             hipMemsetAsync(ws->d_Yg[1], 0, (size_t)n * kGatherBytes[D], ws->st) == hipSuccess &&
             hipStreamSynchronize(ws->st) == hipSuccess;
    }
    if (ok) {
        ws->order.resize(n);
        for (int i = 0; i < n; i++) ws->order[i] = i;
        ok = dalloc(&ws->d_order, n, ws->st) == 0 &&
             hipMemcpyAsync(ws->d_order, ws->order.data(), sizeof(int) * n, hipMemcpyHostToDevice,
                            ws->st) == hipSuccess;
        // renumbering schedule: src/tsne_h2pack.c:1706-1717 (TSNE_REORDER, TSNE_REORDER_ITERS)
        const char *ro = getenv("TSNE_REORDER");
        if (!(ro && ro[0] == '0')) {
            ws->reorder_iters = {30, 100, 250, 400, 600, 800};
            const char *ri = getenv("TSNE_REORDER_ITERS");
            if (ri && *ri) {
                ws->reorder_iters.clear();
                for (const char *c = ri; *c && ws->reorder_iters.size() < 16;) {
                    ws->reorder_iters.push_back(atoi(c));
                    while (*c && *c != ',') c++;
                    if (*c == ',') c++;
                }
            }
        }
    }
    if (ok) {
        // initial embedding, src/tsne_utils.c:1111-1150
        std::vector<double> Y0((size_t)n * D);
        if (params->init_method == 1 && params->init_Y)
            memcpy(Y0.data(), params->init_Y, Y0.size() * sizeof(double));
        else
            init_random(Y0.data(), n, D, (unsigned)params->random_seed);
        ok = upload_embedding(ws, Y0.data()) == 0;
    }
    if (!ok) {
        free_device(ws);
        delete ws;
        return nullptr;
    }
    ws->P = P;
    if (!use_direct(ws)) {
        if (rebuild(ws) != 0) {
            fprintf(stderr, "Error: Failed to build initial H2 matrix\n");
            ws->P = nullptr;  // P stays with the caller on a NULL return
            tsne_destroy(ws);
            return nullptr;
        }
    }
    *taken = true;
    if (params->verbose)
        printf("Initialized t-SNE on the GPU (n=%d, low_dim=%d, nnz=%d, %s repulsion)\n", n, D, nnz,
               use_direct(ws) ? "direct" : "H2");
    return ws;
}

}  // namespace tsne

extern "C" {

int tsne_run_range(tsne_workspace_t *ws, int it_begin, int it_end)
{
    if (!ws) return -1;
    const tsne_params_t &pp = ws->params;
    const int n_iter = pp.n_iter;
    const int direct = use_direct(ws);
    // src/tsne_h2pack.c:1699 (rep_reuse when N >= 2000)
    const int reuse = !direct || (ws->direct_reuse && ws->n >= 2000);
    const int D = ws->dim;
    if (it_begin < 0) it_begin = 0;
    if (it_end > n_iter) it_end = n_iter;
    const bool timing = ws->timing == 1 || pp.enable_profiling;
    // ev: iteration start, after the rebuild, after the product, after Z, after the KL,
    // after the wait for the overlapped attractive pass, end of the update
    hipEvent_t ev[7];
    if (timing)
        for (auto &e : ev) (void)hipEventCreate(&e);
    // timing 2: a pair of events around every attractive launch, read after the last
    // synchronisation, so the run itself is not slowed by host round trips
    std::vector<hipEvent_t> kev;
    std::vector<char> ktiled;  // which timed launches ran the tiled kernel
    std::vector<char> kused;   // which event pairs were recorded (fused launches only)
    if (ws->timing == 2 && it_end > it_begin) {
        kev.resize(2 * (size_t)(it_end - it_begin), nullptr);
        ktiled.assign((size_t)(it_end - it_begin), 0);
        kused.assign((size_t)(it_end - it_begin), 0);
        for (auto &e : kev)
            if (hipEventCreate(&e) != hipSuccess) e = nullptr;
    }
    if (it_begin == 0) {
        ws->t_attr = ws->t_rebuild = ws->t_product = ws->t_other = 0;
        ws->t_attr_kernel = 0;
        ws->n_attr_kernel = 0;
        ws->t_attr_tiled = 0;
        ws->n_attr_tiled = 0;
        ws->n_rebuild = 0;
        if (pp.enable_profiling) ws->prof.assign((size_t)n_iter * kProfCols, 0.0);
    }
    if (pp.verbose && it_begin == 0)
        printf("\nStarting t-SNE optimization (%d iterations, learning rate %.1f)\n", n_iter, ws->lr);
    for (int it = it_begin; it < it_end; it++) {
        const double mom = (it < pp.momentum_switch_iter) ? pp.initial_momentum : pp.momentum;
        double exag;
        if (pp.sklearn_compatible) {
            exag = (it < pp.early_exag_iter) ? pp.early_exag : 1.0;
        } else if (it < pp.early_exag_iter) {
            exag = pp.early_exag;
        } else if (it < pp.early_exag_iter + 100) {
            exag = pp.early_exag - (pp.early_exag - 1.0) * (double)(it - pp.early_exag_iter) / 100.0;
        } else {
            exag = 1.0;
        }
        double lr = ws->lr;
        if (!pp.sklearn_compatible && it >= pp.momentum_switch_iter)
            lr = ws->lr * std::pow(0.995, it - pp.momentum_switch_iter);
        if (timing) (void)hipEventRecord(ev[0], ws->st);
        int rebuilt = (it == 0);
        // split: this iteration's attractive sums go to the second stream while the
        // rebuild and the product run here; the update waits for both
        bool split = false;
        if (!direct && rebuild_iter(ws, it, reuse)) {
            const int nro = (int)ws->reorder_iters.size();
            if (ws->next_reorder < nro && it >= ws->reorder_iters[ws->next_reorder] && ws->h2 &&
                ws->h2_ok) {
                while (ws->next_reorder < nro && it >= ws->reorder_iters[ws->next_reorder])
                    ws->next_reorder++;
                if (renumber(ws) < 0) return -1;
                if (ws->next_reorder >= nro) free_renumber_buffers(ws);  // last one
            }
            if (ws->overlap && ws->st2 && ws->tiles.ntiles == 0) {
                split = true;
                if (ws->attr_after_front && ws->h2)
                    h2_set_front_hook(ws->h2, queue_split_attr, ws);
                else if (queue_split_attr(ws) != 0)
                    return -1;
            }
            if (rebuild(ws) != 0) {
                fprintf(stderr, "ERROR: Failed to rebuild H2 matrix at iteration %d\n", it);
                return -1;
            }
            rebuilt = 1;
            ws->n_rebuild++;
        } else if (direct && reuse && rebuild_iter(ws, it, reuse)) {
            rebuilt = 1;  // check mode: a fresh direct sum where the H2 path rebuilds
        }
        if (timing) (void)hipEventRecord(ev[1], ws->st);
        const bool fresh = !reuse || rebuilt || !ws->rep_valid;
        if (fresh && repulsive_sums(ws, !direct) != 0) return -1;
        if (timing) (void)hipEventRecord(ev[2], ws->st);
        if (fresh) {
            if (repulsive_z(ws) != 0) return -1;
            ws->rep_valid = reuse;
        }
        if (timing) (void)hipEventRecord(ev[3], ws->st);
        if ((pp.verbose && it % 50 == 0) || it == n_iter - 1) {
            if (launch_kl(D, ws->n, ws->d_rowptr, ws->d_col, ws->d_val64, ws->d_Y64, ws->d_scal,
                          exag, ws->d_rowkl, ws->d_partial, ws->nblocks_red, ws->d_scal + SLOT_KL,
                          ws->st) != 0)
                return -1;
        }
        if (timing) (void)hipEventRecord(ev[4], ws->st);
        const int nxt = ws->cur ^ 1;
        hipEvent_t kb = kev.empty() ? nullptr : kev[2 * (size_t)(it - it_begin)];
        hipEvent_t ke = kev.empty() ? nullptr : kev[2 * (size_t)(it - it_begin) + 1];
        if (split) TSNE_HIP_CHECK(hipStreamWaitEvent(ws->st, ws->ev_a, 0));
        if (timing) (void)hipEventRecord(ev[5], ws->st);
        // graph replay of an iteration that reuses the force (no per-launch events, no
        // KL, the per-row kernel)
        const bool kl_now = (pp.verbose && it % 50 == 0) || it == n_iter - 1;
        if (ws->graph_mode && !timing && kev.empty() && !split && !fresh && !kl_now &&
            ws->tiles.ntiles == 0) {
            if (graph_step(ws, exag, mom, lr) != 0) return -1;
            ws->cur ^= 1;
            ws->n_graph_launches++;
            if (pp.verbose && pp.print_freq > 0 && (it % pp.print_freq == 0 || it == n_iter - 1))
                printf("  Iteration %4d/%d: Z=%.3e\n", it + 1, n_iter, read_scalar(ws, SLOT_Z));
            continue;
        }
        const int arc =
            split ? launch_update(D, ws->n, ws->d_A, ws->d_F, ws->d_Y64, ws->d_uY, ws->d_gains,
                                  ws->d_Yg[nxt], ws->d_scal, exag, mom, lr, ws->d_block_sums,
                                  ws->nblocks_upd, ws->st)
            : ws->tiles.ntiles > 0
                ? launch_attr_tile(D, ws->n, ws->tiles, ws->tile_nt, ws->d_rowptr, ws->d_col, ws->d_val32,
                                   ws->d_Yg[ws->cur], ws->d_F, ws->d_Y64, ws->d_uY, ws->d_gains,
                                   ws->d_Yg[nxt], ws->d_scal, exag, mom, lr, ws->d_block_sums,
                                   ws->nblocks_upd, ws->st, kb, ke)
                : launch_attr_update(D, ws->n, ws->d_rowptr, ws->d_col, ws->d_val32,
                                     ws->d_Yg[ws->cur], ws->d_F, ws->d_Y64, ws->d_uY, ws->d_gains,
                                     ws->d_Yg[nxt], ws->d_scal, exag, mom, lr, ws->d_block_sums,
                                     ws->nblocks_upd, ws->st, kb, ke);
        if (arc != 0) return -1;
        if (!ktiled.empty()) {
            ktiled[(size_t)(it - it_begin)] = ws->tiles.ntiles > 0;
            kused[(size_t)(it - it_begin)] = !split;
        }
        ws->cur = nxt;
        if (timing) {
            (void)hipEventRecord(ev[6], ws->st);
            (void)hipEventSynchronize(ev[6]);
            float t[6] = {0, 0, 0, 0, 0, 0};
            for (int k = 0; k < 6; k++) (void)hipEventElapsedTime(&t[k], ev[k], ev[k + 1]);
            // t: rebuild, product, Z, KL, attractive (the wait for the overlapped pass on
            // split iterations, the fused attractive + update pass otherwise), update
            // (split iterations only: the update kernel after the overlapped pass)
            const float attr = split ? t[4] : t[4] + t[5];
            const float upd = split ? t[5] : 0.f;
            ws->t_rebuild += t[0];
            ws->t_product += t[1] + t[2];
            ws->t_other += t[3];
            ws->t_attr += attr + upd;
            if (pp.enable_profiling && (size_t)(it + 1) * kProfCols <= ws->prof.size()) {
                double *r = &ws->prof[(size_t)it * kProfCols];
                r[PROF_ATTR] = attr * 1e-3;
                r[PROF_REBUILD] = t[0] * 1e-3;
                r[PROF_PRODUCT] = t[1] * 1e-3;
                r[PROF_Z] = t[2] * 1e-3;
                r[PROF_UPDATE] = upd * 1e-3;
                r[PROF_OTHER] = t[3] * 1e-3;
            }
        }
        if (pp.verbose && pp.print_freq > 0 && (it % pp.print_freq == 0 || it == n_iter - 1)) {
            const double Z = read_scalar(ws, SLOT_Z);
            printf("  Iteration %4d/%d: Z=%.3e\n", it + 1, n_iter, Z);
            if (it % 50 == 0) printf("  [KL iter %4d] KL divergence: %.5f\n", it, read_scalar(ws, SLOT_KL));
        }
    }
    if (timing)
        for (auto &e : ev) (void)hipEventDestroy(e);
    if (hipStreamSynchronize(ws->st) != hipSuccess) {
        fprintf(stderr, "tsne_run: device failure\n");
        return -1;
    }
    for (size_t k = 0; k + 1 < kev.size(); k += 2) {
        float ms = 0.f;
        if (kused[k / 2] && kev[k] && kev[k + 1] &&
            hipEventElapsedTime(&ms, kev[k], kev[k + 1]) == hipSuccess) {
            ws->t_attr_kernel += ms;
            ws->n_attr_kernel++;
            if (ktiled[k / 2]) {
                ws->t_attr_tiled += ms;
                ws->n_attr_tiled++;
            }
        }
    }
    for (auto &e : kev)
        if (e) (void)hipEventDestroy(e);
    if (it_end == n_iter) ws->kl = read_scalar(ws, SLOT_KL);
    if (read_scalar(ws, SLOT_BAD) != 0.0) {
        fprintf(stderr, "ERROR: Y contains NaN or Inf\n");
        return -1;
    }
    if (read_scalar(ws, SLOT_ZBAD) != 0.0)
        fprintf(stderr, "ERROR: Z <= 0 or non-finite during the run (replaced by 1.0)\n");
    if (pp.enable_profiling && it_end == n_iter && !ws->profile_file.empty()) {
        FILE *f = fopen(ws->profile_file.c_str(), "w");
        if (f) {
            // the reference's CSV schema and summary, src/tsne_h2pack.c:1553-1628.  The
            // product is one fused pass, so h2_matvec_v is 0 as in the reference (:551);
            // on iterations whose update is fused into the attractive pass, gradient is 0
            // and p_matvec holds both
            fprintf(f, "iteration,p_matvec,h2_rebuild,h2_matvec_s,h2_matvec_v,z_computation,"
                       "gradient,other,total\n");
            double tot[kProfCols] = {0, 0, 0, 0, 0, 0}, all = 0.0;
            for (int it = 0; it < n_iter; it++) {
                const double *r = &ws->prof[(size_t)it * kProfCols];
                double t = 0.0;
                for (int k = 0; k < kProfCols; k++) {
                    tot[k] += r[k];
                    t += r[k];
                }
                all += t;
                fprintf(f, "%d,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f\n", it, r[PROF_ATTR],
                        r[PROF_REBUILD], r[PROF_PRODUCT], 0.0, r[PROF_Z], r[PROF_UPDATE],
                        r[PROF_OTHER], t);
            }
            const double pct = all > 0.0 ? 100.0 / all : 0.0;
            fprintf(f, "\n# Summary Statistics\n");
            fprintf(f, "# k-NN time: %.3f seconds\n", ws->knn_time);
            fprintf(f, "# P construction: %.3f seconds\n", ws->p_time);
            fprintf(f, "# Total P-matvec: %.3f seconds (%.1f%%)\n", tot[PROF_ATTR], tot[PROF_ATTR] * pct);
            fprintf(f, "# Total H²-rebuild: %.3f seconds (%.1f%%)\n", tot[PROF_REBUILD],
                    tot[PROF_REBUILD] * pct);
            fprintf(f, "# Total H²-matvec: %.3f seconds (%.1f%%)\n", tot[PROF_PRODUCT],
                    tot[PROF_PRODUCT] * pct);
            fprintf(f, "# Total gradient: %.3f seconds (%.1f%%)\n", tot[PROF_UPDATE],
                    tot[PROF_UPDATE] * pct);
            fprintf(f, "# Total optimization time: %.3f seconds\n", all);
            fclose(f);
        }
    }
    return 0;
}

int tsne_run(tsne_workspace_t *ws)
{
    if (!ws) return -1;
    const int rc = tsne_run_range(ws, 0, ws->params.n_iter);
    if (rc == 0 && ws->params.verbose) printf("\nOptimization complete\n");
    return rc;
}

void tsne_get_embedding(tsne_workspace_t *ws, double *Y_out)
{
    if (!ws || !Y_out) return;
    const int n = ws->n, D = ws->dim;
    std::vector<double> Yi((size_t)n * D);
    double mean[SLOT_COUNT];
    if (hipMemcpyAsync(Yi.data(), ws->d_Y64, Yi.size() * sizeof(double), hipMemcpyDeviceToHost,
                       ws->st) != hipSuccess ||
        hipMemcpyAsync(mean, ws->d_scal, sizeof(mean), hipMemcpyDeviceToHost, ws->st) != hipSuccess ||
        hipStreamSynchronize(ws->st) != hipSuccess) {
        fprintf(stderr, "tsne_get_embedding: device failure\n");
        return;
    }
    for (int d = 0; d < D; d++)
        for (int i = 0; i < n; i++)
            Y_out[(size_t)d * n + ws->order[i]] = Yi[(size_t)i * D + d] - mean[SLOT_MEAN + d];
}

void tsne_destroy(tsne_workspace_t *ws)
{
    if (!ws) return;
    if (ws->st) (void)hipStreamSynchronize(ws->st);
    free_device(ws);
    sparse_matrix_destroy(ws->P);
    delete ws;
}

int tsne_fit_transform(const double *X, int n_point, int pt_dim, int low_dim,
                       const tsne_params_t *params, double *Y_out)
{
    tsne_workspace_t *ws = tsne_init(X, n_point, pt_dim, low_dim, params);
    if (!ws) return -1;
    int rc = tsne_run(ws);
    if (rc == 0) tsne_get_embedding(ws, Y_out);
    tsne_destroy(ws);
    return rc;
}

int tsne_fit_transform_with_P(const double *X, int n_point, int pt_dim, int low_dim,
                              const tsne_params_t *params, sparse_matrix_t *P, double *Y_out)
{
    tsne_workspace_t *ws = tsne_init_with_P(X, n_point, pt_dim, low_dim, params, P);
    if (!ws) return -1;
    int rc = tsne_run(ws);
    if (rc == 0) tsne_get_embedding(ws, Y_out);
    tsne_destroy(ws);
    return rc;
}

void tsne_set_init_profiling_times(tsne_workspace_t *ws, double knn_time, double p_time)
{
    if (!ws || !ws->params.enable_profiling) return;
    ws->knn_time = knn_time;
    ws->p_time = p_time;
}

double compute_kl_divergence(tsne_workspace_t *ws) { return ws ? ws->kl : 0.0; }

int tsne_set_h2_precision(tsne_workspace_t *ws, int mode)
{
    if (!ws || (mode != TSNE_H2_PLAIN && mode != TSNE_H2_MIXED)) return -1;
    ws->precision = mode;
    return 0;
}

int tsne_workspace_info(tsne_workspace_t *ws, int *n_point, int *low_dim, double *lr)
{
    if (!ws) return -1;
    if (n_point) *n_point = ws->n;
    if (low_dim) *low_dim = ws->dim;
    if (lr) *lr = ws->lr;
    return 0;
}

int tsne_set_embedding(tsne_workspace_t *ws, const double *Y)
{
    if (!ws || !Y) return -1;
    return upload_embedding(ws, Y);
}

int tsne_compute_gradient(tsne_workspace_t *ws, double exag, int exact, double *grad_out,
                          double *Z_out)
{
    if (!ws || !grad_out) return -1;
    const int n = ws->n, D = ws->dim;
    double Z = 0.0;
    std::vector<double> rep((size_t)n * D);
    if (exact) {
        double *d_s = nullptr, *d_F = nullptr;
        if (dalloc(&d_s, n, ws->st) || dalloc(&d_F, (size_t)n * D, ws->st)) { dfree(d_s, ws->st); dfree(d_F, ws->st); return -1; }
        int rc = launch_repulsive_direct_f64(D, n, ws->d_Y64, nullptr, n, d_s, d_F, ws->st);
        rc = rc || launch_z_f64(D, n, ws->d_Y64, d_s, d_F, ws->d_scal, ws->d_partial,
                                ws->nblocks_red, ws->d_scal + SLOT_ZEXACT, ws->st);
        rc = rc || hipMemcpyAsync(rep.data(), d_F, rep.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, ws->st) != hipSuccess;
        rc = rc || hipStreamSynchronize(ws->st) != hipSuccess;
        dfree(d_s, ws->st); dfree(d_F, ws->st);
        if (rc) return -1;
        Z = read_scalar(ws, SLOT_ZEXACT);
        for (auto &v : rep) v *= 4.0 / Z;
    } else {
        if (!use_direct(ws) && rebuild(ws) != 0) return -1;
        if (repulsive(ws, !use_direct(ws)) != 0) return -1;
        std::vector<float> F((size_t)n * D);
        if (hipMemcpyAsync(F.data(), ws->d_F, F.size() * sizeof(float), hipMemcpyDeviceToHost,
                           ws->st) != hipSuccess)
            return -1;
        Z = read_scalar(ws, SLOT_Z);
        // as the run: no repulsive force from a product whose Z was invalid
        const double sc = read_scalar(ws, SLOT_ZSKIP) != 0.0 ? 0.0 : 4.0 / Z;
        for (size_t k = 0; k < F.size(); k++) rep[k] = sc * (double)F[k];
        ws->rep_valid = 0;
    }
    // attractive part on the host in FP64 from the FP64 master coordinates, all in the
    // caller's order (P on the host is in that order)
    std::vector<double> Yc((size_t)n * D), Yi((size_t)n * D), rc_((size_t)n * D);
    if (hipMemcpyAsync(Yc.data(), ws->d_Y64, Yc.size() * sizeof(double), hipMemcpyDeviceToHost,
                       ws->st) != hipSuccess ||
        hipStreamSynchronize(ws->st) != hipSuccess)
        return -1;
    for (int k = 0; k < n; k++)
        for (int d = 0; d < D; d++) {
            Yi[(size_t)ws->order[k] * D + d] = Yc[(size_t)k * D + d];
            rc_[(size_t)ws->order[k] * D + d] = rep[(size_t)k * D + d];
        }
    rep.swap(rc_);
    const sparse_matrix_t *P = ws->P;
    for (int i = 0; i < n; i++) {
        double acc[3] = {0, 0, 0};
        for (int e = P->row_ptr[i]; e < P->row_ptr[i + 1]; e++) {
            const int j = P->col_idx[e];
            double r2 = 1.0, df[3];
            for (int d = 0; d < D; d++) {
                df[d] = Yi[(size_t)i * D + d] - Yi[(size_t)j * D + d];
                r2 += df[d] * df[d];
            }
            for (int d = 0; d < D; d++) acc[d] += P->values[e] / r2 * df[d];
        }
        for (int d = 0; d < D; d++)
            grad_out[(size_t)d * n + i] = 4.0 * exag * acc[d] - rep[(size_t)i * D + d];
    }
    if (Z_out) *Z_out = Z;
    return 0;
}

int tsne_exact_repulsive_rows(tsne_workspace_t *ws, const int *rows, int n_rows, double *out,
                              double *Z_out)
{
    if (!ws || !rows || n_rows <= 0 || !out) return -1;
    const int n = ws->n, D = ws->dim;
    for (int r = 0; r < n_rows; r++)
        if (rows[r] < 0 || rows[r] >= n) return -1;
    double *d_s = nullptr, *d_F = nullptr, *d_s2 = nullptr, *d_F2 = nullptr;
    int *d_rows = nullptr;
    int rc = dalloc(&d_s, n, ws->st) || dalloc(&d_F, (size_t)n * D, ws->st) || dalloc(&d_s2, n_rows, ws->st) ||
             dalloc(&d_F2, (size_t)n_rows * D, ws->st) || dalloc(&d_rows, n_rows, ws->st);
    const std::vector<int> inv = inverse_order(ws);
    std::vector<int> cur_rows(n_rows);
    for (int r = 0; r < n_rows; r++) cur_rows[r] = inv[rows[r]];
    rc = rc || hipMemcpyAsync(d_rows, cur_rows.data(), (size_t)n_rows * sizeof(int),
                              hipMemcpyHostToDevice, ws->st) != hipSuccess;
    rc = rc || launch_repulsive_direct_f64(D, n, ws->d_Y64, nullptr, n, d_s, d_F, ws->st);
    rc = rc || launch_z_f64(D, n, ws->d_Y64, d_s, d_F, ws->d_scal, ws->d_partial, ws->nblocks_red,
                            ws->d_scal + SLOT_ZEXACT, ws->st);
    rc = rc || launch_repulsive_direct_f64(D, n, ws->d_Y64, d_rows, n_rows, d_s2, d_F2, ws->st);
    std::vector<double> F((size_t)n_rows * D);
    rc = rc || hipMemcpyAsync(F.data(), d_F2, F.size() * sizeof(double), hipMemcpyDeviceToHost,
                              ws->st) != hipSuccess;
    rc = rc || hipStreamSynchronize(ws->st) != hipSuccess;
    dfree(d_s, ws->st); dfree(d_F, ws->st); dfree(d_s2, ws->st); dfree(d_F2, ws->st); dfree(d_rows, ws->st);
    if (rc) return -1;
    const double Z = read_scalar(ws, SLOT_ZEXACT);
    for (size_t k = 0; k < F.size(); k++) out[k] = 4.0 / Z * F[k];
    if (Z_out) *Z_out = Z;
    return 0;
}

int tsne_exact_repulsive_rows_noalloc(tsne_workspace_t *ws, const int *rows, int n_rows,
                                      double *out, double *Z_out)
{
    if (!ws || !rows || n_rows <= 0 || !out) return -1;
    const int n = ws->n, D = ws->dim;
    for (int r = 0; r < n_rows; r++)
        if (rows[r] < 0 || rows[r] >= n) {
            fprintf(stderr, "tsne_exact_repulsive_rows_noalloc: row %d out of range\n", rows[r]);
            return -1;
        }
    // scratch layout: inverse order, caller rows, current rows, s and F of all points,
    // s and F of the selected rows; each part 256-byte aligned
    const size_t b_inv = align256((size_t)n * sizeof(int));
    const size_t b_rows = align256((size_t)n_rows * sizeof(int));
    const size_t b_s = align256((size_t)n * sizeof(double));
    const size_t b_F = align256((size_t)n * D * sizeof(double));
    const size_t b_s2 = align256((size_t)n_rows * sizeof(double));
    const size_t b_F2 = align256((size_t)n_rows * D * sizeof(double));
    const size_t need = b_inv + 2 * b_rows + b_s + b_F + b_s2 + b_F2;
    if (need > kNoAllocScratchBytes) {
        fprintf(stderr, "tsne_exact_repulsive_rows_noalloc: n=%d, %d rows need %zu bytes of "
                        "scratch, above its %zu-byte maximum (use tsne_exact_repulsive_rows)\n",
                n, n_rows, need, kNoAllocScratchBytes);
        return -1;
    }
    int cur_dev = 0;
    if (hipGetDevice(&cur_dev) != hipSuccess) return -1;
    if (cur_dev != ws->device && hipSetDevice(ws->device) != hipSuccess) return -1;
    int rc = 0;
    char *base = (char *)scratch_ptr(ws->device);
    if (!base) rc = -1;
    if (rc == 0) {
        int *d_inv = (int *)base;
        int *d_rows_in = (int *)(base + b_inv);
        int *d_rows = (int *)(base + b_inv + b_rows);
        double *d_s = (double *)(base + b_inv + 2 * b_rows);
        double *d_F = (double *)((char *)d_s + b_s);
        double *d_s2 = (double *)((char *)d_F + b_F);
        double *d_F2 = (double *)((char *)d_s2 + b_s2);
        hipStream_t st = ws->st;
        rc = hipMemcpyAsync(d_rows_in, rows, (size_t)n_rows * sizeof(int), hipMemcpyHostToDevice,
                            st) != hipSuccess;
        if (rc == 0) {
            hipLaunchKernelGGL(invert_order_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n,
                               ws->d_order, d_inv);
            hipLaunchKernelGGL(map_rows_kernel, dim3((n_rows + 255) / 256), dim3(256), 0, st, n_rows,
                               d_rows_in, d_inv, d_rows);
            rc = hipGetLastError() != hipSuccess;
        }
        // the same kernels and order of operations as tsne_exact_repulsive_rows
        rc = rc || launch_repulsive_direct_f64(D, n, ws->d_Y64, nullptr, n, d_s, d_F, st);
        rc = rc || launch_z_f64(D, n, ws->d_Y64, d_s, d_F, ws->d_scal, ws->d_partial,
                                ws->nblocks_red, ws->d_scal + SLOT_ZEXACT, st);
        rc = rc || launch_repulsive_direct_f64(D, n, ws->d_Y64, d_rows, n_rows, d_s2, d_F2, st);
        rc = rc || hipMemcpyAsync(out, d_F2, (size_t)n_rows * D * sizeof(double),
                                  hipMemcpyDeviceToHost, st) != hipSuccess;
        rc = rc || hipStreamSynchronize(st) != hipSuccess;
        if (rc == 0) {
            const double Z = read_scalar(ws, SLOT_ZEXACT);
            for (size_t k = 0; k < (size_t)n_rows * D; k++) out[k] = 4.0 / Z * out[k];
            if (Z_out) *Z_out = Z;
        }
    }
    if (cur_dev != ws->device) (void)hipSetDevice(cur_dev);
    return rc ? -1 : 0;
}

int tsne_h2_stats(tsne_workspace_t *ws, tsne_h2_stats_t *stats)
{
    if (!ws || !stats || !ws->h2) return -1;
    return h2_stats(ws->h2, stats);
}

int tsne_run_timings(tsne_workspace_t *ws, double *attr, double *reb, double *prod, double *other,
                     int *n_rebuild)
{
    if (!ws) return -1;
    if (attr) *attr = ws->t_attr;
    if (reb) *reb = ws->t_rebuild;
    if (prod) *prod = ws->t_product;
    if (other) *other = ws->t_other;
    if (n_rebuild) *n_rebuild = ws->n_rebuild;
    return 0;
}

int tsne_set_graph_mode(tsne_workspace_t *ws, int on)
{
    if (!ws) return -1;
    ws->graph_mode = on ? 1 : 0;
    if (!on) drop_graphs(ws);
    return 0;
}

int tsne_set_tuning(tsne_workspace_t *ws, const tsne_tuning_t *t)
{
    if (!ws || !t) return -1;
    int cur_dev = 0;
    if (hipGetDevice(&cur_dev) != hipSuccess) return -1;
    if (cur_dev != ws->device && hipSetDevice(ws->device) != hipSuccess) return -1;
    if (t->overlap == 1 && !ws->st2) {
        // the second stream, its events and the split pass's attractive sums, made now
        if (hipStreamCreateWithFlags(&ws->st2, hipStreamNonBlocking) != hipSuccess) {
            ws->st2 = nullptr;
            return -1;
        }
        if (hipEventCreateWithFlags(&ws->ev_y, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ws->ev_a, hipEventDisableTiming) != hipSuccess ||
            (!ws->d_A && dalloc(&ws->d_A, (size_t)ws->n * ws->dim, ws->st) != 0))
            return -1;
        if (ws->h2) h2_set_aux_stream(ws->h2, ws->st2);
    }
    if (t->overlap == 0 || t->overlap == 1) ws->overlap = t->overlap;
    if (t->graph == 0 || t->graph == 1) tsne_set_graph_mode(ws, t->graph);
    if (t->level_batch >= 1 && t->level_batch <= 8) ws->tune_level_batch = t->level_batch;
    if (t->one_class_level >= 0) ws->tune_one_class = t->one_class_level;
    if (ws->h2) h2_set_tuning(ws->h2, ws->tune_level_batch, ws->tune_one_class);
    return 0;
}

int tsne_graph_launches(tsne_workspace_t *ws)
{
    return ws ? ws->n_graph_launches : -1;
}

int tsne_attr_probe(tsne_workspace_t *ws, int mode, int reps, double *avg_ms)
{
    if (!ws || mode < 0 || mode > 2 || reps < 1 || !avg_ms) return -1;
    if (!ws->d_A && dalloc(&ws->d_A, (size_t)ws->n * ws->dim, ws->st) != 0) return -1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TSNE_HIP_CHECK(hipEventCreate(&e0));
    TSNE_HIP_CHECK(hipEventCreate(&e1));
    int rc = launch_attr_probe(ws->dim, ws->n, mode, ws->d_rowptr, ws->d_col, ws->d_val32,
                               ws->d_Yg[ws->cur], ws->d_A, ws->nblocks_upd, ws->st);  // warm-up
    (void)hipEventRecord(e0, ws->st);
    for (int r = 0; r < reps && rc == 0; r++)
        rc = launch_attr_probe(ws->dim, ws->n, mode, ws->d_rowptr, ws->d_col, ws->d_val32,
                               ws->d_Yg[ws->cur], ws->d_A, ws->nblocks_upd, ws->st);
    (void)hipEventRecord(e1, ws->st);
    float ms = 0.f;
    if (rc == 0 && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
        *avg_ms = ms / reps;
    else
        rc = -1;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int tsne_run_state(tsne_workspace_t *ws, int *n_rebuild, int *n_renumber, double *Z, double *kl)
{
    if (!ws) return -1;
    if (n_rebuild) *n_rebuild = ws->n_rebuild;
    if (n_renumber) *n_renumber = ws->n_renumber;
    if (Z) *Z = read_scalar(ws, SLOT_Z);
    if (kl) *kl = read_scalar(ws, SLOT_KL);
    return 0;
}

int tsne_attr_kernel_time_tiled(tsne_workspace_t *ws, double *total_ms, int *launches)
{
    if (!ws) return -1;
    if (total_ms) *total_ms = ws->t_attr_tiled;
    if (launches) *launches = ws->n_attr_tiled;
    return 0;
}

int tsne_attr_tile_stats(tsne_workspace_t *ws, int *tile_rows, int *ntiles, int *umax,
                         long long *ucount, int *n_global)
{
    if (!ws) return -1;
    *tile_rows = ws->tiles.rows;
    *ntiles = ws->tiles.ntiles;
    *umax = ws->tiles.umax;
    *ucount = ws->tiles.ucount;
    *n_global = ws->tiles.n_global;
    return 0;
}

int tsne_attr_kernel_time(tsne_workspace_t *ws, double *total_ms, int *launches)
{
    if (!ws) return -1;
    if (total_ms) *total_ms = ws->t_attr_kernel;
    if (launches) *launches = ws->n_attr_kernel;
    return 0;
}

int tsne_synchronize(tsne_workspace_t *ws)
{
    if (!ws) return -1;
    return hipStreamSynchronize(ws->st) == hipSuccess ? 0 : -1;
}

// Timing of later runs: 0 off, 1 per-phase events (also on with profiling), 2 events
// around the attractive kernel only.
int tsne_enable_timing(tsne_workspace_t *ws, int on)
{
    if (!ws) return -1;
    ws->timing = on;
    return 0;
}

}  // extern "C"
