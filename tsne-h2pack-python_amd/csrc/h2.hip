// h2.hip -- host side of the H2 representation: device allocations, the rebuild
// sequence (root box, Morton sort, level-by-level tree and interaction lists, sorted
// pair lists, per-level ID + upward pass) and the product (coupling, downward sweep,
// leaves).  See h2_internal.h for the layout.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <mutex>
#include <vector>

#include "h2.h"
#include "h2_internal.h"
#include "proxy.h"
#include "radix.h"

namespace tsne {

// timeline marks: build start, front queued, tree queued, ID starts (after the host's
// readback), ID done, sorts done (second stream), build done
enum { TL_START = 0, TL_FRONT, TL_TREE, TL_ID0, TL_ID1, TL_SORTED, TL_END, kTlEvents };

struct H2 {
    int n = 0, D = 2, leaf = 100;
    double tol = 1e-2, ptol = 3e-3, id_tol = 5e-3;
    bool id_mixed = false;  // the mixed path's ID (FP32 Gram-Schmidt at id_tol >= kIdF32MinTol)
    hipStream_t st = nullptr;
    // optional second stream: the NEAR / W pair sorts run there beside the ID
    hipStream_t st_aux = nullptr;
    hipEvent_t ev_tree = nullptr, ev_sorted = nullptr;
    hipEvent_t ev_lvl[2] = {nullptr, nullptr};  // ends of the last two queued level batches
    // diagnostics (TSNE_H2_TIMELINE=1): HIP events at the build's phase boundaries, summed
    // over builds and printed by h2_destroy
    bool tl_on = false;
    hipEvent_t tl_ev[kTlEvents] = {};
    double tl_ms[kTlEvents] = {};
    int tl_builds = 0;
    int (*front_hook)(void *) = nullptr;  // one-shot: queued work after the next build's front
    void *front_arg = nullptr;
    H2Dev d{};
    // buffers not in H2Dev
    uint64_t *keys_alt = nullptr;
    int *perm_alt = nullptr;
    void *cub_tmp = nullptr;
    size_t cub_bytes = 0;
    double *part = nullptr;
    int nblk = 0;
    double *d_proxy = nullptr;
    int *d_proxy_cnt = nullptr;
    int *pin = nullptr;      // pinned host scratch
    int *pin_dev = nullptr;  // the same memory as the device sees it
    std::vector<int> lvl;  // level l nodes are [lvl[l], lvl[l+1])
    std::vector<int> prev_cnt;  // node count per level of the last tree (level grid hints)
    int level_hints = 1;        // TSNE_LEVEL_HINTS at creation (0: off)
    int n_nodes = 0, n_leaves = 0, n_adm = 0, n_x = 0, n_w = 0, n_near = 0;
    int usable = 0;
    int pq = 0;  // proxy slots the ID processes: the table's largest count, rounded up
    int hflags[FLAG_COUNT] = {0};
    int fallbacks = 0;  // builds that left the product to the direct sum, and why (first)
    int rank_capped = 0;  // builds in which some node's ID stopped at RMAX short of id_tol
    // the ID's flags of the last build, copied to pinned memory behind the ID (ev_flags);
    // absorbed at the next build or statistics call, so no host wait follows the ID
    bool flags_pending = false;
    hipEvent_t ev_flags = nullptr;
    // timing of the ID launches of each build (HIP events on the build stream)
    bool timing = false;
    hipEvent_t ev_id0 = nullptr, ev_id1 = nullptr;
    double id_ms = 0.0;   // the last build's ID time
    int last_mixed = 0; // the last product was the mixed-precision one
    // growth factors (grown on overflow)
    double node_f = 8.0, pair_f = 1.0, pool_f = 1.0;
    // launch shapes (h2_set_tuning): 0 / -1 keep the defaults
    int tune_level_batch = 0, tune_one_class = -1;
};

namespace {

constexpr int kLevelBatch = 4;   // tree levels queued between two host checks (default)

// Levels per batch: kLevelBatch, or TSNE_LEVEL_BATCH (1..8) when set; read once.
int level_batch()
{
    static const int b = [] {
        const char *e = getenv("TSNE_LEVEL_BATCH");
        const int v = e ? atoi(e) : kLevelBatch;
        return (v >= 1 && v <= 8) ? v : kLevelBatch;
    }();
    return b;
}
constexpr double kIdF32MinTol = 1e-3;  // smallest ID tolerance of the mixed path's FP32 ID
constexpr int kMaxLevels = 32;   // >= H2Cfg<D>::B + 1 for both dimensions

// Radix sorts of the Morton keys and pair lists.  From kOnesweepMin items on, always the
// onesweep algorithm: the library default sorts up to 2^20 items with a block sort plus
// merge passes, which near 10^6 items costs several times more than the onesweep digit
// passes.  Digits of 10 bits (1024-thread blocks, 8 keys per thread): 6 passes over the
// 60-bit 2D keys instead of 8 with the library's 8-bit default, ~50 us less per 1M-point
// rebuild.  Smaller sorts keep the default, which needs fewer launches (the small jobs of
// a job server are bound by launch latency, not bandwidth).
using SortCfg = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 8>, rocprim::kernel_config<1024, 8>,
                                        10, rocprim::block_radix_rank_algorithm::match>,
    0>;
constexpr int kOnesweepMin = 1 << 18;
// From this many items on, the library's own radix sort (radix.hip: no device-wide
// atomics) sorts the Morton keys and pair lists; below it the library default, which
// needs fewer launches.  TSNE_SORT=rocprim keeps the library sorts at every size.
constexpr int kOwnSortMin = 1 << 15;

bool own_sort(int n)
{
    const char *e = getenv("TSNE_SORT");  // read per sort: a test switches it at run time
    return !(e && strcmp(e, "rocprim") == 0) && n >= kOwnSortMin;
}

template <typename V>
hipError_t sort_pairs64(void *tmp, size_t &bytes, rocprim::double_buffer<uint64_t> &k,
                        rocprim::double_buffer<V> &v, int n, int endbit, hipStream_t st)
{
    return n >= kOnesweepMin
               ? rocprim::radix_sort_pairs<SortCfg>(tmp, bytes, k, v, n, 0, endbit, st)
               : rocprim::radix_sort_pairs(tmp, bytes, k, v, n, 0, endbit, st);
}

hipError_t sort_keys64(void *tmp, size_t &bytes, rocprim::double_buffer<uint64_t> &k, int n,
                       int endbit, hipStream_t st)
{
    return n >= kOnesweepMin ? rocprim::radix_sort_keys<SortCfg>(tmp, bytes, k, n, 0, endbit, st)
                             : rocprim::radix_sort_keys(tmp, bytes, k, n, 0, endbit, st);
}

// Stream-ordered device memory on the structure's stream (see the workspace allocator).
template <typename T> bool dmalloc(T **p, size_t n, hipStream_t st)
{
    if (n == 0) n = 1;
    if (hipMallocAsync((void **)p, n * sizeof(T), st) != hipSuccess) {
        *p = nullptr;
        return false;
    }
    return true;
}

template <typename T> void dfree(T *&p, hipStream_t st)
{
    if (p) (void)hipFreeAsync((void *)p, st);
    p = nullptr;
}

int rmax_of(int D) { return D == 2 ? H2Cfg<2>::RMAX : H2Cfg<3>::RMAX; }
int nch_of(int D) { return D == 2 ? 4 : 8; }
int nearcap_of(int D) { return D == 2 ? H2Cfg<2>::NEAR_CAP : H2Cfg<3>::NEAR_CAP; }
int pmax_of(int D) { return D == 2 ? H2Cfg<2>::PMAX : H2Cfg<3>::PMAX; }

void free_sized(H2 *h)
{
    H2Dev &d = h->d;
    dfree(d.nb, h->st); dfree(d.ne, h->st); dfree(d.level, h->st); dfree(d.parent, h->st); dfree(d.child0, h->st); dfree(d.nchild, h->st);
    dfree(d.cent, h->st); dfree(d.icoord, h->st); dfree(d.bounds, h->st); dfree(d.childcnt, h->st); dfree(d.childoff, h->st);
    dfree(d.near, h->st); dfree(d.near_cnt, h->st); dfree(d.leaves, h->st);
    dfree(d.adm, h->st); dfree(d.xl, h->st); dfree(d.wl, h->st); dfree(d.nl, h->st); dfree(d.pend[0], h->st); dfree(d.pend[1], h->st);
    dfree(d.sort_tmp, h->st);
    dfree(d.adm_beg, h->st); dfree(d.adm_end, h->st); dfree(d.x_beg, h->st); dfree(d.x_end, h->st);
    dfree(d.w_beg, h->st); dfree(d.w_end, h->st); dfree(d.nl_beg, h->st); dfree(d.nl_end, h->st);
    dfree(d.rank, h->st); dfree(d.uoff, h->st); dfree(d.skel, h->st); dfree(d.skelf, h->st); dfree(d.upool, h->st);
    dfree(d.wc, h->st); dfree(d.zc, h->st); dfree(d.nflag, h->st);
    dfree(h->cub_tmp, h->st);
    h->cub_bytes = 0;
}

// (Re)allocate everything whose size depends on the node / pair / pool capacities.
bool alloc_sized(H2 *h)
{
    free_sized(h);
    H2Dev &d = h->d;
    const int D = h->D, RMAX = rmax_of(D), NCH = nch_of(D);
    const long long ncap = std::max(1024LL, (long long)(h->node_f * h->n / h->leaf) + 64);
    if (ncap > (1LL << 30)) return false;
    d.node_cap = (int)ncap;
    const size_t nc = (size_t)ncap;
    const double pf = h->pair_f;
    d.adm_cap = (int)std::min<double>(2e9, pf * nc * (D == 2 ? 32 : 200));
    d.x_cap = (int)std::min<double>(2e9, pf * nc * (D == 2 ? 8 : 32));
    d.w_cap = d.x_cap;
    d.nl_cap = (int)std::min<double>(2e9, pf * nc * (D == 2 ? 16 : 64));
    d.pend_cap = (int)std::min<double>(2e9, pf * nc * (D == 2 ? 8 : 32));
    // U pool: n x 2 RMAX floats for the leaves (placed by point), then the internal
    // nodes' m x 2 RMAX reservations (grown on overflow).  A node's region holds its R
    // entries in FP64 while its ID runs, then its U (m x r floats).
    d.pool_cap = 2LL * h->n * RMAX + (long long)(h->pool_f * (double)nc * (D == 2 ? 24 : 32) * 2 * RMAX);
    const int maxpair = std::max({d.adm_cap, d.x_cap, d.w_cap, d.nl_cap});
    bool ok = dmalloc(&d.nb, nc, h->st) && dmalloc(&d.ne, nc, h->st) && dmalloc(&d.level, nc, h->st) &&
              dmalloc(&d.parent, nc, h->st) && dmalloc(&d.child0, nc, h->st) && dmalloc(&d.nchild, nc, h->st) &&
              dmalloc(&d.cent, nc, h->st) && dmalloc(&d.icoord, nc * 3, h->st) &&
              dmalloc(&d.bounds, nc * (NCH + 1), h->st) && dmalloc(&d.childcnt, nc, h->st) &&
              dmalloc(&d.childoff, std::max(nc, (size_t)kSplitBlocks), h->st) && dmalloc(&d.near, nc * nearcap_of(D), h->st) &&
              dmalloc(&d.near_cnt, nc, h->st) && dmalloc(&d.leaves, nc, h->st) &&
              // the four lists and the sort scratch trade buffers after every sort (the
              // sorted keys stay where the sort left them), so all five have the same size
              dmalloc(&d.adm, (size_t)maxpair, h->st) && dmalloc(&d.xl, (size_t)maxpair, h->st) &&
              dmalloc(&d.wl, (size_t)maxpair, h->st) && dmalloc(&d.nl, (size_t)maxpair, h->st) &&
              dmalloc(&d.pend[0], (size_t)d.pend_cap, h->st) && dmalloc(&d.pend[1], (size_t)d.pend_cap, h->st) &&
              dmalloc(&d.sort_tmp, (size_t)maxpair, h->st) &&
              dmalloc(&d.adm_beg, nc, h->st) && dmalloc(&d.adm_end, nc, h->st) && dmalloc(&d.x_beg, nc, h->st) &&
              dmalloc(&d.x_end, nc, h->st) && dmalloc(&d.w_beg, nc, h->st) && dmalloc(&d.w_end, nc, h->st) &&
              dmalloc(&d.nl_beg, nc, h->st) && dmalloc(&d.nl_end, nc, h->st) && dmalloc(&d.rank, nc, h->st) &&
              dmalloc(&d.uoff, nc, h->st) && dmalloc(&d.skel, nc * RMAX * D, h->st) &&
              dmalloc(&d.skelf, nc * RMAX, h->st) && dmalloc(&d.upool, (size_t)d.pool_cap, h->st) &&
              dmalloc(&d.wc, nc * RMAX, h->st) && dmalloc(&d.zc, nc * RMAX, h->st) &&
              dmalloc(&d.nflag, nc, h->st);
    if (!ok) return false;
    // temporary storage for the sorts (largest of the uses)
    size_t b1 = 0, b3 = 0;
    rocprim::double_buffer<uint64_t> kb(d.keys, h->keys_alt);
    rocprim::double_buffer<int> vb(d.perm, h->perm_alt);
    // temporary storage: the larger of the two algorithms at the sizes in use
    size_t b1a = 0, b3a = 0;
    (void)rocprim::radix_sort_pairs<SortCfg>(nullptr, b1, kb, vb, h->n, 0, 64, h->st);
    (void)rocprim::radix_sort_pairs(nullptr, b1a, kb, vb, h->n, 0, 64, h->st);
    b1 = std::max(b1, b1a);
    rocprim::double_buffer<uint64_t> pb(d.adm, d.sort_tmp);
    (void)rocprim::radix_sort_keys<SortCfg>(nullptr, b3, pb, maxpair, 0, 64, h->st);
    (void)rocprim::radix_sort_keys(nullptr, b3a, pb, maxpair, 0, 64, h->st);
    b3 = std::max(b3, b3a);
    h->cub_bytes = std::max({b1, b3, radix_scratch_bytes(std::max(h->n, maxpair))});
    return hipMallocAsync(&h->cub_tmp, h->cub_bytes, h->st) == hipSuccess;
}

int sort_pairs_list(H2 *h, uint64_t *&buf, int n, int *beg, int *end, hipStream_t s)
{
    H2Dev &d = h->d;
    // beg / end are zero for every node of the tree (set where the node is made)
    if (n <= 0) return 0;
    int bits = 32;
    while ((1LL << (bits - 32)) < d.node_cap) bits++;
    if (own_sort(n)) {
        // on return buf holds the sorted list (buf and sort_tmp exchanged as needed;
        // both hold maxpair entries)
        if (radix_sort_keys_u64(buf, d.sort_tmp, n, 0, bits, h->cub_tmp, h->cub_bytes, s) != 0)
            return -1;
        return launch_pair_offsets(buf, n, beg, end, s);
    }
    rocprim::double_buffer<uint64_t> db(buf, d.sort_tmp);
    size_t bytes = h->cub_bytes;
    if (sort_keys64(h->cub_tmp, bytes, db, n, bits, s) != hipSuccess)
        return -1;
    if (db.current() != buf) {  // the sorted list now lives in the scratch buffer:
        std::swap(buf, d.sort_tmp);  // exchange them (both hold maxpair entries)
    }
    return launch_pair_offsets(buf, n, beg, end, s);
}

// Pinned host blocks of the H2 objects (mapped, coherent; the per-level records and host
// scratch), kept on a process-wide free list per device: freeing pinned memory can
// synchronise the whole device, which would stall every other job on it each time one
// job ends.
constexpr size_t kPinBytes = sizeof(int) * (CNT_COUNT + kPinRec * kMaxLevels + FLAG_COUNT);
constexpr size_t kPinFlags = CNT_COUNT + kPinRec * kMaxLevels;  // the ID's flags (ints)
struct PinBlock {
    int device;
    int *host, *dev;
};
std::mutex g_pin_mu;
std::vector<PinBlock> g_pin_free;

bool pin_take(int **host, int **dev)
{
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return false;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        for (size_t k = 0; k < g_pin_free.size(); k++)
            if (g_pin_free[k].device == device) {
                *host = g_pin_free[k].host;
                *dev = g_pin_free[k].dev;
                g_pin_free.erase(g_pin_free.begin() + (long)k);
                return true;
            }
    }
    if (hipHostMalloc((void **)host, kPinBytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        *host = nullptr;
        return false;
    }
    if (hipHostGetDevicePointer((void **)dev, *host, 0) != hipSuccess) {
        (void)hipHostFree(*host);
        *host = nullptr;
        return false;
    }
    return true;
}

// Back to the free list; the caller has synchronised every stream that wrote the block.
void pin_give(int *host, int *dev)
{
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return;  // kept out of the list (leaked)
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pin_free.push_back({device, host, dev});
}

// The ID's flags of the last build, copied to pinned memory behind it: a rank that
// reached its cap short of the ID tolerance is counted once per build and reported once.
// A U-pool overflow cannot happen when the host skipped its wait (the tree's count of
// children per part bounded every reservation); if it ever did, the next build fails
// with a message.
int absorb_id_flags(H2 *h)
{
    if (!h->flags_pending) return 0;
    if (hipEventSynchronize(h->ev_flags) != hipSuccess) return -1;
    h->flags_pending = false;
    const int *f = h->pin + kPinFlags;
    if (f[FLAG_POOL_OVF]) {
        fprintf(stderr, "tsne_h2pack: an ID overflowed the U pool sized from the tree\n");
        return -1;
    }
    if (f[FLAG_RANK_CAP] && h->rank_capped++ == 0)
        fprintf(stderr, "tsne_h2pack: warning: a skeleton rank reached its cap (%d) before "
                        "the ID tolerance %.3g; the product may exceed h2_tol = %.3g "
                        "(counted in tsne_h2_stats.rank_capped_builds)\n",
                rmax_of(h->D), h->id_tol, h->tol);
    return 0;
}

// The caller's one-shot hook, queued once the build's front (root box, keys, sort, sorted
// gather copy) is on the stream.
int run_front_hook(H2 *h)
{
    int (*fn)(void *) = h->front_hook;
    if (!fn) return 0;
    h->front_hook = nullptr;
    return fn(h->front_arg);
}

// One attempt at a rebuild.  Returns 0 (built), 1 (capacity exceeded: grown, retry),
// -1 (error), -2 (non-finite input).
int build_once(H2 *h, const void *Yg)
{
    H2Dev &d = h->d;
    hipStream_t st = h->st;
    const int D = h->D;
    // counters and flags are cleared by the root-box kernel
    if (h->tl_on) (void)hipEventRecord(h->tl_ev[TL_START], st);
    if (launch_root_box(D, Yg, h->n, d, h->part, h->nblk, st)) return -1;
    if (launch_morton(D, Yg, d, st)) return -1;
    const int endbit = D * (D == 2 ? H2Cfg<2>::B : H2Cfg<3>::B);
    if (own_sort(h->n)) {
        if (radix_sort_pairs_u64(d.keys, h->keys_alt, d.perm, h->perm_alt, h->n, 0, endbit,
                                 h->cub_tmp, h->cub_bytes, st) != 0)
            return -1;
    } else {
        rocprim::double_buffer<uint64_t> kb(d.keys, h->keys_alt);
        rocprim::double_buffer<int> vb(d.perm, h->perm_alt);
        size_t bytes = h->cub_bytes;
        if (sort_pairs64(h->cub_tmp, bytes, kb, vb, h->n, endbit, st) != hipSuccess)
            return -1;
        if (kb.current() != d.keys) std::swap(d.keys, h->keys_alt);
        if (vb.current() != d.perm) std::swap(d.perm, h->perm_alt);
    }
    if (launch_gather_sorted(D, Yg, d, st)) return -1;
    if (run_front_hook(h)) return -1;
    if (h->tl_on) (void)hipEventRecord(h->tl_ev[TL_FRONT], st);
    if (launch_init_root(d, st)) return -1;
    // Levels are queued in batches of kLevelBatch (the first batch: the previous tree's
    // depth, below).  After queuing batch k >= 1 the host waits for batch k - 1 (recorded
    // in ev_lvl), never for the batch it has just queued, and reads that batch's last
    // published level from pinned memory (one record of kPinRec words per level, so later
    // levels cannot overwrite it): whether the tree has ended, and the exact node count
    // that bounds the next grids.  Levels past the last one are empty and cost only their
    // launches.
    const int NCH = (D == 2) ? H2Cfg<2>::NCH : H2Cfg<3>::NCH;
    const int maxlev = (D == 2) ? H2Cfg<2>::B + 1 : H2Cfg<3>::B + 1;  // level B holds leaves only
    // With the previous build's tree as a hint (the points move little between rebuilds),
    // the first batch holds that build's depth and the host reads its last record at
    // once: the tree usually ends there, with one host wait for the whole tree and no
    // empty level queued past it.  1M 2D whole schedule 753 -> 762 it/s; serving mix of
    // 2K-200K-point jobs 8.05 -> 8.25 jobs/s (12 slots); 500K 3D 490 -> 492; the results
    // are the same bit for bit.  TSNE_LEVEL_HINTS=0 turns it off.  Sizing each level's
    // grids from the hint's per-level counts as well (TSNE_LEVEL_GRID_HINTS=1; the
    // kernels' loops cover any count) measured slower at 1M (749 it/s) and is off.
    int *plev = h->pin + CNT_COUNT, *plev_dev = h->pin_dev + CNT_COUNT;
    for (int k = 0; k < maxlev * kPinRec; k++) plev[k] = -1;
    int l = 0, pend = 0;            // the level pass of level l appends to pend[pend]
    const int lb = h->tune_level_batch > 0 ? h->tune_level_batch : level_batch();
    const bool use_hints = h->level_hints != 0;
    static const std::vector<int> no_hint;
    const std::vector<int> &hint = use_hints ? h->prev_cnt : no_hint;
    const int first = hint.empty() ? lb : std::min<int>(maxlev, (int)hint.size());
    int known_l = 0, known_n = 1;   // level known_l has exactly known_n nodes
    hipEvent_t last_ev = nullptr;   // end of the last queued batch
    std::vector<int> bend;          // end level of each queued batch
    for (int batch = 0; l < maxlev; batch++) {
        const int bsz = (batch == 0) ? first : lb;
        for (int k = 0; k < bsz && l < maxlev; k++, l++) {
            long long bound = known_n;
            for (int q = known_l; q < l && bound < d.node_cap; q++) bound *= NCH;
            bound = std::min<long long>(bound, d.node_cap);
            static const bool grid_hints = [] {
                const char *e = getenv("TSNE_LEVEL_GRID_HINTS");
                return e && e[0] == '1';
            }();
            if (l > known_l && !hint.empty() && grid_hints) {
                const long long hb = (l < (int)hint.size()) ? hint[l] + hint[l] / 4 + 64 : 64;
                bound = std::min(bound, hb);
            }
            int nsplit = 1;
            if (launch_level_pass(D, d, l, (int)bound, pend ^ 1, pend, &nsplit, st) ||
                launch_level_write(D, d, l, nsplit, pend ^ 1, plev_dev + kPinRec * l, st))
                return -1;
            pend ^= 1;
        }
        if (hipEventRecord(h->ev_lvl[batch & 1], st) != hipSuccess) return -1;
        last_ev = h->ev_lvl[batch & 1];
        bend.push_back(l);
        if (batch == 0 && hint.empty()) continue;
        // the batch whose records are read now: the first one at once when it holds the
        // hinted depth, else the one before the batch just queued
        const int wb = (batch == 0) ? 0 : batch - 1;
        if (hipEventSynchronize(h->ev_lvl[wb & 1]) != hipSuccess) return -1;
        const int lc = bend[wb] - 1;  // last level of batch wb
        const int *rec = plev + kPinRec * lc;
        if (rec[1] || rec[2]) {  // non-finite input / node arrays full: drain, then report
            if (hipStreamSynchronize(st) != hipSuccess) return -1;
            if (rec[1]) return -2;
            h->node_f *= 2.0;
            return 1;
        }
        known_l = lc + 1;
        known_n = rec[0];
        if (known_n == 0) break;  // level lc + 1 is empty: the tree has ended
    }
    if (h->tl_on) (void)hipEventRecord(h->tl_ev[TL_TREE], st);
    {
        // level offsets, pair counters and flags of the finished tree: one wait for the last
        // batch, then the per-level records (the last queued level's record holds the final
        // counts and flags: no level after the tree's end adds anything)
        if (!last_ev || hipEventSynchronize(last_ev) != hipSuccess) return -1;
        std::vector<int> lv(l + 2);
        lv[0] = 0;
        lv[1] = 1;
        for (int k = 0; k < l; k++) lv[k + 2] = plev[kPinRec * k + 9];
        const int *fin = plev + kPinRec * (l - 1);
        for (int &f : h->hflags) f = 0;
        h->hflags[FLAG_NONFINITE] = fin[1];
        h->hflags[FLAG_NODE_OVF] = fin[2];
        h->hflags[FLAG_NEAR_OVF] = fin[3] & 1;
        h->hflags[FLAG_PAIR_OVF] = (fin[3] >> 1) & 1;
        h->hflags[FLAG_BIG_LEAF] = (fin[3] >> 2) & 1;
        h->pin[CNT_ADM] = fin[4];
        h->pin[CNT_X] = fin[5];
        h->pin[CNT_W] = fin[6];
        h->pin[CNT_NEAR] = fin[7];
        h->pin[CNT_LEAF] = fin[8];
        h->pin[CNT_NEED] = fin[10];  // children of the fullest U-pool part
        h->lvl.assign({0, 1});
        for (int k = 1; k + 1 < (int)lv.size() && lv[k + 1] > lv[k]; k++) h->lvl.push_back(lv[k + 1]);
        h->prev_cnt.resize(h->lvl.size() - 1);
        for (size_t k = 0; k + 1 < h->lvl.size(); k++) h->prev_cnt[k] = h->lvl[k + 1] - h->lvl[k];
    }
    h->n_nodes = h->lvl.back();
    if (h->hflags[FLAG_NONFINITE]) return -2;
    if (h->hflags[FLAG_NODE_OVF]) {
        h->node_f *= 2.0;
        return 1;
    }
    if (h->hflags[FLAG_PAIR_OVF] || h->hflags[FLAG_NEAR_OVF]) {
        if (h->hflags[FLAG_NEAR_OVF]) return -1;  // geometry the fixed near list cannot hold
        h->pair_f *= 2.0;
        return 1;
    }
    h->n_adm = h->pin[CNT_ADM];
    h->n_x = h->pin[CNT_X];
    h->n_w = h->pin[CNT_W];
    h->n_near = h->pin[CNT_NEAR];
    h->n_leaves = h->pin[CNT_LEAF];
    // Whether the U pool holds every reservation the ID can make: an internal node
    // reserves 2 m RMAX floats with m (its candidates) at most RMAX per child, in pool part
    // (its index in the level) mod kPoolParts, and the tree counted the children per part
    // (record word 10).  Then the ID cannot overflow the pool and the product follows the
    // ID without the host reading the flags in between; otherwise (a pool sized below the
    // worst case, as 3D trees can be) the host reads them after the ID and a build that
    // overflowed is retried with a larger pool.
    const bool pool_safe = [&] {
        const int RMAX = rmax_of(D);
        return 2LL * RMAX * RMAX * (long long)std::max(0, h->pin[CNT_NEED]) <= pool_part_cap(d, RMAX);
    }();
    // ADM and X ranges were written per node by the list kernel; NEAR and W pairs are
    // appended by several levels and are grouped by target below.  Only the leaf product
    // reads those lists, so with a second stream the sorts run there, beside the ID.  The
    // ID launches are queued first: queueing the sorts' launches takes the host a few
    // hundred microseconds, which the main stream would otherwise spend idle.
    hipStream_t ss = st;
    if (h->st_aux) {
        if (hipEventRecord(h->ev_tree, st) != hipSuccess ||
            hipStreamWaitEvent(h->st_aux, h->ev_tree, 0) != hipSuccess)
            return -1;
        ss = h->st_aux;
    }
    // ID + upward: every leaf at once (they need nothing from other nodes), then the
    // internal nodes level by level, finest first
    const int nlev = (int)h->lvl.size() - 1;
    const bool f32 = h->id_mixed && h->id_tol >= kIdF32MinTol;
    const bool tid = h->timing && h->ev_id0 && h->ev_id1;
    if (h->tl_on) (void)hipEventRecord(h->tl_ev[TL_ID0], st);
    if (tid && hipEventRecord(h->ev_id0, st) != hipSuccess) return -1;
    if (h->n_nodes > 1 &&
        launch_id_up(D, d, d.leaves, 0, h->n_leaves, h->id_tol, h->pq, f32, h->tune_one_class, st))
        return -1;
    for (int lv = nlev - 1; lv >= 1; lv--)
        if (launch_id_up(D, d, nullptr, h->lvl[lv], h->lvl[lv + 1] - h->lvl[lv], h->id_tol, h->pq,
                         f32, h->tune_one_class, st))
            return -1;
    if (tid && hipEventRecord(h->ev_id1, st) != hipSuccess) return -1;
    if (h->tl_on) (void)hipEventRecord(h->tl_ev[TL_ID1], st);
    if (sort_pairs_list(h, d.wl, h->n_w, d.w_beg, d.w_end, ss) ||
        sort_pairs_list(h, d.nl, h->n_near, d.nl_beg, d.nl_end, ss))
        return -1;
    if (h->st_aux && hipEventRecord(h->ev_sorted, h->st_aux) != hipSuccess) return -1;
    if (h->tl_on) (void)hipEventRecord(h->tl_ev[TL_SORTED], ss);
    if (h->st_aux && hipStreamWaitEvent(st, h->ev_sorted, 0) != hipSuccess) return -1;
    if (hipMemcpyAsync(h->pin + kPinFlags, d.flags, sizeof(int) * FLAG_COUNT,
                       hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipEventRecord(h->ev_flags, st) != hipSuccess)
        return -1;
    h->flags_pending = true;
    if (!pool_safe) {
        if (hipEventSynchronize(h->ev_flags) != hipSuccess) return -1;
        if (h->pin[kPinFlags + FLAG_POOL_OVF]) {
            h->flags_pending = false;
            h->pool_f *= 2.0;
            return 1;
        }
    }
    if (h->tl_on) {
        // the end mark goes in after the host's last wait: the gap to the product's first
        // launch shows in the caller's timing, not here
        (void)hipEventRecord(h->tl_ev[TL_END], st);
        (void)hipEventSynchronize(h->tl_ev[TL_END]);
        for (int k = 1; k < kTlEvents; k++) {
            float ms = 0.f;
            const int from = (k == TL_SORTED) ? TL_ID0 : (k == TL_END ? TL_SORTED : k - 1);
            if (hipEventElapsedTime(&ms, h->tl_ev[from], h->tl_ev[k]) == hipSuccess) h->tl_ms[k] += ms;
        }
        float all = 0.f;
        if (hipEventElapsedTime(&all, h->tl_ev[TL_START], h->tl_ev[TL_END]) == hipSuccess)
            h->tl_ms[0] += all;
        h->tl_builds++;
    }
    if (tid) {
        float ms = 0.f;
        if (hipEventSynchronize(h->ev_id1) == hipSuccess &&
            hipEventElapsedTime(&ms, h->ev_id0, h->ev_id1) == hipSuccess)
            h->id_ms = ms;
    }
    return 0;
}

}  // namespace

H2 *h2_create(int n, int dim, int leaf_size, double tol, hipStream_t st)
{
    H2 *h = new (std::nothrow) H2();
    if (!h) return nullptr;
    h->n = n;
    h->D = dim;
    h->leaf = std::max(1, std::min(leaf_size, kLeafCap));
    h->tol = tol;
    // proxy selection a little tighter than the ID tolerance, clamped as the reference
    // clamps its proxy tolerance to [1e-6, 1e-2] (src/tsne_h2pack.c:966-977)
    h->ptol = std::min(1e-2, std::max(1e-6, 0.15 * tol));
    // the node IDs run at half the requested tolerance: the errors of the nested bases
    // add up over the levels, and the gradient error is what h2_tol bounds
    h->id_tol = 0.5 * tol;
    h->st = st;
    if (const char *e = getenv("TSNE_LEVEL_HINTS")) h->level_hints = (e[0] == '1') ? 1 : 0;
    H2Dev &d = h->d;
    d.n = n;
    d.D = dim;
    d.leaf = h->leaf;
    h->nblk = std::max(1, std::min(512, (n + 4095) / 4096));
    const size_t vb = (dim == 2) ? sizeof(float2) : sizeof(float4);
    bool ok = dmalloc(&d.keys, n, h->st) && dmalloc(&h->keys_alt, n, h->st) && dmalloc(&d.perm, n, h->st) &&
              dmalloc(&h->perm_alt, n, h->st) && dmalloc((char **)&d.Ys, (size_t)n * vb, h->st) &&
              dmalloc(&d.root, 1, h->st) && dmalloc(&d.flags, FLAG_COUNT, h->st) && dmalloc(&d.cnt, CNT_COUNT, h->st) &&
              dmalloc(&d.lvl, kMaxLevels + 4, h->st) &&
              dmalloc(&h->part, (size_t)h->nblk * 7, h->st) &&
              pin_take(&h->pin, &h->pin_dev) &&
              hipEventCreateWithFlags(&h->ev_lvl[0], hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&h->ev_flags, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&h->ev_lvl[1], hipEventDisableTiming) == hipSuccess;
    if (ok && getenv("TSNE_H2_TIMELINE") && getenv("TSNE_H2_TIMELINE")[0] == '1') {
        h->tl_on = true;
        for (hipEvent_t &e : h->tl_ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    }
    if (ok) {
        const ProxyTable &pt = proxy_table(dim, h->ptol, 1.0, pmax_of(dim));
        int cmax = 0;
        for (int c : pt.cnt) cmax = std::max(cmax, c);
        h->pq = cmax;
        ok = dmalloc(&h->d_proxy, pt.pts.size(), h->st) && dmalloc(&h->d_proxy_cnt, pt.cnt.size(), h->st) &&
             hipMemcpyAsync(h->d_proxy, pt.pts.data(), pt.pts.size() * sizeof(double),
                            hipMemcpyHostToDevice, st) == hipSuccess &&
             hipMemcpyAsync(h->d_proxy_cnt, pt.cnt.data(), pt.cnt.size() * sizeof(int),
                            hipMemcpyHostToDevice, st) == hipSuccess &&
             hipStreamSynchronize(st) == hipSuccess;
        d.proxy_pts = h->d_proxy;
        d.proxy_cnt = h->d_proxy_cnt;
    }
    ok = ok && alloc_sized(h);
    if (!ok) {
        fprintf(stderr, "tsne_h2pack: H2 allocation failed (n=%d)\n", n);
        h2_destroy(h);
        return nullptr;
    }
    return h;
}

void h2_set_tuning(H2 *h, int level_batch, int one_class_level)
{
    if (!h) return;
    h->tune_level_batch = (level_batch >= 1 && level_batch <= 8) ? level_batch : 0;
    h->tune_one_class = one_class_level >= 0 ? one_class_level : -1;
}

void h2_set_timing(H2 *h, int on)
{
    if (!h) return;
    if (on && !h->ev_id0) {
        if (hipEventCreate(&h->ev_id0) != hipSuccess || hipEventCreate(&h->ev_id1) != hipSuccess) {
            if (h->ev_id0) (void)hipEventDestroy(h->ev_id0);
            h->ev_id0 = h->ev_id1 = nullptr;
            return;
        }
    }
    h->timing = on != 0;
}

void h2_set_id_mixed(H2 *h, int mixed)
{
    if (h) h->id_mixed = mixed != 0;
}

void h2_set_aux_stream(H2 *h, hipStream_t st_aux)
{
    if (!h || h->st_aux || !st_aux) return;
    if (hipEventCreateWithFlags(&h->ev_tree, hipEventDisableTiming) == hipSuccess &&
        hipEventCreateWithFlags(&h->ev_sorted, hipEventDisableTiming) == hipSuccess) {
        h->st_aux = st_aux;
        return;
    }
    if (h->ev_tree) (void)hipEventDestroy(h->ev_tree);
    if (h->ev_sorted) (void)hipEventDestroy(h->ev_sorted);
    h->ev_tree = h->ev_sorted = nullptr;
}

void h2_destroy(H2 *h)
{
    if (!h) return;
    if (h->st_aux) (void)hipStreamSynchronize(h->st_aux);
    if (h->ev_tree) (void)hipEventDestroy(h->ev_tree);
    if (h->ev_sorted) (void)hipEventDestroy(h->ev_sorted);
    if (h->st) (void)hipStreamSynchronize(h->st);
    if (h->tl_on && h->tl_builds > 0) {
        const double b = h->tl_builds;
        fprintf(stderr, "[h2 timeline] %d builds, ms per build: total %.3f | front %.3f tree %.3f "
                        "readback+host %.3f ID %.3f | ID start to sorts done %.3f, sorts done to end %.3f\n",
                h->tl_builds, h->tl_ms[0] / b, h->tl_ms[TL_FRONT] / b, h->tl_ms[TL_TREE] / b,
                h->tl_ms[TL_ID0] / b, h->tl_ms[TL_ID1] / b, h->tl_ms[TL_SORTED] / b,
                h->tl_ms[TL_END] / b);
    }
    for (hipEvent_t e : h->tl_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : h->ev_lvl)
        if (e) (void)hipEventDestroy(e);
    if (h->ev_id0) (void)hipEventDestroy(h->ev_id0);
    if (h->ev_id1) (void)hipEventDestroy(h->ev_id1);
    if (h->ev_flags) {
        (void)absorb_id_flags(h);  // a rank cap in the last build is still reported
        (void)hipEventDestroy(h->ev_flags);
    }
    free_sized(h);
    H2Dev &d = h->d;
    dfree(d.keys, h->st); dfree(h->keys_alt, h->st); dfree(d.perm, h->st); dfree(h->perm_alt, h->st); dfree(d.Ys, h->st);
    dfree(d.root, h->st); dfree(d.flags, h->st); dfree(d.cnt, h->st); dfree(h->part, h->st);
    dfree(d.lvl, h->st);
    dfree(h->d_proxy, h->st); dfree(h->d_proxy_cnt, h->st);
    if (h->pin) pin_give(h->pin, h->pin_dev);
    delete h;
}

static int build_all(H2 *h, const void *Yg, int *nonfinite);

void h2_set_front_hook(H2 *h, int (*fn)(void *), void *arg)
{
    h->front_hook = fn;
    h->front_arg = arg;
}

int h2_build(H2 *h, const void *Yg, int *nonfinite)
{
    const int rc = build_all(h, Yg, nonfinite);
    const int hr = run_front_hook(h);  // a build that stopped before its front
    return rc != 0 ? rc : (hr != 0 ? -1 : 0);
}

static int build_all(H2 *h, const void *Yg, int *nonfinite)
{
    *nonfinite = 0;
    h->usable = 0;
    if (absorb_id_flags(h) != 0) return -1;  // the previous build's ID
    const char *why = "capacity retries exhausted";
    for (int attempt = 0; attempt < 6; attempt++) {
        const int rc = build_once(h, Yg);
        if (rc == 0) {
            h->usable = !h->hflags[FLAG_BIG_LEAF];
            if (h->usable) return 0;
            why = "a leaf at the deepest level holds more points than a leaf can";
            break;
        }
        if (rc == -2) {
            *nonfinite = 1;
            return -1;
        }
        if (rc < 0) {
            // the structure cannot be built (e.g. a near list beyond its capacity):
            // the caller falls back to the direct sum
            if (hipGetLastError() != hipSuccess) return -1;
            why = h->hflags[FLAG_NEAR_OVF] ? "a near list over its capacity" : "a device error";
            break;
        }
        if (!alloc_sized(h)) return -1;
    }
    if (h->fallbacks++ == 0) {
        fprintf(stderr, "tsne_h2pack: H2 structure not usable (%s); this product and any like it "
                        "use the direct O(N^2) sum\n", why);
        if (getenv("TSNE_H2_DEBUG") && h->hflags[FLAG_BIG_LEAF]) {
            // the largest leaf: its level, size and the extent of its points
            const int nn = h->n_nodes;
            std::vector<int> nb(nn), ne(nn), lv(nn), c0(nn);
            RootInfo root{};
            (void)hipMemcpy(nb.data(), h->d.nb, nn * sizeof(int), hipMemcpyDeviceToHost);
            (void)hipMemcpy(ne.data(), h->d.ne, nn * sizeof(int), hipMemcpyDeviceToHost);
            (void)hipMemcpy(lv.data(), h->d.level, nn * sizeof(int), hipMemcpyDeviceToHost);
            (void)hipMemcpy(c0.data(), h->d.child0, nn * sizeof(int), hipMemcpyDeviceToHost);
            (void)hipMemcpy(&root, h->d.root, sizeof(RootInfo), hipMemcpyDeviceToHost);
            int best = -1, nbig = 0;
            for (int i = 0; i < nn; i++)
                if (c0[i] < 0) {
                    if (best < 0 || ne[i] - nb[i] > ne[best] - nb[best]) best = i;
                    nbig += ne[i] - nb[i] > (h->D == 2 ? H2Cfg<2>::LEAF_CAP : H2Cfg<3>::LEAF_CAP);
                }
            if (best < 0) best = 0;
            fprintf(stderr, "  nodes %d levels %d, leaves over the cap: %d\n", nn,
                    (int)h->lvl.size() - 1, nbig);
            const int m = ne[best] - nb[best];
            const int vs = h->D == 2 ? 2 : 4;
            std::vector<float> y((size_t)m * vs);
            (void)hipMemcpy(y.data(), (const float *)h->d.Ys + (size_t)nb[best] * vs,
                            y.size() * sizeof(float), hipMemcpyDeviceToHost);
            float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
            for (int k = 0; k < m; k++)
                for (int q = 0; q < h->D; q++) {
                    lo[q] = std::min(lo[q], y[(size_t)k * vs + q]);
                    hi[q] = std::max(hi[q], y[(size_t)k * vs + q]);
                }
            fprintf(stderr, "  root H0 %.6g centre (%.6g %.6g %.6g); leaf %d level %d points %d, "
                            "extent %.3g %.3g %.3g at (%.6g %.6g %.6g)\n",
                    root.H0, root.c[0], root.c[1], root.c[2], best, lv[best], m, hi[0] - lo[0],
                    hi[1] - lo[1], hi[2] - lo[2], lo[0], lo[1], lo[2]);
        }
    }
    return 0;
}

int h2_product(H2 *h, const void *Yg, float *s, float *F, int mixed)
{
    (void)Yg;
    if (!h->usable) return -1;
    H2Dev &d = h->d;
    h->last_mixed = mixed;
    if (launch_couple(h->D, d, h->n_nodes, mixed, (float)h->tol, h->st)) return -1;
    const int nlev = (int)h->lvl.size() - 1;
    for (int lv = 2; lv < nlev; lv++)
        if (launch_down(h->D, d, h->lvl[lv], h->lvl[lv + 1], mixed, h->st)) return -1;
    return launch_leaf_eval(h->D, d, h->n_leaves, s, F, h->st);
}

int h2_usable(const H2 *h) { return h->usable; }

const int *h2_order(const H2 *h) { return h->d.perm; }

int h2_stats(H2 *h, tsne_h2_stats_t *st)
{
    (void)absorb_id_flags(h);  // the last build's rank cap is counted
    memset(st, 0, sizeof(*st));
    st->levels = (int)h->lvl.size() - 1;
    st->n_nodes = h->n_nodes;
    st->n_leaves = h->n_leaves;
    st->leaf_points_max = h->leaf;
    st->n_adm_pairs = h->n_adm + h->n_x + h->n_w;
    st->n_near_pairs = h->n_near;
    st->near_point_pairs = -1;
    st->rank_max = 0;
    st->rank_mean = 0.0;
    st->proxy_max = 0;
    st->mixed_fallback_nodes = 0;
    st->direct_fallbacks = h->fallbacks;
    st->rank_capped_builds = h->rank_capped;
    if (!h->usable || h->n_nodes <= 0) return 0;
    // flop counts of one product from the built structure (host copies; diagnostics only)
    const int nn = h->n_nodes;
    std::vector<int> nb(nn), ne(nn), rk(nn), c0(nn), par(nn), fl(nn), nch(nn), lev(nn);
    RootInfo root{};
    std::vector<uint64_t> adm(h->n_adm), xl(h->n_x), wl(h->n_w), nl(h->n_near);
    bool ok = hipMemcpyAsync(nb.data(), h->d.nb, nn * sizeof(int), hipMemcpyDeviceToHost, h->st) == hipSuccess &&
              hipMemcpyAsync(ne.data(), h->d.ne, nn * sizeof(int), hipMemcpyDeviceToHost, h->st) == hipSuccess &&
              hipMemcpyAsync(rk.data(), h->d.rank, nn * sizeof(int), hipMemcpyDeviceToHost, h->st) == hipSuccess &&
              hipMemcpyAsync(c0.data(), h->d.child0, nn * sizeof(int), hipMemcpyDeviceToHost, h->st) == hipSuccess &&
              hipMemcpyAsync(par.data(), h->d.parent, nn * sizeof(int), hipMemcpyDeviceToHost, h->st) == hipSuccess &&
              hipMemcpyAsync(fl.data(), h->d.nflag, nn * sizeof(int), hipMemcpyDeviceToHost, h->st) == hipSuccess &&
              hipMemcpyAsync(nch.data(), h->d.nchild, nn * sizeof(int), hipMemcpyDeviceToHost, h->st) == hipSuccess &&
              hipMemcpyAsync(lev.data(), h->d.level, nn * sizeof(int), hipMemcpyDeviceToHost, h->st) == hipSuccess &&
              hipMemcpyAsync(&root, h->d.root, sizeof(RootInfo), hipMemcpyDeviceToHost, h->st) == hipSuccess;
    auto cp = [&](std::vector<uint64_t> &v, const uint64_t *src) {
        return v.empty() || hipMemcpyAsync(v.data(), src, v.size() * sizeof(uint64_t),
                                           hipMemcpyDeviceToHost, h->st) == hipSuccess;
    };
    ok = ok && cp(adm, h->d.adm) && cp(xl, h->d.xl) && cp(wl, h->d.wl) && cp(nl, h->d.nl) &&
         hipStreamSynchronize(h->st) == hipSuccess;
    if (!ok) return -1;
    // ranks of the non-root nodes, proxies of the levels in use, mixed-product fallbacks
    long long rsum = 0;
    for (int i = 1; i < nn; i++) {
        st->rank_max = std::max(st->rank_max, rk[i]);
        rsum += rk[i];
        if (h->last_mixed) st->mixed_fallback_nodes += fl[i] != 0;
    }
    st->rank_mean = nn > 1 ? (double)rsum / (nn - 1) : 0.0;
    {
        const ProxyTable &pt = proxy_table(h->D, h->ptol, 1.0, pmax_of(h->D));
        const int nlev = (int)h->lvl.size() - 1;
        for (int l = 1; l < nlev; l++) {
            const int kq = std::min(std::max(root.k0 - l, 0), kProxyLevels - 1);
            if (kq < (int)pt.cnt.size()) st->proxy_max = std::max(st->proxy_max, pt.cnt[kq]);
        }
    }
    const double D = h->D;
    auto tgt = [](uint64_t k) { return (int)(k >> 32); };
    auto src = [](uint64_t k) { return (int)(uint32_t)k; };
    auto npt = [&](int i) { return (double)(ne[i] - nb[i]); };
    double near_pairs = 0.0;
    for (uint64_t k : nl) near_pairs += npt(tgt(k)) * npt(src(k));
    // kernel evaluation (3D+3) and accumulation into the D+1 coefficient columns 2(D+1)
    const double per_entry = (3 * D + 3) + 2 * (D + 1);
    double far_entries = 0.0;
    for (uint64_t k : adm) far_entries += (double)rk[tgt(k)] * rk[src(k)];
    for (uint64_t k : xl) far_entries += (double)rk[tgt(k)] * npt(src(k));
    for (uint64_t k : wl) far_entries += npt(tgt(k)) * rk[src(k)];
    // downward transfer (the rows of U_p that belong to child i) and the leaf
    // expansion U z; the root (level 0) has no basis
    double basis = 0.0;
    for (int i = 1; i < nn; i++) {
        if (c0[i] < 0) basis += npt(i) * rk[i];
        if (par[i] > 0) basis += (double)rk[i] * rk[par[i]];
    }
    st->near_point_pairs = (long long)near_pairs;
    st->near_flops = near_pairs * per_entry;
    st->far_flops = far_entries * per_entry + basis * 2 * (D + 1);
    // FP64 work of the rebuild's IDs: per node with m candidates (its points, at most the
    // candidate cap, or its children's skeleton points), p proxies of its level and rank
    // r, the kernel matrix (3D+3) m p and the pivoted Gram-Schmidt 4 m p r; a single-child
    // node keeps its child's skeleton and costs nothing
    {
        const ProxyTable &pt = proxy_table(h->D, h->ptol, 1.0, pmax_of(h->D));
        const int mcap = h->D == 2 ? H2Cfg<2>::MCAP : H2Cfg<3>::MCAP;
        double idf = 0.0;
        for (int i = 1; i < nn; i++) {
            double m;
            if (c0[i] < 0) {
                m = std::min(ne[i] - nb[i], mcap);
            } else {
                if (nch[i] == 1) continue;
                m = 0;
                for (int q = 0; q < nch[i]; q++) m += rk[c0[i] + q];
            }
            const int kq = std::min(std::max(root.k0 - lev[i], 0), kProxyLevels - 1);
            const double p = kq < (int)pt.cnt.size() ? std::min(pt.cnt[kq], h->pq) : h->pq;
            idf += (3 * D + 3) * m * p + 4.0 * m * p * rk[i];
        }
        st->id_flops = idf;
    }
    st->id_ms = h->id_ms;
    return 0;
}

}  // namespace tsne

extern "C" int tsne_proxy_cache_info(int low_dim, double h2_tol, char *path, int path_len,
                                     unsigned long long *digest, int *from_disk)
{
    if ((low_dim != 2 && low_dim != 3) || !(h2_tol > 0.0)) {
        fprintf(stderr, "tsne_proxy_cache_info: low_dim must be 2 or 3 and h2_tol > 0\n");
        return -1;
    }
    // the proxy tolerance of the H2 build for this h2_tol (H2 setup above)
    const double ptol = std::min(1e-2, std::max(1e-6, 0.15 * h2_tol));
    const int pmax = tsne::pmax_of(low_dim);
    const tsne::ProxyTable &t = tsne::proxy_table(low_dim, ptol, 1.0, pmax);
    if (path && path_len > 0) snprintf(path, (size_t)path_len, "%s", tsne::proxy_cache_file(low_dim, ptol, 1.0, pmax).c_str());
    if (digest) *digest = tsne::proxy_table_digest(t);
    if (from_disk) *from_disk = t.from_disk ? 1 : 0;
    return 0;
}
