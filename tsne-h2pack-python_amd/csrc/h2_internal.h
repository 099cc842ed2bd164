// h2_internal.h -- device layout of the H2 representation and the launchers of its
// kernels (h2_tree.hip, h2_id.hip, h2_product.hip); orchestrated by h2.hip.
//
// Tree: adaptive 2^D-tree over the Morton-sorted points.  The root box is the
// bounding square of the embedding with its half-side rounded up onto the proxy ladder
// (H0 = 2^k0 * kProxyBase), so a node at level l is a dyadic box of half-side
// H0 / 2^l whose proxy set is ladder entry k0 - l.  A node is split while it holds more
// than leaf_size points (and its level is below the key depth).  Nodes are stored level
// by level; the points of a node are a contiguous range [nb, ne) of the sorted order.
//
// Interaction lists (per target node, all computed level by level from the parent's
// near list, reference call stack: H2P_build, src/tsne_h2pack.c:908-919):
//   ADM  (i, j) same level, boxes separated by at least one box: skeleton <- skeleton
//   X    (i, j) j a coarser leaf outside i's far region:          skeleton <- points
//   W    (i, j) i a leaf outside j's far region, j finer:          points   <- skeleton
//   NEAR (i, j) both leaves, neither far:                          points   <- points
// Pair lists are sorted by (target, source) so the product is deterministic.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

#include "common.h"

namespace tsne {

template <int D> struct H2Cfg;
template <> struct H2Cfg<2> {
    static constexpr int B = 30;         // key bits per dimension
    static constexpr int NCH = 4;        // children per node
    static constexpr int RMAX = 32;      // maximum skeleton rank
    static constexpr int PMAX = 48;      // maximum proxy points
    static constexpr int MCAP = 128;     // maximum ID candidates (NCH * RMAX, leaf cap)
    static constexpr int NEAR_CAP = 48;  // near-list slots per node
    static constexpr int LEAF_CAP = 128; // points of a leaf at the deepest level (= MCAP)
};
template <> struct H2Cfg<3> {
    static constexpr int B = 21;          // 63 key bits
    static constexpr int NCH = 8;
    static constexpr int RMAX = 48;
    static constexpr int PMAX = 96;
    static constexpr int MCAP = 384;
    static constexpr int NEAR_CAP = 96;
    static constexpr int LEAF_CAP = 384;
};

// Leaf size limit of the build parameter, and the chunk of target / source points the
// leaf kernel holds at a time.  A leaf at the deepest level can hold more (the points
// of a box too small to split further), up to H2Cfg<D>::LEAF_CAP.
constexpr int kLeafCap = 128;

// Error / status flags (device int array H2Dev::flags).
enum H2Flag {
    FLAG_NEAR_OVF = 0,   // a node's near list exceeded NEAR_CAP
    FLAG_PAIR_OVF = 1,   // a pair buffer (ADM/X/W/NEAR/pending) overflowed
    FLAG_POOL_OVF = 2,   // the U pool overflowed
    FLAG_BIG_LEAF = 3,   // a leaf at the maximum depth holds more than LEAF_CAP points
    FLAG_NONFINITE = 4,  // non-finite coordinate
    FLAG_RANK_CAP = 5,   // an ID stopped at RMAX before reaching the tolerance
    FLAG_NODE_OVF = 6,   // the node arrays overflowed
    FLAG_COUNT = 16
};

// Counters of the pair buffers and the U pool (device int array H2Dev::cnt).  Device-wide
// atomics on one cache line serialise across the whole GPU, so every counter has a
// 128-byte line of its own (kCntSlot ints apart), and the internal-node U pool is split
// into kPoolParts regions, each with its own counter.
constexpr int kCntSlot = 32;
constexpr int kPoolParts = 8;
enum H2Count { CNT_ADM = 0, CNT_X = kCntSlot, CNT_W = 2 * kCntSlot, CNT_NEAR = 3 * kCntSlot,
               CNT_PEND0 = 4 * kCntSlot, CNT_PEND1 = 5 * kCntSlot, CNT_LEAF = 6 * kCntSlot,
               CNT_POOL = 7 * kCntSlot,  // kPoolParts counters, one line each
               // per pool part, the children of the internal nodes whose ID reserves there
               // (a node's reservation is at most 2 RMAX^2 floats per child)
               CNT_NEED = (7 + kPoolParts) * kCntSlot,
               CNT_COUNT = (7 + 2 * kPoolParts) * kCntSlot };

struct RootInfo {
    double c[3];   // root box centre
    double H0;     // root half-side (on the proxy ladder)
    int k0;        // ladder index of H0
    int pad;
};

// Pointers to device arrays shared by the kernels.
struct H2Dev {
    int n, D, leaf;
    // sorted points
    uint64_t *keys;
    int *perm;
    void *Ys;          // Vec<D> sorted coordinates
    RootInfo *root;
    // nodes; level l is [lvl[l], lvl[l+1]) (written on the device as the tree grows)
    int *lvl;
    int node_cap;
    int *nb, *ne, *level, *parent, *child0, *nchild;
    float4 *cent;      // centre (xyz) of each node's box, w = half-side
    int *icoord;       // [node][3] integer box coordinates at the node's level
    int *bounds;       // split scratch [node][NCH+1]
    int *childcnt, *childoff;  // per-node child counts; per-chunk totals of a level
    int *near, *near_cnt;  // [node][NEAR_CAP]
    int *leaves;       // leaf node ids (CNT_LEAF of them)
    // pair buffers (target << 32 | source)
    uint64_t *adm, *xl, *wl, *nl, *pend[2];
    int adm_cap, x_cap, w_cap, nl_cap, pend_cap;
    uint64_t *sort_tmp;
    int *adm_beg, *adm_end, *x_beg, *x_end, *w_beg, *w_end, *nl_beg, *nl_end;
    // ID / product
    int *rank, *uoff;
    double *skel;      // [node][RMAX][D]
    float4 *skelf;     // [node][RMAX]
    float *upool;
    long long pool_cap;
    float4 *wc, *zc;   // [node][RMAX] upward / downward coefficients
    int *nflag;        // [node] the last mixed coupling redid this node in FP32
    const double *proxy_pts;  // [levels][PMAX][D]
    const int *proxy_cnt;
    int *flags, *cnt;
};

// Floats per part of the internal nodes' U pool (after the leaves' 2 n RMAX), even so that
// every region starts 8-byte aligned.
__host__ __device__ inline long long pool_part_cap(const H2Dev &h, int rmax)
{
    return ((h.pool_cap - 2LL * h.n * rmax) / kPoolParts) & ~1LL;
}

// h2_tree.hip
int launch_root_box(int D, const void *Yg, int n, const H2Dev &h, double *partial, int nblk,
                    hipStream_t st);
int launch_morton(int D, const void *Yg, const H2Dev &h, hipStream_t st);
int launch_gather_sorted(int D, const void *Yg, const H2Dev &h, hipStream_t st);
int launch_init_root(const H2Dev &h, hipStream_t st);
// Tree levels, two launches each, with no host round trip in between: the kernels take
// the level's node range from h.lvl on the device, and `bound` (an upper bound on the
// level's node count known to the host) only sizes the grid; the loops inside cover any
// count.  The level pass splits the level's nodes (child counts and bounds, per-chunk
// totals in childoff[0..nsplit)), builds their interaction lists and descends the pending
// pairs of pend_in; the level write scans the chunk totals, writes the children and
// publishes the level: lvl[l+2], the reset of the pending counter `pend_next` the next
// level pass appends to, and the pinned words out[0] (children of level l), out[1]
// (non-finite input), out[2] (node overflow), out[3] (near-list / pair-buffer overflow and
// oversized-leaf bits), out[4..8] (ADM, X, W, NEAR pair and leaf counts) and out[9]
// (lvl[l+2]).  kPinRec words per level.
constexpr int kPinRec = 16;
constexpr int kSplitBlocks = 256;  // split workgroups per level at most (childoff entries)
constexpr int kListBlocks = 1024;  // list workgroups per level at most
constexpr int kPendBlocks = 256;   // pending-pair workgroups per level
int launch_level_pass(int D, const H2Dev &h, int l, int bound, int pend_in, int pend_out,
                      int *nsplit, hipStream_t st);
int launch_level_write(int D, const H2Dev &h, int l, int nsplit, int pend_next, int *out,
                       hipStream_t st);
int launch_pair_offsets(const uint64_t *pairs, int n_pairs, int *beg, int *end, hipStream_t st);

// h2_id.hip
// pq: proxy slots to process (>= the largest proxy count of the table); f32: pivoted
// Gram-Schmidt in FP32 (the mixed path at ID tolerances >= 1e-3), else FP64; one_class:
// internal levels of at most this many nodes run in one launch of the widest size class
// (< 0: the default of the dimension).  With the tiled kernels the size class that runs a
// node does not change its result.
int launch_id_up(int D, const H2Dev &h, const int *list, int lb, int count, double tol,
                 int pq, bool f32, int one_class, hipStream_t st);

// h2_product.hip
// mixed: coupling on MFMA with BF16 operands and FP32 accumulation; a node whose
// rounding-error estimate exceeds tau times an entry's scale is flagged (nflag) and
// recomputed in FP32 by the same wavefront.
int launch_couple(int D, const H2Dev &h, int n_nodes, int mixed, float tau, hipStream_t st);
int launch_down(int D, const H2Dev &h, int lb, int le, int mixed, hipStream_t st);
int launch_leaf_eval(int D, const H2Dev &h, int n_leaves, float *s, float *F, hipStream_t st);

}  // namespace tsne
