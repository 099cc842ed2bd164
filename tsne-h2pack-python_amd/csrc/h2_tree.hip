// h2_tree.hip -- H2 rebuild, part 1: root box, Morton keys, adaptive 2^D-tree and
// the interaction lists, built level by level on the device.
//
// Reference counterpart: H2P_partition_points / the admissible-pair search inside
// H2P_build (H2Pack, called at src/tsne_h2pack.c:863 and :908-919) and the root box
// tsne_calc_root_enbox (src/tsne_h2pack.c:651-698).
#include <algorithm>
#include <cmath>

#include "h2_internal.h"
#include "proxy.h"

namespace tsne {

namespace {

template <int D> using VT = typename Vec<D>::T;

__device__ __forceinline__ float vcomp(const float2 &v, int d) { return d == 0 ? v.x : v.y; }
__device__ __forceinline__ float vcomp(const float4 &v, int d)
{
    return d == 0 ? v.x : (d == 1 ? v.y : v.z);
}

// Slots for `count` entries of this lane in a buffer shared through `counter`: one
// atomic per wavefront (all 64 lanes must call it), lanes get consecutive ranges.
__device__ __forceinline__ int wave_alloc(int *counter, int count)
{
    const int lane = threadIdx.x & 63;
    int x = count;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    const int total = __shfl(x, 63, 64);
    int base = 0;
    if (lane == 63 && total > 0) base = atomicAdd(counter, total);
    base = __shfl(base, 63, 64);
    return base + x - count;
}

__device__ __forceinline__ void put_pair(uint64_t *buf, int k, int cap, int t, int s, int *flags)
{
    if (k < cap)
        buf[k] = ((uint64_t)(uint32_t)t << 32) | (uint32_t)s;
    else
        atomicOr(&flags[FLAG_PAIR_OVF], 1);
}

// Separation of the boxes of nodes a (level la) and b (level lb <= la is not assumed),
// measured in box widths of the finer of the two: max over d of the number of whole
// fine boxes between them.  Exact integer arithmetic on the dyadic box coordinates.
template <int D>
__device__ __forceinline__ int box_gap(const int *ic, int a, int la, int b, int lb)
{
    int ga = 0;
    if (la >= lb) {
        const int s = 1 << (la - lb);
        for (int d = 0; d < D; d++) {
            const int x = ic[a * 3 + d];
            const int lo = ic[b * 3 + d] * s, hi = lo + s - 1;
            const int g = max(max(lo - x - 1, x - hi - 1), 0);
            ga = max(ga, g);
        }
    } else {
        const int s = 1 << (lb - la);
        for (int d = 0; d < D; d++) {
            const int x = ic[b * 3 + d];
            const int lo = ic[a * 3 + d] * s, hi = lo + s - 1;
            const int g = max(max(lo - x - 1, x - hi - 1), 0);
            ga = max(ga, g);
        }
    }
    return ga;
}

}  // namespace

// ---------------------------------------------------------------------------------
// Root box: bounding box of the embedding (and a finiteness test), then the root
// square with its half-side rounded up onto the proxy ladder.
// ---------------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256)
bbox_partial_kernel(const VT<D> *__restrict__ Y, int n, int per, double *__restrict__ part)
{
    __shared__ float red[2 * 3][256];
    __shared__ int bad_s[256];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    int bad = 0;
    const int b0 = blockIdx.x * per, b1 = min(n, b0 + per);
    for (int i = b0 + threadIdx.x; i < b1; i += 256) {
        const VT<D> y = Y[i];
        for (int d = 0; d < D; d++) {
            const float v = vcomp(y, d);
            if (!isfinite(v)) bad = 1;
            lo[d] = fminf(lo[d], v);
            hi[d] = fmaxf(hi[d], v);
        }
    }
    for (int d = 0; d < D; d++) { red[d][threadIdx.x] = lo[d]; red[3 + d][threadIdx.x] = hi[d]; }
    bad_s[threadIdx.x] = bad;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            for (int d = 0; d < D; d++) {
                red[d][threadIdx.x] = fminf(red[d][threadIdx.x], red[d][threadIdx.x + off]);
                red[3 + d][threadIdx.x] = fmaxf(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + off]);
            }
            bad_s[threadIdx.x] |= bad_s[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int d = 0; d < D; d++) {
            part[blockIdx.x * 7 + d] = red[d][0];
            part[blockIdx.x * 7 + 3 + d] = red[3 + d][0];
        }
        part[blockIdx.x * 7 + 6] = bad_s[0];
    }
}

template <int D>
__global__ void __launch_bounds__(256)
bbox_final_kernel(int nblk, const double *__restrict__ part, H2Dev h)
{
    __shared__ double slo[256][3], shi[256][3];
    __shared__ int sbad[256];
    const int t = threadIdx.x;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    int bad = 0;
    // the build's counters and flags start at zero (nothing of this build has touched
    // them yet; the non-finite flag below is set after the barrier)
    for (int k = t; k < CNT_COUNT; k += 256) h.cnt[k] = 0;
    for (int k = t; k < FLAG_COUNT; k += 256) h.flags[k] = 0;
    for (int b = t; b < nblk; b += 256) {
        for (int d = 0; d < D; d++) {
            lo[d] = fmin(lo[d], part[b * 7 + d]);
            hi[d] = fmax(hi[d], part[b * 7 + 3 + d]);
        }
        bad |= part[b * 7 + 6] != 0.0;
    }
    for (int d = 0; d < 3; d++) { slo[t][d] = lo[d]; shi[t][d] = hi[d]; }
    sbad[t] = bad;
    __syncthreads();
    if (t != 0) return;
    for (int k = 1; k < 256; k++) {
        for (int d = 0; d < D; d++) {
            lo[d] = fmin(lo[d], slo[k][d]);
            hi[d] = fmax(hi[d], shi[k][d]);
        }
        bad |= sbad[k];
    }
    RootInfo r;
    double ext = 0.0;
    for (int d = 0; d < 3; d++) r.c[d] = 0.0;
    for (int d = 0; d < D; d++) {
        r.c[d] = 0.5 * (lo[d] + hi[d]);
        ext = fmax(ext, hi[d] - lo[d]);
    }
    // half-side with a margin so the extreme points fall strictly inside
    const double H = 0.5 * ext * (1.0 + 1e-6) + 1e-30;
    int k = 0;
    while (k < kProxyLevels - 1 && kProxyBase * exp2((double)k) < H) k++;
    r.k0 = k;
    r.H0 = kProxyBase * exp2((double)k);
    r.pad = 0;
    *h.root = r;
    if (bad || !isfinite(ext)) h.flags[FLAG_NONFINITE] = 1;
}

int launch_root_box(int D, const void *Yg, int n, const H2Dev &h, double *partial, int nblk,
                    hipStream_t st)
{
    const int per = (n + nblk - 1) / nblk;
    if (D == 2) {
        hipLaunchKernelGGL(bbox_partial_kernel<2>, dim3(nblk), dim3(256), 0, st,
                           (const float2 *)Yg, n, per, partial);
        hipLaunchKernelGGL(bbox_final_kernel<2>, dim3(1), dim3(256), 0, st, nblk, partial, h);
    } else {
        hipLaunchKernelGGL(bbox_partial_kernel<3>, dim3(nblk), dim3(256), 0, st,
                           (const float4 *)Yg, n, per, partial);
        hipLaunchKernelGGL(bbox_final_kernel<3>, dim3(1), dim3(256), 0, st, nblk, partial, h);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------------
// Morton keys: coordinates quantised to 2^B cells per dimension of the root box, bits
// interleaved most significant first, so the level-l digit of a key is its l-th group
// of D bits from the top and points of one node are contiguous after sorting.
// ---------------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256)
morton_kernel(const VT<D> *__restrict__ Y, H2Dev h)
{
    constexpr int B = H2Cfg<D>::B;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= h.n) return;
    const RootInfo r = *h.root;
    const VT<D> y = Y[i];
    const double scale = (double)(1u << B) / (2.0 * r.H0);
    uint64_t key = 0;
    for (int d = 0; d < D; d++) {
        double q = floor(((double)vcomp(y, d) - (r.c[d] - r.H0)) * scale);
        q = fmin(fmax(q, 0.0), (double)((1u << B) - 1));
        const uint64_t qi = (uint64_t)q;
        for (int b = 0; b < B; b++) key |= ((qi >> b) & 1ull) << (D * b + d);
    }
    h.keys[i] = key;
    h.perm[i] = i;
}

int launch_morton(int D, const void *Yg, const H2Dev &h, hipStream_t st)
{
    const int g = (h.n + 255) / 256;
    if (D == 2)
        hipLaunchKernelGGL(morton_kernel<2>, dim3(g), dim3(256), 0, st, (const float2 *)Yg, h);
    else
        hipLaunchKernelGGL(morton_kernel<3>, dim3(g), dim3(256), 0, st, (const float4 *)Yg, h);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int D>
__global__ void __launch_bounds__(256)
gather_sorted_kernel(const VT<D> *__restrict__ Y, H2Dev h)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= h.n) return;
    ((VT<D> *)h.Ys)[k] = Y[h.perm[k]];
}

int launch_gather_sorted(int D, const void *Yg, const H2Dev &h, hipStream_t st)
{
    const int g = (h.n + 255) / 256;
    if (D == 2)
        hipLaunchKernelGGL(gather_sorted_kernel<2>, dim3(g), dim3(256), 0, st, (const float2 *)Yg, h);
    else
        hipLaunchKernelGGL(gather_sorted_kernel<3>, dim3(g), dim3(256), 0, st, (const float4 *)Yg, h);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Root node: all points, its own near list {0}; a root that is already a leaf
// interacts with itself directly.
__global__ void init_root_kernel(H2Dev h, int B)
{
    if (threadIdx.x != 0) return;
    const RootInfo r = *h.root;
    h.nb[0] = 0;
    h.ne[0] = h.n;
    h.level[0] = 0;
    h.parent[0] = -1;
    h.child0[0] = -1;
    h.nchild[0] = 0;
    h.cent[0] = make_float4((float)r.c[0], (float)r.c[1], (float)r.c[2], (float)r.H0);
    h.icoord[0] = h.icoord[1] = h.icoord[2] = 0;
    h.near[0] = 0;
    h.near_cnt[0] = 1;
    h.adm_beg[0] = h.adm_end[0] = h.x_beg[0] = h.x_end[0] = 0;
    h.w_beg[0] = h.w_end[0] = h.nl_beg[0] = h.nl_end[0] = 0;
    h.lvl[0] = 0;
    h.lvl[1] = 1;
    if (h.n <= h.leaf || B == 0) {
        h.nl[0] = 0;
        h.cnt[CNT_NEAR] = 1;
    }
}

int launch_init_root(const H2Dev &h, hipStream_t st)
{
    hipLaunchKernelGGL(init_root_kernel, dim3(1), dim3(64), 0, st, h,
                       h.D == 2 ? H2Cfg<2>::B : H2Cfg<3>::B);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------------
// Split of the nodes of level l: child ranges by binary search of the level-(l+1)
// digit inside the node's sorted key range, then (after a scan of the child counts)
// the children are written contiguously after the level.  Leaves are recorded.
// ---------------------------------------------------------------------------------
// NCH consecutive lanes per node: lane v >= 1 finds the first sorted point whose
// level-(l+1) digit is >= v (the NCH - 1 binary searches run side by side instead of one
// after another, which is what a node of the top levels, with up to n points, waits
// for); lane 0 then collects the bounds and records the node.  Slot t of the level
// (node t / NCH, lane t % NCH); every lane of the wavefront takes part (shuffles, a
// ballot).  Slots of [level size, nbound) get a zero child count (the library scan runs
// over nbound counts).
template <int D>
__device__ __forceinline__ int split_count_slot(const H2Dev &h, int l, int lb, int le, int t,
                                                int nbound)
{
    constexpr int B = H2Cfg<D>::B, NCH = H2Cfg<D>::NCH;
    static_assert((64 % NCH) == 0, "lanes per node");
    const int i = lb + t / NCH, v = t % NCH;
    const bool valid = i < le;
    // the scan runs over `nbound` counts: slots past the level's end count no children
    if (!valid && v == 0 && t / NCH < nbound) h.childcnt[t / NCH] = 0;
    int b = 0, e = 0;
    if (valid) {
        b = h.nb[i];
        e = h.ne[i];
    }
    const bool leaf = valid && (e - b <= h.leaf || l >= B);
    int bound = b;
    if (valid && !leaf && v > 0) {
        const int shift = D * (B - l - 1);
        int lo = b, hi = e;  // first index with digit >= v
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const int dg = (int)((h.keys[mid] >> shift) & (uint64_t)(NCH - 1));
            if (dg < v) lo = mid + 1; else hi = mid;
        }
        bound = lo;
    }
    const int g0 = (threadIdx.x & 63) & ~(NCH - 1);  // first lane of this node's group
    int bd[NCH + 1];
#pragma unroll
    for (int q = 0; q < NCH; q++) bd[q] = __shfl(bound, g0 + q, 64);
    bd[NCH] = e;
    // the wavefront's new leaves take consecutive slots of the leaf list (one atomic)
    const bool newleaf = valid && v == 0 && leaf;
    const unsigned long long ml = __ballot(newleaf);
    int lbase = 0;
    if ((threadIdx.x & 63) == 0 && ml) lbase = atomicAdd(&h.cnt[CNT_LEAF], __popcll(ml));
    lbase = __shfl(lbase, 0, 64);
    if (!valid || v != 0) return 0;
    if (leaf) {
        h.childcnt[i - lb] = 0;
        h.child0[i] = -1;
        h.nchild[i] = 0;
        h.leaves[lbase + __popcll(ml & ((1ull << (threadIdx.x & 63)) - 1ull))] = i;
        if (e - b > H2Cfg<D>::LEAF_CAP) atomicOr(&h.flags[FLAG_BIG_LEAF], 1);
        return 0;
    }
    int *bdp = h.bounds + (size_t)(i - lb) * (NCH + 1);
    int c = 0;
#pragma unroll
    for (int q = 0; q <= NCH; q++) bdp[q] = bd[q];
#pragma unroll
    for (int q = 0; q < NCH; q++) c += bd[q + 1] > bd[q];
    h.childcnt[i - lb] = c;
    return c;
}

// After the scan of level l's child counts: the range of level l + 1 (T children after
// the level's end le), the pending counter the next lists / pending pass appends to, and
// the host's copy of T and of the flags.
__device__ __forceinline__ void publish_level(const H2Dev &h, int l, int le, int T, int pend_out,
                                              int *out)
{
    int ne = le + T, ovf = h.flags[FLAG_NODE_OVF];
    if (ne > h.node_cap) {
        ovf = 1;
        ne = h.node_cap;
        h.flags[FLAG_NODE_OVF] = 1;
    }
    h.lvl[l + 2] = ne;
    h.cnt[CNT_PEND0 + pend_out * kCntSlot] = 0;
    out[0] = T;
    out[1] = h.flags[FLAG_NONFINITE];
    out[2] = ovf;
    // the build's state after this level: once no level follows, the finished tree's
    // flags, pair and leaf counts and level end, without a copy from device memory
    out[3] = (h.flags[FLAG_NEAR_OVF] ? 1 : 0) | (h.flags[FLAG_PAIR_OVF] ? 2 : 0) |
             (h.flags[FLAG_BIG_LEAF] ? 4 : 0);
    out[4] = h.cnt[CNT_ADM];
    out[5] = h.cnt[CNT_X];
    out[6] = h.cnt[CNT_W];
    out[7] = h.cnt[CNT_NEAR];
    out[8] = h.cnt[CNT_LEAF];
    out[9] = ne;
    // the most children over the U-pool parts, of the levels whose writes have finished
    // (all of them once no level follows: the last level holds leaves only)
    int need = 0;
    for (int q = 0; q < kPoolParts; q++) need = max(need, h.cnt[CNT_NEED + q * kCntSlot]);
    out[10] = need;
}

// The children of node i of level l (lb its first node), from child slot `base` on.
template <int D>
__device__ __forceinline__ void split_write_node(const H2Dev &h, int l, int lb, int i, int base)
{
    constexpr int NCH = H2Cfg<D>::NCH;
    const int *bd = h.bounds + (size_t)(i - lb) * (NCH + 1);
    const float4 c = h.cent[i];
    const float hc = 0.5f * c.w;
    int k = 0;
    for (int v = 0; v < NCH; v++) {
        if (bd[v + 1] <= bd[v]) continue;
        const int ch = base + k++;
        if (ch >= h.node_cap) continue;  // overflow: flagged by publish_level, rebuilt larger
        h.nb[ch] = bd[v];
        h.ne[ch] = bd[v + 1];
        h.level[ch] = l + 1;
        h.parent[ch] = i;
        h.child0[ch] = -1;
        h.nchild[ch] = 0;
        // empty W / NEAR ranges until the sorted lists' offsets are written
        h.w_beg[ch] = h.w_end[ch] = h.nl_beg[ch] = h.nl_end[ch] = 0;
        float4 cc;
        cc.x = c.x + ((v & 1) ? hc : -hc);
        cc.y = c.y + ((v & 2) ? hc : -hc);
        cc.z = (D == 3) ? c.z + ((v & 4) ? hc : -hc) : 0.f;
        cc.w = hc;
        h.cent[ch] = cc;
        for (int d = 0; d < 3; d++)
            h.icoord[ch * 3 + d] = (d < D) ? 2 * h.icoord[i * 3 + d] + ((v >> d) & 1) : 0;
    }
    h.child0[i] = base;
    h.nchild[i] = k;
}


// ---------------------------------------------------------------------------------
// Interaction lists of the nodes of level l (>= 1).  Candidates are the children of
// the parent's near-list members (a leaf member is its own candidate).  A same-level
// candidate at least one box away is admissible (ADM); a coarser leaf at least one
// box of i away is an X source; the rest stay in i's near list.  For a leaf i the near
// list is final: leaf members become NEAR pairs, others are descended later (pending).
// ---------------------------------------------------------------------------------
// One wavefront per node: lanes enumerate the candidates (children of the parent's
// near-list members, in member order), classify them, and the near list is compacted
// in candidate order with ballots.  The pair buffers are shared by every node of the
// level, so their slots come from global counters: the wavefront first classifies all
// its candidates (the per-round ballot masks are kept in LDS), the workgroup then takes
// the slots of all its pairs with one atomic per category, and the pairs are written.
// (Device-wide atomics on a shared counter are serialised across the GPU: one per node
// and category made this kernel's time grow with the node count.)
template <int D>
__device__ __forceinline__ void lists_block(const H2Dev &h, int l, int pend_out, int blk)
{
    const int lb = h.lvl[l], le = h.lvl[l + 1];
    constexpr int B = H2Cfg<D>::B, NEAR_CAP = H2Cfg<D>::NEAR_CAP, NCH = H2Cfg<D>::NCH;
    constexpr int RMAXR = (NEAR_CAP * NCH + 63) / 64;  // candidate rounds at most
    __shared__ int s_off[4][65], s_c0[4][64];
    __shared__ unsigned long long s_m[4][RMAXR][4];  // ADM, X, NEAR pair, pending masks
    __shared__ int s_cnt[4][4], s_bas[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = lb + blk * 4 + wv;
    const bool valid = i < le;  // uniform over the wavefront
    const unsigned long long lt = (1ull << lane) - 1ull;
    const int p = valid ? h.parent[i] : 0;
    const int np = valid ? h.near_cnt[p] : 0;
    // candidates per member: a leaf member is its own candidate, else its children
    int cc = 0, c0 = 0;
    if (lane < np) {
        const int j = h.near[(size_t)p * NEAR_CAP + lane];
        const bool jl = h.child0[j] < 0;
        c0 = jl ? j : h.child0[j];
        cc = jl ? 1 : h.nchild[j];
    }
    int x = cc;  // inclusive scan over lanes
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    const int T = __shfl(x, 63, 64);
    s_off[wv][lane + 1] = x;
    if (lane == 0) s_off[wv][0] = 0;
    s_c0[wv][lane] = c0;
    __syncthreads();  // reached by every wavefront of the block
    const bool leaf = valid && ((h.ne[i] - h.nb[i]) <= h.leaf || l >= B);
    auto candidate = [&](int k) {  // member a with off[a] <= k < off[a+1]
        int lo = 0, hi = np;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_off[wv][mid] <= k) lo = mid; else hi = mid;
        }
        return s_c0[wv][lo] + (k - s_off[wv][lo]);
    };
    // pass 1: classify (1 ADM, 2 X, 3 near; a leaf's near members become NEAR pairs or
    // pending pairs), compact the near list, count the pairs of each category
    int *my = h.near + (size_t)(valid ? i : 0) * NEAR_CAP;
    int cnt = 0, nA = 0, nX = 0, nN = 0, nP = 0;
    int rounds = valid ? (T + 63) / 64 : 0;
    if (rounds > RMAXR) {  // more candidates than NEAR_CAP members can have: overflow
        if (lane == 0) atomicOr(&h.flags[FLAG_NEAR_OVF], 1);
        rounds = 0;
    }
    for (int r = 0; r < rounds; r++) {
        const int k = r * 64 + lane;
        int cand = -1, cat = 0, sub = 0;
        if (k < T) {
            cand = candidate(k);
            const int lc = h.level[cand];
            if (box_gap<D>(h.icoord, i, l, cand, lc) >= 1) cat = (lc == l) ? 1 : 2;
            else cat = 3;
            if (cat == 3 && leaf)
                sub = ((h.ne[cand] - h.nb[cand]) <= h.leaf || h.level[cand] >= B) ? 1 : 2;
        }
        const unsigned long long mA = __ballot(cat == 1), mX = __ballot(cat == 2),
                                 mN = __ballot(cat == 3), mNL = __ballot(sub == 1),
                                 mP = __ballot(sub == 2);
        if (cat == 3) {
            const int pos = cnt + __popcll(mN & lt);
            if (pos < NEAR_CAP) my[pos] = cand;
            else atomicOr(&h.flags[FLAG_NEAR_OVF], 1);
        }
        cnt += __popcll(mN);
        if (r < RMAXR && lane == 0) {
            s_m[wv][r][0] = mA;
            s_m[wv][r][1] = mX;
            s_m[wv][r][2] = mNL;
            s_m[wv][r][3] = mP;
        }
        nA += __popcll(mA);
        nX += __popcll(mX);
        nN += __popcll(mNL);
        nP += __popcll(mP);
    }
    if (valid && lane == 0) h.near_cnt[i] = min(cnt, NEAR_CAP);
    // one atomic per category for the whole workgroup
    if (lane == 0) {
        s_cnt[wv][0] = nA;
        s_cnt[wv][1] = nX;
        s_cnt[wv][2] = nN;
        s_cnt[wv][3] = nP;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int c = threadIdx.x;
        const int tot = s_cnt[0][c] + s_cnt[1][c] + s_cnt[2][c] + s_cnt[3][c];
        int *ctr = (c == 0) ? &h.cnt[CNT_ADM] : (c == 1) ? &h.cnt[CNT_X]
                 : (c == 2) ? &h.cnt[CNT_NEAR] : &h.cnt[CNT_PEND0 + pend_out * kCntSlot];
        s_bas[c] = tot ? atomicAdd(ctr, tot) : 0;
    }
    __syncthreads();
    if (!valid) return;
    int bA = s_bas[0], bX = s_bas[1], bN = s_bas[2], bP = s_bas[3];
    for (int w = 0; w < wv; w++) {
        bA += s_cnt[w][0];
        bX += s_cnt[w][1];
        bN += s_cnt[w][2];
        bP += s_cnt[w][3];
    }
    // this node's ADM and X pairs are contiguous, in candidate order: their ranges are
    // final (no sort; only the NEAR and W lists, which later levels append to, are sorted)
    if (lane == 0) {
        h.adm_beg[i] = bA;
        h.adm_end[i] = bA + nA;
        h.x_beg[i] = bX;
        h.x_end[i] = bX + nX;
    }
    // pass 2: write the pairs in candidate order
    for (int r = 0; r < rounds; r++) {
        const int k = r * 64 + lane;
        const unsigned long long mA = s_m[wv][r][0], mX = s_m[wv][r][1], mNL = s_m[wv][r][2],
                                 mP = s_m[wv][r][3];
        const unsigned long long me = 1ull << lane;
        if (k < T && ((mA | mX | mNL | mP) & me)) {
            const int cand = candidate(k);
            if (mA & me) put_pair(h.adm, bA + __popcll(mA & lt), h.adm_cap, i, cand, h.flags);
            if (mX & me) put_pair(h.xl, bX + __popcll(mX & lt), h.x_cap, i, cand, h.flags);
            if (mNL & me) put_pair(h.nl, bN + __popcll(mNL & lt), h.nl_cap, i, cand, h.flags);
            if (mP & me)
                put_pair(h.pend[pend_out], bP + __popcll(mP & lt), h.pend_cap, i, cand, h.flags);
        }
        bA += __popcll(mA);
        bX += __popcll(mX);
        bN += __popcll(mNL);
        bP += __popcll(mP);
    }
}

// Pending pairs (leaf i, non-leaf c): each child d of c is a W source when i lies at
// least one box of d away, a NEAR source when d is a leaf, otherwise pending again.
// Pair k of the pending buffer pend_in (every lane of the wavefront takes part).
template <int D>
__device__ __forceinline__ void pending_pair(const H2Dev &h, int pend_in, int pend_out, int k,
                                             int n_pend)
{
    constexpr int B = H2Cfg<D>::B;
    const bool valid = k < n_pend;
    int i = 0, c = 0, li = 0, nc = 0, c0 = 0;
    if (valid) {
        const uint64_t pr = h.pend[pend_in][k];
        i = (int)(pr >> 32);
        c = (int)(pr & 0xffffffffu);
        li = h.level[i];
        nc = h.nchild[c];
        c0 = h.child0[c];
    }
    int nw = 0, nn = 0, np = 0;
    for (int q = 0; q < nc; q++) {
        const int d = c0 + q;
        const int ld = h.level[d];
        if (box_gap<D>(h.icoord, i, li, d, ld) >= 1) nw++;
        else if ((h.ne[d] - h.nb[d]) <= h.leaf || ld >= B) nn++;
        else np++;
    }
    int kw = wave_alloc(&h.cnt[CNT_W], nw);
    int kn = wave_alloc(&h.cnt[CNT_NEAR], nn);
    int kp = wave_alloc(&h.cnt[CNT_PEND0 + pend_out * kCntSlot], np);
    for (int q = 0; q < nc; q++) {
        const int d = c0 + q;
        const int ld = h.level[d];
        if (box_gap<D>(h.icoord, i, li, d, ld) >= 1)
            put_pair(h.wl, kw++, h.w_cap, i, d, h.flags);
        else if ((h.ne[d] - h.nb[d]) <= h.leaf || ld >= B)
            put_pair(h.nl, kn++, h.nl_cap, i, d, h.flags);
        else
            put_pair(h.pend[pend_out], kp++, h.pend_cap, i, d, h.flags);
    }
}

// Workgroup-wide sum over 256 threads (every thread gets it); s_red: 4 ints of LDS.
__device__ __forceinline__ int block_sum256(int x, int *s_red)
{
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    __syncthreads();  // s_red free (a previous call may still be reading it)
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = x;
    __syncthreads();
    return s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

// The nodes of level l split into `nsplit` contiguous chunks, one per split workgroup of
// the level pass and of the level write: chunk b is [b C, min(cnt, (b + 1) C)).
__device__ __forceinline__ void split_chunk(int cnt, int nsplit, int b, int &c0, int &c1)
{
    const int C = (cnt + nsplit - 1) / nsplit;
    c0 = min(cnt, b * C);
    c1 = min(cnt, c0 + C);
}

// Level pass of level l, one launch.  Workgroups [0, nsplit): the child counts and child
// bounds of a chunk of the level's nodes, and the chunk's total (childoff[b]).
// Workgroups [nsplit, nsplit + nlists): the interaction lists of the level's nodes, four
// per workgroup round (l >= 1; the root's lists come from init_root).  The last
// kPendBlocks workgroups: the pending pairs of pend_in.  The three parts read nothing
// the others write: the lists read the children of the parents' near-list members and
// the pending pairs the children of coarser nodes, all made by earlier levels' writes.
template <int D>
__global__ void __launch_bounds__(256)
level_pass_kernel(H2Dev h, int l, int nsplit, int nlists, int pend_in, int pend_out)
{
    constexpr int NCH = H2Cfg<D>::NCH;
    __shared__ int s_red[4];
    if (h.flags[FLAG_NODE_OVF]) return;  // uniform over the grid; the build is redone
    const int lb = h.lvl[l], le = h.lvl[l + 1];
    const int cnt = le - lb;
    const int b = blockIdx.x;
    if (b < nsplit) {
        int c0, c1;
        split_chunk(cnt, nsplit, b, c0, c1);
        int mine = 0;
        // slots NCH per node, 256 per round; the trip count is uniform over the workgroup
        for (int t = c0 * NCH; t < c1 * NCH; t += 256)
            mine += split_count_slot<D>(h, l, lb, lb + c1, t + (int)threadIdx.x, 0);
        const int tot = block_sum256(mine, s_red);
        if (threadIdx.x == 0) h.childoff[b] = tot;
    } else if (b < nsplit + nlists) {
        if (l == 0) return;
        for (int blk = b - nsplit; blk * 4 < cnt; blk += nlists) {
            lists_block<D>(h, l, pend_out, blk);
            __syncthreads();  // the next round reuses the workgroup's LDS
        }
    } else {
        const int n_pend = min(h.cnt[CNT_PEND0 + pend_in * kCntSlot], h.pend_cap);
        for (int base = (b - nsplit - nlists) * 256; base < n_pend; base += kPendBlocks * 256)
            pending_pair<D>(h, pend_in, pend_out, base + threadIdx.x, n_pend);
    }
}

// Level write of level l: workgroup b takes its chunk's first child slot from the chunk
// totals before it (childoff[0..b)), scans its nodes' child counts and writes their
// children; workgroup 0 publishes the level (level l + 1's end, the reset of the pending
// counter the next level pass appends to, and the pinned words out[0..2]).
template <int D>
__global__ void __launch_bounds__(256)
level_write_kernel(H2Dev h, int l, int nsplit, int pend_next, int *out)
{
    __shared__ int s_red[4], s_wsum[4];
    const int lb = h.lvl[l], le = h.lvl[l + 1];
    const int cnt = le - lb;
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, w = t >> 6;
    const bool ovf = h.flags[FLAG_NODE_OVF];
    // chunk totals: the prefix of this chunk and the level's total
    int pre = 0, tot = 0;
    for (int k = t; k < nsplit; k += 256) {
        const int x = ovf ? 0 : h.childoff[k];
        tot += x;
        if (k < b) pre += x;
    }
    pre = block_sum256(pre, s_red);
    tot = block_sum256(tot, s_red);
    if (b == 0 && t == 0) publish_level(h, l, le, ovf ? 0 : tot, pend_next, out);
    if (ovf) return;
    int c0, c1;
    split_chunk(cnt, nsplit, b, c0, c1);
    int run = le + pre;  // first child slot of the round
    // children of this thread's nodes: the ID of node k of the level reserves its U in
    // pool part k mod kPoolParts (one workgroup per node, blockIdx = k), the same part
    // for every round since 256 is a multiple of kPoolParts
    int kids = 0;
    for (int r0 = c0; r0 < c1; r0 += 256) {  // uniform trip count
        const int k = r0 + t;
        const int c = (k < c1) ? h.childcnt[k] : 0;
        kids += c;
        int x = c;  // inclusive scan over the wavefront, then over the four wavefronts
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        __syncthreads();
        if (lane == 63) s_wsum[w] = x;
        __syncthreads();
        int wp = 0;
        for (int q = 0; q < w; q++) wp += s_wsum[q];
        if (c > 0) split_write_node<D>(h, l, lb, lb + k, run + wp + x - c);
        run += s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
    }
    // per-part sums: lanes with the same lane mod 8 share a part (lanes ^8, ^16, ^32),
    // then the four wavefronts
    static_assert(kPoolParts == 8, "parts by lane mod 8");
    for (int o = 8; o < 64; o <<= 1) kids += __shfl_xor(kids, o, 64);
    __shared__ int s_kids[4][8];
    __syncthreads();
    if (lane < 8) s_kids[w][lane] = kids;
    __syncthreads();
    if (t < 8) {
        const int v = s_kids[0][t] + s_kids[1][t] + s_kids[2][t] + s_kids[3][t];
        if (v) atomicAdd(&h.cnt[CNT_NEED + ((c0 + t) % kPoolParts) * kCntSlot], v);
    }
}

int launch_level_pass(int D, const H2Dev &h, int l, int bound, int pend_in, int pend_out,
                      int *nsplit_out, hipStream_t st)
{
    const int NCH = (D == 2) ? H2Cfg<2>::NCH : H2Cfg<3>::NCH;
    const int nsplit = std::max(1, std::min(kSplitBlocks, (int)(((long long)bound * NCH + 255) / 256)));
    const int nlists = (l == 0) ? 0 : std::max(1, std::min(kListBlocks, (bound + 3) / 4));
    const int g = nsplit + nlists + ((l == 0) ? 0 : kPendBlocks);
    if (D == 2)
        hipLaunchKernelGGL(level_pass_kernel<2>, dim3(g), dim3(256), 0, st, h, l, nsplit, nlists,
                           pend_in, pend_out);
    else
        hipLaunchKernelGGL(level_pass_kernel<3>, dim3(g), dim3(256), 0, st, h, l, nsplit, nlists,
                           pend_in, pend_out);
    *nsplit_out = nsplit;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_level_write(int D, const H2Dev &h, int l, int nsplit, int pend_next, int *out,
                       hipStream_t st)
{
    if (D == 2)
        hipLaunchKernelGGL(level_write_kernel<2>, dim3(nsplit), dim3(256), 0, st, h, l, nsplit,
                           pend_next, out);
    else
        hipLaunchKernelGGL(level_write_kernel<3>, dim3(nsplit), dim3(256), 0, st, h, l, nsplit,
                           pend_next, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// CSR offsets of (target << 32 | source) pairs sorted by key: beg/end per target.
__global__ void __launch_bounds__(256)
pair_offsets_kernel(const uint64_t *__restrict__ pairs, int n, int *__restrict__ beg,
                    int *__restrict__ end)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int t = (int)(pairs[k] >> 32);
    if (k == 0 || (int)(pairs[k - 1] >> 32) != t) beg[t] = k;
    if (k == n - 1 || (int)(pairs[k + 1] >> 32) != t) end[t] = k + 1;
}

int launch_pair_offsets(const uint64_t *pairs, int n_pairs, int *beg, int *end, hipStream_t st)
{
    if (n_pairs <= 0) return 0;
    hipLaunchKernelGGL(pair_offsets_kernel, dim3((n_pairs + 255) / 256), dim3(256), 0, st, pairs,
                       n_pairs, beg, end);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace tsne
