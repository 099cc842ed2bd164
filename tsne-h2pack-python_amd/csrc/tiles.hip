// tiles.hip -- row tiles of the joint P for the attractive pass.
//
// Once the points are numbered along the tree order, the rows of a tile (consecutive
// points, a small region of the embedding) share most of their neighbours: a tile of
// 256 rows with ~150 nonzeros each references a few thousand distinct points.  The
// attractive kernel stages those points' gather records in LDS once per tile and reads
// every neighbour from LDS through a 16-bit index, instead of issuing one scattered HBM /
// L2 read per nonzero; the index stream is also half the size of the int32 columns.
//
// Build (after each renumbering; the P pattern is fixed in between):
//   1. each tile's columns, sorted (segmented radix sort, one segment per tile),
//   2. distinct counts per tile, their exclusive scan -> off,
//   3. the distinct columns compacted in order -> cols,
//   4. for each nonzero, the position of its column in its tile's list -> lidx.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "common.h"
#include "tiles.h"

namespace tsne {

namespace {

constexpr int kLidxStage = 16384;  // tile lists up to this size are searched in LDS

__global__ void __launch_bounds__(256)
tile_seg_kernel(int n, int ntiles, int rows, const int *__restrict__ rp, int *__restrict__ seg)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b <= ntiles) seg[b] = rp[min(n, b * rows)];
}

// Distinct values of each sorted segment: cnt[b].
__global__ void __launch_bounds__(256)
tile_count_kernel(const int *__restrict__ seg, const int *__restrict__ sorted, int *__restrict__ cnt)
{
    __shared__ int s_w[4];
    const int b = blockIdx.x;
    const int s0 = seg[b], s1 = seg[b + 1];
    int c = 0;
    for (int e = s0 + threadIdx.x; e < s1; e += 256) c += (e == s0 || sorted[e] != sorted[e - 1]);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) cnt[b] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Distinct values of each sorted segment, compacted in order into cols[off[b] ...].
__global__ void __launch_bounds__(256)
tile_write_kernel(const int *__restrict__ seg, const int *__restrict__ sorted,
                  const int *__restrict__ off, int *__restrict__ cols)
{
    __shared__ int s_w[4];
    const int b = blockIdx.x;
    const int s0 = seg[b], s1 = seg[b + 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    int base = off[b];
    for (int c0 = s0; c0 < s1; c0 += 256) {
        const int e = c0 + threadIdx.x;
        const bool f = e < s1 && (e == s0 || sorted[e] != sorted[e - 1]);
        const unsigned long long m = __ballot(f);
        if (lane == 0) s_w[wv] = __popcll(m);
        __syncthreads();
        int before = 0, total = 0;
        for (int w = 0; w < 4; w++) {
            before += (w < wv) ? s_w[w] : 0;
            total += s_w[w];
        }
        if (f) cols[base + before + __popcll(m & lt)] = sorted[e];
        base += total;
        __syncthreads();  // s_w is rewritten by the next chunk
    }
}

// lidx[e] = position of col[e] in its tile's distinct list (tiles up to ucap only).
__global__ void __launch_bounds__(256)
tile_lidx_kernel(const int *__restrict__ seg, const int *__restrict__ col,
                 const int *__restrict__ off, const int *__restrict__ cols, int ucap,
                 uint16_t *__restrict__ lidx)
{
    __shared__ int s_c[kLidxStage];
    const int b = blockIdx.x;
    const int u0 = off[b], U = off[b + 1] - u0;
    if (U > ucap) return;  // this tile gathers from HBM
    const bool staged = U <= kLidxStage;
    if (staged)
        for (int u = threadIdx.x; u < U; u += 256) s_c[u] = cols[u0 + u];
    __syncthreads();
    const int s0 = seg[b], s1 = seg[b + 1];
    for (int e = s0 + threadIdx.x; e < s1; e += 256) {
        const int v = col[e];
        int lo = 0, hi = U;  // first position with value >= v (present by construction)
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const int cv = staged ? s_c[mid] : cols[u0 + mid];
            if (cv < v) lo = mid + 1; else hi = mid;
        }
        lidx[e] = (uint16_t)lo;
    }
}

template <typename T> void sfree(T *&p, hipStream_t st)
{
    if (p) (void)hipFreeAsync((void *)p, st);
    p = nullptr;
}

}  // namespace

void tiles_free(AttrTiles &t, hipStream_t st)
{
    sfree(t.off, st);
    sfree(t.cols, st);
    sfree(t.lidx, st);
    t.ntiles = 0;
    t.ucount = 0;
    t.umax = 0;
    t.n_global = 0;
}

int tiles_build(AttrTiles &t, int n, int nnz, const int *rp, const int *col, int rows, int D,
                int ucap, hipStream_t st)
{
    (void)D;
    tiles_free(t, st);
    if (n <= 0 || rows <= 0) return -1;
    const int ntiles = (n + rows - 1) / rows;
    ucap = std::min(ucap, 65536);
    int *seg = nullptr, *sorted = nullptr, *cnt = nullptr;
    void *tmp = nullptr;
    int rc = -1;
    size_t b1 = 0, b2 = 0;
    (void)hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, b1, (const int *)nullptr,
                                                     (int *)nullptr, std::max(nnz, 1), ntiles,
                                                     (const int *)nullptr, (const int *)nullptr,
                                                     0, 32, st);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b2, (const int *)nullptr, (int *)nullptr,
                                           ntiles + 1, st);
    const size_t tbytes = std::max(b1, b2);
    int bits = 1;
    while (bits < 31 && (1LL << bits) < (long long)n) bits++;
    std::vector<int> hc(ntiles + 1);
    t.rows = rows;
    t.ntiles = ntiles;
    if (hipMallocAsync((void **)&seg, sizeof(int) * (ntiles + 1), st) != hipSuccess ||
        hipMallocAsync((void **)&sorted, sizeof(int) * std::max(nnz, 1), st) != hipSuccess ||
        hipMallocAsync((void **)&cnt, sizeof(int) * (ntiles + 1), st) != hipSuccess ||
        hipMallocAsync(&tmp, std::max<size_t>(tbytes, 1), st) != hipSuccess ||
        hipMallocAsync((void **)&t.off, sizeof(int) * (ntiles + 1), st) != hipSuccess ||
        hipMallocAsync((void **)&t.lidx, sizeof(uint16_t) * ((size_t)nnz + 8), st) != hipSuccess)
        goto done;
    hipLaunchKernelGGL(tile_seg_kernel, dim3((ntiles + 1 + 255) / 256), dim3(256), 0, st, n, ntiles,
                       rows, rp, seg);
    if (nnz > 0) {
        size_t bytes = tbytes;
        if (hipcub::DeviceSegmentedRadixSort::SortKeys(tmp, bytes, col, sorted, nnz, ntiles, seg,
                                                       seg + 1, 0, bits, st) != hipSuccess)
            goto done;
    }
    if (hipMemsetAsync(cnt + ntiles, 0, sizeof(int), st) != hipSuccess) goto done;
    hipLaunchKernelGGL(tile_count_kernel, dim3(ntiles), dim3(256), 0, st, seg, sorted, cnt);
    {
        size_t bytes = tbytes;
        if (hipcub::DeviceScan::ExclusiveSum(tmp, bytes, cnt, t.off, ntiles + 1, st) != hipSuccess)
            goto done;
    }
    if (hipMemcpyAsync(hc.data(), cnt, sizeof(int) * (ntiles + 1), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        goto done;
    t.ucount = 0;
    t.umax = 0;
    t.n_global = 0;
    for (int b = 0; b < ntiles; b++) {
        t.ucount += hc[b];
        if (hc[b] > ucap) t.n_global++;
        else t.umax = std::max(t.umax, hc[b]);
    }
    if (hipMallocAsync((void **)&t.cols, sizeof(int) * std::max<long long>(t.ucount, 1), st) !=
        hipSuccess)
        goto done;
    hipLaunchKernelGGL(tile_write_kernel, dim3(ntiles), dim3(256), 0, st, seg, sorted, t.off, t.cols);
    hipLaunchKernelGGL(tile_lidx_kernel, dim3(ntiles), dim3(256), 0, st, seg, col, t.off, t.cols,
                       ucap, t.lidx);
    if (hipMemsetAsync(t.lidx + nnz, 0, sizeof(uint16_t) * 8, st) != hipSuccess) goto done;
    rc = hipGetLastError() == hipSuccess ? 0 : -1;
done:
    sfree(seg, st);
    sfree(sorted, st);
    sfree(cnt, st);
    if (tmp) (void)hipFreeAsync(tmp, st);
    if (rc != 0) tiles_free(t, st);
    return rc;
}

}  // namespace tsne
