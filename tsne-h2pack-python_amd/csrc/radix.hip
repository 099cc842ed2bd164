// radix.hip -- stable LSD radix sort of 64-bit keys (optionally with int32 values) for
// the H2 rebuild's Morton keys and pair lists, without device-wide atomics.
//
// Each pass sorts by one R-bit digit in three launches:
//   count    every workgroup histograms the digit over its tile of kTile items in LDS
//            (runs of equal digits in a wavefront add once) and stores its counts with
//            plain writes, bin-major: counts[bin][tile];
//   scan     one wavefront per bin turns counts[bin][*] into the exclusive offsets of the
//            tiles inside the bin and writes the bin's total;
//   scatter  every workgroup ranks its tile stably -- one wavefront per 1024 consecutive
//            items, 64 items per step, the lanes of equal digit found with R ballots and
//            counted with running per-wavefront counters in LDS -- and writes each item
//            to  exclusive_scan(totals)[bin] + offset[bin][tile] + rank.
// Stable LSD passes give the one stable order of the keys, so the result is the same as
// that of any stable radix sort over the same bits.  No pass uses global atomics or a
// cross-workgroup wait: the library onesweep sort's histogram pass adds every workgroup's
// counts into one global histogram with device-scope atomics, which across the eight XCDs
// cost ~0.6-1.2 ms per 10^6-key sort.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>

#include "radix.h"

namespace tsne {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kItems = 8;                  // per lane
constexpr int kTile = kThreads * kItems;   // items per workgroup
constexpr int kWaveItems = kTile / kWaves; // consecutive items per wavefront

template <int R>
__global__ void __launch_bounds__(kThreads)
radix_count_kernel(const uint64_t *__restrict__ keys, int n, int shift, uint32_t mask,
                   uint32_t *__restrict__ counts, int ntiles)
{
    constexpr int BINS = 1 << R;
    __shared__ uint32_t hist[BINS];
    for (int i = threadIdx.x; i < BINS; i += kThreads) hist[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long base = (long long)blockIdx.x * kTile + (long long)w * kWaveItems;
    // all of the lane's keys are loaded before any is counted (one memory round trip)
    uint32_t dg[kItems];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        const long long idx = base + j * 64 + lane;
        dg[j] = idx < n ? (uint32_t)(keys[idx] >> shift) & mask : 0xffffffffu;
    }
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        const long long idx = base + j * 64 + lane;
        const bool ok = idx < n;
        const uint32_t d = dg[j];
        const uint32_t prev = __shfl_up(d, 1, 64);
        const bool head = ok && (lane == 0 || prev != d);
        const unsigned long long hm = __ballot(head);
        const int nvalid = __popcll(__ballot(ok));  // the valid lanes are a prefix
        if (head) {
            const unsigned long long above = (lane == 63) ? 0ull : (hm >> (lane + 1));
            const int len = above ? __builtin_ctzll(above) + 1 : nvalid - lane;
            atomicAdd(&hist[d], (uint32_t)len);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < BINS; b += kThreads)
        counts[(size_t)b * ntiles + blockIdx.x] = hist[b];
}

// One wavefront per bin: counts[bin][*] -> exclusive offsets inside the bin; the total.
// Lane l owns the contiguous chunk [l*C, (l+1)*C) of the row (C = ceil(ntiles / 64), at
// most kScanChunk; longer rows are walked in super-chunks): its entries are loaded at
// once, summed, scanned across the wavefront, and written back as offsets.
constexpr int kScanChunk = 16;
__global__ void __launch_bounds__(kThreads)
radix_scan_kernel(uint32_t *__restrict__ counts, int ntiles, int bins,
                  uint32_t *__restrict__ totals)
{
    const int lane = threadIdx.x & 63;
    const int bin = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (bin >= bins) return;
    uint32_t *row = counts + (size_t)bin * ntiles;
    uint32_t run = 0;
    for (int s0 = 0; s0 < ntiles; s0 += 64 * kScanChunk) {
        const int len = min(ntiles - s0, 64 * kScanChunk);
        const int C = (len + 63) / 64;
        const int b0 = s0 + lane * C;
        uint32_t v[kScanChunk];
        uint32_t sum = 0;
#pragma unroll
        for (int q = 0; q < kScanChunk; q++) {
            const int b = b0 + q;
            v[q] = (q < C && b < s0 + len) ? row[b] : 0u;
            sum += v[q];
        }
        uint32_t x = sum;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        uint32_t acc = run + x - sum;
#pragma unroll
        for (int q = 0; q < kScanChunk; q++) {
            const int b = b0 + q;
            if (q < C && b < s0 + len) {
                row[b] = acc;
                acc += v[q];
            }
        }
        run += __shfl(x, 63, 64);
    }
    if (lane == 0) totals[bin] = run;
}

template <int R, bool VALS>
__global__ void __launch_bounds__(kThreads)
radix_scatter_kernel(const uint64_t *__restrict__ kin, const int *__restrict__ vin,
                     uint64_t *__restrict__ kout, int *__restrict__ vout, int n, int shift,
                     uint32_t mask, const uint32_t *__restrict__ offs,
                     const uint32_t *__restrict__ totals, int ntiles)
{
    constexpr int BINS = 1 << R;
    constexpr int PER = (BINS + kThreads - 1) / kThreads;  // bins per thread in the scans
    __shared__ uint32_t wcnt[kWaves][BINS];  // running counts, then in-tile positions
    __shared__ uint32_t s_gdelta[BINS];      // global position - in-tile position, per bin
    __shared__ uint32_t s_wsum[2][kWaves];
    __shared__ uint64_t s_key[kTile];
    __shared__ int s_val[VALS ? kTile : 1];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int i = t; i < kWaves * BINS; i += kThreads) (&wcnt[0][0])[i] = 0;
    // exclusive scan of the bin totals: thread t owns bins [t*PER, (t+1)*PER); its own
    // sum, scanned over the wavefront, then over the wavefront sums
    uint32_t tot[PER];
    uint32_t mine = 0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = t * PER + q;
        tot[q] = b < BINS ? totals[b] : 0u;
        mine += tot[q];
    }
    uint32_t xs = mine;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(xs, off, 64);
        if (lane >= off) xs += y;
    }
    if (lane == 63) s_wsum[0][w] = xs;
    // all of the lane's items are loaded (one memory round trip), then ranked through LDS
    const long long tile0 = (long long)blockIdx.x * kTile;
    const long long base = tile0 + (long long)w * kWaveItems;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint64_t key[kItems];
    int val[kItems];
    uint32_t rank[kItems];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        const long long idx = base + j * 64 + lane;
        const bool ok = idx < n;
        key[j] = ok ? kin[idx] : 0ull;
        if constexpr (VALS) val[j] = ok ? vin[idx] : 0;
    }
    __syncthreads();  // counters zeroed, wavefront sums written
    uint32_t gexcl = xs - mine;
    for (int k = 0; k < w; k++) gexcl += s_wsum[0][k];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        const long long idx = base + j * 64 + lane;
        const bool ok = idx < n;
        const uint32_t d = (uint32_t)(key[j] >> shift) & mask;
        unsigned long long peers = __ballot(ok);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const unsigned long long b = __ballot((d >> r) & 1u);
            peers &= ((d >> r) & 1u) ? b : ~b;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lt);
        // the lanes of one wavefront read the counter before its lowest lane of the
        // digit writes it back (LDS operations of a wavefront run in program order)
        const uint32_t c = ok ? wcnt[w][d] : 0u;
        rank[j] = c + below;
        if (ok && below == 0) wcnt[w][d] = c + (uint32_t)__popcll(peers);
    }
    __syncthreads();  // every wavefront's counts final
    // in-tile positions: bins in order, wavefronts in order inside a bin; the global
    // position of in-tile position p of bin b is p + gdelta[b]
    uint32_t cnt[PER], lsum = 0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = t * PER + q;
        cnt[q] = 0;
        if (b < BINS)
#pragma unroll
            for (int k = 0; k < kWaves; k++) cnt[q] += wcnt[k][b];
        lsum += cnt[q];
    }
    uint32_t xl = lsum;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(xl, off, 64);
        if (lane >= off) xl += y;
    }
    if (lane == 63) s_wsum[1][w] = xl;
    __syncthreads();
    uint32_t lexcl = xl - lsum;
    for (int k = 0; k < w; k++) lexcl += s_wsum[1][k];
    {
        uint32_t grun = gexcl, lrun = lexcl;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int b = t * PER + q;
            if (b < BINS) {
                s_gdelta[b] = grun + offs[(size_t)b * ntiles + blockIdx.x] - lrun;
                uint32_t p = lrun;
#pragma unroll
                for (int k = 0; k < kWaves; k++) {
                    const uint32_t a = wcnt[k][b];
                    wcnt[k][b] = p;
                    p += a;
                }
            }
            grun += tot[q];
            lrun += cnt[q];
        }
    }
    __syncthreads();
    // the items in in-tile order in LDS, then written out: consecutive threads write
    // consecutive positions of a bin's run
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        const long long idx = base + j * 64 + lane;
        if (idx < n) {
            const uint32_t d = (uint32_t)(key[j] >> shift) & mask;
            const uint32_t p = wcnt[w][d] + rank[j];
            s_key[p] = key[j];
            if constexpr (VALS) s_val[p] = val[j];
        }
    }
    __syncthreads();
    const int tn = (int)min((long long)kTile, (long long)n - tile0);
    for (int p = t; p < tn; p += kThreads) {
        const uint64_t k = s_key[p];
        const uint32_t pos = s_gdelta[(uint32_t)(k >> shift) & mask] + (uint32_t)p;
        kout[pos] = k;
        if constexpr (VALS) vout[pos] = s_val[p];
    }
}

template <int R>
int sort_passes(uint64_t *&keys, uint64_t *&keys_alt, int *vals_buf[2], bool vals, int n,
                int begin_bit, int end_bit, uint32_t *counts, uint32_t *totals, hipStream_t st)
{
    constexpr int BINS = 1 << R;
    const int ntiles = (n + kTile - 1) / kTile;
    int cur = 0;
    uint64_t *kb[2] = {keys, keys_alt};
    for (int shift = begin_bit; shift < end_bit; shift += R) {
        const int bits = end_bit - shift < R ? end_bit - shift : R;
        const uint32_t mask = (1u << bits) - 1u;
        hipLaunchKernelGGL(radix_count_kernel<R>, dim3(ntiles), dim3(kThreads), 0, st, kb[cur], n,
                           shift, mask, counts, ntiles);
        hipLaunchKernelGGL(radix_scan_kernel, dim3((BINS + kWaves - 1) / kWaves), dim3(kThreads),
                           0, st, counts, ntiles, BINS, totals);
        if (vals)
            hipLaunchKernelGGL((radix_scatter_kernel<R, true>), dim3(ntiles), dim3(kThreads), 0, st,
                               kb[cur], vals_buf[cur], kb[cur ^ 1], vals_buf[cur ^ 1], n, shift, mask,
                               counts, totals, ntiles);
        else
            hipLaunchKernelGGL((radix_scatter_kernel<R, false>), dim3(ntiles), dim3(kThreads), 0, st,
                               kb[cur], nullptr, kb[cur ^ 1], nullptr, n, shift, mask, counts,
                               totals, ntiles);
        cur ^= 1;
    }
    if (hipGetLastError() != hipSuccess) return -1;
    if (cur) {  // the sorted data sit in the alternate buffers: exchange the pointers
        keys = kb[1];
        keys_alt = kb[0];
    }
    return cur;
}


// Digit width of the passes: kRadixBits, or TSNE_RADIX_BITS (8, 10 or 11) when set.
int radix_bits()
{
    const char *e = getenv("TSNE_RADIX_BITS");
    const int b = e ? atoi(e) : kRadixBits;
    return (b == 10 || b == 11) ? b : 8;
}

template <int R>
int sort_any(uint64_t *&keys, uint64_t *&keys_alt, int *vb[2], bool vals, int n, int begin_bit,
             int end_bit, void *scratch, hipStream_t st)
{
    const size_t ntiles = ((size_t)n + kTile - 1) / kTile;
    uint32_t *counts = (uint32_t *)scratch;
    uint32_t *totals = counts + ntiles * (1u << R);
    return sort_passes<R>(keys, keys_alt, vb, vals, n, begin_bit, end_bit, counts, totals, st);
}

int sort_dispatch(uint64_t *&keys, uint64_t *&keys_alt, int *vb[2], bool vals, int n,
                  int begin_bit, int end_bit, void *scratch, hipStream_t st)
{
    switch (radix_bits()) {
    case 10: return sort_any<10>(keys, keys_alt, vb, vals, n, begin_bit, end_bit, scratch, st);
    case 11: return sort_any<11>(keys, keys_alt, vb, vals, n, begin_bit, end_bit, scratch, st);
    default: return sort_any<8>(keys, keys_alt, vb, vals, n, begin_bit, end_bit, scratch, st);
    }
}

}  // namespace

size_t radix_scratch_bytes(int n)
{
    const size_t ntiles = ((size_t)n + kTile - 1) / kTile;
    constexpr size_t kMaxBins = 1u << 11;
    return (ntiles * kMaxBins + kMaxBins) * sizeof(uint32_t) + 256;
}

int radix_sort_pairs_u64(uint64_t *&keys, uint64_t *&keys_alt, int *&vals, int *&vals_alt, int n,
                         int begin_bit, int end_bit, void *scratch, size_t scratch_bytes,
                         hipStream_t st)
{
    if (n <= 0 || begin_bit >= end_bit) return 0;
    if (scratch_bytes < radix_scratch_bytes(n)) return -1;
    int *vb[2] = {vals, vals_alt};
    const int swapped = sort_dispatch(keys, keys_alt, vb, true, n, begin_bit, end_bit, scratch, st);
    if (swapped < 0) return -1;
    if (swapped) {
        vals = vb[1];
        vals_alt = vb[0];
    }
    return 0;
}

int radix_sort_keys_u64(uint64_t *&keys, uint64_t *&keys_alt, int n, int begin_bit, int end_bit,
                        void *scratch, size_t scratch_bytes, hipStream_t st)
{
    if (n <= 0 || begin_bit >= end_bit) return 0;
    if (scratch_bytes < radix_scratch_bytes(n)) return -1;
    int *vb[2] = {nullptr, nullptr};
    return sort_dispatch(keys, keys_alt, vb, false, n, begin_bit, end_bit, scratch, st) < 0 ? -1 : 0;
}

}  // namespace tsne
