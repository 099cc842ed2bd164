// knn.hip -- exact Euclidean k-nearest-neighbour graph on the device.
//
// Replaces the reference's exact search (tsne_h2pack/_knn_methods.py:71-119, ExactKNN:
// sklearn NearestNeighbors(algorithm='brute', metric='euclidean').kneighbors_graph(
// mode='distance'), float32 distances).  Result per point: its k nearest other points
// in increasing distance (ties by index), the point itself removed as sklearn removes
// it (kneighbors with X=None: k+1 neighbours are searched; the point's own entry is
// dropped, or the first entry when the point is not among them because of duplicates),
// distances sqrt(d^2) rounded to float32.
//
// The neighbour sets are exact in FP64.  Three steps:
//
//  1. Screening on MFMA.  Points are centred (FP64 mean) and split into two BF16 terms,
//     x = xh + xl + O(2^-16 |x|).  d^2 = |x|^2 + |y|^2 - 2 x.y with x.y taken as
//     xh.yh + xh.yl + xl.yh: one BF16 dot product of packed rows [xh | xh | xl | 1 1 1] and
//     [yh | yl | yh | -|y|^2/2 in three BF16 parts] (3D + 3 terms, zero padded to a
//     multiple of 32) on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  Every estimate
//     is within eps * (|x|^2 + |y|^2) of the exact d^2 (eps from the error analysis in
//     knn_eps below); the screening key of candidate y is the estimate minus eps |y|^2, so
//     key - eps |x|^2 <= d^2 <= key + eps |x|^2.
//     One workgroup owns 128 queries (4 waves of 32, the operands of 2 x 4 MFMA tiles held
//     in registers) and streams every candidate through LDS in chunks of 64 (direct
//     global-to-LDS loads, two chunks ahead).  Each query keeps a threshold and a buffer
//     in HBM: candidates whose key is below the threshold are appended (one compare per
//     pair against the accumulator; keys are formed only in the rare chunks where some
//     pair passes); a buffer that fills is cut to its L smallest keys (radix select over
//     the key bits by the owning wave), and the threshold drops to the L-th key.
//  2. Refinement in FP64.  The L survivors of each query get exact FP64 distances
//     (differences, fixed fma order); the k+1 smallest, ordered by (distance, index),
//     give the row.  The row is certain when its (k+1)-th exact d^2 is below
//     threshold - eps |x|^2: every rejected candidate is then farther.
//  3. Fallback for uncertain rows (none on well-conditioned data): every point within
//     the row's (k+1)-th refined distance -- an upper bound of the true one -- is
//     collected in FP64 and ranked exactly.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "common.h"
#include "joint_p.h"
#include "kernels.h"
#include "knn.h"
#include "../../include/tsne_h2pack.h"

namespace tsne {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kQBMax = 256;    // largest query block of a workgroup (8 waves x 2 x 16 rows)
constexpr int kCC = 64;        // candidates per chunk
constexpr int kLMax = 1024;    // largest survivor count L
constexpr int kFCap = 4096;    // fallback: largest number of points within the bound

struct KEntry {
    float key;
    int idx;
};

struct DEntry {
    double d2;
    int idx;
};

// Rigorous bound of |estimate - exact d^2| / (|x|^2 + |y|^2) for a packed dot product of
// K terms (K = 3 D padded): representation x - (xh + xl) <= 2^-16 |x| per component (two
// BF16 roundings) gives 2^-14 through the difference and 2^-16 through the dropped xl.yl;
// FP32 summation of the K exact products (including |y|^2 / 2) <= K * 2^-24 * 1.02
// (|x||y| + |y|^2 / 2); norms and the final
// combination add a few 2^-24.  A 10% margin covers the second-order terms.
double knn_eps(int kterms)
{
    return 1.1 * (std::ldexp(1.0, -14) + std::ldexp(1.0, -16) + 1.1 * kterms * std::ldexp(1.0, -24) +
                  8.0 * std::ldexp(1.0, -24));
}

// exact squared distance in a fixed order (the refinement and the fallback must agree)
__device__ __forceinline__ double exact_d2(const double *__restrict__ a, const double *__restrict__ b,
                                           int D)
{
    double s = 0.0;
    for (int f = 0; f < D; f++) {
        const double t = a[f] - b[f];
        s = fma(t, t, s);
    }
    return s;
}

// ---------------------------------------------------------------- preparation

// per-feature sums of 1024-row chunks (fixed order: reproducible)
__global__ void __launch_bounds__(256)
col_partial_kernel(const double *__restrict__ X, int n, int D, double *__restrict__ part)
{
    const int chunk = blockIdx.x;
    const int r0 = chunk * 1024, r1 = min(n, r0 + 1024);
    for (int f = threadIdx.x; f < D; f += 256) {
        double s = 0.0;
        for (int r = r0; r < r1; r++) s += X[(size_t)r * D + f];
        part[(size_t)chunk * D + f] = s;
    }
}

__global__ void __launch_bounds__(256)
col_mean_kernel(const double *__restrict__ part, int nchunk, int n, int D, double *__restrict__ mean)
{
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (f >= D) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; c++) s += part[(size_t)c * D + f];
    mean[f] = s / (double)n;
}

// Packed BF16 rows of point i (K = 32 * KT * nkb terms each, zero padded):
// A (queries) = [xh | xh | xl | 1 1 1], B (candidates) = [xh | xl | xh | -|y|^2 (1-eps)/2 in
// three parts], so that A_i . B_j = xh_i.yh_j + xh_i.yl_j + xl_i.yh_j - |y_j|^2 (1-eps)/2 and
// the screening key |x|^2 + |y|^2 (1-eps) - 2 x.y is |x|^2 - 2 A_i.B_j.  nq = |xh + xl|^2.
__global__ void __launch_bounds__(256)
split_kernel(const double *__restrict__ X, const double *__restrict__ mean, int n, int npad, int D,
             int K, float eps, __bf16 *__restrict__ XA, __bf16 *__restrict__ XB, float *__restrict__ nq)
{
    const int i = blockIdx.x;
    if (i >= npad) return;
    __shared__ double red[256];
    __shared__ float sh[1024], sl[1024];  // D <= 1024 per pass
    double s = 0.0;
    for (int f0 = 0; f0 < D; f0 += 1024) {
        const int nf = min(1024, D - f0);
        for (int f = threadIdx.x; f < nf; f += 256) {
            float hf = 0.f, lf = 0.f;
            if (i < n) {
                const double x = X[(size_t)i * D + f0 + f] - mean[f0 + f];
                hf = (float)(__bf16)(float)x;
                lf = (float)(__bf16)(float)(x - (double)hf);
            }
            sh[f] = hf;
            sl[f] = lf;
            const double r = (double)hf + (double)lf;
            s = fma(r, r, s);
        }
        __syncthreads();
        for (int f = threadIdx.x; f < nf; f += 256) {
            const size_t row = (size_t)i * K;
            XA[row + f0 + f] = (__bf16)sh[f];
            XA[row + D + f0 + f] = (__bf16)sh[f];
            XA[row + 2 * D + f0 + f] = (__bf16)sl[f];
            XB[row + f0 + f] = (__bf16)sh[f];
            XB[row + D + f0 + f] = (__bf16)sl[f];
            XB[row + 2 * D + f0 + f] = (__bf16)sh[f];
        }
        __syncthreads();
    }
    for (int f = 3 * D + 3 + threadIdx.x; f < K; f += 256) {
        XA[(size_t)i * K + f] = (__bf16)0.f;
        XB[(size_t)i * K + f] = (__bf16)0.f;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        nq[i] = (float)red[0];
        // terms 3D .. 3D+2: A = 1, B = -|y|^2 (1 - eps) / 2 as three BF16 parts, so that the
        // MFMA sum is x.y - |y|^2 (1 - eps) / 2 (padding rows: -1e38, never below a finite
        // threshold)
        const double v = (i < n) ? -0.5 * red[0] * (1.0 - (double)eps) : -1e38;
        const __bf16 h1 = (__bf16)(float)v;
        const double r1 = v - (double)(float)h1;
        const __bf16 h2 = (__bf16)(float)r1;
        const __bf16 h3 = (__bf16)(float)(r1 - (double)(float)h2);
        const size_t o3 = (size_t)i * K + 3 * D;
        XA[o3] = XA[o3 + 1] = XA[o3 + 2] = (__bf16)1.f;
        XB[o3] = h1;
        XB[o3 + 1] = h2;
        XB[o3 + 2] = h3;
    }
}

// ---------------------------------------------------------------- wave helpers

// Lanes of one wave exchanging data through LDS or global memory: the fences make the
// compiler treat the exchange as synchronisation (no value a lane stored earlier is
// forwarded to its later load of a location other lanes may have written).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int l = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (l >= o) v += u;
    }
    return v;
}

__device__ __forceinline__ unsigned long long lanes_below()
{
    const int l = threadIdx.x & 63;
    return (l == 0) ? 0ull : (~0ull >> (64 - l));
}

// Buffer entries are read back by the wave that wrote them (other lanes, after earlier
// cuts rewrote them): these reads go to L2 (agent-scope relaxed atomic loads), never to a
// line this CU's L1 may still hold from before the rewrite.
__device__ __forceinline__ KEntry read_entry(const KEntry *p)
{
    const unsigned long long v =
        __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return KEntry{__uint_as_float((unsigned)(v & 0xffffffffull)), (int)(unsigned)(v >> 32)};
}

// Cut the m entries of buffer b to the L smallest keys (one wave).  Keys are non-negative
// floats, so their bit patterns order like the values; a radix select over 4 digits of 8
// bits finds the L-th key T, then entries < T and the first (in buffer order) entries
// == T are kept, compacted in place.  Returns T.
__device__ __forceinline__ float wave_select_cut(KEntry *b, int m, int L, int *__restrict__ hist)
{
    const int l = threadIdx.x & 63;
    unsigned prefix = 0u, mask = 0u;
    int rem = L;
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 24 - 8 * pass;
        for (int j = 0; j < 4; j++) hist[4 * l + j] = 0;
        wave_sync();
        for (int e = l; e < m; e += 64) {
            const unsigned u = __float_as_uint(read_entry(b + e).key);
            if ((u & mask) == prefix) atomicAdd(&hist[(u >> shift) & 255u], 1);
        }
        wave_sync();
        const int h0 = hist[4 * l], h1 = hist[4 * l + 1], h2 = hist[4 * l + 2], h3 = hist[4 * l + 3];
        const int tot = h0 + h1 + h2 + h3;
        const int incl = wave_incl_scan(tot), excl = incl - tot;
        const bool here = excl < rem && rem <= incl;
        int dig = 0, before = 0;
        if (here) {
            const int r = rem - excl;
            if (r <= h0) { dig = 4 * l; before = excl; }
            else if (r <= h0 + h1) { dig = 4 * l + 1; before = excl + h0; }
            else if (r <= h0 + h1 + h2) { dig = 4 * l + 2; before = excl + h0 + h1; }
            else { dig = 4 * l + 3; before = excl + h0 + h1 + h2; }
        }
        const unsigned long long bal = __ballot(here);
        const int src = bal ? (__ffsll((long long)bal) - 1) : 0;
        dig = __shfl(dig, src, 64);
        before = __shfl(before, src, 64);
        rem -= before;
        prefix |= (unsigned)dig << shift;
        mask |= 255u << shift;
        wave_sync();
    }
    const unsigned T = prefix;
    int wpos = 0, eqseen = 0;
    for (int base = 0; base < m; base += 64) {
        const int e = base + l;
        KEntry v = {0.f, -1};
        if (e < m) v = read_entry(b + e);
        const unsigned u = __float_as_uint(v.key);
        const bool lt = e < m && u < T, eq = e < m && u == T;
        const unsigned long long eqm = __ballot(eq);
        const int eqrank = eqseen + __popcll(eqm & lanes_below());
        const bool keep = lt || (eq && eqrank < rem);
        const unsigned long long km = __ballot(keep);
        const int pos = wpos + __popcll(km & lanes_below());
        if (keep) b[pos] = v;
        wpos += __popcll(km);
        eqseen += __popcll(eqm);
    }
    return __uint_as_float(T);
}

// ---------------------------------------------------------------- screening (MFMA)

// Workgroup b owns queries q0 + QB b .. (wave w: QW of them).  Buffers: entry
// e of local query j at buf[j * cap + e].  On exit cnt[q] <= L entries and thr[q] hold
// the survivors and the final threshold (+inf when nothing was ever cut).  KT: MFMA
// k-steps (32 terms each) per K block; the packed rows hold nkb blocks.
template <int KT, bool DMA, int RB, int NW>
__global__ void __launch_bounds__(64 * NW)
knn_screen_kernel(const __bf16 *__restrict__ XA, const __bf16 *__restrict__ XB,
                  const float *__restrict__ nq, int n, int nkb,
                  int q0, int L, int cap, KEntry *__restrict__ buf, int *__restrict__ cnt_out,
                  float *__restrict__ thr_out)
{
    constexpr int QW = 16 * RB;        // queries per wave (RB 16-row MFMA blocks)
    constexpr int QB = NW * QW;        // queries per workgroup (NW waves)
    constexpr int NT = 64 * NW;
    constexpr int SEG = 32 * KT;       // bf16 per row per K block
    constexpr int LROW = SEG + 8;      // LDS row (16-byte pad: conflict-free b128 reads)
    constexpr int UNITS = SEG / 8;     // 16-byte units per candidate per K block
    constexpr int PER_T = (kCC * UNITS + NT - 1) / NT;  // 16-byte units each thread stages
    // direct-to-LDS staging: rounds of NT slots of 16 bytes (row stride UNITS + 1 slots)
    constexpr int ROUNDS = (kCC * (UNITS + 1) + NT - 1) / NT;
    constexpr int BSZ = DMA ? ROUNDS * NT * 8 : kCC * LROW;
    // DMA: three buffers, chunks fetched two stages ahead; registers: two buffers
    constexpr int NBUF = DMA ? 3 : 2;
    __shared__ __attribute__((aligned(16))) __bf16 Bs[NBUF][BSZ];
    __shared__ int cnt_lds[NW][QW];
    __shared__ float thr_lds[NW][QW];
    __shared__ int hist[NW][256];

    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int qw0 = q0 + blockIdx.x * QB + w * QW;  // first query of this wave
    const int rowlen = SEG * nkb;
    const int nchunk = (n + kCC - 1) / kCC;
    const int nstage = nchunk * nkb;
    KEntry *const wbuf = buf + (size_t)(blockIdx.x * QB + w * QW) * cap;

    if (l < QW) {
        cnt_lds[w][l] = 0;
        thr_lds[w][l] = INFINITY;
    }

    // query operands: row rb*16 + l%16, terms s*32 + (l/16)*8 .. +7 of the block
    bf16x8 a[RB][KT];
    float nqr[RB][4], htq[RB][4];  // htq = (|x|^2 - threshold) / 2: the pass test is A.B > htq
    auto load_a = [&](int kb) {
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const __bf16 *p = XA + (size_t)(qw0 + rb * 16 + (l & 15)) * rowlen + kb * SEG + (l >> 4) * 8;
#pragma unroll
            for (int ks = 0; ks < KT; ks++) a[rb][ks] = *(const bf16x8 *)(p + ks * 32);
        }
    };
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            nqr[rb][i] = nq[qw0 + rb * 16 + (l >> 4) * 4 + i];
            htq[rb][i] = -INFINITY;
        }
    if (nkb == 1) load_a(0);

    // stage s = chunk s / nkb, K block s % nkb
    uint4 pre[DMA ? 1 : PER_T];
    // DMA: slot r*NT + 64w + l of a buffer holds unit u of candidate cand (slot =
    // cand * (UNITS + 1) + u); padding slots fetch a valid unit that is never read.  The
    // per-lane element offsets within a chunk are fixed: computed once.
    int doff[DMA ? ROUNDS : 1];
    if constexpr (DMA) {
#pragma unroll
        for (int r = 0; r < ROUNDS; r++) {
            const int slot = r * NT + w * 64 + l;
            int cand = slot / (UNITS + 1), u = slot % (UNITS + 1);
            if (cand >= kCC || u >= UNITS) {
                cand = 0;
                u = 0;
            }
            doff[r] = cand * rowlen + u * 8;
        }
    }
    auto fetch = [&](int s) {
        const int c0 = (s / nkb) * kCC, kb = s % nkb;
        if constexpr (DMA) {
            const int b = s % NBUF;
            const __bf16 *base = XB + (size_t)c0 * rowlen + kb * SEG;  // uniform
#pragma unroll
            for (int r = 0; r < ROUNDS; r++)
                __builtin_amdgcn_global_load_lds(
                    (const void *)(base + doff[r]),
                    (void __attribute__((address_space(3))) *)&Bs[b][(r * NT + w * 64) * 8], 16, 0, 0);
        } else {
#pragma unroll
            for (int r = 0; r < PER_T; r++) {
                const int u = r * NT + t, cand = u / UNITS, part = u % UNITS;
                if (u < kCC * UNITS)
                    pre[r] = *(const uint4 *)(XB + (size_t)(c0 + cand) * rowlen + kb * SEG + part * 8);
            }
        }
    };
    auto stash = [&](int b) {
        if constexpr (!DMA) {
#pragma unroll
            for (int r = 0; r < PER_T; r++) {
                const int u = r * NT + t, cand = u / UNITS, part = u % UNITS;
                if (u < kCC * UNITS) *(uint4 *)(&Bs[b][cand * LROW + part * 8]) = pre[r];
            }
        }
    };
    fetch(0);
    stash(0);
    if constexpr (DMA) {
        if (nstage > 1) fetch(1);
        __builtin_amdgcn_s_waitcnt(0);
    }
    __syncthreads();

    f32x4 acc[RB][4];
    for (int s = 0; s < nstage; s++) {
        const int b = s % NBUF;
        const bool more = s + 1 < nstage;
        if constexpr (DMA) {
            if (s + 2 < nstage) fetch(s + 2);
        } else {
            if (more) fetch(s + 1);
        }
        const int kb = s % nkb;
        if (kb == 0) {
#pragma unroll
            for (int rb = 0; rb < RB; rb++)
#pragma unroll
                for (int cb = 0; cb < 4; cb++) acc[rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
        if (nkb > 1) load_a(kb);
#pragma unroll
        for (int ks = 0; ks < KT; ks++) {
            bf16x8 bv[4];
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                bv[cb] = *(const bf16x8 *)&Bs[b][(cb * 16 + (l & 15)) * LROW + (l >> 4) * 8 + ks * 32];
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int rb = 0; rb < RB; rb++)
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rb][ks], bv[cb], acc[rb][cb], 0, 0, 0);
        }
        if (kb == nkb - 1) {
            // D layout: lane l holds column l%16, rows (l/16)*4 + i of each 16 x 16 tile.
            // Pass test per pair: A.B > (|x|^2 - threshold) / 2, one compare.
            const int c0 = (s / nkb) * kCC;
            unsigned long long anym = 0;  // compare results go to scalar masks
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int rb = 0; rb < RB; rb++)
#pragma unroll
                    for (int i = 0; i < 4; i++) anym |= __ballot(acc[rb][cb][i] > htq[rb][i]);
            if (anym) {
#pragma unroll
                for (int cb = 0; cb < 4; cb++) {
                    const int c = c0 + cb * 16 + (l & 15);
#pragma unroll
                    for (int rb = 0; rb < RB; rb++)
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            if (acc[rb][cb][i] > htq[rb][i] && c < n) {
                                float key = fmaf(-2.f, acc[rb][cb][i], nqr[rb][i]);
                                key = key > 0.f ? key : 0.f;
                                const int j = rb * 16 + (l >> 4) * 4 + i;
                                const int pos = atomicAdd(&cnt_lds[w][j], 1);
                                wbuf[(size_t)j * cap + pos] = KEntry{key, c};
                            }
                        }
                }
                // cut every buffer that could overflow in the next chunk (cap >= L + 64)
                __threadfence_block();
                wave_sync();
                unsigned long long full = __ballot(l < QW && cnt_lds[w][l] > cap - kCC);
                while (full) {
                    const int j = __ffsll((long long)full) - 1;
                    full &= full - 1;
                    const float T = wave_select_cut(wbuf + (size_t)j * cap, cnt_lds[w][j], L, hist[w]);
                    __threadfence_block();
                    wave_sync();
                    if (l == 0) {
                        cnt_lds[w][j] = L;
                        thr_lds[w][j] = T;
                    }
                    wave_sync();
                }
#pragma unroll
                for (int rb = 0; rb < RB; rb++)
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        htq[rb][i] = 0.5f * (nqr[rb][i] - thr_lds[w][rb * 16 + (l >> 4) * 4 + i]);
            }
        }
        if (more) stash(b ^ 1);
        if constexpr (DMA) {
            // the loads of stage s + 1 have landed; those of stage s + 2 (ROUNDS, issued
            // last) may still be in flight.  vmcnt <= ROUNDS, no wait on lgkm / exp.
            if (s + 2 < nstage)
                __builtin_amdgcn_s_waitcnt(0xF70 | ROUNDS);
            else
                __builtin_amdgcn_s_waitcnt(0);
        }
        __syncthreads();
    }
    // final cut to L survivors
    __threadfence_block();
    wave_sync();
    unsigned long long over = __ballot(l < QW && cnt_lds[w][l] > L);
    while (over) {
        const int j = __ffsll((long long)over) - 1;
        over &= over - 1;
        const float T = wave_select_cut(wbuf + (size_t)j * cap, cnt_lds[w][j], L, hist[w]);
        wave_sync();
        if (l == 0) {
            cnt_lds[w][j] = L;
            thr_lds[w][j] = T;
        }
        wave_sync();
    }
    const int q = qw0 + l;
    if (l < QW && q < n) {
        cnt_out[q] = cnt_lds[w][l];
        thr_out[q] = thr_lds[w][l];
    }
}

// ---------------------------------------------------------------- refinement (FP64)

// Write row q from m exact entries (LDS, one wave): ranks by (d2, idx), the point itself
// removed (or rank 0 when it is not among the first k+1), k outputs.  Returns the
// (k+1)-th d2 (broadcast).
__device__ __forceinline__ double wave_emit_row(const double *d2s, const int *ids, int m, int q, int k,
                                int *__restrict__ out_idx, float *__restrict__ out_dist)
{
    const int l = threadIdx.x & 63;
    const int K1 = k + 1;
    auto rank_of = [&](int e) {
        const double d = d2s[e];
        const int c = ids[e];
        int r = 0;
        for (int f = 0; f < m; f++) {
            const double g = d2s[f];
            r += (g < d) || (g == d && ids[f] < c);
        }
        return r;
    };
    int self_r = -1;
    double dk1 = 0.0;
    for (int e = l; e < m; e += 64) {
        const int r = rank_of(e);
        if (ids[e] == q && r < K1) self_r = r;
        if (r == K1 - 1) dk1 = d2s[e];
    }
    // the self rank (at most one lane) and the (k+1)-th d2 (exactly one lane)
    for (int o = 32; o > 0; o >>= 1) {
        self_r = max(self_r, __shfl_xor(self_r, o, 64));
        dk1 = fmax(dk1, __shfl_xor(dk1, o, 64));
    }
    const int drop = self_r >= 0 ? self_r : 0;
    for (int e = l; e < m; e += 64) {
        const int r = rank_of(e);
        if (r < K1 && r != drop) {
            const int slot = r - (r > drop ? 1 : 0);
            out_idx[(size_t)q * k + slot] = ids[e];
            out_dist[(size_t)q * k + slot] = (float)sqrt(d2s[e]);
        }
    }
    return dk1;
}

__global__ void __launch_bounds__(256)
knn_refine_kernel(const double *__restrict__ X, int n, int D, int k, int q0, int nq_batch, int cap,
                  const KEntry *__restrict__ buf, const int *__restrict__ cnt,
                  const float *__restrict__ thr, const float *__restrict__ nq, double eps,
                  int *__restrict__ out_idx, float *__restrict__ out_dist, double *__restrict__ bound,
                  int *__restrict__ n_unc, int *__restrict__ unc)
{
    __shared__ double d2s[4][kLMax];
    __shared__ int ids[4][kLMax];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + w;  // local query
    const int q = q0 + j;
    if (j >= nq_batch || q >= n) return;  // whole wave
    const int m = cnt[q];
    const KEntry *b = buf + (size_t)j * cap;
    const double *xq = X + (size_t)q * D;
    for (int e = l; e < m; e += 64) {
        const int c = b[e].idx;
        ids[w][e] = c;
        d2s[w][e] = exact_d2(xq, X + (size_t)c * D, D);
    }
    wave_sync();
    const double dk1 = wave_emit_row(d2s[w], ids[w], m, q, k, out_idx, out_dist);
    if (l == 0) {
        bound[q] = dk1;
        const float T = thr[q];
        // rejected candidates have d^2 >= T - eps |x|^2 up to the FP32 rounding of the
        // pass test (a few 2^-24 of |x|^2 + T)
        const double nqd = (double)nq[q], Td = (double)T;
        const bool certain = isinf(T) || dk1 < Td - eps * nqd * 1.0001 - std::ldexp(1.0, -20) * (nqd + Td);
        if (!certain) unc[atomicAdd(n_unc, 1)] = q;
    }
}

// ---------------------------------------------------------------- fallback (FP64)

// Block b: candidates (b / nu) * 256 + t of uncertain row b % nu.  Collects every point
// with d2 <= bound (the row's (k+1)-th refined d2, itself from real points).
__global__ void __launch_bounds__(256)
knn_fb_collect_kernel(const double *__restrict__ X, int n, int D, const int *__restrict__ rows,
                      int nu, const double *__restrict__ bound, int *__restrict__ fcnt,
                      DEntry *__restrict__ fbuf)
{
    const int b = blockIdx.x;
    const int u = b % nu, c = (b / nu) * 256 + threadIdx.x;
    if (c >= n) return;
    const int q = rows[u];
    const double d = exact_d2(X + (size_t)q * D, X + (size_t)c * D, D);
    if (d <= bound[q]) {
        const int pos = atomicAdd(&fcnt[u], 1);
        if (pos < kFCap) fbuf[(size_t)u * kFCap + pos] = DEntry{d, c};
    }
}

__global__ void __launch_bounds__(64)
knn_fb_finish_kernel(int k, const int *__restrict__ rows, const int *__restrict__ fcnt,
                     const DEntry *__restrict__ fbuf, int *__restrict__ out_idx,
                     float *__restrict__ out_dist, int *__restrict__ n_ovf, int *__restrict__ ovf)
{
    __shared__ double d2s[kFCap];
    __shared__ int ids[kFCap];
    const int u = blockIdx.x, l = threadIdx.x;
    const int q = rows[u];
    const int m = fcnt[u];
    if (m > kFCap || m < k + 1) {  // too many inside the bound (or, defensively, too few)
        if (l == 0) ovf[atomicAdd(n_ovf, 1)] = q;
        return;
    }
    for (int e = l; e < m; e += 64) {
        d2s[e] = fbuf[(size_t)u * kFCap + e].d2;
        ids[e] = fbuf[(size_t)u * kFCap + e].idx;
    }
    wave_sync();
    (void)wave_emit_row(d2s, ids, m, q, k, out_idx, out_dist);
}

// last resort for one row: every distance, sorted
__global__ void __launch_bounds__(256)
knn_row_all_kernel(const double *__restrict__ X, int n, int D, int q, unsigned long long *__restrict__ key,
                   int *__restrict__ val)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const double d = exact_d2(X + (size_t)q * D, X + (size_t)c * D, D);
    key[c] = __double_as_longlong(d);  // d >= 0: bit order is value order
    val[c] = c;
}

__global__ void __launch_bounds__(64)
knn_row_emit_kernel(int q, int k, const unsigned long long *__restrict__ key, const int *__restrict__ val,
                    int *__restrict__ out_idx, float *__restrict__ out_dist)
{
    __shared__ double d2s[kLMax];
    __shared__ int ids[kLMax];
    const int l = threadIdx.x, m = k + 1;
    for (int e = l; e < m; e += 64) {
        d2s[e] = __longlong_as_double(key[e]);
        ids[e] = val[e];
    }
    wave_sync();
    (void)wave_emit_row(d2s, ids, m, q, k, out_idx, out_dist);
}

template <typename T> int kalloc(T **p, size_t count, hipStream_t st)
{
    *p = nullptr;
    return hipMallocAsync((void **)p, std::max<size_t>(count, 1) * sizeof(T), st) == hipSuccess ? 0 : -1;
}

template <typename T> void kfree(T *&p, hipStream_t st)
{
    if (p) (void)hipFreeAsync((void *)p, st);
    p = nullptr;
}

template <int KT, int NW>
void launch_screen_nw(bool dma, int nqb, hipStream_t st, const __bf16 *XA, const __bf16 *XB, const float *nq,
                      int n, int nkb, int q0, int L, int cap, KEntry *buf, int *cnt, float *thr)
{
    const int grid = nqb / (NW * 32);
    if (dma)
        hipLaunchKernelGGL((knn_screen_kernel<KT, true, 2, NW>), dim3(grid), dim3(64 * NW), 0, st, XA, XB, nq,
                           n, nkb, q0, L, cap, buf, cnt, thr);
    else
        hipLaunchKernelGGL((knn_screen_kernel<KT, false, 2, NW>), dim3(grid), dim3(64 * NW), 0, st, XA, XB, nq,
                           n, nkb, q0, L, cap, buf, cnt, thr);
}

// 2 row blocks (32 queries) per wave; 4 waves per workgroup (TSNE_KNN_WAVES=8: 8 waves,
// 256 queries per staged chunk, one workgroup per CU: slower), candidate chunks staged
// straight into LDS (TSNE_KNN_DMA=0 stages through registers).
template <int KT>
void launch_screen(int nqb, hipStream_t st, const __bf16 *XA, const __bf16 *XB, const float *nq, int n,
                   int nkb, int q0, int L, int cap, KEntry *buf, int *cnt, float *thr)
{
    static const bool dma = !getenv("TSNE_KNN_DMA") || atoi(getenv("TSNE_KNN_DMA")) != 0;
    static const int nw = (getenv("TSNE_KNN_WAVES") && atoi(getenv("TSNE_KNN_WAVES")) == 8) ? 8 : 4;
    if (nw == 4)
        launch_screen_nw<KT, 4>(dma, nqb, st, XA, XB, nq, n, nkb, q0, L, cap, buf, cnt, thr);
    else
        launch_screen_nw<KT, 8>(dma, nqb, st, XA, XB, nq, n, nkb, q0, L, cap, buf, cnt, thr);
}

// Host check of the survivor buffers of one screening launch (TSNE_KNN_DEBUG): counts,
// distinct candidates, keys below the threshold.
void debug_buffers(int n, int q0, int nqb, int L, int cap, const KEntry *d_buf, const int *d_cnt,
                   const float *d_thr, hipStream_t st)
{
    const int nq = std::min(nqb, n - q0);
    std::vector<KEntry> b((size_t)nq * cap);
    std::vector<int> cnt(nq);
    std::vector<float> thr(nq);
    if (hipMemcpyAsync(b.data(), d_buf, b.size() * sizeof(KEntry), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(cnt.data(), d_cnt + q0, nq * sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(thr.data(), d_thr + q0, nq * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return;
    int bad_cnt = 0, dup = 0, above = 0, shown = 0;
    for (int j = 0; j < nq; j++) {
        const int m = cnt[j];
        if (m < 0 || m > L) { bad_cnt++; continue; }
        std::vector<int> ids;
        bool ab = false;
        for (int e = 0; e < m; e++) {
            ids.push_back(b[(size_t)j * cap + e].idx);
            ab |= !(b[(size_t)j * cap + e].key <= thr[j]);
        }
        std::sort(ids.begin(), ids.end());
        const bool d = std::adjacent_find(ids.begin(), ids.end()) != ids.end();
        dup += d;
        above += ab;
        if ((d || ab) && shown < 3) {
            shown++;
            fprintf(stderr, "knn debug: query %d m %d thr %.9g entries:", q0 + j, m, thr[j]);
            for (int e = 0; e < m; e++)
                fprintf(stderr, " (%d %.6g)", b[(size_t)j * cap + e].idx, b[(size_t)j * cap + e].key);
            fprintf(stderr, "\n");
        }
    }
    fprintf(stderr, "knn debug: batch q0 %d: %d queries, bad counts %d, with duplicates %d, "
                    "keys above threshold %d\n", q0, nq, bad_cnt, dup, above);
}

}  // namespace

int knn_survivors(int k) { return std::max(k + 1 + 32, ((k + 1) * 5 + 3) / 4); }

int knn_exact_device(const double *d_X, int n, int D, int k, int *d_idx, float *d_dist,
                     KnnStats *stats, hipStream_t st)
{
    KnnStats S{};
    const int L = knn_survivors(k);
    if (n < 2 || D < 1 || k < 1 || k >= n || L > kLMax) return -2;
    // packed rows of 3 D + 3 terms: nkb K blocks of KT <= 8 MFMA k-steps
    const int ksteps = (3 * D + 3 + 31) / 32;
    const int nkb = (ksteps + 7) / 8;
    const int KT = (ksteps + nkb - 1) / nkb;
    const int kpad = 32 * KT * nkb;
    const int npad = (n + kQBMax - 1) / kQBMax * kQBMax;
    const int cap = std::max(4 * L, L + 2 * kCC);
    const double eps = knn_eps(kpad);
    // queries per screening launch: the survivor buffers stay under ~1 GiB
    const long long per_q = (long long)cap * sizeof(KEntry);
    int qbatch = (int)std::min<long long>(npad, std::max<long long>(kQBMax, (1ll << 30) / per_q / kQBMax * kQBMax));
    const int nchunk = (n + 1023) / 1024;

    double *d_part = nullptr, *d_mean = nullptr, *d_bound = nullptr;
    __bf16 *d_XA = nullptr, *d_XB = nullptr;
    float *d_nq = nullptr, *d_thr = nullptr;
    int *d_cnt = nullptr, *d_ctr = nullptr, *d_unc = nullptr, *d_ovf = nullptr;
    KEntry *d_buf = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    const long long nchunk_c = (n + kCC - 1) / kCC;
    int rc = -1;
    if (kalloc(&d_part, (size_t)nchunk * D, st) || kalloc(&d_mean, D, st) ||
        kalloc(&d_XA, (size_t)npad * kpad, st) || kalloc(&d_XB, (size_t)npad * kpad, st) || kalloc(&d_nq, npad, st) ||
        kalloc(&d_thr, n, st) || kalloc(&d_cnt, n, st) || kalloc(&d_bound, n, st) || kalloc(&d_ctr, 64, st) ||
        kalloc(&d_unc, n, st) || kalloc(&d_ovf, n, st) || kalloc(&d_buf, (size_t)qbatch * cap, st))
        goto done;
    if (hipMemsetAsync(d_ctr, 0, 64 * sizeof(int), st) != hipSuccess) goto done;
    for (int e = 0; e < 4; e++)
        if (hipEventCreate(&ev[e]) != hipSuccess) goto done;
    (void)hipEventRecord(ev[0], st);
    hipLaunchKernelGGL(col_partial_kernel, dim3(nchunk), dim3(256), 0, st, d_X, n, D, d_part);
    hipLaunchKernelGGL(col_mean_kernel, dim3((D + 255) / 256), dim3(256), 0, st, d_part, nchunk, n, D, d_mean);
    hipLaunchKernelGGL(split_kernel, dim3(npad), dim3(256), 0, st, d_X, d_mean, n, npad, D, kpad,
                       (float)eps, d_XA, d_XB, d_nq);
    for (int q0 = 0; q0 < npad; q0 += qbatch) {
        const int nqb = std::min(qbatch, npad - q0);
        (void)hipEventRecord(ev[2], st);
        switch (KT) {
#define KNN_SCREEN_CASE(v)                                                                          \
    case v:                                                                                         \
        launch_screen<v>(nqb, st, d_XA, d_XB, d_nq, n, nkb, q0, L, cap, d_buf, d_cnt, d_thr); \
        break;
            KNN_SCREEN_CASE(1) KNN_SCREEN_CASE(2) KNN_SCREEN_CASE(3) KNN_SCREEN_CASE(4)
            KNN_SCREEN_CASE(5) KNN_SCREEN_CASE(6) KNN_SCREEN_CASE(7)
            default: launch_screen<8>(nqb, st, d_XA, d_XB, d_nq, n, nkb, q0, L, cap, d_buf, d_cnt, d_thr);
#undef KNN_SCREEN_CASE
        }
        (void)hipEventRecord(ev[3], st);
        {
            float ms = 0.f;
            if (hipEventSynchronize(ev[3]) == hipSuccess && hipEventElapsedTime(&ms, ev[2], ev[3]) == hipSuccess)
                S.screen_ms += ms;
        }
        if (getenv("TSNE_KNN_DEBUG")) debug_buffers(n, q0, nqb, L, cap, d_buf, d_cnt, d_thr, st);
        hipLaunchKernelGGL(knn_refine_kernel, dim3((nqb + 3) / 4), dim3(256), 0, st, d_X, n, D, k, q0, nqb,
                           cap, d_buf, d_cnt, d_thr, d_nq, eps, d_idx, d_dist, d_bound, d_ctr, d_unc);
    }
    {
        int nunc = 0;
        if (hipMemcpyAsync(&nunc, d_ctr, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            goto done;
        S.uncertain_rows = nunc;
        // fallback in groups of rows: every point within the row's bound, ranked exactly
        const int ncb = (n + 255) / 256;
        const int grp = std::max(1, std::min(nunc, std::max(1, (int)std::min<long long>(1024, (1ll << 31) / ncb - 1))));
        DEntry *d_fbuf = nullptr;
        int *d_fcnt = nullptr;
        if (nunc > 0 && (kalloc(&d_fbuf, (size_t)grp * kFCap, st) || kalloc(&d_fcnt, grp, st))) {
            kfree(d_fbuf, st);
            goto done;
        }
        for (int u0 = 0; u0 < nunc; u0 += grp) {
            const int nu = std::min(grp, nunc - u0);
            if (hipMemsetAsync(d_fcnt, 0, sizeof(int) * nu, st) != hipSuccess) break;
            hipLaunchKernelGGL(knn_fb_collect_kernel, dim3(ncb * nu), dim3(256), 0, st, d_X, n, D, d_unc + u0,
                               nu, d_bound, d_fcnt, d_fbuf);
            hipLaunchKernelGGL(knn_fb_finish_kernel, dim3(nu), dim3(64), 0, st, k, d_unc + u0, d_fcnt, d_fbuf,
                               d_idx, d_dist, d_ctr + 16, d_ovf);
        }
        kfree(d_fbuf, st);
        kfree(d_fcnt, st);
        int novf = 0;
        if (hipMemcpyAsync(&novf, d_ctr + 16, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            goto done;
        S.overflow_rows = novf;
        if (novf > 0) {
            std::vector<int> rows(novf);
            if (hipMemcpyAsync(rows.data(), d_ovf, sizeof(int) * novf, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                goto done;
            unsigned long long *k1 = nullptr, *k2 = nullptr;
            int *v1 = nullptr, *v2 = nullptr;
            void *tmp = nullptr;
            size_t tb = 0;
            bool ok = true;
            {
                rocprim::double_buffer<unsigned long long> kb(k1, k2);
                rocprim::double_buffer<int> vb(v1, v2);
                ok = rocprim::radix_sort_pairs(nullptr, tb, kb, vb, n, 0, 64, st) == hipSuccess;
            }
            ok = ok && !kalloc(&k1, n, st) && !kalloc(&k2, n, st) && !kalloc(&v1, n, st) && !kalloc(&v2, n, st) &&
                 hipMallocAsync(&tmp, std::max<size_t>(tb, 1), st) == hipSuccess;
            for (int r = 0; ok && r < novf; r++) {
                hipLaunchKernelGGL(knn_row_all_kernel, dim3(ncb), dim3(256), 0, st, d_X, n, D, rows[r], k1, v1);
                rocprim::double_buffer<unsigned long long> kb(k1, k2);
                rocprim::double_buffer<int> vb(v1, v2);
                size_t bytes = tb;
                ok = rocprim::radix_sort_pairs(tmp, bytes, kb, vb, n, 0, 64, st) == hipSuccess;
                if (ok)
                    hipLaunchKernelGGL(knn_row_emit_kernel, dim3(1), dim3(64), 0, st, rows[r], k, kb.current(),
                                       vb.current(), d_idx, d_dist);
            }
            kfree(k1, st); kfree(k2, st); kfree(v1, st); kfree(v2, st);
            if (tmp) (void)hipFreeAsync(tmp, st);
            if (!ok) goto done;
        }
    }
    (void)hipEventRecord(ev[1], st);
    rc = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess ? 0 : -1;
    if (rc == 0) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) S.device_ms = ms;
    }
    S.survivors = L;
    S.eps = eps;
    S.screen_flops = 2LL * kpad * (long long)npad * (long long)(nchunk_c * kCC);
done:
    for (int e = 0; e < 4; e++)
        if (ev[e]) (void)hipEventDestroy(ev[e]);
    kfree(d_part, st); kfree(d_mean, st); kfree(d_bound, st); kfree(d_XA, st); kfree(d_XB, st); kfree(d_nq, st);
    kfree(d_thr, st); kfree(d_cnt, st); kfree(d_ctr, st); kfree(d_unc, st);
    kfree(d_ovf, st); kfree(d_buf, st);
    if (stats) *stats = S;
    return rc;
}

namespace {

// Row i of a k-NN graph (k entries, distance order) in the reference's neighbour indexing:
// columns increasing, distances as float64 (one wave per row; indices of a row are
// distinct, so an entry's rank is the number of smaller indices).
__global__ void __launch_bounds__(256)
row_colsort_kernel(int n, int k, const int *__restrict__ idx, const float *__restrict__ dist,
                   int *__restrict__ col, double *__restrict__ dd)
{
    __shared__ int s_i[4][kLMax];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + w;
    if (row >= n) return;  // whole wave
    const size_t o = (size_t)row * k;
    for (int e = l; e < k; e += 64) s_i[w][e] = idx[o + e];
    wave_sync();
    for (int e = l; e < k; e += 64) {
        const int c = s_i[w][e];
        int r = 0;
        for (int f = 0; f < k; f++) r += s_i[w][f] < c;
        col[o + r] = c;
        dd[o + r] = (double)dist[o + e];
    }
}

__global__ void __launch_bounds__(256) arange_kernel(int n, int k, int *__restrict__ ptr)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i <= n) ptr[i] = i * k;
}

}  // namespace

}  // namespace tsne

extern "C" sparse_matrix_t *tsne_joint_P_from_points(const double *X, int n_samples, int n_features, int k,
                                                     double perplexity, tsne_knn_stats_t *stats)
{
    if (stats) *stats = tsne_knn_stats_t{};
    if (!X || n_samples < 2 || n_features < 1 || k < 1 || k >= n_samples || tsne::knn_survivors(k) > 1024 ||
        !(perplexity > 0.0) || (long long)n_samples * k > 0x7fffffffLL) {
        fprintf(stderr, "tsne_joint_P_from_points: invalid arguments (need 1 <= k < n_samples, k <= 900, "
                        "perplexity > 0)\n");
        return nullptr;
    }
    const size_t nx = (size_t)n_samples * n_features;
    for (size_t i = 0; i < nx; i++)
        if (!std::isfinite(X[i])) {
            fprintf(stderr, "tsne_joint_P_from_points: X holds a value that is not finite\n");
            return nullptr;
        }
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        fprintf(stderr, "tsne_joint_P_from_points: no HIP device available\n");
        return nullptr;
    }
    const int n = n_samples, nnz = n_samples * k;
    double *d_X = nullptr, *d_dd = nullptr, *d_p = nullptr;
    int *d_idx = nullptr, *d_col = nullptr, *d_ptr = nullptr;
    float *d_dist = nullptr;
    tsne::KnnStats S{};
    tsne::JointP J{};
    sparse_matrix_t *P = nullptr;
    bool ok = hipMallocAsync((void **)&d_X, nx * sizeof(double), st) == hipSuccess &&
              hipMallocAsync((void **)&d_idx, (size_t)nnz * sizeof(int), st) == hipSuccess &&
              hipMallocAsync((void **)&d_dist, (size_t)nnz * sizeof(float), st) == hipSuccess &&
              hipMemcpyAsync(d_X, X, nx * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              tsne::knn_exact_device(d_X, n, n_features, k, d_idx, d_dist, &S, st) == 0;
    if (d_X) (void)hipFreeAsync(d_X, st);
    ok = ok && hipMallocAsync((void **)&d_col, (size_t)nnz * sizeof(int), st) == hipSuccess &&
         hipMallocAsync((void **)&d_dd, (size_t)nnz * sizeof(double), st) == hipSuccess &&
         hipMallocAsync((void **)&d_ptr, ((size_t)n + 1) * sizeof(int), st) == hipSuccess &&
         hipMallocAsync((void **)&d_p, (size_t)nnz * sizeof(double), st) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(tsne::row_colsort_kernel, dim3((n + 3) / 4), dim3(256), 0, st, n, k, d_idx, d_dist,
                           d_col, d_dd);
        hipLaunchKernelGGL(tsne::arange_kernel, dim3((n + 1 + 255) / 256), dim3(256), 0, st, n, k, d_ptr);
        ok = hipGetLastError() == hipSuccess &&
             tsne::launch_perplexity(n, d_dd, d_ptr, perplexity, d_p, st) == 0 &&
             tsne::joint_P_device(n, nnz, d_ptr, d_col, d_p, &J, st) == 0;
    }
    if (ok) {
        P = sparse_matrix_create(n, n, J.nnz);
        ok = P && hipMemcpyAsync(P->row_ptr, J.rp, ((size_t)n + 1) * sizeof(int), hipMemcpyDeviceToHost, st) ==
                      hipSuccess &&
             (J.nnz == 0 ||
              (hipMemcpyAsync(P->col_idx, J.col, (size_t)J.nnz * sizeof(int), hipMemcpyDeviceToHost, st) ==
                   hipSuccess &&
               hipMemcpyAsync(P->values, J.val, (size_t)J.nnz * sizeof(double), hipMemcpyDeviceToHost, st) ==
                   hipSuccess)) &&
             hipStreamSynchronize(st) == hipSuccess;
    }
    for (void *q : {(void *)d_idx, (void *)d_dist, (void *)d_col, (void *)d_dd, (void *)d_ptr, (void *)d_p,
                    (void *)J.rp, (void *)J.col, (void *)J.val})
        if (q) (void)hipFreeAsync(q, st);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (!ok) {
        fprintf(stderr, "tsne_joint_P_from_points: device failure\n");
        if (P) sparse_matrix_destroy(P);
        return nullptr;
    }
    if (stats) {
        stats->survivors = S.survivors;
        stats->uncertain_rows = S.uncertain_rows;
        stats->overflow_rows = S.overflow_rows;
        stats->eps = S.eps;
        stats->device_ms = S.device_ms;
        stats->screen_ms = S.screen_ms;
        stats->screen_flops = S.screen_flops;
    }
    return P;
}

extern "C" int tsne_knn_exact(const double *X, int n_samples, int n_features, int k, int *indices,
                              float *distances, tsne_knn_stats_t *stats)
{
    if (stats) *stats = tsne_knn_stats_t{};
    if (!X || !indices || !distances || n_samples < 2 || n_features < 1 || k < 1 || k >= n_samples ||
        tsne::knn_survivors(k) > 1024) {
        fprintf(stderr, "tsne_knn_exact: invalid arguments (need 1 <= k < n_samples, k <= 900)\n");
        return -1;
    }
    const size_t nx = (size_t)n_samples * n_features;
    for (size_t i = 0; i < nx; i++)
        if (!std::isfinite(X[i])) {
            fprintf(stderr, "tsne_knn_exact: X holds a value that is not finite\n");
            return -1;
        }
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        fprintf(stderr, "tsne_knn_exact: no HIP device available\n");
        return -1;
    }
    double *d_X = nullptr;
    int *d_idx = nullptr;
    float *d_dist = nullptr;
    tsne::KnnStats S{};
    const size_t nk = (size_t)n_samples * k;
    bool ok = hipMallocAsync((void **)&d_X, nx * sizeof(double), st) == hipSuccess &&
              hipMallocAsync((void **)&d_idx, nk * sizeof(int), st) == hipSuccess &&
              hipMallocAsync((void **)&d_dist, nk * sizeof(float), st) == hipSuccess &&
              hipMemcpyAsync(d_X, X, nx * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              tsne::knn_exact_device(d_X, n_samples, n_features, k, d_idx, d_dist, &S, st) == 0 &&
              hipMemcpyAsync(indices, d_idx, nk * sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
              hipMemcpyAsync(distances, d_dist, nk * sizeof(float), hipMemcpyDeviceToHost, st) == hipSuccess &&
              hipStreamSynchronize(st) == hipSuccess;
    if (d_X) (void)hipFreeAsync(d_X, st);
    if (d_idx) (void)hipFreeAsync(d_idx, st);
    if (d_dist) (void)hipFreeAsync(d_dist, st);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (!ok) {
        fprintf(stderr, "tsne_knn_exact: device failure\n");
        return -1;
    }
    if (stats) {
        stats->survivors = S.survivors;
        stats->uncertain_rows = S.uncertain_rows;
        stats->overflow_rows = S.overflow_rows;
        stats->eps = S.eps;
        stats->device_ms = S.device_ms;
        stats->screen_ms = S.screen_ms;
        stats->screen_flops = S.screen_flops;
    }
    return 0;
}
