// quality.hip -- embedding quality on the device: the reference's common-footing score
// (benchmarks/quality.py:56-81, score()).
//
//   kNN-preservation@k: for each query point, the share of its k high-dimensional
//     neighbours (the first k of the shared graph's rows, distance order) that are among
//     its k nearest neighbours in the embedding.  The reference takes the k+1 nearest
//     embedding points (sklearn KDTree, exact) and drops the first column (the point
//     itself); here the k+1 nearest are ranked exactly in FP64 by (d^2, index) and the
//     first is dropped the same way.
//   KL = sum p log p + sum p log1p(d^2) + log Z over the nonzeros of the common P, with
//     Z = n * mean over anchor rows a of sum_{j != a} 1 / (1 + |y_a - y_j|^2).
//
// One workgroup per query scans every point (FP64 differences); each thread keeps its 16
// nearest in registers (an unrolled insertion network, pairs ordered by (d^2, index)), and
// the workgroup merges the 256 sorted lists by repeated block-wide minimum.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "../../include/tsne_h2pack.h"

namespace tsne {

namespace {

constexpr int kTop = 16;  // neighbours kept per thread / largest k + 1

__device__ __forceinline__ bool pair_less(double da, int ia, double db, int ib)
{
    return da < db || (da == db && ia < ib);
}

__global__ void __launch_bounds__(256)
lowdim_knn_kernel(const double *__restrict__ Y, int n, int dim, const int *__restrict__ queries,
                  int k1, int *__restrict__ out)
{
    __shared__ double rd[256];
    __shared__ int ri[256];
    __shared__ int rt[256];
    const int t = threadIdx.x;
    const int q = queries[blockIdx.x];
    double bd[kTop];
    int bi[kTop];
#pragma unroll
    for (int j = 0; j < kTop; j++) {
        bd[j] = INFINITY;
        bi[j] = 0x7fffffff;
    }
    for (int c = t; c < n; c += 256) {
        double s = 0.0;
        for (int f = 0; f < dim; f++) {
            const double u = Y[(size_t)q * dim + f] - Y[(size_t)c * dim + f];
            s = fma(u, u, s);
        }
        if (pair_less(s, c, bd[kTop - 1], bi[kTop - 1])) {
            bd[kTop - 1] = s;
            bi[kTop - 1] = c;
#pragma unroll
            for (int j = kTop - 1; j > 0; j--)
                if (pair_less(bd[j], bi[j], bd[j - 1], bi[j - 1])) {
                    const double td = bd[j]; bd[j] = bd[j - 1]; bd[j - 1] = td;
                    const int ti = bi[j]; bi[j] = bi[j - 1]; bi[j - 1] = ti;
                }
        }
    }
    // k1 rounds: block minimum of the list heads; the owner pops its head
    for (int r = 0; r < k1; r++) {
        rd[t] = bd[0];
        ri[t] = bi[0];
        rt[t] = t;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (t < o && pair_less(rd[t + o], ri[t + o], rd[t], ri[t])) {
                rd[t] = rd[t + o];
                ri[t] = ri[t + o];
                rt[t] = rt[t + o];
            }
            __syncthreads();
        }
        const int win = rt[0];
        if (t == 0) out[(size_t)blockIdx.x * k1 + r] = ri[0];
        if (t == win) {
#pragma unroll
            for (int j = 0; j < kTop - 1; j++) {
                bd[j] = bd[j + 1];
                bi[j] = bi[j + 1];
            }
            bd[kTop - 1] = INFINITY;
            bi[kTop - 1] = 0x7fffffff;
        }
        __syncthreads();
    }
}

// |{hd[i, 0..k-1]} & {ld[r, 1..k]}| per query r (ld holds k+1 columns, the first dropped)
__global__ void __launch_bounds__(256)
overlap_kernel(const int *__restrict__ queries, int nq, const int *__restrict__ hd, int hd_k, int k,
               const int *__restrict__ ld, int *__restrict__ hits)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nq) return;
    const int i = queries[r];
    int h = 0;
    for (int a = 0; a < k; a++) {
        const int v = hd[(size_t)i * hd_k + a];
        bool in = false;
        for (int b = 1; b <= k; b++) in |= ld[(size_t)r * (k + 1) + b] == v;
        h += in;
    }
    hits[r] = h;
}

// sum_{j != a} 1 / (1 + d^2) for anchor rows (block per anchor, fixed reduction order)
__global__ void __launch_bounds__(256)
anchor_sum_kernel(const double *__restrict__ Y, int n, int dim, const int *__restrict__ anchors,
                  double *__restrict__ sums)
{
    __shared__ double red[256];
    const int a = anchors[blockIdx.x], t = threadIdx.x;
    double s = 0.0;
    for (int c = t; c < n; c += 256) {
        if (c == a) continue;
        double d2 = 0.0;
        for (int f = 0; f < dim; f++) {
            const double u = Y[(size_t)a * dim + f] - Y[(size_t)c * dim + f];
            d2 = fma(u, u, d2);
        }
        s += 1.0 / (1.0 + d2);
    }
    red[t] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) sums[blockIdx.x] = red[0];
}

// per row i: sum_j p log p and sum_j p log1p(d^2) (block partials of 256 rows)
__global__ void __launch_bounds__(256)
kl_terms_kernel(const double *__restrict__ Y, int n, int dim, const int *__restrict__ rp,
                const int *__restrict__ col, const double *__restrict__ val, double *__restrict__ part)
{
    __shared__ double r0[256], r1[256];
    const int t = threadIdx.x, i = blockIdx.x * 256 + t;
    double a = 0.0, b = 0.0;
    if (i < n) {
        for (int e = rp[i]; e < rp[i + 1]; e++) {
            const double p = val[e];
            const int j = col[e];
            double d2 = 0.0;
            for (int f = 0; f < dim; f++) {
                const double u = Y[(size_t)i * dim + f] - Y[(size_t)j * dim + f];
                d2 = fma(u, u, d2);
            }
            a += p * log(fmax(p, 1e-300));
            b += p * log1p(d2);
        }
    }
    r0[t] = a;
    r1[t] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
            r0[t] += r0[t + o];
            r1[t] += r1[t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
        part[2 * blockIdx.x] = r0[0];
        part[2 * blockIdx.x + 1] = r1[0];
    }
}

template <typename T> bool qalloc(T **p, size_t n, hipStream_t st)
{
    *p = nullptr;
    return hipMallocAsync((void **)p, std::max<size_t>(n, 1) * sizeof(T), st) == hipSuccess;
}

template <typename T> bool up(T *d, const T *h, size_t n, hipStream_t st)
{
    return n == 0 || hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, st) == hipSuccess;
}

}  // namespace

}  // namespace tsne

extern "C" int tsne_quality_score(const double *Y, int n, int dim, const int *hd_idx, int hd_k, int k,
                                  const int *queries, int n_queries, const int *P_row_ptr,
                                  const int *P_col_idx, const double *P_values, const int *anchors,
                                  int n_anchors, tsne_quality_t *out)
{
    using namespace tsne;
    if (!Y || !out || n < 2 || dim < 1 || k < 1 || k + 1 > kTop || k > hd_k || !hd_idx || !queries ||
        n_queries < 1 || !P_row_ptr || !anchors || n_anchors < 1 || k + 1 > n) {
        fprintf(stderr, "tsne_quality_score: invalid arguments (1 <= k <= 15, k <= hd_k)\n");
        return -1;
    }
    const long long nnz = P_row_ptr[n];
    bool valid = P_row_ptr[0] == 0 && nnz >= 0 && (nnz == 0 || (P_col_idx && P_values));
    for (int r = 0; valid && r < n_queries; r++) valid = queries[r] >= 0 && queries[r] < n;
    for (int r = 0; valid && r < n_anchors; r++) valid = anchors[r] >= 0 && anchors[r] < n;
    for (int i = 0; valid && i < n; i++) valid = P_row_ptr[i] <= P_row_ptr[i + 1];
    for (long long e = 0; valid && e < nnz; e++) valid = P_col_idx[e] >= 0 && P_col_idx[e] < n;
    for (long long e = 0; valid && e < (long long)n * hd_k; e++) valid = hd_idx[e] >= 0 && hd_idx[e] < n;
    for (long long e = 0; valid && e < (long long)n * dim; e++) valid = std::isfinite(Y[e]);
    if (!valid) {
        fprintf(stderr, "tsne_quality_score: index out of range, unordered row pointers or a "
                        "coordinate that is not finite\n");
        return -1;
    }
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        fprintf(stderr, "tsne_quality_score: no HIP device available\n");
        return -1;
    }
    const int nb = (n + 255) / 256;
    double *dY = nullptr, *dval = nullptr, *dsum = nullptr, *dpart = nullptr;
    int *dhd = nullptr, *dq = nullptr, *da = nullptr, *dld = nullptr, *dhits = nullptr, *drp = nullptr,
        *dcol = nullptr;
    std::vector<double> sums(n_anchors), part(2 * (size_t)nb);
    std::vector<int> hits(n_queries);
    bool ok = qalloc(&dY, (size_t)n * dim, st) && qalloc(&dhd, (size_t)n * hd_k, st) &&
              qalloc(&dq, n_queries, st) && qalloc(&da, n_anchors, st) &&
              qalloc(&dld, (size_t)n_queries * (k + 1), st) && qalloc(&dhits, n_queries, st) &&
              qalloc(&dsum, n_anchors, st) && qalloc(&dpart, 2 * (size_t)nb, st) &&
              qalloc(&drp, (size_t)n + 1, st) && qalloc(&dcol, nnz, st) && qalloc(&dval, nnz, st) &&
              up(dY, Y, (size_t)n * dim, st) && up(dhd, hd_idx, (size_t)n * hd_k, st) &&
              up(dq, queries, n_queries, st) && up(da, anchors, n_anchors, st) &&
              up(drp, P_row_ptr, (size_t)n + 1, st) && up(dcol, P_col_idx, nnz, st) &&
              up(dval, P_values, nnz, st);
    if (ok) {
        hipLaunchKernelGGL(lowdim_knn_kernel, dim3(n_queries), dim3(256), 0, st, dY, n, dim, dq, k + 1, dld);
        hipLaunchKernelGGL(overlap_kernel, dim3((n_queries + 255) / 256), dim3(256), 0, st, dq, n_queries, dhd,
                           hd_k, k, dld, dhits);
        hipLaunchKernelGGL(anchor_sum_kernel, dim3(n_anchors), dim3(256), 0, st, dY, n, dim, da, dsum);
        hipLaunchKernelGGL(kl_terms_kernel, dim3(nb), dim3(256), 0, st, dY, n, dim, drp, dcol, dval, dpart);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(hits.data(), dhits, sizeof(int) * n_queries, hipMemcpyDeviceToHost, st) ==
                 hipSuccess &&
             hipMemcpyAsync(sums.data(), dsum, sizeof(double) * n_anchors, hipMemcpyDeviceToHost, st) ==
                 hipSuccess &&
             hipMemcpyAsync(part.data(), dpart, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, st) ==
                 hipSuccess &&
             hipStreamSynchronize(st) == hipSuccess;
    }
    for (void *p : {(void *)dY, (void *)dval, (void *)dsum, (void *)dpart, (void *)dhd, (void *)dq, (void *)da,
                    (void *)dld, (void *)dhits, (void *)drp, (void *)dcol})
        if (p) (void)hipFreeAsync(p, st);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (!ok) {
        fprintf(stderr, "tsne_quality_score: device failure\n");
        return -1;
    }
    long long h = 0;
    for (int r = 0; r < n_queries; r++) h += hits[r];
    double s = 0.0, plogp = 0.0, plog1p = 0.0;
    for (int a = 0; a < n_anchors; a++) s += sums[a];
    for (int b = 0; b < nb; b++) {
        plogp += part[2 * b];
        plog1p += part[2 * b + 1];
    }
    const double Z = (double)n * (s / n_anchors);
    out->knn_preservation = ((double)h / n_queries) / k;  // mean of the counts, then / k
    out->Z = Z;
    out->kl = plogp + plog1p + std::log(Z);
    double span = 0.0;
    for (int f = 0; f < dim; f++) {
        double lo = Y[f], hi = Y[f];
        for (int i = 1; i < n; i++) {
            lo = std::min(lo, Y[(size_t)i * dim + f]);
            hi = std::max(hi, Y[(size_t)i * dim + f]);
        }
        span = std::max(span, hi - lo);
    }
    out->span = span;
    return 0;
}
