// joint_p.hip -- the joint P of t-SNE on the device: conditional P from the k-NN
// distances (perplexity bisection), then the symmetrisation P = Pc + Pc^T and the
// normalisation P /= sum(P), bit for bit what the reference computes on the host with
// scipy (tsne_h2pack/_backend.py:695-718):
//
//   * Pc + Pc.T: Pc.T is converted to CSR (rows in increasing column order) and the two
//     canonical matrices are merged row by row; entries present in both are added, an
//     entry whose result is exactly zero is dropped (scipy's csr_binop_csr_canonical).
//   * P.sum(): (P @ ones) -- each row summed sequentially in storage order starting from
//     0.0 -- then numpy's sum of the N row sums: blocks of 8192 (the reduction buffer),
//     each summed pairwise (8 running sums up to 128 elements, halves above), and the
//     block sums added in order.  The total is formed on the host from the row sums.
//   * P.data /= total: IEEE division per element.
//
// Device steps: expand the row of every nonzero, sort (column, row) keys (the
// transpose, rows sorted), count the merged rows, scan, merge and sum rows, divide.
#include <rocprim/device/device_radix_sort.hpp>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "common.h"
#include "joint_p.h"
#include "kernels.h"

namespace tsne {

namespace {

__global__ void __launch_bounds__(256)
row_of_kernel(int n, const int *__restrict__ rp, int *__restrict__ row)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int e = rp[i]; e < rp[i + 1]; e++) row[e] = i;
}

// key = column << rbits | row: sorting the keys orders the nonzeros by column, then row
__global__ void __launch_bounds__(256)
transpose_keys_kernel(int nnz, int rbits, const int *__restrict__ col, const int *__restrict__ row,
                      uint64_t *__restrict__ key, int *__restrict__ src)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nnz) return;
    key[e] = ((uint64_t)(uint32_t)col[e] << rbits) | (uint32_t)row[e];
    src[e] = e;
}

// rows of the transpose: rpT[c] = first sorted key whose column is >= c
__global__ void __launch_bounds__(256)
transpose_rows_kernel(int n, int nnz, int rbits, const uint64_t *__restrict__ key,
                      int *__restrict__ rpT)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e > nnz) return;
    const int c = (e < nnz) ? (int)(key[e] >> rbits) : n;
    const int cp = (e > 0) ? (int)(key[e - 1] >> rbits) : -1;
    for (int q = cp + 1; q <= c; q++) rpT[q] = e;
}

// the transpose's columns (original rows) and values
__global__ void __launch_bounds__(256)
transpose_fill_kernel(int nnz, int rbits, const uint64_t *__restrict__ key,
                      const int *__restrict__ src, const double *__restrict__ val,
                      int *__restrict__ colT, double *__restrict__ valT)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nnz) return;
    colT[e] = (int)(key[e] & ((1ull << rbits) - 1ull));
    valT[e] = val[src[e]];
}

// Merge of row i of A and of B (both with increasing columns): pass 0 counts the kept
// entries, pass 1 writes them and the row's sequential sum.
template <int PASS>
__global__ void __launch_bounds__(256)
merge_kernel(int n, const int *__restrict__ rpA, const int *__restrict__ colA,
             const double *__restrict__ valA, const int *__restrict__ rpB,
             const int *__restrict__ colB, const double *__restrict__ valB,
             const int *__restrict__ rpC, int *__restrict__ cnt, int *__restrict__ colC,
             double *__restrict__ valC, double *__restrict__ rowsum)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int a = rpA[i], b = rpB[i];
    const int ae = rpA[i + 1], be = rpB[i + 1];
    int k = (PASS == 1) ? rpC[i] : 0, m = 0;
    double s = 0.0;
    auto emit = [&](int c, double v) {
        if (v == 0.0) return;  // scipy drops results that are exactly zero
        if constexpr (PASS == 1) {
            colC[k] = c;
            valC[k] = v;
            k++;
            s += v;
        }
        m++;
    };
    while (a < ae && b < be) {
        const int ca = colA[a], cb = colB[b];
        if (ca == cb) {
            emit(ca, valA[a] + valB[b]);
            a++;
            b++;
        } else if (ca < cb) {
            emit(ca, valA[a] + 0.0);
            a++;
        } else {
            emit(cb, 0.0 + valB[b]);
            b++;
        }
    }
    for (; a < ae; a++) emit(colA[a], valA[a] + 0.0);
    for (; b < be; b++) emit(colB[b], 0.0 + valB[b]);
    if constexpr (PASS == 0) cnt[i] = m;
    if constexpr (PASS == 1) rowsum[i] = s;
}

__global__ void __launch_bounds__(256)
divide_kernel(int nnz, double total, double *__restrict__ v)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < nnz) v[e] = v[e] / total;
}

template <typename T> bool dmal(T **p, size_t n, hipStream_t st)
{
    return hipMallocAsync((void **)p, std::max<size_t>(n, 1) * sizeof(T), st) == hipSuccess;
}

template <typename T> void dfre(T *&p, hipStream_t st)
{
    if (p) (void)hipFreeAsync((void *)p, st);
    p = nullptr;
}

// numpy's pairwise sum of one block (at most 8192 elements)
double pairwise(const double *a, long long n)
{
    if (n < 8) {
        double r = -0.0;
        for (long long i = 0; i < n; i++) r += a[i];
        return r;
    }
    if (n <= 128) {
        double r[8];
        for (int j = 0; j < 8; j++) r[j] = a[j];
        long long i = 8;
        for (; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; j++) r[j] += a[i + j];
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; i++) res += a[i];
        return res;
    }
    long long n2 = n / 2;
    n2 -= n2 % 8;
    return pairwise(a, n2) + pairwise(a + n2, n - n2);
}

}  // namespace

double numpy_sum(const double *a, long long n)
{
    double r = 0.0;
    for (long long s = 0; s < n; s += 8192) r = r + pairwise(a + s, std::min<long long>(8192, n - s));
    return r;
}

int joint_P_device(int n, int nnz, const int *d_rp, const int *d_col, const double *d_val,
                   JointP *out, hipStream_t st)
{
    *out = JointP{};
    int rbits = 1;
    while (rbits < 31 && (1LL << rbits) < (long long)n) rbits++;
    const int cbits = rbits;  // columns are < n as well
    int *row = nullptr, *src = nullptr, *src2 = nullptr, *rpT = nullptr, *colT = nullptr,
        *cnt = nullptr;
    uint64_t *key = nullptr, *key2 = nullptr;
    double *valT = nullptr, *rowsum = nullptr;
    void *tmp = nullptr;
    size_t tb1 = 0, tb2 = 0;
    int rc = -1;
    std::vector<double> hs;
    {
        rocprim::double_buffer<uint64_t> kb(key, key2);
        rocprim::double_buffer<int> vb(src, src2);
        (void)rocprim::radix_sort_pairs(nullptr, tb1, kb, vb, std::max(nnz, 1), 0, rbits + cbits, st);
        (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, (const int *)nullptr, (int *)nullptr,
                                               n + 1, st);
    }
    const size_t tbytes = std::max(tb1, tb2);
    const int g = (n + 255) / 256, ge = (nnz + 255) / 256;
    if (!(dmal(&row, nnz, st) && dmal(&src, nnz, st) && dmal(&src2, nnz, st) &&
          dmal(&key, nnz, st) && dmal(&key2, nnz, st) && dmal(&rpT, (size_t)n + 1, st) &&
          dmal(&colT, nnz, st) && dmal(&valT, nnz, st) && dmal(&cnt, (size_t)n + 1, st) &&
          dmal(&rowsum, n, st) && hipMallocAsync(&tmp, std::max<size_t>(tbytes, 1), st) == hipSuccess &&
          dmal(&out->rp, (size_t)n + 1, st)))
        goto done;
    // transpose
    if (nnz > 0) {
        hipLaunchKernelGGL(row_of_kernel, dim3(g), dim3(256), 0, st, n, d_rp, row);
        hipLaunchKernelGGL(transpose_keys_kernel, dim3(ge), dim3(256), 0, st, nnz, rbits, d_col, row,
                           key, src);
        rocprim::double_buffer<uint64_t> kb(key, key2);
        rocprim::double_buffer<int> vb(src, src2);
        size_t bytes = tbytes;
        if (rocprim::radix_sort_pairs(tmp, bytes, kb, vb, nnz, 0, rbits + cbits, st) != hipSuccess)
            goto done;
        if (kb.current() != key) std::swap(key, key2);
        if (vb.current() != src) std::swap(src, src2);
    }
    hipLaunchKernelGGL(transpose_rows_kernel, dim3((nnz + 1 + 255) / 256), dim3(256), 0, st, n, nnz,
                       rbits, key, rpT);
    if (nnz > 0)
        hipLaunchKernelGGL(transpose_fill_kernel, dim3(ge), dim3(256), 0, st, nnz, rbits, key, src,
                           d_val, colT, valT);
    // merged rows: counts, offsets, entries and row sums
    hipLaunchKernelGGL(merge_kernel<0>, dim3(g), dim3(256), 0, st, n, d_rp, d_col, d_val, rpT, colT,
                       valT, nullptr, cnt, nullptr, nullptr, nullptr);
    if (hipMemsetAsync(cnt + n, 0, sizeof(int), st) != hipSuccess) goto done;
    {
        size_t bytes = tbytes;
        if (hipcub::DeviceScan::ExclusiveSum(tmp, bytes, cnt, out->rp, n + 1, st) != hipSuccess)
            goto done;
    }
    if (hipMemcpyAsync(&out->nnz, out->rp + n, sizeof(int), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        goto done;
    if (!dmal(&out->col, out->nnz, st) || !dmal(&out->val, out->nnz, st)) goto done;
    hipLaunchKernelGGL(merge_kernel<1>, dim3(g), dim3(256), 0, st, n, d_rp, d_col, d_val, rpT, colT,
                       valT, out->rp, nullptr, out->col, out->val, rowsum);
    hs.resize(n);
    if (hipMemcpyAsync(hs.data(), rowsum, sizeof(double) * n, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        goto done;
    out->total = numpy_sum(hs.data(), n);
    if (out->nnz > 0)
        hipLaunchKernelGGL(divide_kernel, dim3((out->nnz + 255) / 256), dim3(256), 0, st, out->nnz,
                           out->total, out->val);
    rc = hipGetLastError() == hipSuccess ? 0 : -1;
done:
    dfre(row, st); dfre(src, st); dfre(src2, st); dfre(key, st); dfre(key2, st); dfre(rpT, st);
    dfre(colT, st); dfre(valT, st); dfre(cnt, st); dfre(rowsum, st);
    if (tmp) (void)hipFreeAsync(tmp, st);
    if (rc != 0) {
        dfre(out->rp, st); dfre(out->col, st); dfre(out->val, st);
    }
    return rc;
}

}  // namespace tsne
