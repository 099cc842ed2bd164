// renumber.hip -- renumbering of the points along the H2 tree order (device side).
//
// Reference: tsne_renumber_points (src/tsne_reorder.c:112-167): new point k is old point
// ord[k]; P is permuted by rows and its columns are relabelled (the reference does it
// with two CSR transposes, which leaves every row's columns sorted, src/tsne_reorder.c:
// 28-108); every per-point array is permuted; a cumulative order maps the current
// numbering to the caller's, and results are reported in the caller's order.  Here the
// rows are copied with relabelled columns and each row is then sorted by column with a
// segmented radix sort, which gives the same sorted-column CSR.
#include <hipcub/hipcub.hpp>

#include "common.h"
#include "renumber.h"

namespace tsne {

__global__ void __launch_bounds__(256)
invert_perm_kernel(int n, const int *__restrict__ ord, int *__restrict__ inv)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) inv[ord[k]] = k;
}

__global__ void __launch_bounds__(256)
row_len_kernel(int n, const int *__restrict__ ord, const int *__restrict__ rp,
               int *__restrict__ len)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) len[k] = rp[ord[k] + 1] - rp[ord[k]];
}

// One wavefront per new row: copy the old row's entries with relabelled columns and
// remember where each came from.
__global__ void __launch_bounds__(256)
copy_rows_kernel(int n, const int *__restrict__ ord, const int *__restrict__ inv,
                 const int *__restrict__ rp, const int *__restrict__ col,
                 const int *__restrict__ rp_new, int *__restrict__ col_new,
                 int *__restrict__ src_new)
{
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= n) return;
    const int i = ord[k];
    const int b = rp[i], e = rp[i + 1], o = rp_new[k];
    for (int q = b + lane; q < e; q += 64) {
        col_new[o + (q - b)] = inv[col[q]];
        src_new[o + (q - b)] = q;
    }
}

__global__ void __launch_bounds__(256)
gather_values_kernel(int nnz, const int *__restrict__ src, const float *__restrict__ v32,
                     const double *__restrict__ v64, float *__restrict__ v32_new,
                     double *__restrict__ v64_new)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nnz) return;
    const int s = src[e];
    v32_new[e] = v32[s];
    v64_new[e] = v64[s];
}

template <typename T>
__global__ void __launch_bounds__(256)
permute_rows_kernel(int n, int width, const int *__restrict__ ord, const T *__restrict__ src,
                    T *__restrict__ dst)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const T *s = src + (size_t)ord[k] * width;
    T *d = dst + (size_t)k * width;
    for (int w = 0; w < width; w++) d[w] = s[w];
}

int renumber_csr(const RenumberCsr &in, const int *ord, RenumberCsr &out,
                 const RenumberScratch &sc, hipStream_t st)
{
    const int n = in.n, nnz = in.nnz;
    const int g = (n + 255) / 256;
    hipLaunchKernelGGL(invert_perm_kernel, dim3(g), dim3(256), 0, st, n, ord, sc.inv);
    hipLaunchKernelGGL(row_len_kernel, dim3(g), dim3(256), 0, st, n, ord, in.rp, sc.len);
    size_t bytes = sc.tmp_bytes;
    if (hipMemsetAsync(out.rp, 0, sizeof(int), st) != hipSuccess) return -1;
    if (hipcub::DeviceScan::InclusiveSum(sc.tmp, bytes, sc.len, out.rp + 1, n, st) != hipSuccess)
        return -1;
    hipLaunchKernelGGL(copy_rows_kernel, dim3((n + 3) / 4), dim3(256), 0, st, n, ord, sc.inv,
                       in.rp, in.col, out.rp, sc.col_tmp, sc.src);
    // sort each row by its new column index; the source positions ride along
    bytes = sc.tmp_bytes;
    if (hipcub::DeviceSegmentedRadixSort::SortPairs(sc.tmp, bytes, sc.col_tmp, out.col, sc.src,
                                                    sc.src_sorted, nnz, n, out.rp, out.rp + 1, 0,
                                                    32, st) != hipSuccess)
        return -1;
    hipLaunchKernelGGL(gather_values_kernel, dim3((nnz + 255) / 256), dim3(256), 0, st, nnz,
                       sc.src_sorted, in.v32, in.v64, out.v32, out.v64);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

size_t renumber_tmp_bytes(int n, int nnz)
{
    size_t a = 0, b = 0;
    (void)hipcub::DeviceScan::InclusiveSum(nullptr, a, (int *)nullptr, (int *)nullptr, n);
    (void)hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, b, (int *)nullptr, (int *)nullptr,
                                                      (int *)nullptr, (int *)nullptr, nnz, n,
                                                      (int *)nullptr, (int *)nullptr, 0, 32);
    return a > b ? a : b;
}

int permute_points(int n, int width, const int *ord, const void *src, void *dst, int elem_bytes,
                   hipStream_t st)
{
    const int g = (n + 255) / 256;
    if (elem_bytes == 8)
        hipLaunchKernelGGL(permute_rows_kernel<double>, dim3(g), dim3(256), 0, st, n, width, ord,
                           (const double *)src, (double *)dst);
    else if (elem_bytes == 4)
        hipLaunchKernelGGL(permute_rows_kernel<float>, dim3(g), dim3(256), 0, st, n, width, ord,
                           (const float *)src, (float *)dst);
    else
        return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace tsne
