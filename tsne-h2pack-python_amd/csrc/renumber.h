// renumber.h -- point renumbering on the device (see renumber.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace tsne {

struct RenumberCsr {
    int n = 0, nnz = 0;
    int *rp = nullptr, *col = nullptr;
    float *v32 = nullptr;
    double *v64 = nullptr;
};

// Scratch: inv (n), len (n), src (nnz), src_sorted (nnz), col_tmp (nnz), tmp (tmp_bytes).
struct RenumberScratch {
    int *inv = nullptr, *len = nullptr, *src = nullptr, *src_sorted = nullptr, *col_tmp = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
};

size_t renumber_tmp_bytes(int n, int nnz);
// out = P renumbered so that new point k is old point ord[k]; columns sorted per row.
int renumber_csr(const RenumberCsr &in, const int *ord, RenumberCsr &out,
                 const RenumberScratch &s, hipStream_t st);
// dst[k] = src[ord[k]] for rows of `width` elements of elem_bytes (4 or 8).
int permute_points(int n, int width, const int *ord, const void *src, void *dst, int elem_bytes,
                   hipStream_t st);

}  // namespace tsne
