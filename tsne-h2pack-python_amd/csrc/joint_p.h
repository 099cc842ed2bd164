// joint_p.h -- joint P on the device (see joint_p.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace tsne {

// Device CSR of the joint P (stream-ordered allocations owned by the caller).
struct JointP {
    int nnz = 0;
    int *rp = nullptr, *col = nullptr;
    double *val = nullptr;
    double total = 0.0;  // the normalisation sum(Pc + Pc^T)
};

// From the conditional P (device CSR with increasing columns per row, n rows):
// P = (Pc + Pc^T) / sum, with the host library's arithmetic reproduced exactly.
int joint_P_device(int n, int nnz, const int *d_rp, const int *d_col, const double *d_val,
                   JointP *out, hipStream_t st);

// numpy's float64 sum of a contiguous array (buffer blocks of 8192, pairwise inside).
double numpy_sum(const double *a, long long n);

}  // namespace tsne
