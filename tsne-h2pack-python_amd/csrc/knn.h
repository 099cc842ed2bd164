// knn.h -- exact Euclidean k-NN graph on the device (knn.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace tsne {

struct KnnStats {
    int survivors;       // L: candidates per point kept from the BF16 screening
    int uncertain_rows;  // rows re-ranked by the FP64 fallback
    int overflow_rows;   // of those, rows ranked by a full sort of their distances
    double eps;          // screening error bound relative to |x|^2 + |y|^2
    double device_ms;    // HIP events: first preparation kernel to the last kernel
    double screen_ms;    // HIP events around the screening launches (summed)
    long long screen_flops;  // BF16 MFMA flops of the screening (3 products per pair)
};

// survivors kept per point for k neighbours
int knn_survivors(int k);

// d_X: n x D float64 row-major on the device.  Writes n x k neighbour indices and float32
// distances (rows in increasing distance, ties by index, the point itself removed).
// 0 on success, -2 for invalid sizes, -1 on a device failure.  Synchronises st.
int knn_exact_device(const double *d_X, int n, int D, int k, int *d_idx, float *d_dist,
                     KnnStats *stats, hipStream_t st);

}  // namespace tsne
