// h2.h -- device H2 representation of the repulsive kernel K(x, y) = (1 + |x - y|^2)^-2:
// rebuild (Morton-sorted cluster tree, admissible / near pair lists, proxy-point
// interpolative decomposition per node) and the fused product that returns
//   s_i = sum_{j != i} K_ij   and   F_i = sum_{j != i} K_ij (y_i - y_j).
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/tsne_h2pack.h"

namespace tsne {

struct H2;

// n points of dimension dim (2 or 3); leaf_size is the maximum number of points per
// leaf; tol the relative tolerance of the interpolative decompositions.
H2 *h2_create(int n, int dim, int leaf_size, double tol, hipStream_t st);
void h2_destroy(H2 *h);

// A second stream (default priority) for work of the build that nothing before the
// product reads (the NEAR / W pair sorts): it runs there beside the ID.  The build joins
// it before returning.  The stream must outlive the H2 object.
void h2_set_aux_stream(H2 *h, hipStream_t st_aux);
// Mixed path: the ID's pivoted Gram-Schmidt runs in FP32 (when the ID tolerance, half of
// h2_tol, is >= 1e-3); R entries, the solve and the coefficients stay FP64.
void h2_set_id_mixed(H2 *h, int mixed);
// HIP events around each build's ID launches (their time of the last build in
// tsne_h2_stats.id_ms).
void h2_set_timing(H2 *h, int on);
// Launch shapes of later builds (none changes a result): tree levels queued between two
// host checks (1..8; <= 0 keeps the default) and the largest internal level whose IDs run
// in one launch of the widest size class (< 0 keeps the default).
void h2_set_tuning(H2 *h, int level_batch, int one_class_level);

// One-shot hook for the next h2_build: called (on the build's host thread) once the
// build's front -- root box, Morton keys, sort, sorted gather copy -- is queued on the
// build stream, so that work the caller queues on another stream from the hook overlaps
// the tree, ID and sorts rather than the front; a build that fails before its front
// still calls it before returning.  A non-zero return fails the build.
void h2_set_front_hook(H2 *h, int (*fn)(void *), void *arg);
// Rebuild from the float32 gather copy Yg (caller order).  Returns 0, or -1 on a
// device error or a non-finite coordinate (reported in *nonfinite).
int h2_build(H2 *h, const void *Yg, int *nonfinite);

// Product: s (n) and F (n x dim, interleaved), caller order.  mixed selects the
// BF16/MFMA far field.
int h2_product(H2 *h, const void *Yg, float *s, float *F, int mixed);

// Whether the last build produced a usable H2 structure (a tree too shallow for any
// admissible pair falls back to the direct sum, as the reference does).
int h2_usable(const H2 *h);

int h2_stats(H2 *h, tsne_h2_stats_t *st);

// Device array (n): the point at Morton position k of the last build, i.e. the tree order.
const int *h2_order(const H2 *h);

}  // namespace tsne
