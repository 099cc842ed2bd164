// h2.hip -- placeholder H2 module: reports no usable structure, so the product falls
// back to the direct sum.
#include "h2.h"
#include <new>

namespace tsne {

struct H2 {
    int n, dim;
};

H2 *h2_create(int n, int dim, int, double, hipStream_t) { return new (std::nothrow) H2{n, dim}; }
void h2_destroy(H2 *h) { delete h; }
int h2_build(H2 *, const void *, int *nonfinite) { *nonfinite = 0; return 0; }
int h2_product(H2 *, const void *, float *, float *, int) { return -1; }
int h2_usable(const H2 *) { return 0; }
int h2_stats(const H2 *, tsne_h2_stats_t *) { return -1; }

}  // namespace tsne
