// api_internal.h -- entry points shared between the C ABI files of the library.
#pragma once
#include "../../include/tsne_h2pack.h"

namespace tsne {

// tsne_init_with_P with an explicit report of whether P now belongs to the workspace
// (set once the workspace exists, even if the initial H2 build then fails).  shapes, if
// not NULL: launch shapes that replace the tuning table's (fields out of range keep the
// table's choice; the TSNE_OVERLAP / TSNE_GRAPH environment variables still win).
tsne_workspace_t *init_with_P(int n_point, int pt_dim, int low_dim, const tsne_params_t *params,
                              sparse_matrix_t *P, bool *taken,
                              const tsne_tuning_t *shapes = nullptr);

}  // namespace tsne
