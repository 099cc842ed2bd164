// tune.h -- launch-shape tuning table of the library (tune.cpp).
#pragma once
#include <string>

#include "../../include/tsne_h2pack.h"

namespace tsne {

// Size bucket of a job: round(log2 n_point).
int tune_bucket(int n_point);
// Built-in launch shapes of a job (no table entry).
tsne_tuning_t tune_defaults(int low_dim, int n_point);
// The table entry of the job's (dimension, bucket), if the device's table holds one.
bool tune_lookup(int low_dim, int n_point, tsne_tuning_t *out);
// Records an entry and rewrites the table file.
void tune_store(int low_dim, int n_point, const tsne_tuning_t &e);
// Path of the device's table file ("" when the cache is off or there is no device).
std::string tune_path();

}  // namespace tsne
