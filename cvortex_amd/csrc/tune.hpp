// tune.hpp -- launch-shape selection.  A shape comes from the autotuner's
// cache when a measurement exists for the problem's bucket, otherwise from
// default_cfg().
#pragma once

#include "kernels.hpp"

namespace cvtx {

enum Kind : int { KIND_P3D = 0, KIND_P2D = 1, KIND_F3D = 2, KIND_BATCH = 3 };

LaunchCfg choose_cfg(int kind, int reg, int mask, int acc, int n_src, int n_tgt, int n_cu);

// Record a measured best shape for the bucket of (n_src, n_tgt).
void record_cfg(int kind, int reg, int mask, int acc, int n_src, int n_tgt, int T, int n_split);

// The launch shape with T targets per lane and about n_split source
// partitions of whole 256-source tiles.
LaunchCfg launch_shape(int T, int n_split, int n_src);

// Shape-cache bucket: floor(log2) of the sizes.
int size_bucket(int n);

// Persistent multi-cloud launches (batched.hip): the cache key of a batch
// and its tuned shape.
struct BatchSignature {
  int reg, mask, acc;
  long long pairs;  // sum of the clouds' squared sizes
  int n_clouds;
};
struct BatchShape {
  int T;    // targets per lane
  int wg;   // workgroup size
  int wpc;  // resident workgroups per CU
  float ms;
};
bool batch_shape_lookup(const BatchSignature& g, BatchShape* out);
void batch_shape_record(const BatchSignature& g, const BatchShape& s);

}  // namespace cvtx
