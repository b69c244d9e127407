// tune.hpp -- launch-shape selection.  A shape comes from the autotuner's
// cache when a measurement exists for the problem's bucket, otherwise from
// default_cfg().
#pragma once

#include "kernels.hpp"

namespace cvtx {

enum Kind : int { KIND_P3D = 0, KIND_P2D = 1, KIND_F3D = 2 };

LaunchCfg choose_cfg(int kind, int reg, int mask, int acc, int n_src, int n_tgt, int n_cu);

// Record a measured best shape for the bucket of (n_src, n_tgt).
void record_cfg(int kind, int reg, int mask, int acc, int n_src, int n_tgt, int T, int n_split);

// The launch shape with T targets per lane and about n_split source
// partitions of whole 256-source tiles.
LaunchCfg launch_shape(int T, int n_split, int n_src);

// Shape-cache bucket: floor(log2) of the sizes.
int size_bucket(int n);

}  // namespace cvtx
