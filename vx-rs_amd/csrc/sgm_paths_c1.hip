// Aggregation kernels of lane-layout class 1 (D <= 128).
#include "sgm_paths_kernels.h"

VXH_DEFINE_PATHS_LAUNCHER(1)
