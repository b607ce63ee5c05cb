// Aggregation kernels of lane-layout class 0 (D <= 64).
#include "sgm_paths_kernels.h"

VXH_DEFINE_PATHS_LAUNCHER(0)
