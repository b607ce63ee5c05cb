// Aggregation kernels of lane-layout class 2 (D <= 256).
#include "sgm_paths_kernels.h"

VXH_DEFINE_PATHS_LAUNCHER(2)
