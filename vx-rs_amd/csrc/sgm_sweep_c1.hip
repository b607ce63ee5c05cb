// Sweep kernels of lane count 32 (layout class 1).
#include "sgm_sweep.h"

VXH_DEFINE_SWEEP_LAUNCHER(1)
