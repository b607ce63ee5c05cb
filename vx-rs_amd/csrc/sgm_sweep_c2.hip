// Sweep kernels of lane count 64 (layout class 2).
#include "sgm_sweep.h"

VXH_DEFINE_SWEEP_LAUNCHER(2)
