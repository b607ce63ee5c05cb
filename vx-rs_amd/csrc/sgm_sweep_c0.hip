// Sweep kernels of lane count 16 (layout class 0).
#include "sgm_sweep.h"

VXH_DEFINE_SWEEP_LAUNCHER(0)
