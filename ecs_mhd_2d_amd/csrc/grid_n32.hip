// Kernels and drivers of the 32 x 32 grid.
#include "engine.h"

ECS_GRID(32)
