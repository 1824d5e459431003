// Kernels and drivers of the 64 x 64 grid.
#include "engine.h"

ECS_GRID(64)
