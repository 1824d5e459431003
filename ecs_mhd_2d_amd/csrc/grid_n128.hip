// Kernels and drivers of the 128 x 128 grid.
#include "engine.h"

ECS_GRID(128)
