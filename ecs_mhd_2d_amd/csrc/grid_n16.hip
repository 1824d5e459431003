// Kernels and drivers of the 16 x 16 grid.
#include "engine.h"

ECS_GRID(16)
