// Kernels and drivers of the 256 x 256 grid.
#include "engine.h"

ECS_GRID(256)
