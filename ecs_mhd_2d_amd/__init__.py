"""MI355X-native periodic-orbit hot path of 2D MHD (mgolden30/ECS_MHD_2D, jax_scripts/lib).

Hand-written HIP kernels for gfx950 behind a C-ABI (include/ecsmhd.h); this
package mirrors the reference's Python call shapes on torch CUDA tensors.
"""
