// Krylov vector kernels: deterministic dot products and one modified
// Gram-Schmidt sweep of an Arnoldi step (linalg.py:65-85), FP64.
//
// Reductions are two-level and fixed-order (no floating-point atomics), so
// results are bitwise reproducible run to run: each block writes a partial
// sum, and the next kernel sums the partials in index order.  The sweep
// fuses "w -= h_j q_j" with the partial dot for q_{j+1} (or |w|^2 after the
// last column), so step k costs k + 3 launches and one read of each column.
#pragma once
#include <hip/hip_runtime.h>

namespace ecs {

constexpr int VNT = 256;      // threads per block
constexpr int VMAXG = 256;    // max blocks of a reduction

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum; the result is valid in every thread.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[VNT / 64];
  v = wave_sum(v);
  const int tid = threadIdx.x;
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < VNT / 64; ++w) s += red[w];
  return s;
}

__device__ __forceinline__ double sum_partials(const double* part, int G) {
  // every thread sums the same G values in the same order
  double s = 0.0;
  for (int g = 0; g < G; ++g) s += part[g];
  return s;
}

__device__ __forceinline__ void chunk(int n, int G, int& lo, int& hi) {
  const int ch = (n + G - 1) / G;
  lo = blockIdx.x * ch;
  hi = min(n, lo + ch);
}

__global__ __launch_bounds__(VNT) void dot_partial_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                          int n, double* __restrict__ part) {
  int lo, hi;
  chunk(n, gridDim.x, lo, hi);
  double s = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += VNT) s += a[i] * b[i];
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(VNT) void dot_final_kernel(const double* __restrict__ part, int G, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = sum_partials(part, G);
}

// h[j] = sum(part_in); w -= h[j] q_j; part_out = partial <w, q_{j+1}> (j < k) or <w, w> (j == k)
__global__ __launch_bounds__(VNT) void mgs_update_kernel(double* __restrict__ w, const double* __restrict__ Q, int ldq,
                                                         int n, int j, int k, double* __restrict__ h,
                                                         const double* __restrict__ part_in,
                                                         double* __restrict__ part_out) {
  const int G = gridDim.x;
  const double hj = sum_partials(part_in, G);
  if (blockIdx.x == 0 && threadIdx.x == 0) h[j] = hj;
  int lo, hi;
  chunk(n, G, lo, hi);
  const double* qj = Q + (size_t)j * ldq;
  const double* qn = (j < k) ? Q + (size_t)(j + 1) * ldq : nullptr;
  double s = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += VNT) {
    const double wi = w[i] - hj * qj[i];
    w[i] = wi;
    s += (j < k) ? wi * qn[i] : wi * wi;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part_out[blockIdx.x] = s;
}

// h[k+1] = |w|; Q[:, k+1] = w / h[k+1]
__global__ __launch_bounds__(VNT) void mgs_finish_kernel(const double* __restrict__ w, double* __restrict__ Q, int ldq,
                                                         int n, int k, double* __restrict__ h,
                                                         const double* __restrict__ part) {
  const int G = gridDim.x;
  const double nrm = sqrt(sum_partials(part, G));
  if (blockIdx.x == 0 && threadIdx.x == 0) h[k + 1] = nrm;
  int lo, hi;
  chunk(n, G, lo, hi);
  double* q = Q + (size_t)(k + 1) * ldq;
  for (int i = lo + threadIdx.x; i < hi; i += VNT) q[i] = w[i] / nrm;
}

}  // namespace ecs
