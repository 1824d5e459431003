// Recurrence diagram of a turbulent time series (generate_turbulence.py:140-151):
// dist[i][j] = || |fs_i| - |fs_j| ||_2 over all modes of the spectra, FP64.
//
// Each block computes one 16 x 16 tile (i-block <= j-block only; the mirror
// tile is written from the same sums, so dist is exactly symmetric).  The
// magnitudes of 16 snapshots x 32 modes per side are staged in LDS (rows
// padded to 33 doubles so the 16 column reads of a step fall in distinct
// banks), and each thread accumulates sum (a_i - a_j)^2 over the chunk.
// Differences are formed directly, never through |a|^2 + |b|^2 - 2 a.b,
// so near-recurrences (small distances) keep their relative accuracy.
#pragma once
#include <hip/hip_runtime.h>

namespace ecs {

constexpr int REC_TS = 16;   // tile edge (snapshots)
constexpr int REC_CH = 32;   // modes per LDS chunk

__global__ __launch_bounds__(256) void recurrence_kernel(const double2* __restrict__ fs, int S, long long M,
                                                         double* __restrict__ dist) {
  const int bi = blockIdx.y, bj = blockIdx.x;
  if (bj < bi) return;
  __shared__ double As[REC_TS][REC_CH + 1];
  __shared__ double Bs[REC_TS][REC_CH + 1];
  const int tid = threadIdx.x;
  const int ti = tid / REC_TS, tj = tid % REC_TS;
  const int i0 = bi * REC_TS, j0 = bj * REC_TS;
  double acc = 0.0;
  for (long long m0 = 0; m0 < M; m0 += REC_CH) {
    for (int idx = tid; idx < REC_TS * REC_CH; idx += 256) {
      const int r = idx / REC_CH, c = idx % REC_CH;
      const long long m = m0 + c;
      double a = 0.0, b = 0.0;
      if (m < M) {
        if (i0 + r < S) {
          const double2 z = fs[(long long)(i0 + r) * M + m];
          a = sqrt(z.x * z.x + z.y * z.y);
        }
        if (j0 + r < S) {
          const double2 z = fs[(long long)(j0 + r) * M + m];
          b = sqrt(z.x * z.x + z.y * z.y);
        }
      }
      As[r][c] = a;
      Bs[r][c] = b;
    }
    __syncthreads();
#pragma unroll 8
    for (int c = 0; c < REC_CH; ++c) {
      const double d = As[ti][c] - Bs[tj][c];
      acc += d * d;
    }
    __syncthreads();
  }
  const int i = i0 + ti, j = j0 + tj;
  if (i < S && j < S) {
    const double v = sqrt(acc);
    dist[(long long)i * S + j] = v;
    dist[(long long)j * S + i] = v;
  }
}

}  // namespace ecs
