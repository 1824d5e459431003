// Batched complex FFT over sequences staged in LDS (FP64, gfx950).
//
// Stockham auto-sort, radix-4 passes plus one radix-2 pass when log2(N) is
// odd.  All threads of the workgroup cooperate on NSEQ sequences of length N;
// sequence s lives at buf + s * SSTRIDE (in double2 units).  Each pass reads
// every butterfly's inputs into registers, synchronises, then writes, so it
// runs in place.  Output is in natural order and unnormalised:
//   forward  X[k] = sum_n x[n] exp(-2 pi i n k / N)
//   inverse  x[n] = sum_k X[k] exp(+2 pi i n k / N)
// The twiddle table tw[j] = exp(-2 pi i j / N), j in [0, N), lives in global
// memory (built on the host from long-double sin/cos).
#pragma once
#include <hip/hip_runtime.h>

namespace ecs {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by i
__device__ __forceinline__ double2 cmuli(double2 a) { return make_double2(-a.y, a.x); }

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

template <int N, int NSEQ, int SSTRIDE, int NT, bool INV>
__device__ __forceinline__ void fft_lds(double2* __restrict__ buf, const double2* __restrict__ tw) {
  constexpr int LOG2N = ilog2c(N);
  constexpr int N4 = LOG2N / 2;           // number of radix-4 passes
  constexpr bool HAS2 = (LOG2N & 1) != 0;  // one trailing radix-2 pass
  const int tid = threadIdx.x;

  // ---------------- radix-4 passes ----------------
  {
    constexpr int NB = NSEQ * (N / 4);           // butterflies per pass
    constexpr int BPT = (NB + NT - 1) / NT;      // butterflies per thread
    int Ns = 1;
#pragma unroll
    for (int p = 0; p < N4; ++p) {
      double2 v[BPT][4];
#pragma unroll
      for (int q = 0; q < BPT; ++q) {
        const int bfly = tid + q * NT;
        if (bfly < NB) {
          const int s = bfly / (N / 4);
          const int j = bfly - s * (N / 4);
          const double2* src = buf + s * SSTRIDE;
          const int k = j & (Ns - 1);
          const int tstep = (N / 4) / Ns;  // = N / (Ns * 4)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            double2 x = src[j + r * (N / 4)];
            if (r > 0) {
              double2 w = tw[(r * k * tstep) & (N - 1)];
              if (INV) w.y = -w.y;
              x = cmul(x, w);
            }
            v[q][r] = x;
          }
          // radix-4 DFT
          double2 a0 = cadd(v[q][0], v[q][2]);
          double2 a1 = csub(v[q][0], v[q][2]);
          double2 a2 = cadd(v[q][1], v[q][3]);
          double2 a3 = csub(v[q][1], v[q][3]);
          a3 = INV ? make_double2(-a3.y, a3.x) : make_double2(a3.y, -a3.x);
          v[q][0] = cadd(a0, a2);
          v[q][1] = cadd(a1, a3);
          v[q][2] = csub(a0, a2);
          v[q][3] = csub(a1, a3);
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < BPT; ++q) {
        const int bfly = tid + q * NT;
        if (bfly < NB) {
          const int s = bfly / (N / 4);
          const int j = bfly - s * (N / 4);
          double2* dst = buf + s * SSTRIDE;
          const int k = j & (Ns - 1);
          const int d = (j - k) * 4 + k;
#pragma unroll
          for (int r = 0; r < 4; ++r) dst[d + r * Ns] = v[q][r];
        }
      }
      __syncthreads();
      Ns *= 4;
    }
  }
  // ---------------- radix-2 pass ----------------
  if constexpr (HAS2) {
    constexpr int NB = NSEQ * (N / 2);
    constexpr int BPT = (NB + NT - 1) / NT;
    constexpr int Ns = N / 2;
    double2 v[BPT][2];
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
      const int bfly = tid + q * NT;
      if (bfly < NB) {
        const int s = bfly / (N / 2);
        const int j = bfly - s * (N / 2);
        const double2* src = buf + s * SSTRIDE;
        double2 x0 = src[j];
        double2 x1 = src[j + N / 2];
        double2 w = tw[j];  // k = j (j < Ns), tstep = 1
        if (INV) w.y = -w.y;
        x1 = cmul(x1, w);
        v[q][0] = cadd(x0, x1);
        v[q][1] = csub(x0, x1);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
      const int bfly = tid + q * NT;
      if (bfly < NB) {
        const int s = bfly / (N / 2);
        const int j = bfly - s * (N / 2);
        double2* dst = buf + s * SSTRIDE;
        dst[j] = v[q][0];
        dst[j + Ns] = v[q][1];
      }
    }
    __syncthreads();
  }
}

}  // namespace ecs
