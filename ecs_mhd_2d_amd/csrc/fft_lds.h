// Batched complex FFT over sequences staged in LDS (FP64, gfx950).
//
// Stockham auto-sort with mixed radices 16/8/4/2 (plans listed at Plan).
// All threads of the workgroup cooperate on NSEQ sequences of length N;
// sequence s lives at buf + s * SSTRIDE (double2 units).  Each pass reads its
// butterflies' inputs into registers, synchronises, then writes, so it runs
// in place; output is in natural order and unnormalised:
//   forward  X[k] = sum_n x[n] exp(-2 pi i n k / N)
//   inverse  x[n] = sum_k X[k] exp(+2 pi i n k / N)
// Element i of sequence s lives at buf[s * SSTRIDE + pidx(i)].
// Twiddles come from per-pass LDS tables built from tw[j] = exp(-2 pi i j / N)
// (computed on the host from long-double sin/cos); radix-8/16 internal
// constants are the exact cos/sin of multiples of pi/8 to double precision.
#pragma once
#include <hip/hip_runtime.h>

// No multiply-add contraction in the device code of the library (this header
// comes first in every kernel translation unit): whether the compiler fuses
// a product into an FMA depends on the surrounding schedule, which differs
// between template instances (tile sizes), and would make a candidate's
// rounding depend on the batch size it runs at.  With contraction off every
// instance computes bit-identical results.
#pragma clang fp contract(off)

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

// LDS sequences are padded by one element per 16 (element i of a sequence
// lives at i + i/16), which keeps the strided butterfly writes of the
// radix-8/16 passes free of bank conflicts.  Seq<N>::S, the stride between
// sequences, is one element more than the padded length: an odd number of
// 16-byte elements, so lanes that touch the same element of neighbouring
// sequences (the column kernel's tiles of several columns) fall in
// different banks instead of all in the same ones.
template <int N>
struct Seq {
  static constexpr int S = N + N / 16 + 1;
};
__device__ __forceinline__ int pidx(int i) { return i + (i >> 4); }

// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ double2 rot_q(double2 a) {
  return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

constexpr double C8 = 0.70710678118654752440;   // cos(pi/4)
constexpr double C16 = 0.92387953251128675613;  // cos(pi/8)
constexpr double S16 = 0.38268343236508977173;  // sin(pi/8)

// cos(pi e / 8) and sin(pi e / 8) for e in [0, 16)
__host__ __device__ constexpr double cos16(int e) {
  return e == 0 ? 1.0 : e == 1 ? C16 : e == 2 ? C8 : e == 3 ? S16 : e == 4 ? 0.0 : e == 5 ? -S16 : e == 6 ? -C8
       : e == 7 ? -C16 : e == 8 ? -1.0 : e == 9 ? -C16 : e == 10 ? -C8 : e == 11 ? -S16 : e == 12 ? 0.0
       : e == 13 ? S16 : e == 14 ? C8 : C16;
}
__host__ __device__ constexpr double sin16(int e) { return cos16((e + 12) & 15); }

// multiply by W^E where W = exp(-2 pi i / 16) (forward) or its conjugate (inverse)
template <int E, bool INV>
__device__ __forceinline__ double2 w16(double2 a) {
  constexpr int e = E & 15;
  if constexpr (e == 0) return a;
  else if constexpr (e == 4) return rot_q<INV>(a);
  else if constexpr (e == 8) return make_double2(-a.x, -a.y);
  else {
    // W^e = cos(pi e/8) - i sin(pi e/8)
    constexpr double c = cos16(e);
    constexpr double sg = INV ? sin16(e) : -sin16(e);
    return make_double2(a.x * c - a.y * sg, a.x * sg + a.y * c);
  }
}

template <bool INV>
__device__ __forceinline__ void dft2(double2& a, double2& b) {
  const double2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(double2& v0, double2& v1, double2& v2, double2& v3) {
  const double2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3);
  const double2 a3 = rot_q<INV>(csub(v1, v3));
  v0 = cadd(a0, a2);
  v1 = cadd(a1, a3);
  v2 = csub(a0, a2);
  v3 = csub(a1, a3);
}

template <int R, bool INV> struct Dft;
template <bool INV> struct Dft<2, INV> {
  __device__ __forceinline__ static void run(double2* v) { dft2<INV>(v[0], v[1]); }
};
template <bool INV> struct Dft<4, INV> {
  __device__ __forceinline__ static void run(double2* v) { dft4<INV>(v[0], v[1], v[2], v[3]); }
};
template <bool INV> struct Dft<8, INV> {
  // X[k] = E[k] + W8^k O[k], X[k+4] = E[k] - W8^k O[k]
  __device__ __forceinline__ static void run(double2* v) {
    dft4<INV>(v[0], v[2], v[4], v[6]);
    dft4<INV>(v[1], v[3], v[5], v[7]);
    const double2 o1 = w16<2, INV>(v[3]), o2 = w16<4, INV>(v[5]), o3 = w16<6, INV>(v[7]);
    const double2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
  }
};
template <bool INV> struct Dft<16, INV> {
  // n = 4m + q: A_q = DFT4 over m, twiddle W16^{q k'}, then DFT4 over q:
  // X[k' + 4l] = sum_q W4^{q l} (W16^{q k'} A_q[k'])
  __device__ __forceinline__ static void run(double2* v) {
    dft4<INV>(v[0], v[4], v[8], v[12]);
    dft4<INV>(v[1], v[5], v[9], v[13]);
    dft4<INV>(v[2], v[6], v[10], v[14]);
    dft4<INV>(v[3], v[7], v[11], v[15]);
    // A_q[k'] sits at v[q + 4k']
    v[5] = w16<1, INV>(v[5]);  v[9] = w16<2, INV>(v[9]);   v[13] = w16<3, INV>(v[13]);
    v[6] = w16<2, INV>(v[6]);  v[10] = w16<4, INV>(v[10]); v[14] = w16<6, INV>(v[14]);
    v[7] = w16<3, INV>(v[7]);  v[11] = w16<6, INV>(v[11]); v[15] = w16<9, INV>(v[15]);
    dft4<INV>(v[0], v[1], v[2], v[3]);
    dft4<INV>(v[4], v[5], v[6], v[7]);
    dft4<INV>(v[8], v[9], v[10], v[11]);
    dft4<INV>(v[12], v[13], v[14], v[15]);
    // now X[k' + 4l] at v[4k' + l]: transpose the 4 x 4 block into natural order
    double2 t[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) t[i] = v[i];
#pragma unroll
    for (int kp = 0; kp < 4; ++kp)
#pragma unroll
      for (int l = 0; l < 4; ++l) v[kp + 4 * l] = t[4 * kp + l];
  }
};

// Radix plans, selected per kernel by the template argument P:
//   P = 16 (wide):   16 = 16, 32 = 8x4, 64 = 8x8, 128 = 16x8, 256 = 16x16
//   P = 8 (narrow):  16 = 4x4, 32 = 4x4x2, 64 = 4x4x4, 128 = 8x4x4, 256 = 8x8x4
// Wide plans take fewer passes (fewer LDS round trips and barriers); narrow
// ones keep more threads busy with shorter butterflies when a block holds few
// sequences.  The autotuner picks one per grid size and kernel family.
template <int N, int P> struct Plan;
template <> struct Plan<16, 16> { static constexpr int R[3] = {16, 1, 1}; };
template <> struct Plan<32, 16> { static constexpr int R[3] = {8, 4, 1}; };
template <> struct Plan<64, 16> { static constexpr int R[3] = {8, 8, 1}; };
template <> struct Plan<128, 16> { static constexpr int R[3] = {16, 8, 1}; };
template <> struct Plan<256, 16> { static constexpr int R[3] = {16, 16, 1}; };
template <> struct Plan<16, 8> { static constexpr int R[3] = {4, 4, 1}; };
template <> struct Plan<32, 8> { static constexpr int R[3] = {4, 4, 2}; };
template <> struct Plan<64, 8> { static constexpr int R[3] = {4, 4, 4}; };
template <> struct Plan<128, 8> { static constexpr int R[3] = {8, 4, 4}; };
template <> struct Plan<256, 8> { static constexpr int R[3] = {8, 8, 4}; };
// P = 9: the narrow plans reversed, the forward transform of the row pass
// whose first pass takes the values the inverse's last narrow pass leaves in
// registers (row_kernel with P = 8)
template <> struct Plan<16, 9> { static constexpr int R[3] = {4, 4, 1}; };
template <> struct Plan<32, 9> { static constexpr int R[3] = {2, 4, 4}; };
template <> struct Plan<64, 9> { static constexpr int R[3] = {4, 4, 4}; };
template <> struct Plan<128, 9> { static constexpr int R[3] = {4, 4, 8}; };
template <> struct Plan<256, 9> { static constexpr int R[3] = {4, 8, 8}; };

// Twiddles in LDS, one table per pass with NS > 1: entry (r - 1) * NS + k
// holds exp(-2 pi i r k / (NS R)) = tw[r k N / (NS R)], the factor of input r
// of a butterfly whose position in its sub-transform is k.  Lanes with
// consecutive k then read consecutive words (a table indexed r k N/(NS R)
// strides through the banks and conflicts 2-4 ways).  Pass 2 uses offset 0,
// pass 3 offset (R1 - 1) R0; the tables fit in N entries for every plan.
template <int N, int P>
struct TwLayout {
  static constexpr int R0 = Plan<N, P>::R[0], R1 = Plan<N, P>::R[1], R2 = Plan<N, P>::R[2];
  static constexpr int OFF2 = 0;
  static constexpr int OFF3 = (R1 > 1 ? (R1 - 1) * R0 : 0);
  static constexpr int SIZE = OFF3 + (R2 > 1 ? (R2 - 1) * R0 * R1 : 0);
  static_assert(SIZE <= N, "per-pass twiddle tables must fit the N-entry LDS table");
};

template <int N, int R, int NS, int NSEQ, int SSTRIDE, int NT, bool INV, int TWOFF>
__device__ __forceinline__ void stockham_pass(double2* __restrict__ buf, const double2* __restrict__ twl) {
  constexpr int NB = NSEQ * (N / R);
  constexpr int BPT = (NB + NT - 1) / NT;
  const int tid = threadIdx.x;
  double2 v[BPT][R];
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int bfly = tid + q * NT;
    if (bfly < NB) {
      const int s = bfly / (N / R);
      const int j = bfly - s * (N / R);
      const double2* src = buf + s * SSTRIDE;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = src[pidx(j + r * (N / R))];
      if (NS > 1) {
        const int k = j & (NS - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          double2 w = twl[TWOFF + (r - 1) * NS + k];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      Dft<R, INV>::run(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int bfly = tid + q * NT;
    if (bfly < NB) {
      const int s = bfly / (N / R);
      const int j = bfly - s * (N / R);
      double2* dst = buf + s * SSTRIDE;
      const int k = j & (NS - 1);
      const int d = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[pidx(d + r * NS)] = v[q][r];
    }
  }
  __syncthreads();
}

// twl must hold the plan's twiddle tables in LDS (TwFetch<N, NT, P>::store + a barrier).
template <int N, int NSEQ, int SSTRIDE, int NT, bool INV, int P = 16>
__device__ __forceinline__ void fft_lds(double2* __restrict__ buf, const double2* __restrict__ twl) {
  using TL = TwLayout<N, P>;
  constexpr int R0 = TL::R0, R1 = TL::R1, R2 = TL::R2;
  static_assert(R0 * R1 * R2 == N, "radix plan must factor N");
  stockham_pass<N, R0, 1, NSEQ, SSTRIDE, NT, INV, 0>(buf, twl);
  if constexpr (R1 > 1) stockham_pass<N, R1, R0, NSEQ, SSTRIDE, NT, INV, TL::OFF2>(buf, twl);
  if constexpr (R2 > 1) stockham_pass<N, R2, R0 * R1, NSEQ, SSTRIDE, NT, INV, TL::OFF3>(buf, twl);
}

// The same transform with the first pass reading its inputs through
// gather(s, i) (element i of sequence s, typically computed from global
// memory) instead of from LDS: the caller skips staging the inputs, and the
// first pass (NS = 1, no twiddles) writes the first LDS copy.  The twiddle
// tables must be in LDS by the barrier after the first pass: written before
// the call, or by before_barrier(), which runs once the first pass's inputs
// have arrived.  SEQFAST maps neighbouring lanes to
// neighbouring sequences at one butterfly (coalesced when the sequences are
// neighbouring columns of one array) instead of neighbouring butterflies of
// one sequence; the arithmetic is the same either way.
template <int N, int NSEQ, int SSTRIDE, int NT, bool INV, int P, bool SEQFAST = false, class F, class H>
__device__ __forceinline__ void fft_lds_gather(double2* __restrict__ buf, const double2* __restrict__ twl,
                                               F&& gather, H&& before_barrier) {
  using TL = TwLayout<N, P>;
  constexpr int R0 = TL::R0, R1 = TL::R1, R2 = TL::R2;
  static_assert(R0 * R1 * R2 == N, "radix plan must factor N");
  constexpr int NB = NSEQ * (N / R0);
  constexpr int BPT = (NB + NT - 1) / NT;
  const int tid = threadIdx.x;
  double2 v[BPT][R0];
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int bfly = tid + q * NT;
    if (bfly < NB) {
      const int s = SEQFAST ? bfly % NSEQ : bfly / (N / R0);
      const int j = SEQFAST ? bfly / NSEQ : bfly - s * (N / R0);
#pragma unroll
      for (int r = 0; r < R0; ++r) v[q][r] = gather(s, j + r * (N / R0));
    }
  }
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int bfly = tid + q * NT;
    if (bfly < NB) {
      const int s = SEQFAST ? bfly % NSEQ : bfly / (N / R0);
      const int j = SEQFAST ? bfly / NSEQ : bfly - s * (N / R0);
      Dft<R0, INV>::run(v[q]);
      double2* dst = buf + s * SSTRIDE;
#pragma unroll
      for (int r = 0; r < R0; ++r) dst[pidx(j * R0 + r)] = v[q][r];
    }
  }
  before_barrier();
  __syncthreads();
  if constexpr (R1 > 1) stockham_pass<N, R1, R0, NSEQ, SSTRIDE, NT, INV, TL::OFF2>(buf, twl);
  if constexpr (R2 > 1) stockham_pass<N, R2, R0 * R1, NSEQ, SSTRIDE, NT, INV, TL::OFF3>(buf, twl);
}

// The transform's last pass run in registers, its outputs handed to
// emit(s, i, value) (element i of sequence s) instead of written back to LDS.
// The last pass of a Stockham plan puts each output where its butterfly read
// its inputs (positions j + r N / R), so the arithmetic and the values are
// those of fft_lds.  Lanes are dealt GRP sequences fastest (sequences g GRP +
// 0 .. GRP-1 at one butterfly), then butterflies, then groups: with GRP
// neighbouring columns per group, neighbouring lanes emit neighbouring words.
// fft_lds_head<N, ...> must have run the earlier passes (its last barrier
// orders their writes; for a one-pass plan it runs none, and the caller's
// barrier after writing the inputs does); the caller orders later writes to
// buf after a barrier.
template <int N, int NSEQ, int SSTRIDE, int NT, bool INV, int P>
__device__ __forceinline__ void fft_lds_head(double2* __restrict__ buf, const double2* __restrict__ twl) {
  using TL = TwLayout<N, P>;
  constexpr int R0 = TL::R0, R1 = TL::R1, R2 = TL::R2;
  static_assert(R0 * R1 * R2 == N, "radix plan must factor N");
  if constexpr (R1 > 1) stockham_pass<N, R0, 1, NSEQ, SSTRIDE, NT, INV, 0>(buf, twl);
  if constexpr (R2 > 1) stockham_pass<N, R1, R0, NSEQ, SSTRIDE, NT, INV, TL::OFF2>(buf, twl);
}

template <int N, int NSEQ, int SSTRIDE, int NT, bool INV, int P, int GRP, class E>
__device__ __forceinline__ void fft_lds_tail(const double2* __restrict__ buf, const double2* __restrict__ twl,
                                             E&& emit) {
  using TL = TwLayout<N, P>;
  constexpr int RL = TL::R2 > 1 ? TL::R2 : (TL::R1 > 1 ? TL::R1 : TL::R0);
  constexpr int TWOFF = TL::R2 > 1 ? TL::OFF3 : TL::OFF2;
  constexpr int NSL = N / RL;  // the pass's NS: its butterflies' sub-transform length
  static_assert(NSEQ % GRP == 0, "sequence groups must tile the sequences");
  constexpr int NB = NSEQ * NSL;
  constexpr int BPT = (NB + NT - 1) / NT;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int bfly = tid + q * NT;
    if (bfly < NB) {
      const int g = bfly % GRP;
      const int rest = bfly / GRP;
      const int j = rest % NSL;
      const int s = (rest / NSL) * GRP + g;
      const double2* src = buf + s * SSTRIDE;
      double2 v[RL];
#pragma unroll
      for (int r = 0; r < RL; ++r) v[r] = src[pidx(j + r * NSL)];
      if (NSL > 1) {
#pragma unroll
        for (int r = 1; r < RL; ++r) {
          double2 w = twl[TWOFF + (r - 1) * NSL + j];
          if (INV) w.y = -w.y;
          v[r] = cmul(v[r], w);
        }
      }
      Dft<RL, INV>::run(v);
#pragma unroll
      for (int r = 0; r < RL; ++r) emit(s, j + r * NSL, v[r]);
    }
  }
}

// fft_lds_gather without the last pass: the gather pass and, for a
// three-pass plan, the middle pass.  The caller runs the last pass itself.
// SEQFAST as in fft_lds_gather.
template <int N, int NSEQ, int SSTRIDE, int NT, bool INV, int P, bool SEQFAST = false, class F, class H>
__device__ __forceinline__ void fft_lds_gather_head(double2* __restrict__ buf, const double2* __restrict__ twl,
                                                    F&& gather, H&& before_barrier) {
  using TL = TwLayout<N, P>;
  constexpr int R0 = TL::R0, R1 = TL::R1, R2 = TL::R2;
  static_assert(R1 > 1, "a head needs a plan of at least two passes");
  constexpr int NB = NSEQ * (N / R0);
  constexpr int BPT = (NB + NT - 1) / NT;
  const int tid = threadIdx.x;
  double2 v[BPT][R0];
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int bfly = tid + q * NT;
    if (bfly < NB) {
      const int s = SEQFAST ? bfly % NSEQ : bfly / (N / R0);
      const int j = SEQFAST ? bfly / NSEQ : bfly - s * (N / R0);
#pragma unroll
      for (int r = 0; r < R0; ++r) v[q][r] = gather(s, j + r * (N / R0));
    }
  }
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int bfly = tid + q * NT;
    if (bfly < NB) {
      const int s = SEQFAST ? bfly % NSEQ : bfly / (N / R0);
      const int j = SEQFAST ? bfly / NSEQ : bfly - s * (N / R0);
      Dft<R0, INV>::run(v[q]);
      double2* dst = buf + s * SSTRIDE;
#pragma unroll
      for (int r = 0; r < R0; ++r) dst[pidx(j * R0 + r)] = v[q][r];
    }
  }
  before_barrier();
  __syncthreads();
  if constexpr (R2 > 1) stockham_pass<N, R1, R0, NSEQ, SSTRIDE, NT, INV, TL::OFF2>(buf, twl);
}

template <int N, int NSEQ, int SSTRIDE, int NT, bool INV, int P, bool SEQFAST = false, class F>
__device__ __forceinline__ void fft_lds_gather(double2* __restrict__ buf, const double2* __restrict__ twl,
                                               F&& gather) {
  fft_lds_gather<N, NSEQ, SSTRIDE, NT, INV, P, SEQFAST>(buf, twl, gather, [] {});
}

// The per-pass twiddle tables of plan P, fetched into registers by fetch()
// (at most one load per thread and table, issued before anything else a
// kernel waits on) and written to LDS by store() once the first pass has consumed its
// inputs, so the table's memory latency overlaps the kernel's other loads
// instead of preceding them.  The barrier that follows the first pass orders
// the writes for the second.
template <int N, int NT, int P = 16>
struct TwFetch {
  using TL = TwLayout<N, P>;
  static constexpr int N2 = TL::R1 > 1 ? (TL::R1 - 1) * TL::R0 : 0;
  static constexpr int N3 = TL::R2 > 1 ? (TL::R2 - 1) * TL::R0 * TL::R1 : 0;
  static_assert(N2 <= NT && N3 <= NT, "one twiddle entry per thread and table");
  // plain doubles: a double2 member is copied with memcpy, which keeps the
  // object in scratch memory instead of registers
  double w2x, w2y, w3x, w3y;

  __device__ __forceinline__ void fetch(const double2* __restrict__ tw) {
    const int i = threadIdx.x;
    if (N2 > 0 && i < N2) {
      constexpr int NS = TL::R0, TSTEP = N / (NS * TL::R1);
      const int r = i / NS + 1, k = i - (r - 1) * NS;
      const double2 t = tw[(r * k * TSTEP) & (N - 1)];
      w2x = t.x;
      w2y = t.y;
    }
    if (N3 > 0 && i < N3) {
      constexpr int NS = TL::R0 * TL::R1, TSTEP = N / (NS * TL::R2);
      const int r = i / NS + 1, k = i - (r - 1) * NS;
      const double2 t = tw[(r * k * TSTEP) & (N - 1)];
      w3x = t.x;
      w3y = t.y;
    }
  }

  __device__ __forceinline__ void store(double2* __restrict__ twl) const {
    const int i = threadIdx.x;
    if (N2 > 0 && i < N2) twl[TL::OFF2 + i] = make_double2(w2x, w2y);
    if (N3 > 0 && i < N3) twl[TL::OFF3 + i] = make_double2(w3x, w3y);
  }
};

}  // namespace ecs
