// Fused pseudo-spectral MHD kernels for gfx950 (FP64).
//
// Data layouts (all double2 = complex128, row-major):
//   compact spectral state  S[b][c][r][ky]   c in {omega, j}, r in [0,R), ky in [0,KC)
//       kept rows r <= KMAX hold kx = r, rows r > KMAX hold kx = r - R (negative);
//       KMAX = (N-1)/3, R = 2 KMAX + 1, KC = KMAX + 1.  These are exactly the
//       modes of the 2/3 mask |kx|,|ky| < N/3 (mhd_jax.py:65-70); every
//       derivative and the RHS vanish outside them (kx, ky, to_u, to_v are
//       pre-masked, mhd_jax.py:87-90).
//   column-pass output I[b][g][x][ky]  x in [0,N): 8 derived fields (16 with a
//       tangent) after the inverse x-FFT, g = set*8 + 4c + d with
//       d = 0: i kx f, 1: i ky f, 2: i to_u f, 3: i to_v f.
//   row-pass output    O[b][o][x][ky]  o = set*2 + c: the y-r2c of the two
//       real-space nonlinear terms (omega tendency, (u x B)_z).
//
// One RHS evaluation = one column kernel (x-FFT of the previous row output,
// Lawson stage glue, derived spectra, inverse x-FFT) + one row kernel
// (Hermitian c2r along y of 4 (8) packed pairs, products, r2c along y).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fft_lds.h"

namespace ecs {

template <int N>
struct Geo {
  static constexpr int KMAX = (N - 1) / 3;
  static constexpr int KC = KMAX + 1;
  static constexpr int R = 2 * KMAX + 1;
  static constexpr int NH = N / 2 + 1;
  // row stride of the x-space arrays [..][N][KP] the column and row passes
  // exchange: KC padded to a multiple of 8, so every row starts on a 128-byte
  // line and a tile of 8 columns writes whole lines
  static constexpr int KP = (KC + 7) & ~7;
  __device__ __host__ static constexpr int row_kx(int r) { return r <= KMAX ? r : r - R; }
  __device__ __host__ static constexpr int row_x(int r) { return r <= KMAX ? r : r - R + N; }
};

// Logical block coordinates of the launch grid, remapped for the 8 XCDs of
// an MI355X: the dispatcher deals workgroups to the XCDs round-robin by linear
// id (x fastest), so linear id L runs on XCD L % 8.  The logical tiles are
// dealt to the XCDs in contiguous chunks instead, so tiles with consecutive
// logical x (neighbouring columns or rows, which share cache lines) run on one
// XCD at about the same time and its L2 serves the shared lines once.  A
// bijection of the grid onto itself: every tile runs exactly once.
__device__ __forceinline__ void xcd_tile(int& lx, int& ly, int& lz) {
  const int gx = (int)gridDim.x, gy = (int)gridDim.y;
  const int total = gx * gy * (int)gridDim.z;
  const int L = (int)blockIdx.x + gx * ((int)blockIdx.y + gy * (int)blockIdx.z);
  const int q = total / 8, r = total % 8;
  const int x = L % 8, loc = L / 8;
  const int T = x < r ? x * (q + 1) + loc : r * (q + 1) + (x - r) * q + loc;
  lx = T % gx;
  const int t = T / gx;
  ly = t % gy;
  lz = t / gy;
}

// Per-attempt record of the adaptive controller, one per candidate and attempt.
// Entry i describes attempt i; entry i+1 also records whether attempt i was
// accepted (prev_acc), because that decision is taken by the first kernel of
// attempt i+1.
struct Hist {
  double s;                 // pseudo-time at the start of the attempt
  double h;                 // step tried by the attempt
  unsigned long long err;   // bit pattern of max|err| (non-negative double, atomicMax)
  int active;               // attempt performed (integration not yet complete)
  int prev_acc;             // attempt i-1 was accepted
  int nacc, nrej;           // counters before this attempt
};

// One stage combination out = sum_t coef_t h^hp_t E^ep_t in_t (with its
// tangent), E = exp(h L / ediv); read from device memory by the column
// kernels' glue (C_GEN, C_TD2, td_col_b; each takes the term whose input is its
// own output pointer from registers).  C_GEN reads a null h as the launch's Args::dt.
constexpr int COMB_MAX = 10;

struct Comb {
  const double2* in[COMB_MAX];   // null: the term has no primal part (tangent-only)
  const double2* din[COMB_MAX];  // tangent inputs (null: zero)
  double coef[COMB_MAX];
  int hp[COMB_MAX];              // power of h (0, 1 or 2)
  int ep[COMB_MAX];              // power of E (0..6)
  int nterm;
  double ediv;
  int masked;
  const double* h;               // [B]
  const double* dh;              // [B] (tangent)
  double nu[2];
  double2* out;
  double2* dout;
  size_t elems;                  // compact elements per candidate
};

// Column-kernel modes.
enum ColMode : int {
  C_LOAD = 0,   // y = x0
  C_RHS = 1,    // out = v(x0) [+ L x0]; no inverse pass
  C_S1F = 2,    // fixed step: v1 = v(x0) from the row output; stage-1 glue -> y2
  C_S2 = 3,     // k2 -> y3
  C_S3 = 4,     // k3 -> y4
  C_S4 = 5,     // k4 -> xf (fixed step: x0 = xf) -> y = xf
  C_S5 = 6,     // adaptive: k5, err spectrum (tangent replay: accept xf)
  C_A0 = 7,     // adaptive, first attempt: v1 = v(x0) from the row output; stage-1 glue
  C_A = 8,      // adaptive: apply previous decision, new h, stage-1 glue from stored v1
  C_FIN = 9,    // adaptive: apply the last decision only
  C_S5A = 10,   // tangent replay: accept this attempt (C_S5) and run the next step's C_A in the same kernel
  C_GEN = 11,   // generic explicit stepper: out = v(x0) [+ L x0], then the stage combination *gc -> xf -> y
  // two-derivative RK4 (primal): C_TD1 (no tangent input) stores out = v(x0) and transforms the
  // pair (x0, v) as a tangent pass's input; C_TD2 (tangent input) stores dout = J v (the tangent
  // output) and forms the next stage input (or the step's result) from *gc
  C_TD1 = 12,
  C_TD2 = 13,
};

// Row-kernel modes.
enum RowMode : int { R_RHS = 0, R_ERR = 1 };

// adaptive flag values
enum AdaptKind : int { AD_FIXED = 0, AD_PRIMAL = 1, AD_TANGENT = 2 };

struct Args {
  const double2* tw;       // twiddles [N]
  const double2* fhat;     // masked forcing spectrum, compact [R][KC]
  double nu[2];
  double b0x, b0y;
  double2* rowout;         // O
  double2* inter;          // I
  // compact state [B][2][R][KC]
  double2* x0;
  double2* acc;
  double2* v1;
  double2* v4;
  double2* v5;
  double2* xf;
  double2* out;            // C_RHS / C_GEN output
  const Comb* gc;          // C_GEN: stage combination giving the next stage input (device memory)
  // tangents
  double2* dx0;
  double2* dacc;
  double2* dv1;
  double2* dv4;
  double2* dxf;
  double2* dout;
  // per-candidate scalars
  const double* dt;        // fixed step: dt[b]; tangent: ddt[b]
  const double* ddt;
  const double* Tb;        // adaptive: T[b]
  const double* dTb;       // adaptive tangent: dT[b]
  double atol;
  Hist* hist;              // [B][hstride]
  int hstride;
  int attempt;             // adaptive attempt index (primal) / accepted step (tangent)
  // graph-replayable index: when cnt is set, the index is cnt[cslot] and the
  // first kernel of an attempt writes cnt[cslot ^ 1] = index + 1 (a captured
  // chunk of an even number of attempts then replays unchanged)
  int* cnt;
  int cslot;
  int budget;              // adaptive primal: attempts at or beyond it only apply the last decision
  // tangent replay schedule of the adaptive stepper, [B][sstride]
  const double* s_h;       // step of the accepted attempt
  const int* s_rs;         // start of its preceding rejections in s_rh
  const int* s_rc;         // number of preceding rejections
  const double* s_rh;      // rejected steps, [B][rstride]
  const int* s_nacc;       // number of accepted attempts per candidate
  int sstride, rstride;
  double h0;               // adaptive: first step (pseudo-time)
  // tape mode of the forward (checkpointed reverse pass): when non-null the
  // stage glue stores v(y_s) at vtape[ci] and the step's x0 at xtape[ci]
  double2* vtape;
  double2* xtape;
  // reverse pass (adjoint kernels)
  double2* xb;             // state cotangent
  double2* k1b;            // stage-increment cotangents
  double2* k2b;
  double2* k3b;
  double* hb;              // per-mode accumulator of d/dh, compact real [B][2][R][KC]
  const double2* tx;       // tape: x0 of the current step
  const double2* tv;       // tape: v1..v4 of the current step, stage stride cstride
  const double2* txp;      // tape of the previous step (fused entry)
  const double2* tvp;
  size_t cstride;          // elements of one compact batch array
  double2* aout;           // adjoint column output AO [B][2][N][KP]
  double2* ain;            // adjoint row output AI [B][8][N][KP]
  const double2* prim;     // primal column output for the row-adjoint [B][8][N][KP]
  const double2* gin;      // cotangent input of a standalone RHS adjoint
  // C_GEN, primal: when pdt is set, the stage RHS is stored raw at vraw and
  // scaled by exp(pdt[b] L) before it is stored at out and combined (the
  // RK4(3) reverse pass's carried-k1 factor on stage 1)
  const double* pdt;
  double2* vraw;
  // fixed step with per-candidate step counts: candidate b takes part in the
  // launches of global step `step` only while step < nstep[b] (null: all do)
  const int* nstep;
  int step;
  // row passes: candidates b >= b0_from see b0 = 0 (the second-derivative
  // pass of the two-derivative RK4 tangent, batched with the other passes)
  int b0_from;
  // flags, one byte each (the struct is every kernel's argument block; it is
  // kept under 512 bytes)
  int8_t include_diss;
  int8_t adaptive;         // AdaptKind
  int8_t masked_e;         // fixed step: e = exp(dt/2 L) * mask
  int8_t skip_inv;         // 1: do not produce the inverse pass (last stage of a run)
  // tangent replay with the index as an argument: 1 when every candidate runs
  // accepted step `attempt` (resp. has a step after it), so the kernels need
  // not read s_nacc before their gathers
  int8_t run_all;
  int8_t next_all;
  int8_t out_acc;          // standalone RHS adjoint: add its result to out instead of storing it
  // two-derivative RK4 tangent: J(z) v, J(z) dv and H[dz, v] of the half step
  double2* tdw;
  double2* tdjdv;
  double2* tdh;
};
// Every kernel takes Args by value.  Above 512 bytes the launches of the
// latency-bound batch-1 kernels measured 1.5-3.5% slower (520 against 496
// bytes: GMRES matvecs 14.57 against 14.78/s, Lawson RK4 at N=128 22.8k
// against 23.6k steps/s, alternating runs on one MI355X).
static_assert(sizeof(Args) <= 512, "kernel argument block above 512 bytes");

}  // namespace ecs
