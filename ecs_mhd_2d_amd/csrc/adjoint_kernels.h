// Reverse-mode (adjoint) kernels of the fused RHS and of the Lawson RK4 step.
//
// Cotangent convention: a spectral cotangent holds d/dRe in its real part and
// d/dIm in its imaginary part (the real inner product on stored entries).
// The adjoint RHS  g -> J(y)^T g  is the transpose of the forward chain
//   f -> (i m_d / N^2) f -> inverse x-FFT -> c2r along y -> products
//     -> r2c along y -> x-FFT -> mask (omega) / k^2 mask (j),
// i.e. (column-adjoint) m g -> inverse x-FFT -> (row-adjoint) weighted c2r
// along y -> transposed products -> weighted r2c along y -> (column-adjoint)
// x-FFT -> sum_d (-i m_d / N^2) G_d.
//
// One Lawson RK4 step  y2 = e x0 + e k1/2, y3 = e x0 + k2/2, y4 = e e x0 + e k3,
// x' = e e x0 + e e k1/6 + e k2/3 + e k3/3 + k4/6,  k_s = dt v(y_s),
// is reversed stage by stage; d/d(dt) collects <v_s, kbar_s> and the e terms
// (de/d(dt) = L/2 e) per mode into hb, summed once at the end.
#pragma once
#include "rhs_kernels.h"
#include "comb_kernels.h"

namespace ecs {

enum AdjMode : int {
  B_G = 0,      // standalone: gbar input -> column-adjoint output
  B_OUT = 1,    // standalone: ybar -> out (+ L g)
  B_ENTRY = 2,  // entry of the last step of a segment (xb = cotangent of its output)
  B_S4 = 3,     // ybar4 -> ..., next g = dt k3bar
  B_S3 = 4,
  B_S2 = 5,
  B_S1E = 6,    // ybar1 -> xb; fused entry of the previous step
  B_S1 = 7,     // ybar1 -> xb; no previous step in this segment
  B_OUTG = 8,   // B_OUT, then (after the fused combinations) the next pass's B_G
};

template <int N, int MODE>
__device__ __forceinline__ void adj_glue(const Args& a, int b, size_t ci, int r, int c, int kx, int ky, double2 yb,
                                         double2& g) {
  const double k2 = (double)(kx * kx + ky * ky);
  const bool dc = (kx == 0 && ky == 0);
  const double L = -k2 * a.nu[c];
  g = make_double2(0.0, 0.0);
  if (MODE == B_G) {
    g = a.gin[ci];
    return;
  }
  if (MODE == B_OUT) {
    double2 o = yb;
    if (a.include_diss) o = cadd(o, cscale(L, a.gin[ci]));
    // added as the combination kernel adds a term of weight 1; g returns the
    // value stored (the input of a combination fused after it)
    g = a.out_acc ? cadd(a.out[ci], cscale(1.0, o)) : o;
    a.out[ci] = g;
    return;
  }
  const double dt = a.dt[b];
  double e = exp((dt * 0.5) * L);
  if (a.masked_e && dc) e = 0.0;
  const double dedh = 0.5 * L * e;  // de/d(dt)
  const size_t cs = a.cstride;
  double2 x0 = make_double2(0.0, 0.0);
  double hacc = 0.0;
  if (MODE != B_ENTRY) x0 = a.tx[ci];
  if (MODE == B_S4) {
    const double2 k3 = cscale(dt, a.tv[2 * cs + ci]);
    a.xb[ci] = cadd(a.xb[ci], cscale(e * e, yb));
    const double2 k3b = cadd(a.k3b[ci], cscale(e, yb));
    const double2 dy = cadd(cscale(2.0 * e, x0), k3);
    hacc += (yb.x * dy.x + yb.y * dy.y) * dedh;
    const double2 v3 = a.tv[2 * cs + ci];
    hacc += v3.x * k3b.x + v3.y * k3b.y;
    g = cscale(dt, k3b);
  }
  if (MODE == B_S3) {
    a.xb[ci] = cadd(a.xb[ci], cscale(e, yb));
    const double2 k2b = cadd(a.k2b[ci], cdivs(yb, 2.0));
    hacc += (yb.x * x0.x + yb.y * x0.y) * dedh;
    const double2 v2 = a.tv[1 * cs + ci];
    hacc += v2.x * k2b.x + v2.y * k2b.y;
    g = cscale(dt, k2b);
  }
  if (MODE == B_S2) {
    const double2 k1 = cscale(dt, a.tv[ci]);
    a.xb[ci] = cadd(a.xb[ci], cscale(e, yb));
    const double2 k1b = cadd(a.k1b[ci], cdivs(cscale(e, yb), 2.0));
    const double2 dy = cadd(x0, cdivs(k1, 2.0));
    hacc += (yb.x * dy.x + yb.y * dy.y) * dedh;
    const double2 v1 = a.tv[ci];
    hacc += v1.x * k1b.x + v1.y * k1b.y;
    g = cscale(dt, k1b);
  }
  if (MODE == B_S1 || MODE == B_S1E) {
    a.xb[ci] = cadd(a.xb[ci], yb);
  }
  if (MODE != B_ENTRY) a.hb[ci] += hacc;
  if (MODE == B_ENTRY || MODE == B_S1E) {
    // entry of a step: xbar' known; tape of that step
    const double2* tx = (MODE == B_ENTRY) ? a.tx : a.txp;
    const double2* tv = (MODE == B_ENTRY) ? a.tv : a.tvp;
    const double2 xbp = a.xb[ci];
    const double2 x0s = tx[ci];
    const double2 k1 = cscale(dt, tv[ci]);
    const double2 k2 = cscale(dt, tv[cs + ci]);
    const double2 k3 = cscale(dt, tv[2 * cs + ci]);
    const double2 v4 = tv[3 * cs + ci];
    const double2 k4b = cdivs(xbp, 6.0);
    double h = v4.x * k4b.x + v4.y * k4b.y;
    a.xb[ci] = cscale(e * e, xbp);
    a.k1b[ci] = cdivs(cscale(e * e, xbp), 6.0);
    a.k2b[ci] = cdivs(cscale(e, xbp), 3.0);
    a.k3b[ci] = cdivs(cscale(e, xbp), 3.0);
    // d x'/de = 2 e x0 + 2 e k1/6 + k2/3 + k3/3
    const double2 dxe = cadd(cadd(cscale(2.0 * e, x0s), cdivs(cscale(2.0 * e, k1), 6.0)),
                             cadd(cdivs(k2, 3.0), cdivs(k3, 3.0)));
    h += (xbp.x * dxe.x + xbp.y * dxe.y) * dedh;
    a.hb[ci] += h;
    g = cscale(dt, k4b);
  }
}

// Column-adjoint kernel: [x-FFT of the row-adjoint output -> ybar] -> glue ->
// [m g -> inverse x-FFT -> AO].  `before(b, ci, c, kx, ky)` runs on every
// mode before a B_G glue reads its input there; `after(b, ci, c, kx, ky, v)`
// runs on every mode after a B_OUT glue has stored v there and returns the
// input of B_OUTG's inverse part there (the next pass's B_G input).
template <int N, int TY, int NT, int MODE, int PL, class P, class F>
__device__ __forceinline__ void adj_col_body(const Args& a, P&& before, F&& after) {
  using G = Geo<N>;
  constexpr int KC = G::KC, R = G::R, KMAX = G::KMAX;
  constexpr bool FWD = !(MODE == B_G || MODE == B_ENTRY);
  constexpr bool INV = !(MODE == B_OUT || MODE == B_S1);
  constexpr int GM = MODE == B_OUTG ? B_OUT : MODE;  // the glue of the mode
  constexpr int NSEQ = FWD ? 4 * TY : TY;
  constexpr int SS = Seq<N>::S;
  __shared__ double2 sm[NSEQ * SS];
  __shared__ double2 twl[N];
  TwFetch<N, NT, PL> twf;
  twf.fetch(a.tw);
  int lx, ly, lz;
  xcd_tile(lx, ly, lz);
  const int b = lz, c = ly, ky0 = lx * TY, tid = threadIdx.x;
  if (a.nstep && a.step >= a.nstep[b]) return;  // candidate past its own step count
  if (FWD) {
    const double2* AI = a.ain + (size_t)b * 8 * N * G::KP;
    // first x-FFT pass straight from the row-adjoint output
    fft_lds_gather<N, 4 * TY, SS, NT, false, PL, (TY > 1)>(
        sm, twl,
        [&](int s, int x) -> double2 {
          const int d = s / TY, ky = ky0 + s % TY;
          return ky < KC ? AI[((size_t)(4 * c + d) * N + x) * G::KP + ky] : make_double2(0.0, 0.0);
        },
        [&] { twf.store(twl); });
  } else {
    twf.store(twl);  // ordered by the barrier before the inverse FFT
  }
  constexpr double invN2 = 1.0 / ((double)N * (double)N);
  for (int idx = tid; idx < R * TY; idx += NT) {
    const int cy = idx % TY;
    const int r = idx / TY;
    const int ky = ky0 + cy;
    if (ky >= KC) continue;
    const int kx = G::row_kx(r);
    const int xr = G::row_x(r);
    const size_t ci = (((size_t)b * 2 + c) * R + r) * KC + ky;
    double2 yb = make_double2(0.0, 0.0);
    if (FWD) {
      const bool dc = (kx == 0 && ky == 0);
      const double k2s = dc ? 1.0 : (double)(kx * kx + ky * ky);
      const double mult[4] = {(double)kx, (double)ky, (double)ky / k2s, -(double)kx / k2s};
      // ybar = sum_d conj(i m_d / N^2) Gbar_d = sum_d (-i m_d / N^2) Gbar_d
      double2 s = make_double2(0.0, 0.0);
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const double2 gd = sm[(d * TY + cy) * SS + pidx(xr)];
        s = cadd(s, cscale(mult[d] * invN2, make_double2(gd.y, -gd.x)));
      }
      yb = s;
    }
    double2 g;
    if constexpr (MODE == B_G) before(b, ci, c, kx, ky);
    adj_glue<N, GM>(a, b, ci, r, c, kx, ky, yb, g);
    if constexpr (MODE == B_OUT || MODE == B_OUTG) g = after(b, ci, c, kx, ky, g);
    if (INV) {
      const bool dc = (kx == 0 && ky == 0);
      const double m = dc ? 0.0 : (c == 0 ? 1.0 : (double)(kx * kx + ky * ky));
      sm[(0 * TY + cy) * SS + pidx(xr)] = cscale(m, g);
    }
  }
  if (!INV) return;
  // gap rows and the other 3 (unused) sequence slots are not needed: only slot 0
  constexpr int GAP = N - R;
  for (int idx = tid; idx < TY * GAP; idx += NT) {
    const int s = idx / GAP;
    const int gg = idx - s * GAP;
    sm[s * SS + pidx(KMAX + 1 + gg)] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  // the last pass in registers, stored straight to the x-space cotangent
  fft_lds_head<N, TY, SS, NT, true, PL>(sm, twl);
  double2* AO = a.aout + (size_t)b * 2 * N * G::KP;
  fft_lds_tail<N, TY, SS, NT, true, PL, TY>(sm, twl, [&](int cy, int x, double2 v) {
    const int ky = ky0 + cy;
    if (ky < KC) AO[((size_t)c * N + x) * G::KP + ky] = v;
  });
}

template <int N, int TY, int NT, int MODE, int PL>
__global__ __launch_bounds__(NT) void adj_col_kernel(Args a) {
  adj_col_body<N, TY, NT, MODE, PL>(a, [](int, size_t, int, int, int) {},
                                    [](int, size_t, int, int, int, double2 v) { return v; });
}

// The first kernel of an RHS-VJP pass (B_G) with the transposed stage
// combination that produces its input applied before it in the same launch,
// mode by mode (the combination's outputs include gin; the thread that
// writes an entry of gin reads it): one launch instead of two, the same
// arithmetic.
template <int N, int TY, int NT, int PL>
__global__ __launch_bounds__(NT) void adj_comb_g_kernel(Args a, CombAdj cb) {
  adj_col_body<N, TY, NT, B_G, PL>(
      a, [&](int b, size_t ci, int c, int kx, int ky) { comb_adj_elem(cb, b, ci, c, kx, ky, cb.yb[ci]); },
      [](int, size_t, int, int, int, double2 v) { return v; });
}

// The last kernel of an RHS-VJP pass (B_OUT) with the transposed stage
// combination that reads its output applied in the same launch, mode by
// mode: one launch instead of two, the same arithmetic.  The table comes as a
// second argument (the Args block stays within its 512 bytes).
template <int N, int TY, int NT, int PL>
__global__ __launch_bounds__(NT) void adj_out_comb_kernel(Args a, CombAdj cb) {
  adj_col_body<N, TY, NT, B_OUT, PL>(a, [](int, size_t, int, int, int) {},
                                     [&](int b, size_t ci, int c, int kx, int ky, double2 v) {
                                       comb_adj_elem(cb, b, ci, c, kx, ky, v);
                                       return v;
                                     });
}

// The last kernel of one RHS-VJP pass and the first of the next in one
// launch: B_OUT, the transposed stage combination `then` that reads its
// output and the combination `pre` that produces the next pass's input (each
// when it has terms), then that pass's B_G on gnext (B_OUT's output, one of
// the combinations' outputs, or an input already complete; the thread that
// writes an entry reads it) into AO.  Up to two launches fewer than
// adj_out_comb_kernel, comb_adj_kernel and adj_col_kernel<B_G>, the same
// arithmetic.
template <int N, int TY, int NT, int PL>
__global__ __launch_bounds__(NT) void adj_out_g_kernel(Args a, CombAdj then, CombAdj pre, const double2* gnext) {
  adj_col_body<N, TY, NT, B_OUTG, PL>(a, [](int, size_t, int, int, int) {},
                                      [&](int b, size_t ci, int c, int kx, int ky, double2 v) {
                                        if (then.nterm) comb_adj_elem(then, b, ci, c, kx, ky, v);
                                        if (pre.nterm) comb_adj_elem(pre, b, ci, c, kx, ky, pre.yb[ci]);
                                        return gnext[ci];
                                      });
}

// Row-adjoint kernel: weighted c2r of (Pbar_omega, Pbar_j), transposed
// products with the primal real fields, weighted r2c of the 8 cotangents.
template <int N, int TX, int NT, int PL>
__global__ __launch_bounds__(NT) void adj_row_kernel(Args a) {
  using G = Geo<N>;
  constexpr int KC = G::KC;
  constexpr int NP = 5;  // 4 primal pairs + 1 adjoint pair
  constexpr int NHALF = N / 2 + 1;
  constexpr int SS = Seq<N>::S;
  __shared__ double2 sm[TX * NP * SS];
  __shared__ double2 twl[N];
  TwFetch<N, NT, PL> twf;
  twf.fetch(a.tw);
  int lx, ly, lz;
  xcd_tile(lx, ly, lz);
  const int b = lz, xbase = lx * TX, tid = threadIdx.x;
  if (a.nstep && a.step >= a.nstep[b]) return;  // candidate past its own step count
  const double2* P = a.prim + (size_t)b * 8 * N * G::KP;
  const double2* AO = a.aout + (size_t)b * 2 * N * G::KP;
  // Hermitian extension of the packed pairs, formed while the first inverse
  // pass loads its inputs: 4 primal pairs, then the adjoint pair with the
  // r2c^T weights (pbar[y] = c2r of (Pbar_0, Pbar_k / 2)).  Every lane loads
  // (gap lanes re-read bin 0 and discard it), so all loads of the pass are in
  // flight before the first wait.
  fft_lds_gather<N, TX * NP, SS, NT, true, PL>(
      sm, twl,
      [&](int s, int m) -> double2 {
        const int p = s % NP;
        const int x = xbase + s / NP;
        const bool lo = m <= G::KMAX, hi = m >= N - G::KMAX;
        const int k = lo ? m : (hi ? N - m : 0);
        const double2* src = (p < 4) ? P + (size_t)(2 * p) * N * G::KP : AO;
        double2 A = src[((size_t)x) * G::KP + k];
        double2 B = src[((size_t)N + x) * G::KP + k];
        const double w = (p < 4 || k == 0) ? 1.0 : 0.5;
        A = cscale(w, A);
        B = cscale(w, B);
        if (m == 0) return make_double2(A.x, B.x);
        const double2 vlo = make_double2(A.x - B.y, A.y + B.x), vhi = make_double2(A.x + B.y, B.x - A.y);
        return lo ? vlo : (hi ? vhi : make_double2(0.0, 0.0));
      },
      [&] { twf.store(twl); });
  constexpr int EPT = (TX * N + NT - 1) / NT;
  double2 q[EPT][4];
  const double b0x = a.b0x, b0y = a.b0y;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int idx = tid + e * NT;
    if (idx < TX * N) {
      const int xl = idx / N, yy = idx - xl * N;
      const double2* zb = sm + xl * NP * SS + pidx(yy);
      const double2 z0 = zb[0], z1 = zb[SS], z2 = zb[2 * SS], z3 = zb[3 * SS], z4 = zb[4 * SS];
      const double wx = z0.x, wy = z0.y, u = z1.x, v = z1.y, jx = z2.x, jy = z2.y;
      const double Bx = z3.x + b0x, By = z3.y + b0y;
      const double pw = z4.x, pj = z4.y;
      // p_w = -(u wx + v wy) + (Bx jx + By jy),  p_j = u By - v Bx
      q[e][0] = make_double2(-pw * u, -pw * v);                       // (wx, wy)
      q[e][1] = make_double2(-pw * wx + pj * By, -pw * wy - pj * Bx);  // (u, v)
      q[e][2] = make_double2(pw * Bx, pw * By);                        // (jx, jy)
      q[e][3] = make_double2(pw * jx - pj * v, pw * jy + pj * u);      // (a, b)
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int idx = tid + e * NT;
    if (idx < TX * N) {
      const int xl = idx / N, yy = idx - xl * N;
#pragma unroll
      for (int p = 0; p < 4; ++p) sm[(xl * 4 + p) * SS + pidx(yy)] = q[e][p];
    }
  }
  __syncthreads();
  fft_lds<N, TX * 4, SS, NT, false, PL>(sm, twl);
  double2* AI = a.ain + (size_t)b * 8 * N * G::KP;
  for (int idx = tid; idx < TX * 4 * KC; idx += NT) {
    const int k = idx % KC;
    const int t = idx / KC;
    const int p = t % 4;
    const int xl = t / 4;
    const int x = xbase + xl;
    const double2* Z = sm + (xl * 4 + p) * SS;
    const double2 z = Z[pidx(k)], zm = Z[pidx((N - k) & (N - 1))];
    // c2r^T: Dbar_k = w'_k sum_y dbar e^{-i theta}, w'_0 = 1, w'_k = 2
    const double w = (k == 0) ? 0.5 : 1.0;
    const double2 da = make_double2(w * (z.x + zm.x), w * (z.y - zm.y));
    const double2 db = make_double2(w * (z.y + zm.y), -w * (z.x - zm.x));
    AI[((size_t)(2 * p) * N + x) * G::KP + k] = da;
    AI[((size_t)(2 * p + 1) * N + x) * G::KP + k] = db;
  }
}

// sum of hb over the compact modes of each candidate (fixed order)
template <int N>
__global__ __launch_bounds__(256) void hb_sum_kernel(const double* __restrict__ hb, double* __restrict__ out) {
  using G = Geo<N>;
  constexpr int M = 2 * G::R * G::KC;
  const int b = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < M; i += 256) s += hb[(size_t)b * M + i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[b] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace ecs
