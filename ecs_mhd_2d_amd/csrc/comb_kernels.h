// Stage combinations of the generic explicit steppers (classical RK4,
// Lawson RK6, two-derivative RK4, and the Lawson RK4 steps of the RK4(3)
// reverse pass) over compact spectra [B][2][R][KC].  The forward combination
// is evaluated in the column kernels' glue (C_GEN, C_TD2, td_col_b) from a
// table in device memory; this file holds its transpose for the reverse pass.
//
// out = sum_t coef_t * h^hp_t * E^ep_t * in_t,   E = exp(h L / ediv)
// with L = -(kx^2 + ky^2) [nu; eta] and h the candidate's step.  A Lawson
// scheme that multiplies its state and stage increments by e = exp(h/m L)
// after some stages (timestepping.py:83-125) puts the accumulated power of e
// on each term here: stage s reads x0 E^{p_s} and k_j E^{p_s - p_j}.  With
// `masked`, E^p for p > 0 carries the 2/3 mask, whose only zero inside the
// compact layout is the mean mode (mhd_jax.py:65-70).
// The tangent adds d/dh of every weight times dh:
//   dout = sum_t w_t din_t + dh * sum_t (hp_t / h + ep_t L / ediv) w_t in_t.
#pragma once
#include "kernels.h"

namespace ecs {

// Transpose of one stage combination, for the reverse pass:
//   out_t += c_t h^a_t E^p_t yb          (each target once per call; = for the
//                                          terms marked in `assign`, the first
//                                          contribution a target gets)
//   hb    += sum_t d(c_t h^a_t E^p_t)/dh <yb, in_t>   (per mode, real inner product)
// where in_t is the primal operand of term t (null: no d/dh term) and the
// spectral cotangent convention holds d/dRe, d/dIm in the real, imag parts.
struct CombAdj {
  const double2* yb;
  double2* out[COMB_MAX];
  const double2* in[COMB_MAX];
  double coef[COMB_MAX];
  int hp[COMB_MAX];
  int ep[COMB_MAX];
  int nterm;
  unsigned assign;               // bit t: term t overwrites out_t instead of adding
  double ediv;
  int masked;
  const double* h;
  double nu[2];
  double* hb;                    // per-mode accumulator [B][2][R][KC] (null: no d/dh)
  const double* hscale;          // [B] multiplier of the d/dh terms (null: 1), e.g. dh/dT
  size_t elems;
};

// The transposed combination at one mode (c, kx, ky) of candidate b, compact
// index i, for the stage cotangent value yb (cb.yb[i]).  The combination
// kernel below and the RHS-VJP pass's last kernel (which produces yb and
// applies the combination that follows it in the same launch) share it.
__device__ __forceinline__ void comb_adj_elem(const CombAdj& cb, int b, size_t i, int c, int kx, int ky,
                                              double2 yb) {
  const bool dc = (kx == 0 && ky == 0);
  const double L = -(double)(kx * kx + ky * ky) * cb.nu[c];
  const double h = cb.h[b];
  const double E1 = exp(h / cb.ediv * L);
  double hacc = 0.0;
  for (int t = 0; t < cb.nterm; ++t) {
    double w = cb.coef[t];
    for (int q = 0; q < cb.hp[t]; ++q) w *= h;
    double em = 1.0;
    if (cb.ep[t] > 0) {
      double Ep = E1;
      for (int q = 1; q < cb.ep[t]; ++q) Ep *= E1;
      em = (cb.masked && dc) ? 0.0 : Ep;
      w *= em;
    }
    if (cb.out[t]) cb.out[t][i] = ((cb.assign >> t) & 1u) ? cscale(w, yb) : cadd(cb.out[t][i], cscale(w, yb));
    if (cb.in[t] && cb.hb) {
      double dw = cb.coef[t] * (double)cb.hp[t];
      for (int q = 1; q < cb.hp[t]; ++q) dw *= h;
      if (cb.ep[t] > 0) dw = dw * em + w * ((double)cb.ep[t] * L / cb.ediv);
      const double2 v = cb.in[t][i];
      hacc += dw * (yb.x * v.x + yb.y * v.y);
    }
  }
  if (cb.hb) cb.hb[i] += cb.hscale ? cb.hscale[b] * hacc : hacc;
}

template <int N>
__global__ __launch_bounds__(256) void comb_adj_kernel(CombAdj cb, int B) {
  using G = Geo<N>;
  constexpr int KC = G::KC, R = G::R;
  const size_t tot = (size_t)B * cb.elems;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(i / cb.elems);
    const size_t e = i - (size_t)b * cb.elems;
    const int ky = (int)(e % KC);
    const int r = (int)((e / KC) % R);
    const int c = (int)(e / ((size_t)KC * R));
    comb_adj_elem(cb, b, i, c, G::row_kx(r), ky, cb.yb[i]);
  }
}

}  // namespace ecs
