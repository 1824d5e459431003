// Column kernels of the two-derivative RK4 tangent step (timestepping.py:127-149
// under jax.jvp).  Per half step at the point z with tangent dz the stepper
// needs v = f(z), dv = J(z) dz, w = J(z) v, J(z) dv and H[dz, v] (the second
// derivative, f's nonlinearity being a symmetric bilinear form plus b0 terms).
// With these two kernels a half step is four launches:
//   tangent row pass at (z, dz)                       -> td_col_a
//   td_col_a: v, dv; the inputs (z, v), (z, dv), (dz, v) of three tangent
//             passes, batched as candidates b, B + b, 2B + b
//   tangent row pass over 3B candidates (b0 = 0 from 2B on: the H pass)
//   td_col_b: w, J dv, H; the stage combination *gc -> next (z, dz) -> inverse
// One column per block, grid (KC, 2 fields, B); the FFT plan is the tangent
// column family's.
#pragma once
#include "rhs_kernels.h"

namespace ecs {

// derived output d of a stage value y at mode (kx, ky): i m_d y / N^2 (the
// column kernel's multipliers: d/dx, d/dy, u = d/dy psi, v = -d/dx psi)
template <int N>
__device__ __forceinline__ double2 td_derived(int d, int kx, int ky, double2 y) {
  constexpr double invN2 = 1.0 / ((double)N * (double)N);
  const bool dc = (kx == 0 && ky == 0);
  const double k2s = dc ? 1.0 : (double)(kx * kx + ky * ky);
  const double md = d == 0 ? (double)kx : d == 1 ? (double)ky : d == 2 ? (double)ky / k2s : -(double)kx / k2s;
  return cscale(md * invN2, cmuli(y));
}

// Writes the derived spectra of the pair (p, t) of this block's rows into the
// 8 LDS sequences (derived d of p in sequence 2 d, of t in 2 d + 1), zeroes
// the dealiased gap, runs the inverse x-FFT and stores the 16 derived arrays
// of candidate `cand` in the row pass's input layout.  Starts with a barrier,
// which orders the previous call's last reads of the LDS sequences; the last
// transform pass reads them without a barrier after it.
template <int N, int NT, int P, int EPT>
__device__ __forceinline__ void td_emit_pair(const Args& a, double2* sm, const double2* twl, int c, int ky, int cand,
                                             const double2 (&p)[EPT], const double2 (&t)[EPT]) {
  using G = Geo<N>;
  constexpr int KC = G::KC, R = G::R, KMAX = G::KMAX, SS = Seq<N>::S, GAP = N - R;
  const int tid = threadIdx.x;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int r = tid + q * NT;
    if (r >= R) continue;
    const int kx = G::row_kx(r), xr = G::row_x(r);
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      sm[(2 * d) * SS + pidx(xr)] = td_derived<N>(d, kx, ky, p[q]);
      sm[(2 * d + 1) * SS + pidx(xr)] = td_derived<N>(d, kx, ky, t[q]);
    }
  }
  for (int idx = tid; idx < 8 * GAP; idx += NT) {
    const int s = idx / GAP;
    sm[s * SS + pidx(KMAX + 1 + (idx - s * GAP))] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  // the last pass in registers, stored straight to the derived arrays
  fft_lds_head<N, 8, SS, NT, true, P>(sm, twl);
  fft_lds_tail<N, 8, SS, NT, true, P, 1>(sm, twl, [&](int s2, int x, double2 v) {
    const int d = s2 / 2, f = s2 % 2;
    const int g = f * 8 + 4 * c + d;
    a.inter[(((size_t)cand * 16 + g) * N + x) * G::KP + ky] = v;
  });
}

template <int N, int NT, int P>
__global__ __launch_bounds__(NT) void td_col_a(Args a, int B) {
  using G = Geo<N>;
  constexpr int KC = G::KC, R = G::R, SS = Seq<N>::S;
  constexpr int EPT = (R + NT - 1) / NT;
  __shared__ double2 sm[8 * SS];
  __shared__ double2 twl[N];
  TwFetch<N, NT, P> twf;
  twf.fetch(a.tw);
  int lx, ly, lz;
  xcd_tile(lx, ly, lz);
  const int b = lz, c = ly, ky = lx, tid = threadIdx.x;
  double2 z[EPT], dz[EPT], F[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int r = tid + q * NT;
    z[q] = dz[q] = F[q] = make_double2(0.0, 0.0);
    if (r < R) {
      const size_t ci = (((size_t)b * 2 + c) * R + r) * KC + ky;
      z[q] = a.x0[ci];
      dz[q] = a.dx0[ci];
      if (c == 0) F[q] = a.fhat[(size_t)r * KC + ky];
    }
  }
  const double2* O = a.rowout + (size_t)b * 4 * N * G::KP;
  fft_lds_gather<N, 2, SS, NT, false, P>(
      sm, twl, [&](int s, int x) -> double2 { return O[((size_t)(s * 2 + c) * N + x) * G::KP + ky]; },
      [&] { twf.store(twl); });
  // v = f(z), dv = J(z) dz (mask and forcing as in the column kernel's glue)
  double2 v[EPT], dv[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int r = tid + q * NT;
    v[q] = dv[q] = make_double2(0.0, 0.0);
    if (r >= R) continue;
    const int kx = G::row_kx(r), xr = G::row_x(r);
    const double k2 = (double)(kx * kx + ky * ky);
    const bool dc = (kx == 0 && ky == 0);
    const double m = dc ? 0.0 : (c == 0 ? 1.0 : k2);
    const double L = -k2 * a.nu[c];
    const double2 rh = sm[0 * SS + pidx(xr)], drh = sm[1 * SS + pidx(xr)];
    double2 o = make_double2(m * rh.x + F[q].x, m * rh.y + F[q].y);
    double2 od = cscale(m, drh);
    if (a.include_diss) {
      o = cadd(o, cscale(L, z[q]));
      od = cadd(od, cscale(L, dz[q]));
    }
    const size_t ci = (((size_t)b * 2 + c) * R + r) * KC + ky;
    a.out[ci] = o;
    a.dout[ci] = od;
    v[q] = o;
    dv[q] = od;
  }
  td_emit_pair<N, NT, P, EPT>(a, sm, twl, c, ky, b, z, v);
  td_emit_pair<N, NT, P, EPT>(a, sm, twl, c, ky, B + b, z, dv);
  td_emit_pair<N, NT, P, EPT>(a, sm, twl, c, ky, 2 * B + b, dz, v);
}

template <int N, int NT, int P>
__global__ __launch_bounds__(NT) void td_col_b(Args a, int B) {
  using G = Geo<N>;
  constexpr int KC = G::KC, R = G::R, SS = Seq<N>::S;
  constexpr int EPT = (R + NT - 1) / NT;
  __shared__ double2 sm[8 * SS];
  __shared__ double2 twl[N];
  TwFetch<N, NT, P> twf;
  twf.fetch(a.tw);
  int lx, ly, lz;
  xcd_tile(lx, ly, lz);
  const int b = lz, c = ly, ky = lx, tid = threadIdx.x;
  double2 v[EPT], dv[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int r = tid + q * NT;
    v[q] = dv[q] = make_double2(0.0, 0.0);
    if (r < R) {
      const size_t ci = (((size_t)b * 2 + c) * R + r) * KC + ky;
      v[q] = a.out[ci];
      dv[q] = a.dout[ci];
    }
  }
  // the tangent outputs of the three passes: candidates b, B + b, 2B + b
  fft_lds_gather<N, 3, SS, NT, false, P>(
      sm, twl,
      [&](int s, int x) -> double2 {
        const double2* O = a.rowout + (size_t)(s * B + b) * 4 * N * G::KP;
        return O[((size_t)(1 * 2 + c) * N + x) * G::KP + ky];
      },
      [&] { twf.store(twl); });
  double2 y[EPT], dy[EPT];
  const bool inv = !a.skip_inv;
  const Comb* cm = a.gc;
  const double h = cm->h[b], dh = cm->dh[b];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int r = tid + q * NT;
    y[q] = dy[q] = make_double2(0.0, 0.0);
    if (r >= R) continue;
    const int kx = G::row_kx(r), xr = G::row_x(r);
    const double k2 = (double)(kx * kx + ky * ky);
    const bool dc = (kx == 0 && ky == 0);
    const double m = dc ? 0.0 : (c == 0 ? 1.0 : k2);
    const double L = -k2 * a.nu[c];
    // w = J(z) v and J(z) dv carry the dissipation; H[dz, v] does not
    double2 w = cscale(m, sm[0 * SS + pidx(xr)]);
    double2 jdv = cscale(m, sm[1 * SS + pidx(xr)]);
    const double2 hh = cscale(m, sm[2 * SS + pidx(xr)]);
    if (a.include_diss) {
      w = cadd(w, cscale(L, v[q]));
      jdv = cadd(jdv, cscale(L, dv[q]));
    }
    const size_t ci = (((size_t)b * 2 + c) * R + r) * KC + ky;
    a.tdw[ci] = w;
    a.tdjdv[ci] = jdv;
    a.tdh[ci] = hh;
    // the combination, term by term in table order (no E factors)
    double2 acc = make_double2(0.0, 0.0), dacc = acc;
    for (int t = 0; t < cm->nterm; ++t) {
      double wt = cm->coef[t];
      for (int k = 0; k < cm->hp[t]; ++k) wt *= h;
      const double2 val = cm->in[t] ? (cm->in[t] == a.tdw ? w : cm->in[t][ci]) : make_double2(0.0, 0.0);
      acc = cadd(acc, cscale(wt, val));
      double dw = cm->coef[t] * (double)cm->hp[t];
      for (int k = 1; k < cm->hp[t]; ++k) dw *= h;
      dacc = cadd(dacc, cscale(dh * dw, val));
      if (cm->din[t]) {
        const double2 dval = cm->din[t] == a.tdjdv ? jdv : (cm->din[t] == a.tdh ? hh : cm->din[t][ci]);
        dacc = cadd(dacc, cscale(wt, dval));
      }
    }
    a.xf[ci] = acc;
    a.dxf[ci] = dacc;
    y[q] = acc;
    dy[q] = dacc;
  }
  if (inv) td_emit_pair<N, NT, P, EPT>(a, sm, twl, c, ky, b, y, dy);
}

}  // namespace ecs
