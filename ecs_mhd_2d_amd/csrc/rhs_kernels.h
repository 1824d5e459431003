// Column and row kernels of the fused RHS / Lawson stepper (see kernels.h).
#pragma once
#include "kernels.h"

namespace ecs {

// Block-uniform control values.
struct Blk {
  bool run;      // block has any work
  bool stage;    // stage glue runs and produces a next-stage input
  bool pend;     // adaptive: a previous attempt awaits its decision
  bool pacc;     // adaptive: that attempt is accepted
  double hprev;  // adaptive: step of that attempt
  double T, h;   // adaptive: period and pseudo-time step
  double dt;     // stage scale (physical time step)
  double ddt;    // tangent of dt per unit tangent
  bool next;     // C_S5A: the next accepted step exists (its stage-1 glue runs)
  double h2, dt2, ddt2;  // C_S5A: its step, scale and tangent
  double dT;     // tangent replay: dT[b]
  // tangent replay: rejections preceding the attempt whose stage 1 runs here
  // (unsigned: widening it needs no sign extension at the load, which would
  // make the prologue wait for it)
  unsigned rs;
  int rc;
};

// The index comes from the argument unless a device counter is set.  The
// empty asm makes the argument an opaque register value: otherwise the
// compiler turns the choice into one load through a selected pointer, and the
// argument path would wait on a memory load too.
__device__ __forceinline__ int attempt_of(const Args& a) {
  int att = a.attempt;
  asm volatile("" : "+s"(att));
  return a.cnt ? a.cnt[a.cslot] : att;
}

__device__ __forceinline__ double relu_nan(double x) { return (x != x) ? x : (x > 0.0 ? x : 0.0); }

// Controller decision of timestepping.py:288-313, from the record of attempt i-1.
// AK (an AdaptKind) is a template argument, so each kernel instance has one
// straight-line prologue and its loads are not forced to complete where the
// control paths of the other kinds would merge.  `lead` marks one block per
// candidate, the one that writes the candidate's controller record.
template <int MODE, bool TAN, int AK>
__device__ __forceinline__ Blk prologue(const Args& a, int b, bool lead) {
  Blk bk;
  bk.run = true; bk.stage = true; bk.pend = false; bk.pacc = false;
  bk.hprev = 0.0; bk.T = 0.0; bk.h = 0.0; bk.dt = 0.0; bk.ddt = 0.0;
  bk.next = false; bk.h2 = 0.0; bk.dt2 = 0.0; bk.ddt2 = 0.0;
  bk.dT = 0.0; bk.rs = 0u; bk.rc = 0;
  if (AK == AD_FIXED) {
    if (a.nstep && a.step >= a.nstep[b]) {
      bk.run = bk.stage = false;
      return bk;
    }
    bk.dt = a.dt ? a.dt[b] : 0.0;
    bk.ddt = (TAN && a.ddt) ? a.ddt[b] : 0.0;
    return bk;
  }
  bk.T = a.Tb[b];
  const int att = attempt_of(a);
  if ((MODE == C_A || MODE == C_A0) && a.cnt && lead && b == 0 &&
      threadIdx.x == 0)
    a.cnt[a.cslot ^ 1] = att + 1;
  if (AK == AD_TANGENT) {
    // only the run test waits here: the schedule entries below are loaded
    // now and used after the column gathers (finish() forms the scales)
    const int m = att;
    const bool known = a.run_all && (MODE != C_S5A || a.next_all);
    const int nacc = known ? 0 : a.s_nacc[b];
    bk.dT = a.dTb[b];
    bk.run = bk.stage = a.run_all || m < nacc;
    if (!bk.run) return bk;
    const size_t so = (size_t)b * a.sstride + m;
    bk.h = a.s_h[so];
    if (MODE == C_S5A) {
      // loaded unconditionally (from entry m when there is no next step, and
      // then unused) so no branch has to wait for them
      bk.next = a.next_all || m + 1 < nacc;
      const size_t sn = bk.next ? so + 1 : so;
      bk.h2 = a.s_h[sn];
      bk.rs = (unsigned)a.s_rs[sn];
      bk.rc = a.s_rc[sn];
    }
    if (MODE == C_A || MODE == C_A0) {
      bk.rs = (unsigned)a.s_rs[so];
      bk.rc = a.s_rc[so];
    }
    return bk;
  }
  // AD_PRIMAL
  const int i = att;
  Hist* H = a.hist + (size_t)b * a.hstride;
  if (MODE == C_A0) {
    bk.h = a.h0;
    bk.dt = bk.h * bk.T;
    if (lead && threadIdx.x == 0) {
      Hist w; w.s = 0.0; w.h = a.h0; w.err = 0ull; w.active = 1; w.prev_acc = 0; w.nacc = 0; w.nrej = 0;
      H[0] = w;
    }
    return bk;
  }
  if (MODE == C_A || MODE == C_FIN) {
    const Hist hp = H[i - 1];
    double s = hp.s, hn = hp.h;
    int nacc = hp.nacc, nrej = hp.nrej;
    bk.pend = hp.active != 0;
    bk.hprev = hp.h;
    if (bk.pend) {
      const double err = __longlong_as_double((long long)hp.err);
      bk.pacc = err < a.atol;
      if (bk.pacc) { s = hp.s + hp.h; nacc += 1; } else { nrej += 1; }
      if (err != err) hn = hp.h / 2;
      else hn = hp.h * 0.9 * pow(a.atol / err, 0.25);
      hn = hn - relu_nan(s + hn - 1.0);
    }
    const bool act = (MODE == C_A) && (a.budget <= 0 || i < a.budget) && !(fabs(s - 1.0) < 1e-10);
    bk.stage = act;
    bk.run = bk.pend || act;
    bk.h = hn;
    bk.dt = hn * bk.T;
    if (lead && threadIdx.x == 0) {
      Hist w; w.s = s; w.h = hn; w.err = 0ull; w.active = act ? 1 : 0; w.prev_acc = bk.pacc ? 1 : 0;
      w.nacc = nacc; w.nrej = nrej;
      H[i] = w;
    }
    return bk;
  }
  const Hist hc = H[i];
  bk.run = bk.stage = hc.active != 0;
  bk.h = hc.h;
  bk.dt = bk.h * bk.T;
  return bk;
}

// Scales of the tangent replay from the schedule entries the prologue loaded.
template <int AK>
__device__ __forceinline__ void finish(Blk& bk) {
  if (AK != AD_TANGENT) return;
  bk.dt = bk.h * bk.T;
  bk.ddt = bk.h * bk.dT;
  if (bk.next) {
    bk.dt2 = bk.h2 * bk.T;
    bk.ddt2 = bk.h2 * bk.dT;
  }
}

template <int AK>
__device__ __forceinline__ bool row_active(const Args& a, int b) {
  if (AK == AD_FIXED) return !(a.nstep && a.step >= a.nstep[b]);
  if (AK == AD_PRIMAL) return a.hist[(size_t)b * a.hstride + attempt_of(a)].active != 0;
  return a.run_all || attempt_of(a) < a.s_nacc[b];
}

__device__ __forceinline__ double2 cdivs(double2 a, double s) { return make_double2(a.x / s, a.y / s); }

// Glue operands loaded before the column FFT so their latency overlaps it.
struct Pre {
  double2 F;              // masked forcing spectrum (omega only)
  double2 x0, dx0;        // stage-1 state (modes that read it before writing)
  double2 acc, dacc;      // RK accumulator
  double2 xf, dxf;        // tangent replay C_S5A: the accepted state
  double2 v1, dv1;        // tangent replay C_A: the carried k1
  // C_GEN with one glue element per thread: the operands of the combination's
  // terms (in_t, din_t), loaded with the rest; the terms reading this
  // kernel's own output (in_t == out) are formed in the glue instead
  double2 gv[COMB_MAX], gdv[COMB_MAX];
};

// The C_GEN operands of Pre (element ci), when each thread has one glue element.
template <bool TAN>
__device__ __forceinline__ void prefetch_comb(const Args& a, size_t ci, Pre& p) {
  const Comb* cm = a.gc;
  const int nt = cm->nterm;
#pragma unroll
  for (int t = 0; t < COMB_MAX; ++t) {
    p.gv[t] = make_double2(0.0, 0.0);
    p.gdv[t] = p.gv[t];
    if (t < nt) {
      const double2* in = cm->in[t];
      if (in && in != a.out) p.gv[t] = in[ci];
      if (TAN) {
        const double2* din = cm->din[t];
        if (din && din != a.dout) p.gdv[t] = din[ci];
      }
    }
  }
}

template <int MODE, bool TAN, int AK>
__device__ __forceinline__ Pre prefetch(const Args& a, size_t ci, size_t fi, int c) {
  Pre p;
  const double2 z = make_double2(0.0, 0.0);
  p.F = (c == 0 && MODE != C_LOAD) ? a.fhat[fi] : z;
  constexpr bool X0 = MODE == C_LOAD || MODE == C_RHS || MODE == C_S1F || MODE == C_S2 || MODE == C_S3 ||
                      MODE == C_A0 || MODE == C_GEN || MODE == C_TD1 || MODE == C_TD2;
  constexpr bool ACC = MODE == C_S2 || MODE == C_S3 || MODE == C_S4;
  p.x0 = X0 ? a.x0[ci] : z;
  p.dx0 = (X0 && TAN) ? a.dx0[ci] : z;
  p.acc = ACC ? a.acc[ci] : z;
  p.dacc = (ACC && TAN) ? a.dacc[ci] : z;
  constexpr bool rep = TAN && AK == AD_TANGENT;
  p.xf = (MODE == C_S5A && rep) ? a.xf[ci] : z;
  p.dxf = (MODE == C_S5A && rep) ? a.dxf[ci] : z;
  p.v1 = (MODE == C_A && rep) ? a.v1[ci] : z;
  p.dv1 = (MODE == C_A && rep) ? a.dv1[ci] : z;
  return p;
}

// Stage-1 glue of one Lawson step (timestepping.py:70-74, :263-270):
// acc = e e x0 + e e k1/6, y2 = e x0 + e k1/2 with k1 = dt v1, and tangents.
template <bool TAN>
__device__ __forceinline__ void stage1(const Args& a, bool wr, size_t ci, double e, double de, double dt, double ddt,
                                       double2 x0, double2 dx0, double2 v1, double2 dv1, double2& y, double2& dy) {
  const double de2 = 2.0 * e * de;
  const double2 k1 = cscale(dt, v1);
  const double2 ek1 = cscale(e, k1);
  const double2 ex0 = cscale(e, x0);
  if (wr) a.acc[ci] = cadd(cscale(e, ex0), cdivs(cscale(e, ek1), 6.0));
  y = cadd(ex0, cdivs(ek1, 2.0));
  if (TAN) {
    const double2 dk1 = cadd(cscale(ddt, v1), cscale(dt, dv1));
    if (wr)
      a.dacc[ci] = cadd(cadd(cscale(de2, x0), cscale(e * e, dx0)),
                        cdivs(cadd(cscale(de2, k1), cscale(e * e, dk1)), 6.0));
    dy = cadd(cadd(cscale(de, x0), cscale(e, dx0)), cdivs(cadd(cscale(de, k1), cscale(e, dk1)), 2.0));
  }
}

// Rejections preceding accepted step m of the tangent replay: the carried k1
// becomes e (e k1) per rejected step h_j (timestepping.py:267,276,297).
__device__ __forceinline__ void replay_rejections(const Args& a, int bb, unsigned rs, int rc, double dT, double TL,
                                                  double L, double2& v1, double2& dv1) {
  for (int q = 0; q < rc; ++q) {
    const double hj = a.s_rh[(size_t)bb * a.rstride + rs + q];
    const double ep = exp((TL * hj) * 0.5);
    v1 = cscale(ep, cscale(ep, v1));
    dv1 = cadd(cscale(ep, cscale(ep, dv1)), cscale(L * (hj * dT), v1));
  }
}

// Stage glue for one kept mode.  rh/drh: row-pass spectrum after the x-FFT
// (nonlinear term before masking).  Produces the next stage input y (dy).
// Global writes happen only when wr is set (one of the blocks that share a
// column when its output sequences are split over several blocks).
// Lawson RK4 in the form of timestepping.py:70-80 and :263-282:
//   y2 = e x0 + e k1/2, y3 = e x0 + k2/2, y4 = e e x0 + e k3,
//   xf = e e x0 + e e k1/6 + e k2/3 + e k3/3 + k4/6,   k_s = dt v(y_s).
template <int N, int MODE, bool TAN, int AK, bool GPF = false>
__device__ __forceinline__ void glue(const Args& a, bool wr, const Blk& bk, const Pre& p, int bb, size_t ci, int r,
                                     int c, int kx, int ky, double2 rh, double2 drh, double2& y, double2& dy) {
  const double k2 = (double)(kx * kx + ky * ky);
  const bool dc = (kx == 0 && ky == 0);
  // mask (omega) and k^2 mask (j) of mhd_jax.py:160,164; forcing added spectrally
  const double m = dc ? 0.0 : (c == 0 ? 1.0 : k2);
  const double L = -k2 * a.nu[c];
  const double2 vh = make_double2(m * rh.x + p.F.x, m * rh.y + p.F.y);
  const double2 dvh = cscale(m, drh);
  y = make_double2(0.0, 0.0);
  dy = y;
  if (MODE == C_LOAD) {
    y = p.x0;
    if (TAN) dy = p.dx0;
    return;
  }
  if (MODE == C_TD1) {
    // v = RHS(x0); the next pass is the tangent pass at x0 along v
    double2 o = vh;
    if (a.include_diss) o = cadd(o, cscale(L, p.x0));
    a.out[ci] = o;
    y = p.x0;
    dy = o;
    return;
  }
  if (MODE == C_TD2) {
    // w = J(x0) dx0 (the tangent output), then the combination of *gc, whose
    // term reading dout takes w from registers
    double2 od = dvh;
    if (a.include_diss) od = cadd(od, cscale(L, p.dx0));
    a.dout[ci] = od;
    const Comb* cm = a.gc;
    const double h = cm->h[bb];
    double2 acc = make_double2(0.0, 0.0);
    for (int t = 0; t < cm->nterm; ++t) {
      double w = cm->coef[t];
      for (int q = 0; q < cm->hp[t]; ++q) w *= h;
      const double2 v = cm->in[t] == a.dout ? od : cm->in[t][ci];
      acc = cadd(acc, cscale(w, v));
    }
    a.xf[ci] = acc;
    y = acc;
    return;
  }
  if (MODE == C_GEN) {
    // this stage's RHS (and tangent), then the next stage input from the
    // combination, term by term in table order
    double2 o = vh, od = dvh;
    if (a.include_diss) {
      o = cadd(o, cscale(L, p.x0));
      if (TAN) od = cadd(od, cscale(L, p.dx0));
    }
    if (!TAN && a.pdt) {
      // the combination kernel's weight of a one-term table with ediv 1
      a.vraw[ci] = o;
      o = cscale(exp(a.pdt[bb] / 1.0 * L), o);
    }
    a.out[ci] = o;
    if (TAN) a.dout[ci] = od;
    const Comb* cm = a.gc;
    // a table without its own step array takes the launch's (per-step schedules)
    const double h = cm->h ? cm->h[bb] : a.dt[bb];
    const double dh = TAN ? cm->dh[bb] : 0.0;
    const double E1 = exp(h / cm->ediv * L);
    double2 acc = make_double2(0.0, 0.0), dacc = acc;
    // term t with its operands prefetched (vpre, dvpre; GPF) or loaded here
    auto term = [&](int t, double2 vpre, double2 dvpre) {
      double w = cm->coef[t];
      for (int q = 0; q < cm->hp[t]; ++q) w *= h;
      if (cm->ep[t] > 0) {
        double Ep = E1;
        for (int q = 1; q < cm->ep[t]; ++q) Ep *= E1;
        w *= (cm->masked && dc) ? 0.0 : Ep;
      }
      const double2 v = cm->in[t] == a.out ? o : (GPF ? vpre : cm->in[t][ci]);
      acc = cadd(acc, cscale(w, v));
      if (TAN) {
        double dw = cm->coef[t] * (double)cm->hp[t];
        for (int q = 1; q < cm->hp[t]; ++q) dw *= h;
        if (cm->ep[t] > 0) {
          double Ep = E1;
          for (int q = 1; q < cm->ep[t]; ++q) Ep *= E1;
          const double em = (cm->masked && dc) ? 0.0 : Ep;
          dw = dw * em + w * ((double)cm->ep[t] * L / cm->ediv);
        }
        dacc = cadd(dacc, cscale(dh * dw, v));
        if (cm->din[t]) dacc = cadd(dacc, cscale(w, cm->din[t] == a.dout ? od : (GPF ? dvpre : cm->din[t][ci])));
      }
    };
    const int nt = cm->nterm;
    if constexpr (GPF) {
      // unrolled, so the prefetched operands stay in registers
#pragma unroll
      for (int t = 0; t < COMB_MAX; ++t)
        if (t < nt) term(t, p.gv[t], p.gdv[t]);
    } else {
      for (int t = 0; t < nt; ++t) term(t, make_double2(0.0, 0.0), make_double2(0.0, 0.0));
    }
    a.xf[ci] = acc;
    y = acc;
    if (TAN) {
      a.dxf[ci] = dacc;
      dy = dacc;
    }
    return;
  }
  if (MODE == C_RHS) {
    double2 o = vh;
    if (a.include_diss) o = cadd(o, cscale(L, p.x0));
    if (wr) a.out[ci] = o;
    if (TAN) {
      double2 d = dvh;
      if (a.include_diss) d = cadd(d, cscale(L, p.dx0));
      if (wr) a.dout[ci] = d;
    }
    return;
  }
  const double TL = bk.T * L;
  if (MODE == C_S5A) {
    // accept this attempt: k1 <- k5, x0 <- xf; then the next step's C_A glue
    const double2 x0 = p.xf, dx0 = p.dxf;
    if (wr) {
      a.x0[ci] = x0;
      a.dx0[ci] = dx0;
    }
    double2 v1 = vh, dv1 = dvh;
    if (bk.next) replay_rejections(a, bb, bk.rs, bk.rc, bk.dT, TL, L, v1, dv1);
    if (wr) {
      a.v1[ci] = v1;
      a.dv1[ci] = dv1;
    }
    if (!bk.next) return;
    const double e = exp((TL * bk.h2) * 0.5);
    const double de = (bk.ddt2 * 0.5) * L * e;
    stage1<TAN>(a, wr, ci, e, de, bk.dt2, bk.ddt2, x0, dx0, v1, dv1, y, dy);
    return;
  }
  // v1 of the stage-1 glue
  double2 v1 = make_double2(0.0, 0.0), dv1 = v1;
  if (MODE == C_S1F) {
    v1 = vh;
    dv1 = dvh;
  }
  if (MODE == C_A0 || MODE == C_A || MODE == C_FIN) {
    if (AK == AD_PRIMAL) {
      if (MODE == C_A0) {
        v1 = vh;
        if (wr) a.v1[ci] = v1;
      } else {
        v1 = a.v1[ci];
        if (bk.pend) {
          if (bk.pacc) {
            if (wr) a.x0[ci] = a.xf[ci];
            v1 = a.v5[ci];
          } else {
            // rejected: carried k1 is e (e k1) (timestepping.py:267,276,297)
            const double ep = exp((TL * bk.hprev) * 0.5);
            v1 = cscale(ep, cscale(ep, v1));
          }
          if (wr) a.v1[ci] = v1;
        }
      }
    } else {  // tangent replay: rejections preceding this accepted attempt
      if (MODE == C_A0) {
        v1 = vh;
        dv1 = dvh;
      } else {
        v1 = p.v1;
        dv1 = p.dv1;
      }
      replay_rejections(a, bb, bk.rs, bk.rc, bk.dT, TL, L, v1, dv1);
      if (wr) {
        a.v1[ci] = v1;
        a.dv1[ci] = dv1;
      }
    }
    if (!bk.stage) return;
  }
  // Lawson factor
  double e;
  if (AK == AD_FIXED) {
    e = exp((bk.dt * 0.5) * L);
    if (a.masked_e && dc) e = 0.0;
  } else {
    e = exp((TL * bk.h) * 0.5);
  }
  const double de = TAN ? (bk.ddt * 0.5) * L * e : 0.0;
  const double de2 = 2.0 * e * de;
  // C_A may have just replaced x0 (accepted previous attempt): read it back
  const double2 x0 = (MODE == C_A) ? a.x0[ci] : p.x0;
  const double2 dx0 = (MODE == C_A && TAN) ? a.dx0[ci] : p.dx0;
  if (MODE == C_S1F && a.xtape && wr) a.xtape[ci] = x0;
  if ((MODE == C_S1F || MODE == C_S2 || MODE == C_S3 || MODE == C_S4) && a.vtape && wr) a.vtape[ci] = vh;
  if (MODE == C_S1F || MODE == C_A0 || MODE == C_A) {
    stage1<TAN>(a, wr, ci, e, de, bk.dt, bk.ddt, x0, dx0, v1, dv1, y, dy);
    return;
  }
  const double2 ks = cscale(bk.dt, vh);
  const double2 dks = TAN ? cadd(cscale(bk.ddt, vh), cscale(bk.dt, dvh)) : make_double2(0.0, 0.0);
  if (MODE == C_S2) {
    if (wr) a.acc[ci] = cadd(p.acc, cdivs(cscale(e, ks), 3.0));
    y = cadd(cscale(e, x0), cdivs(ks, 2.0));
    if (TAN) {
      if (wr) a.dacc[ci] = cadd(p.dacc, cdivs(cadd(cscale(de, ks), cscale(e, dks)), 3.0));
      dy = cadd(cadd(cscale(de, x0), cscale(e, dx0)), cdivs(dks, 2.0));
    }
    return;
  }
  if (MODE == C_S3) {
    if (wr) a.acc[ci] = cadd(p.acc, cdivs(cscale(e, ks), 3.0));
    y = cadd(cscale(e, cscale(e, x0)), cscale(e, ks));
    if (TAN) {
      if (wr) a.dacc[ci] = cadd(p.dacc, cdivs(cadd(cscale(de, ks), cscale(e, dks)), 3.0));
      dy = cadd(cadd(cscale(de2, x0), cscale(e * e, dx0)), cadd(cscale(de, ks), cscale(e, dks)));
    }
    return;
  }
  if (MODE == C_S4) {
    const double2 xf = cadd(p.acc, cdivs(ks, 6.0));
    y = xf;
    double2 dxf = make_double2(0.0, 0.0);
    if (TAN) {
      dxf = cadd(p.dacc, cdivs(dks, 6.0));
      dy = dxf;
    }
    if (!wr) return;
    if (AK == AD_FIXED) {
      a.x0[ci] = xf;
      if (TAN) a.dx0[ci] = dxf;
    } else {
      a.xf[ci] = xf;
      if (TAN) a.dxf[ci] = dxf;
      else a.v4[ci] = vh;
    }
    return;
  }
  if (MODE == C_S5) {
    if (TAN) {  // replay: this attempt is accepted; k1 <- k5, x0 <- xf
      a.x0[ci] = a.xf[ci];
      a.dx0[ci] = a.dxf[ci];
      a.v1[ci] = vh;
      a.dv1[ci] = dvh;
    } else {
      a.v5[ci] = vh;
      // err = h (k5 - k4) / 6 (timestepping.py:288)
      y = cdivs(cscale(bk.dt, csub(vh, a.v4[ci])), 6.0);
    }
    return;
  }
}

// Modes whose column work may be split over several blocks: each block of a
// column runs the forward FFT and the glue, and produces NDO / SPL of the
// derived output sequences; only split 0 writes the glue's global state.  A
// mode qualifies when no block reads a value another block of its column
// writes in the same launch (C_A reads back the state it replaces).
template <int MODE>
constexpr bool col_splittable() {
  return MODE == C_LOAD || MODE == C_S1F || MODE == C_S2 || MODE == C_S3 || MODE == C_S4 || MODE == C_A0 ||
         MODE == C_S5A;
}

template <int N, int TY, int NT, int MODE, bool TAN, int P, int AK, int SPL>
__global__ __launch_bounds__(NT) void col_kernel(Args a) {
  using G = Geo<N>;
  constexpr int KC = G::KC, KP = G::KP, R = G::R, KMAX = G::KMAX;
  // fields of the row output read (NFI) and of the derived output written
  // (NFO): primal only, or primal and tangent; they differ only in the
  // two-derivative RK4 modes
  constexpr int NF = TAN ? 2 : 1;
  constexpr int NFI = NF;
  constexpr int NFO = MODE == C_TD1 ? 2 : (MODE == C_TD2 ? 1 : NF);
  constexpr int NDO = (MODE == C_S5 && !TAN) ? 1 : 4;
  constexpr bool FWD = MODE == C_RHS || MODE == C_S1F || MODE == C_S2 || MODE == C_S3 || MODE == C_S4 ||
                       MODE == C_S5 || MODE == C_A0 || MODE == C_S5A || MODE == C_GEN || MODE == C_TD1 ||
                       MODE == C_TD2;
  constexpr bool HASINV = !(MODE == C_RHS || MODE == C_FIN || (MODE == C_S5 && TAN));
  static_assert(SPL == 1 || (col_splittable<MODE>() && NDO % SPL == 0), "column split not allowed for this mode");
  constexpr int DPS = NDO / SPL;               // derived outputs per block
  constexpr int EPT = (R * TY + NT - 1) / NT;  // glue elements per thread
  // C_GEN loads its combination operands before the FFT when a thread has one
  // glue element (the small tiles of small batches, where a kernel is a chain
  // of dependent memory round trips); wider tiles load them in the glue
  constexpr bool GPF = MODE == C_GEN && EPT == 1;
  // A block's derived outputs are transformed in TCH chunks of at most 8
  // sequences that reuse one LDS region (the glue results wait in registers),
  // so tiles of 2-4 columns keep the LDS of a block small enough for several
  // blocks per CU.
  constexpr int NSF = DPS * NFO * TY;
  constexpr int TCH = (HASINV && NDO == 4 && NSF > 8) ? (NSF / 8 < DPS ? NSF / 8 : DPS) : 1;
  static_assert(DPS % TCH == 0, "chunks must divide the derived outputs");
  constexpr int DPC = DPS / TCH;  // derived outputs per chunk
  constexpr int NSEQ = HASINV ? DPC * NFO * TY : NFI * TY;
  constexpr int NSM = NSEQ > NFI * TY ? NSEQ : NFI * TY;
  constexpr int ND = (NDO == 1) ? 2 : 8 * NFO;
  constexpr int SS = Seq<N>::S;
  __shared__ double2 sm[NSM * SS];
  __shared__ double2 twl[N];
  TwFetch<N, NT, P> twf;
  twf.fetch(a.tw);
  int lx, ly, lz;
  xcd_tile(lx, ly, lz);
  const int b = lz, ky0 = lx * TY, tid = threadIdx.x;
  const int c = SPL > 1 ? ly % 2 : ly;
  const int sp = SPL > 1 ? ly / 2 : 0;
  const bool wr = sp == 0;
  Blk bk = prologue<MODE, TAN, AK>(a, b, lx == 0 && ly == 0);
  if (!bk.run) return;
  // issue the glue operand loads first: their latency overlaps the x-FFT
  Pre pre[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int idx = tid + q * NT;
    const int cy = idx % TY;
    const int r = idx / TY;
    const int ky = ky0 + cy;
    if (idx < R * TY && ky < KC) {
      const size_t ci = (((size_t)b * 2 + c) * R + r) * KC + ky;
      pre[q] = prefetch<MODE, TAN, AK>(a, ci, (size_t)r * KC + ky, c);
      if constexpr (GPF) prefetch_comb<TAN>(a, ci, pre[q]);
    }
  }

  if (FWD) {
    const double2* O = a.rowout + (size_t)b * (2 * NFI) * N * KP;
    // the first x-FFT pass reads the row output directly; with several
    // columns per block neighbouring lanes take neighbouring ky of one row
    fft_lds_gather<N, NFI * TY, SS, NT, false, P, (TY > 1)>(
        sm, twl,
        [&](int s, int x) -> double2 {
          const int f = s / TY, ky = ky0 + s % TY;
          return ky < KC ? O[((size_t)(f * 2 + c) * N + x) * KP + ky] : make_double2(0.0, 0.0);
        },
        [&] { twf.store(twl); });
  } else {
    twf.store(twl);  // ordered by the barrier before the inverse FFT
  }
  finish<AK>(bk);

  const bool inv = HASINV && (MODE == C_S5A ? bk.next : bk.stage) && !a.skip_inv;
  constexpr double invN2 = 1.0 / ((double)N * (double)N);
  // derived output d of stage input y at mode (kx, ky): i m_d y / N^2
  auto derived = [&](int d, int kx, int ky, double2 y) -> double2 {
    const bool dc = (kx == 0 && ky == 0);
    const double k2s = dc ? 1.0 : (double)(kx * kx + ky * ky);
    const double md = d == 0 ? (double)kx : d == 1 ? (double)ky : d == 2 ? (double)ky / k2s : -(double)kx / k2s;
    return cscale(md * invN2, cmuli(y));
  };
  double2 yv[TCH > 1 ? EPT : 1], dyv[TCH > 1 ? EPT : 1];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int idx = tid + q * NT;
    const int cy = idx % TY;
    const int r = idx / TY;
    const int ky = ky0 + cy;
    if (idx >= R * TY || ky >= KC) continue;
    const int kx = G::row_kx(r);
    const int xr = G::row_x(r);
    const size_t ci = (((size_t)b * 2 + c) * R + r) * KC + ky;
    double2 rh = make_double2(0.0, 0.0), drh = rh;
    if (FWD) {
      rh = sm[(0 * TY + cy) * SS + pidx(xr)];
      if (NFI == 2) drh = sm[(1 * TY + cy) * SS + pidx(xr)];
    }
    double2 y, dy;
    glue<N, MODE, TAN, AK, GPF>(a, wr, bk, pre[q], b, ci, r, c, kx, ky, rh, drh, y, dy);
    if (TCH > 1) {
      yv[q] = y;
      dyv[q] = dy;
    }
    if (HASINV && inv) {
      if (NDO == 1) {
        sm[(0 * TY + cy) * SS + pidx(xr)] = cscale(invN2, y);
      } else {
#pragma unroll
        for (int dl = 0; dl < DPC; ++dl) {
          const int d = sp * DPS + dl;
          sm[((dl * NFO + 0) * TY + cy) * SS + pidx(xr)] = derived(d, kx, ky, y);
          if (NFO == 2) sm[((dl * NFO + 1) * TY + cy) * SS + pidx(xr)] = derived(d, kx, ky, dy);
        }
      }
    }
  }
  if (!HASINV || !inv) return;
  constexpr int GAP = N - R;
  for (int ch = 0; ch < TCH; ++ch) {
    if (ch > 0) {
      __syncthreads();  // the previous chunk's sequences have been stored
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
        const int idx = tid + q * NT;
        const int cy = idx % TY;
        const int r = idx / TY;
        const int ky = ky0 + cy;
        if (idx >= R * TY || ky >= KC) continue;
        const int kx = G::row_kx(r);
        const int xr = G::row_x(r);
#pragma unroll
        for (int dl = 0; dl < DPC; ++dl) {
          const int d = sp * DPS + ch * DPC + dl;
          sm[((dl * NFO + 0) * TY + cy) * SS + pidx(xr)] = derived(d, kx, ky, yv[q]);
          if (NFO == 2) sm[((dl * NFO + 1) * TY + cy) * SS + pidx(xr)] = derived(d, kx, ky, dyv[q]);
        }
      }
    }
    for (int idx = tid; idx < NSEQ * GAP; idx += NT) {
      const int s = idx / GAP;
      const int g = idx - s * GAP;
      sm[s * SS + pidx(KMAX + 1 + g)] = make_double2(0.0, 0.0);
    }
    __syncthreads();
    // all passes but the last through LDS; the last in registers, its outputs
    // stored straight to the derived arrays (neighbouring lanes: neighbouring ky)
    fft_lds_head<N, NSEQ, SS, NT, true, P>(sm, twl);
    fft_lds_tail<N, NSEQ, SS, NT, true, P, TY>(sm, twl, [&](int s, int x, double2 v) {
      const int cy = s % TY, s2 = s / TY;
      const int ky = ky0 + cy;
      if (ky >= KC) return;
      const int d = sp * DPS + ch * DPC + s2 / NFO, f = s2 % NFO;
      const int g = (NDO == 1) ? c : (f * 8 + 4 * c + d);
      a.inter[(((size_t)b * ND + g) * N + x) * KP + ky] = v;
    });
  }
}

// Modes the column kernel with its middle in registers (col_kernel_fm) runs:
// primal, forward x-FFT, glue and four derived outputs transformed back.
template <int MODE>
constexpr bool col_fm_mode() {
  return MODE == C_S1F || MODE == C_S2 || MODE == C_S3 || MODE == C_S4 || MODE == C_A0 || MODE == C_A ||
         MODE == C_GEN;
}

// Column kernel for tiles of TY columns (primal, three-pass plan) with its
// middle in registers.  The forward x-FFT's last pass (radix RF, in place:
// butterfly jf reads and writes positions jf + r N/RF) and the inverse x-FFT's
// first pass (radix R0, butterfly j reads positions j + m N/R0) meet when one
// thread takes the G = R0/RF forward butterflies jf = j + g N/R0: it then holds
// the R0 positions p_m = j + m N/R0 (m = g + G r) of its column, applies the
// stage glue to them and, per derived output, the inverse's first butterfly,
// and writes only that butterfly's outputs to LDS.  The forward output never
// goes through LDS and the derived inputs are never read back from it: per
// block one LDS round trip and its barriers fewer for the forward transform
// and for each of the four derived outputs.  Every butterfly, twiddle and
// glue operation is the one col_kernel performs, in the same order, so the
// results are the same bits.  Positions in the dealiased gap are zero (slots
// m whose whole range lies in the gap are skipped at compile time).
template <int N, int TY, int NT, int MODE, int P, int AK>
__global__ __launch_bounds__(NT) void col_kernel_fm(Args a) {
  using G = Geo<N>;
  using TL = TwLayout<N, P>;
  constexpr int KC = G::KC, KP = G::KP, KMAX = G::KMAX;
  constexpr int RF = TL::R2, R0 = TL::R0, R1 = TL::R1;
  static_assert(RF > 1, "a plan of three passes");
  static_assert(R0 % RF == 0, "the forward's last radix must divide the inverse's first");
  static_assert(col_fm_mode<MODE>(), "mode without the register middle");
  constexpr int GG = R0 / RF;    // forward butterflies per work item
  constexpr int NI = N / R0;     // inverse first-pass butterflies per sequence
  constexpr int NSF = N / RF;    // the forward last pass's sub-transform length (its NS)
  static_assert(TY * NI <= NT, "one work item per thread");
  constexpr int SS = Seq<N>::S;
  __shared__ double2 sm[TY * SS];
  __shared__ double2 twl[N];
  TwFetch<N, NT, P> twf;
  twf.fetch(a.tw);
  int lx, ly, lz;
  xcd_tile(lx, ly, lz);
  const int b = lz, c = ly, ky0 = lx * TY, tid = threadIdx.x;
  Blk bk = prologue<MODE, false, AK>(a, b, lx == 0 && ly == 0);
  if (!bk.run) return;
  // work item: column cy, inverse butterfly j
  const int cy = tid % TY, j = tid / TY;
  const int ky = ky0 + cy;
  const bool act = tid < TY * NI && ky < KC;
  // slot m of the item may hold a kept mode (p_m <= KMAX or p_m >= N - KMAX for some j)
  auto maybe = [](int m) constexpr { return m * NI <= KMAX || m * NI + NI - 1 >= N - KMAX; };
  auto kept = [&](int m) { const int p = j + m * NI; return act && (p <= KMAX || p >= N - KMAX); };
  auto crow = [&](int m) { const int p = j + m * NI; return p <= KMAX ? p : p - N + G::R; };
  Pre pre[R0];
#pragma unroll
  for (int m = 0; m < R0; ++m) {
    if (!maybe(m) || !kept(m)) continue;
    const int r = crow(m);
    const size_t ci = (((size_t)b * 2 + c) * G::R + r) * KC + ky;
    pre[m] = prefetch<MODE, false, AK>(a, ci, (size_t)r * KC + ky, c);
  }
  const double2* O = a.rowout + (size_t)b * 2 * N * KP;
  fft_lds_gather_head<N, TY, SS, NT, false, P, (TY > 1)>(
      sm, twl,
      [&](int s, int x) -> double2 {
        const int kys = ky0 + s;
        return kys < KC ? O[((size_t)c * N + x) * KP + kys] : make_double2(0.0, 0.0);
      },
      [&] { twf.store(twl); });
  finish<AK>(bk);
  // the forward's last pass, G butterflies of this item's column, into registers
  double2 vf[R0];
  if (tid < TY * NI) {
    const double2* src = sm + cy * SS;
#pragma unroll
    for (int g = 0; g < GG; ++g) {
      const int jf = j + g * NI;
      double2 v[RF];
#pragma unroll
      for (int r = 0; r < RF; ++r) v[r] = src[pidx(jf + r * NSF)];
#pragma unroll
      for (int r = 1; r < RF; ++r) v[r] = cmul(v[r], twl[TL::OFF3 + (r - 1) * NSF + jf]);
      Dft<RF, false>::run(v);
#pragma unroll
      for (int r = 0; r < RF; ++r) vf[g + GG * r] = v[r];
    }
  }
  const bool inv = bk.stage && !a.skip_inv;
  constexpr double invN2 = 1.0 / ((double)N * (double)N);
  auto derived = [&](int d, int kx, int kyy, double2 y) -> double2 {
    const bool dc = (kx == 0 && kyy == 0);
    const double k2s = dc ? 1.0 : (double)(kx * kx + kyy * kyy);
    const double md = d == 0 ? (double)kx : d == 1 ? (double)kyy : d == 2 ? (double)kyy / k2s : -(double)kx / k2s;
    return cscale(md * invN2, cmuli(y));
  };
  // glue on the kept modes of the item
  double2 yv[R0];
#pragma unroll
  for (int m = 0; m < R0; ++m) {
    yv[m] = make_double2(0.0, 0.0);
    if (!maybe(m) || !kept(m)) continue;
    const int r = crow(m);
    const int kx = G::row_kx(r);
    const size_t ci = (((size_t)b * 2 + c) * G::R + r) * KC + ky;
    double2 dy;
    glue<N, MODE, false, AK>(a, true, bk, pre[m], b, ci, r, c, kx, ky, vf[m], make_double2(0.0, 0.0), yv[m], dy);
  }
  if (!inv) return;
  __syncthreads();  // every forward value is in registers before the buffer is rewritten
  for (int d = 0; d < 4; ++d) {
    if (d > 0) __syncthreads();  // the previous derived output's last pass has read the buffer
    if (tid < TY * NI) {
      // the inverse's first pass (no twiddles) on this item's butterfly j
      double2 u[R0];
#pragma unroll
      for (int m = 0; m < R0; ++m) {
        u[m] = make_double2(0.0, 0.0);
        if (maybe(m) && kept(m)) u[m] = derived(d, G::row_kx(crow(m)), ky, yv[m]);
      }
      Dft<R0, true>::run(u);
      double2* dst = sm + cy * SS;
#pragma unroll
      for (int m = 0; m < R0; ++m) dst[pidx(j * R0 + m)] = u[m];
    }
    __syncthreads();
    stockham_pass<N, R1, R0, TY, SS, NT, true, TL::OFF2>(sm, twl);
    fft_lds_tail<N, TY, SS, NT, true, P, TY>(sm, twl, [&](int s, int x, double2 v) {
      const int kys = ky0 + s;
      if (kys >= KC) return;
      a.inter[(((size_t)b * 8 + 4 * c + d) * N + x) * KP + kys] = v;
    });
  }
}

// -advection[0] + advection[1] and u x B (mhd_jax.py:152-163) at one point
// from the packed real fields (wx, wy), (u, v), (jx, jy), (Bx, By) and, with
// TAN, their linearisation along the tangent fields z4..z7
template <bool TAN>
__device__ __forceinline__ void rhs_products(double2 z0, double2 z1, double2 z2, double2 z3, double2 z4, double2 z5,
                                             double2 z6, double2 z7, double b0x, double b0y, double2& q, double2& dq) {
  const double wx = z0.x, wy = z0.y, u = z1.x, v = z1.y, jx = z2.x, jy = z2.y;
  const double Bx = z3.x + b0x, By = z3.y + b0y;
  q = make_double2(-(u * wx + v * wy) + (Bx * jx + By * jy), u * By - v * Bx);
  if (TAN) {
    const double twx = z4.x, twy = z4.y, tu = z5.x, tv = z5.y, tjx = z6.x, tjy = z6.y, ta = z7.x, tb = z7.y;
    dq = make_double2(-(tu * wx + u * twx + tv * wy + v * twy) + (ta * jx + Bx * tjx + tb * jy + By * tjy),
                      tu * By + u * tb - tv * Bx - v * ta);
  }
}

// Row pass.  With the narrow plan (P = 8) the RHS row pass is fused across its
// middle: the inverse transform's last pass, the products and the forward
// transform's first pass run in registers, one thread per row position j and
// all of the row's sequences, and the forward transform takes the narrow plan
// reversed (P = 9), whose first radix is the inverse's last.  That saves the
// LDS round trips and barriers of three steps; the forward arithmetic is that
// of the reversed plan, the same for every tile and batch.
template <int N, int TX, int NT, int RMODE, bool TAN, int P, int AK>
__global__ __launch_bounds__(NT) void row_kernel(Args a) {
  using G = Geo<N>;
  constexpr int KC = G::KC, KP = G::KP;
  constexpr int NP = (RMODE == R_ERR) ? 1 : (TAN ? 8 : 4);
  constexpr int NF = TAN ? 2 : 1;
  constexpr int ND = (RMODE == R_ERR) ? 2 : 8 * NF;
  constexpr int NHALF = N / 2 + 1;
  constexpr int SS = Seq<N>::S;
  constexpr bool FUSED = P == 8 && RMODE == R_RHS;
  constexpr int PF = FUSED ? 9 : P;
  __shared__ double2 sm[TX * NP * SS];
  __shared__ double2 twl[N];
  TwFetch<N, NT, P> twf;
  twf.fetch(a.tw);
  [[maybe_unused]] TwFetch<N, NT, PF> twff;
  if constexpr (FUSED) twff.fetch(a.tw);
  int lx, ly, lz;
  xcd_tile(lx, ly, lz);
  const int b = lz, xbase = lx * TX, tid = threadIdx.x;
  if (!row_active<AK>(a, b)) return;
  const double2* I = a.inter + (size_t)b * ND * N * KP;
  // Hermitian extension of the packed pairs (A + iB), formed while the first
  // inverse-FFT pass loads its inputs from I; the imaginary parts of the ky = 0
  // bin are dropped, as numpy's c2r does (ky = N/2 lies in the dealiased gap).
  // Every lane loads (gap lanes re-read bin 0 and discard it), so all loads of
  // the pass are in flight before the first wait.
  auto gather = [&](int s, int m) -> double2 {
    const int p = s % NP;
    const int x = xbase + s / NP;
    const bool lo = m <= G::KMAX, hi = m >= N - G::KMAX;
    const int k = lo ? m : (hi ? N - m : 0);
    const double2 A = I[((size_t)(2 * p) * N + x) * KP + k];
    const double2 B = I[((size_t)(2 * p + 1) * N + x) * KP + k];
    const double2 vlo = m == 0 ? make_double2(A.x, B.x) : make_double2(A.x - B.y, A.y + B.x);
    const double2 vhi = make_double2(A.x + B.y, B.x - A.y);
    return lo ? vlo : (hi ? vhi : make_double2(0.0, 0.0));
  };
  if constexpr (FUSED)
    fft_lds_gather_head<N, TX * NP, SS, NT, true, P>(sm, twl, gather, [&] { twf.store(twl); });
  else
    fft_lds_gather<N, TX * NP, SS, NT, true, P>(sm, twl, gather, [&] { twf.store(twl); });

  if (RMODE == R_ERR) {
    unsigned long long mx = 0ull;
    for (int idx = tid; idx < TX * N; idx += NT) {
      const double2 z = sm[(idx / N) * SS + pidx(idx % N)];
      const unsigned long long u1 = (unsigned long long)__double_as_longlong(fabs(z.x));
      const unsigned long long u2 = (unsigned long long)__double_as_longlong(fabs(z.y));
      mx = u1 > mx ? u1 : mx;
      mx = u2 > mx ? u2 : mx;
    }
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(mx, off, 64);
      mx = o > mx ? o : mx;
    }
    __shared__ unsigned long long red[NT / 64];
    if ((tid & 63) == 0) red[tid >> 6] = mx;
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < NT / 64; ++w) mx = red[w] > mx ? red[w] : mx;
      atomicMax(&a.hist[(size_t)b * a.hstride + attempt_of(a)].err, mx);
    }
    return;
  }

  const bool nob0 = a.b0_from > 0 && b >= a.b0_from;
  const double b0x = nob0 ? 0.0 : a.b0x, b0y = nob0 ? 0.0 : a.b0y;
  const double2 zz = make_double2(0.0, 0.0);
  if constexpr (FUSED) {
    using TL = TwLayout<N, P>;
    using TF = TwLayout<N, PF>;
    constexpr int RL = TL::R2 > 1 ? TL::R2 : TL::R1;        // the inverse's last radix
    constexpr int TWL = TL::R2 > 1 ? TL::OFF3 : TL::OFF2;   // its twiddle table
    constexpr int NSL = N / RL;                              // its sub-transform length
    static_assert(TF::R0 == RL, "the reversed plan starts with the inverse's last radix");
    constexpr int NPAIR = TX * NSL;
    constexpr int EPP = (NPAIR + NT - 1) / NT;
    double2 fq[EPP][RL], fdq[EPP][TAN ? RL : 1];
#pragma unroll
    for (int e = 0; e < EPP; ++e) {
      const int pair = tid + e * NT;
      if (pair < NPAIR) {
        const int xl = pair / NSL, j = pair - xl * NSL;
        // the inverse's last pass over the row's sequences: outputs at
        // positions j + r NSL (the pass's output index, with k = j)
        double2 z[NP][RL];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          const double2* src = sm + (xl * NP + p) * SS;
#pragma unroll
          for (int r = 0; r < RL; ++r) z[p][r] = src[pidx(j + r * NSL)];
#pragma unroll
          for (int r = 1; r < RL; ++r) {
            double2 w = twl[TWL + (r - 1) * NSL + j];
            w.y = -w.y;
            z[p][r] = cmul(z[p][r], w);
          }
          Dft<RL, true>::run(z[p]);
        }
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          double2 dqv = zz;
          if constexpr (TAN)
            rhs_products<true>(z[0][r], z[1][r], z[2][r], z[3][r], z[4][r], z[5][r], z[6][r], z[7][r], b0x, b0y,
                               fq[e][r], dqv);
          else
            rhs_products<false>(z[0][r], z[1][r], z[2][r], z[3][r], zz, zz, zz, zz, b0x, b0y, fq[e][r], dqv);
          if (TAN) fdq[e][TAN ? r : 0] = dqv;
        }
        // the forward's first pass (inputs at j + r NSL = j + r N / RL)
        Dft<RL, false>::run(fq[e]);
        if constexpr (TAN) Dft<RL, false>::run(fdq[e]);
      }
    }
    __syncthreads();  // every read of the inverse's sequences and twiddles is done
#pragma unroll
    for (int e = 0; e < EPP; ++e) {
      const int pair = tid + e * NT;
      if (pair < NPAIR) {
        const int xl = pair / NSL, j = pair - xl * NSL;
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          sm[(xl * NF + 0) * SS + pidx(j * RL + r)] = fq[e][r];
          if constexpr (TAN) sm[(xl * NF + 1) * SS + pidx(j * RL + r)] = fdq[e][r];
        }
      }
    }
    twff.store(twl);
    __syncthreads();
    if constexpr (TF::R1 > 1) stockham_pass<N, TF::R1, TF::R0, TX * NF, SS, NT, false, TF::OFF2>(sm, twl);
    if constexpr (TF::R2 > 1) stockham_pass<N, TF::R2, TF::R0 * TF::R1, TX * NF, SS, NT, false, TF::OFF3>(sm, twl);
  } else {
    constexpr int EPT = (TX * N + NT - 1) / NT;
    double2 q[EPT], dq[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int idx = tid + e * NT;
      if (idx < TX * N) {
        const int xl = idx / N, yy = idx - xl * N;
        const double2* zb = sm + xl * NP * SS + pidx(yy);
        if constexpr (TAN)
          rhs_products<true>(zb[0], zb[SS], zb[2 * SS], zb[3 * SS], zb[4 * SS], zb[5 * SS], zb[6 * SS], zb[7 * SS],
                             b0x, b0y, q[e], dq[e]);
        else
          rhs_products<false>(zb[0], zb[SS], zb[2 * SS], zb[3 * SS], zz, zz, zz, zz, b0x, b0y, q[e], dq[e]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int idx = tid + e * NT;
      if (idx < TX * N) {
        const int xl = idx / N, yy = idx - xl * N;
        sm[(xl * NF + 0) * SS + pidx(yy)] = q[e];
        if (TAN) sm[(xl * NF + 1) * SS + pidx(yy)] = dq[e];
      }
    }
    __syncthreads();
    fft_lds<N, TX * NF, SS, NT, false, P>(sm, twl);
  }
  double2* O = a.rowout + (size_t)b * (2 * NF) * N * KP;
  for (int idx = tid; idx < TX * NF * KC; idx += NT) {
    const int k = idx % KC;
    const int t = idx / KC;
    const int f = t % NF;
    const int xl = t / NF;
    const int x = xbase + xl;
    const double2* Z = sm + (xl * NF + f) * SS;
    const double2 z = Z[pidx(k)], zm = Z[pidx((N - k) & (N - 1))];
    const double2 pw = make_double2(0.5 * (z.x + zm.x), 0.5 * (z.y - zm.y));
    const double2 pj = make_double2(0.5 * (z.y + zm.y), -0.5 * (z.x - zm.x));
    O[((size_t)(f * 2 + 0) * N + x) * KP + k] = pw;
    O[((size_t)(f * 2 + 1) * N + x) * KP + k] = pj;
  }
}

// full [B][2][N][N/2+1] <-> compact [B][2][R][KC]
template <int N>
__global__ void pack_kernel(const double2* __restrict__ full, double2* __restrict__ comp, int B) {
  using G = Geo<N>;
  const size_t tot = (size_t)B * 2 * G::R * G::KC;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const int ky = (int)(i % G::KC);
    const size_t t = i / G::KC;
    const int r = (int)(t % G::R);
    const size_t bc = t / G::R;
    comp[i] = full[(bc * N + G::row_x(r)) * G::NH + ky];
  }
}

template <int N>
__global__ void unpack_kernel(const double2* __restrict__ comp, double2* __restrict__ full, int B, int zero_outside) {
  using G = Geo<N>;
  if (zero_outside) {
    const size_t tot = (size_t)B * 2 * N * G::NH;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
      const int ky = (int)(i % G::NH);
      const size_t t = i / G::NH;
      const int x = (int)(t % N);
      const size_t bc = t / N;
      double2 v = make_double2(0.0, 0.0);
      const int r = x <= G::KMAX ? x : (x >= N - G::KMAX ? x - N + G::R : -1);
      if (r >= 0 && ky < G::KC) v = comp[(bc * G::R + r) * G::KC + ky];
      full[i] = v;
    }
  } else {
    const size_t tot = (size_t)B * 2 * G::R * G::KC;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
      const int ky = (int)(i % G::KC);
      const size_t t = i / G::KC;
      const int r = (int)(t % G::R);
      const size_t bc = t / G::R;
      full[(bc * N + G::row_x(r)) * G::NH + ky] = comp[i];
    }
  }
}

}  // namespace ecs
