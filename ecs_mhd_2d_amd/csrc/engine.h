// Grid-size-templated orchestration of the fused kernels: launches, the
// autotuner, and the stepper / tangent / adjoint drivers behind the C-ABI.
// Included once per supported N by grid_n*.hip, which instantiate it through
// ECS_GRID(N).
#pragma once
#include "rhs_kernels.h"
#include "adjoint_kernels.h"
#include "comb_kernels.h"
#include "td_kernels.h"
#include "host.h"

#include <functional>

// ------------------------------------------------------------------ launches
// default variant indices (column: TY = 1 << v, row: TX = 1 << v; wide plan)
template <int N> struct DefTile;
template <> struct DefTile<16> { static constexpr int COL = 1, ROW = 1; };
template <> struct DefTile<32> { static constexpr int COL = 2, ROW = 1; };
template <> struct DefTile<64> { static constexpr int COL = 2, ROW = 1; };
template <> struct DefTile<128> { static constexpr int COL = 1, ROW = 0; };
template <> struct DefTile<256> { static constexpr int COL = 1, ROW = 0; };

template <int N, int TY, int MODE, bool TAN, int P, int AK, int SPL = 1>
static int col_v(const Args& a, int B, hipStream_t st) {
  // modes that cannot split run unsplit under a split variant, one column per
  // block: the split variants serve small batches, which want the most blocks
  constexpr int SP = col_splittable<MODE>() ? SPL : 1;
  constexpr int T = (SPL > 1 && SP == 1) ? 1 : TY;
  dim3 grid((Geo<N>::KC + T - 1) / T, 2 * SP, B);
  hipLaunchKernelGGL((col_kernel<N, T, NT, MODE, TAN, P, AK, SP>), grid, dim3(NT), 0, st, a);
  LAUNCHCHK();
  return 0;
}
template <int N, int TX, int RMODE, bool TAN, int P, int AK>
static int row_v(const Args& a, int B, hipStream_t st) {
  dim3 grid(N / TX, 1, B);
  hipLaunchKernelGGL((row_kernel<N, TX, NT, RMODE, TAN, P, AK>), grid, dim3(NT), 0, st, a);
  LAUNCHCHK();
  return 0;
}
// Variants (var_* in host.h): v < 6: tile 1 << (v % 3), plan wide (v < 3) or
// narrow; column kernels also v = 6..9: one column per block, its output
// sequences split over 2 (even v) or 4 (odd v) blocks, plan wide (6, 7) or
// narrow (8, 9); for N >= 128 v = 10, 11: tiles of 8 columns, plan wide (10)
// or narrow (11); v = 12..15: tiles of 2 columns split over 2 (12, 14) or 4
// (13, 15) blocks, plan wide (12, 13) or narrow (14, 15); for N >= 128 and the
// primal family v = 16: tiles of 8 columns, narrow plan, the middle in
// registers (col_kernel_fm; variant 11 for the modes it does not cover).
template <int N, int MODE, bool TAN, int AK>
static int col_var(int v, const Args& a, int B, hipStream_t st) {
  switch (v) {
    case 0: return col_v<N, 1, MODE, TAN, 16, AK>(a, B, st);
    case 1: return col_v<N, 2, MODE, TAN, 16, AK>(a, B, st);
    case 2: return col_v<N, 4, MODE, TAN, 16, AK>(a, B, st);
    case 3: return col_v<N, 1, MODE, TAN, 8, AK>(a, B, st);
    case 4: return col_v<N, 2, MODE, TAN, 8, AK>(a, B, st);
    case 5: return col_v<N, 4, MODE, TAN, 8, AK>(a, B, st);
    case 6: return col_v<N, 1, MODE, TAN, 16, AK, 2>(a, B, st);
    case 7: return col_v<N, 1, MODE, TAN, 16, AK, 4>(a, B, st);
    case 8: return col_v<N, 1, MODE, TAN, 8, AK, 2>(a, B, st);
    case 9: return col_v<N, 1, MODE, TAN, 8, AK, 4>(a, B, st);
    default:
      if constexpr (N >= 128) {
        if (v == 10) return col_v<N, 8, MODE, TAN, 16, AK>(a, B, st);
        if (v == 11) return col_v<N, 8, MODE, TAN, 8, AK>(a, B, st);
        if (v == 16) {
          // tiles of 8 columns, narrow plan, the middle in registers where the
          // mode allows it (the same arithmetic as variant 11 otherwise)
          if constexpr (!TAN && col_fm_mode<MODE>()) {
            dim3 grid((Geo<N>::KC + 7) / 8, 2, B);
            hipLaunchKernelGGL((col_kernel_fm<N, 8, NT, MODE, 8, AK>), grid, dim3(NT), 0, st, a);
            LAUNCHCHK();
            return 0;
          } else {
            return col_v<N, 8, MODE, TAN, 8, AK>(a, B, st);
          }
        }
      }
      switch (v) {
        case 12: return col_v<N, 2, MODE, TAN, 16, AK, 2>(a, B, st);
        case 13: return col_v<N, 2, MODE, TAN, 16, AK, 4>(a, B, st);
        case 14: return col_v<N, 2, MODE, TAN, 8, AK, 2>(a, B, st);
        case 15: return col_v<N, 2, MODE, TAN, 8, AK, 4>(a, B, st);
        default: break;
      }
      return fail(1, "column variant " + std::to_string(v) + " does not exist for N = " + std::to_string(N));
  }
}
template <int N, int RMODE, bool TAN, int AK>
static int row_var(int v, const Args& a, int B, hipStream_t st) {
  switch (v) {
    case 0: return row_v<N, 1, RMODE, TAN, 16, AK>(a, B, st);
    case 1: return row_v<N, 2, RMODE, TAN, 16, AK>(a, B, st);
    case 2: return row_v<N, 4, RMODE, TAN, 16, AK>(a, B, st);
    case 3: return row_v<N, 1, RMODE, TAN, 8, AK>(a, B, st);
    case 4: return row_v<N, 2, RMODE, TAN, 8, AK>(a, B, st);
    default: return row_v<N, 4, RMODE, TAN, 8, AK>(a, B, st);
  }
}

// Stepper kinds each kernel mode runs under: fixed-step passes (RHS, classical
// and Lawson RK4 stages), the adaptive primal (no tangent) and the tangent
// replay of an adaptive schedule.  Only these instances are compiled.
template <int MODE, bool TAN, int AK>
constexpr bool col_kind_ok() {
  if (AK == AD_FIXED)
    return MODE == C_LOAD || MODE == C_RHS || MODE == C_S1F || MODE == C_S2 || MODE == C_S3 || MODE == C_S4 ||
           MODE == C_GEN || (MODE == C_TD1 && !TAN) || (MODE == C_TD2 && TAN);
  if (AK == AD_PRIMAL)
    return !TAN && (MODE == C_S2 || MODE == C_S3 || MODE == C_S4 || MODE == C_S5 || MODE == C_A0 ||
                    MODE == C_A || MODE == C_FIN);
  return TAN && (MODE == C_S2 || MODE == C_S3 || MODE == C_S4 || MODE == C_S5 || MODE == C_S5A ||
                 MODE == C_A0 || MODE == C_A);
}
template <int RMODE, bool TAN, int AK>
constexpr bool row_kind_ok() {
  if (RMODE == R_ERR) return AK == AD_PRIMAL && !TAN;
  if (AK == AD_PRIMAL) return !TAN;
  if (AK == AD_TANGENT) return TAN;
  return true;
}
template <int N, int MODE, bool TAN>
static int col(const Args& a, int B, hipStream_t st) {
  const int v = tuned(N, !TAN ? K_COL : a.adaptive == AD_TANGENT ? K_COLR : K_COLT, B, DefTile<N>::COL);
  switch (a.adaptive) {
    case AD_FIXED:
      if constexpr (col_kind_ok<MODE, TAN, AD_FIXED>()) return col_var<N, MODE, TAN, AD_FIXED>(v, a, B, st);
      break;
    case AD_PRIMAL:
      if constexpr (col_kind_ok<MODE, TAN, AD_PRIMAL>()) return col_var<N, MODE, TAN, AD_PRIMAL>(v, a, B, st);
      break;
    case AD_TANGENT:
      if constexpr (col_kind_ok<MODE, TAN, AD_TANGENT>()) return col_var<N, MODE, TAN, AD_TANGENT>(v, a, B, st);
      break;
    default:
      break;
  }
  return fail(3, "column kernel mode " + std::to_string(MODE) + " has no instance for stepper kind " +
                     std::to_string(a.adaptive));
}
template <int N, int RMODE, bool TAN>
static int row(const Args& a, int B, hipStream_t st) {
  const int v = tuned(N, !TAN ? K_ROW : a.adaptive == AD_TANGENT ? K_ROWR : K_ROWT, B, DefTile<N>::ROW);
  switch (a.adaptive) {
    case AD_FIXED:
      if constexpr (row_kind_ok<RMODE, TAN, AD_FIXED>()) return row_var<N, RMODE, TAN, AD_FIXED>(v, a, B, st);
      break;
    case AD_PRIMAL:
      if constexpr (row_kind_ok<RMODE, TAN, AD_PRIMAL>()) return row_var<N, RMODE, TAN, AD_PRIMAL>(v, a, B, st);
      break;
    case AD_TANGENT:
      if constexpr (row_kind_ok<RMODE, TAN, AD_TANGENT>()) return row_var<N, RMODE, TAN, AD_TANGENT>(v, a, B, st);
      break;
    default:
      break;
  }
  return fail(3, "row kernel mode " + std::to_string(RMODE) + " has no instance for stepper kind " +
                     std::to_string(a.adaptive));
}

template <int N>
static int pack(const double* full, double2* comp, int B, hipStream_t st) {
  const size_t tot = (size_t)B * 2 * Geo<N>::R * Geo<N>::KC;
  int blocks = (int)std::min<size_t>((tot + 255) / 256, 4096);
  hipLaunchKernelGGL((pack_kernel<N>), dim3(blocks), dim3(256), 0, st, (const double2*)full, comp, B);
  LAUNCHCHK();
  return 0;
}
template <int N>
static int unpack(const double2* comp, double* full, int B, int zero_outside, hipStream_t st) {
  const size_t tot = zero_outside ? (size_t)B * 2 * N * Geo<N>::NH : (size_t)B * 2 * Geo<N>::R * Geo<N>::KC;
  int blocks = (int)std::min<size_t>((tot + 255) / 256, 4096);
  hipLaunchKernelGGL((unpack_kernel<N>), dim3(blocks), dim3(256), 0, st, comp, (double2*)full, B, zero_outside);
  LAUNCHCHK();
  return 0;
}

// ------------------------------------------------------------------ autotuner
// Average time (ms) of variants vs[0..nv) of one kernel family at batch B,
// timed with HIP events on `st` over back-to-back launches on zeroed
// workspace (the workspace must hold B candidates).
// The replay kinds are timed in pairs as the replay runs them, the column
// kernel then the row kernel, with `partner` the variant of the other one:
// launched alone, one after another, variants can rank differently (at
// N = 256, batch 1 the replay's tangent column variants 3 and 14).
template <int N>
static int time_variants(ecs_ctx* c, int B, int kind, const int* vs, int nv, float* t, hipStream_t st,
                         int partner = 0) {
  const size_t ce = c->compact_elems() * B * sizeof(double2);
  std::vector<double> dt(B, 1e-3);
  for (int slot = 0; slot < 4; ++slot) TRY(upload_scalars(c, slot, dt.data(), B, st));
  Args a = base_args(c);
  // the replay instances: every candidate active, attempt 0, its step size
  // read from the schedule (here the B step sizes of scalar slot 0)
  Args ar = a;
  ar.adaptive = AD_TANGENT;
  ar.run_all = 1;
  ar.next_all = 1;
  ar.attempt = 0;
  ar.s_h = c->scal;
  ar.sstride = 1;
  HIPCHK(hipMemsetAsync(c->x0, 0, ce, st));
  HIPCHK(hipMemsetAsync(c->dx0, 0, ce, st));
  HIPCHK(hipMemsetAsync(c->acc, 0, ce, st));
  HIPCHK(hipMemsetAsync(c->dacc, 0, ce, st));
  TRY((col_var<N, C_LOAD, true, AD_FIXED>(1, a, B, st)));
  TRY((row_var<N, R_RHS, true, AD_FIXED>(1, a, B, st)));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // three interleaved rounds over the variants, keeping each variant's
  // fastest round: one slow round (a clock change, another process) then
  // cannot decide the choice
  const int reps = std::max(8, 256 / B);
  int rc = 0;
  for (int i = 0; i < nv; ++i) t[i] = 1e30f;
  for (int round = 0; round < 3 && !rc; ++round) {
    for (int iv = 0; iv < nv && !rc; ++iv) {
      const int v = vs[iv];
      auto launch = [&]() -> int {
        switch (kind) {
          case K_COL: return col_var<N, C_S2, false, AD_FIXED>(v, a, B, st);
          case K_COLT: return col_var<N, C_S2, true, AD_FIXED>(v, a, B, st);
          case K_ROW: return row_var<N, R_RHS, false, AD_FIXED>(v, a, B, st);
          case K_ROWT: return row_var<N, R_RHS, true, AD_FIXED>(v, a, B, st);
          case K_COLR:
            TRY((col_var<N, C_S2, true, AD_TANGENT>(v, ar, B, st)));
            return row_var<N, R_RHS, true, AD_TANGENT>(partner, ar, B, st);
          default:
            TRY((col_var<N, C_S2, true, AD_TANGENT>(partner, ar, B, st)));
            return row_var<N, R_RHS, true, AD_TANGENT>(v, ar, B, st);
        }
      };
      rc = launch();
      for (int i = 0; i < reps && !rc; ++i) {
        if (i == 0 && hipEventRecord(e0, st) != hipSuccess) rc = fail(2, "hipEventRecord");
        if (!rc) rc = launch();
      }
      if (!rc && hipEventRecord(e1, st) != hipSuccess) rc = fail(2, "hipEventRecord");
      if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = fail(2, "hipEventSynchronize");
      float ms = 0.f;
      if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(2, "hipEventElapsedTime");
      t[iv] = std::min(t[iv], ms / reps);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// Fixes, per kernel family (column, tangent column, row, tangent row), the
// FFT radix plan of a grid size and picks the tile per batch bucket (powers
// of two), for the tangent families once in the fixed-step instance and once
// in the tangent replay's.  The plan decides the arithmetic, so it comes from a fixed table
// (plan_table: the plan that timed fastest on MI355X at batch 1 and 16), and
// every process, batch size and run computes the same bits; a candidate's
// result does not depend on the batch it rides in or on timing noise.  Tiles
// and splits do not change the arithmetic and are timed per bucket within the
// fixed plan.  Runs at the start of an entry point, before any input is
// packed (it overwrites the workspace and may grow it to 16 candidates).
// ECSMHD_NO_TUNE=1 keeps the default tiles in the table's plans;
// ECSMHD_PLAN=wide|narrow fixes the plan of every family; ECSMHD_PLAN=tune
// chooses it by timing (both plans at batch 1 and 16, the plan with the
// smaller worst-case slowdown wins), which is how the table was measured;
// ECSMHD_VARIANTS="N:c,ct,r,rt[;N:...]" fixes the variants (var_* in host.h)
// of the four families for grid N at every batch (the replay instances take
// the tangent families' ct and rt), so a profiled run can use the
// configuration an unprofiled run tuned.
static bool forced_variants(int n, int v[4]) {
  const char* s = getenv("ECSMHD_VARIANTS");
  while (s && *s) {
    int nn, a, b, c, d;
    if (sscanf(s, "%d:%d,%d,%d,%d", &nn, &a, &b, &c, &d) == 5 && nn == n) {
      const int t[4] = {a, b, c, d};
      for (int k = 0; k < 4; ++k) {
        if (!var_ok(k, n, t[k])) return false;
        v[k] = t[k];
      }
      return true;
    }
    s = strchr(s, ';');
    if (s) ++s;
  }
  return false;
}

template <int N>
static int autotune(ecs_ctx* c, int B, hipStream_t st) {
  if (getenv("ECSMHD_NO_TUNE")) return 0;
  int fv[4];
  if (forced_variants(N, fv)) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (int kind = 0; kind < 4; ++kind) {
      g_plan[plan_key(N, kind)] = var_plan(fv[kind]);
      g_tune[tune_key(N, kind, B)] = fv[kind];
    }
    g_tune[tune_key(N, K_COLR, B)] = fv[K_COLT];
    g_tune[tune_key(N, K_ROWR, B)] = fv[K_ROWT];
    return 0;
  }
  static std::mutex run_mu;  // one tuning pass per grid size at a time
  std::lock_guard<std::mutex> rl(run_mu);
  const char* force = getenv("ECSMHD_PLAN");
  const bool timed = force && strcmp(force, "tune") == 0;
  const int forced = force && !timed ? (strcmp(force, "narrow") == 0 ? 1 : 0) : -1;
  int plan[4] = {-1, -1, -1, -1};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (g_tune.count(tune_key(N, K_ROWR, B))) return 0;
    for (int kind = 0; kind < 4; ++kind) {
      auto it = g_plan.find(plan_key(N, kind));
      if (it != g_plan.end()) plan[kind] = it->second;
      if (plan[kind] < 0 && !timed) {
        plan[kind] = forced >= 0 ? forced : plan_table(N, kind);
        g_plan[plan_key(N, kind)] = plan[kind];
      }
    }
  }
  if (plan[0] < 0 || plan[1] < 0 || plan[2] < 0 || plan[3] < 0) {
    constexpr int PB[2] = {1, 16};
    TRY(ensure_cap(c, PB[1]));
    float t[2][4][VMAX];
    for (int p = 0; p < 2; ++p)
      for (int kind = 0; kind < 4; ++kind) {
        int vs[VMAX];
        float tl[VMAX];
        const int nv = var_list(kind, N, -1, vs);
        TRY(time_variants<N>(c, PB[p], kind, vs, nv, tl, st));
        for (int v = 0; v < VMAX; ++v) t[p][kind][v] = 1e30f;
        for (int i = 0; i < nv; ++i) t[p][kind][vs[i]] = tl[i];
      }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (int kind = 0; kind < 4; ++kind) {
      int pl = plan[kind];
      if (pl < 0) pl = forced;
      if (pl < 0) {
        float worst[2] = {0.f, 0.f};
        for (int p = 0; p < 2; ++p) {
          float bp[2] = {1e30f, 1e30f};
          for (int v = 0; v < VMAX; ++v)
            if (var_ok(kind, N, v)) bp[var_plan(v)] = std::min(bp[var_plan(v)], t[p][kind][v]);
          for (int q = 0; q < 2; ++q) worst[q] = std::max(worst[q], bp[q] / std::min(bp[0], bp[1]));
        }
        pl = worst[1] < worst[0] ? 1 : 0;
      }
      plan[kind] = pl;
      g_plan[plan_key(N, kind)] = pl;
      for (int p = 0; p < 2; ++p) {
        int v = -1;
        for (int q = 0; q < VMAX; ++q)
          if (var_ok(kind, N, q) && var_plan(q) == pl && (v < 0 || t[p][kind][q] < t[p][kind][v])) v = q;
        if (!g_tune.count(tune_key(N, kind, PB[p]))) g_tune[tune_key(N, kind, PB[p])] = v;
      }
    }
    if (g_tune.count(tune_key(N, K_ROWR, B))) return 0;
  }
  int pick[NKIND];
  for (int kind = 0; kind < NKIND; ++kind) {
    int vs[VMAX];
    float t[VMAX];
    const int nv = var_list(kind, N, plan[base_kind(kind)], vs);
    // the replay column kernel paired with the fixed-step tangent row pick,
    // the replay row kernel with the replay column pick
    const int partner = kind == K_COLR ? pick[K_ROWT] : kind == K_ROWR ? pick[K_COLR] : 0;
    TRY(time_variants<N>(c, B, kind, vs, nv, t, st, partner));
    int best = 0;
    for (int i = 1; i < nv; ++i)
      if (t[i] < t[best]) best = i;
    pick[kind] = vs[best];
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int kind = 0; kind < NKIND; ++kind) g_tune[tune_key(N, kind, B)] = pick[kind];
  return 0;
}

// ------------------------------------------------------------------ state_vel
template <int N>
static int run_state_vel(ecs_ctx* c, const double* in, const double* din, double* out, double* dout, int B,
                         int diss, hipStream_t st) {
  TRY(ensure_cap(c, B));
  TRY(autotune<N>(c, B, st));
  Args a = base_args(c);
  a.include_diss = diss;
  TRY(pack<N>(in, c->x0, B, st));
  if (din) {
    TRY(pack<N>(din, c->dx0, B, st));
    TRY((col<N, C_LOAD, true>(a, B, st)));
    TRY((row<N, R_RHS, true>(a, B, st)));
    TRY((col<N, C_RHS, true>(a, B, st)));
    TRY(unpack<N>(c->dout, dout, B, 1, st));
  } else {
    TRY((col<N, C_LOAD, false>(a, B, st)));
    TRY((row<N, R_RHS, false>(a, B, st)));
    TRY((col<N, C_RHS, false>(a, B, st)));
  }
  TRY(unpack<N>(c->out, out, B, 1, st));
  return 0;
}

// state_vel with its workspace in caller-provided scratch (B candidates):
// [x0: B comp][out: B comp][I: B 8 N KP][O: B 2 N KP]; no autotuning, no
// allocation.  The three kernels (C_LOAD, R_RHS, C_RHS) and pack/unpack read
// and write only these arrays, the twiddles and the forcing spectrum.  Kernel variants are the tuned ones where a call of the same
// batch bucket has tuned them, else the default tile in the grid size's fixed
// FFT plan, so the arithmetic is that of run_state_vel.
template <int N>
static int run_state_vel_noalloc(ecs_ctx* c, const double* in, double* out, int B, int diss, double2* scr,
                                 size_t scr_elems, hipStream_t st) {
  using G = Geo<N>;
  const size_t comp = c->compact_elems();
  const size_t per = 2 * comp + (size_t)10 * N * G::KP;
  if ((size_t)B * per > scr_elems)
    return fail(1, "ecs_state_vel_noalloc: batch " + std::to_string(B) + " at N=" + std::to_string(N) +
                       " needs " + std::to_string((size_t)B * per * sizeof(double2)) + " bytes of scratch; the limit is " +
                       std::to_string(scr_elems * sizeof(double2)) + " bytes (at most " +
                       std::to_string(scr_elems / per) + " candidates)");
  Args a = base_args(c);
  a.include_diss = diss;
  // the context's scalar table may hold fewer than B candidates (or none);
  // these kernels take no step size
  a.dt = nullptr;
  a.ddt = nullptr;
  a.x0 = scr;
  a.out = scr + (size_t)B * comp;
  a.inter = scr + (size_t)2 * B * comp;
  a.rowout = a.inter + (size_t)B * 8 * N * G::KP;
  TRY(pack<N>(in, a.x0, B, st));
  TRY((col<N, C_LOAD, false>(a, B, st)));
  TRY((row<N, R_RHS, false>(a, B, st)));
  TRY((col<N, C_RHS, false>(a, B, st)));
  TRY(unpack<N>(a.out, out, B, 1, st));
  return 0;
}

// ------------------------------------------------------------------ Lawson RK4
template <int N, bool TAN>
static int rk4_step(const Args& a, int B, hipStream_t st, bool last) {
  TRY((row<N, R_RHS, TAN>(a, B, st)));
  TRY((col<N, C_S1F, TAN>(a, B, st)));
  TRY((row<N, R_RHS, TAN>(a, B, st)));
  TRY((col<N, C_S2, TAN>(a, B, st)));
  TRY((row<N, R_RHS, TAN>(a, B, st)));
  TRY((col<N, C_S3, TAN>(a, B, st)));
  TRY((row<N, R_RHS, TAN>(a, B, st)));
  Args a4 = a;
  a4.skip_inv = last ? 1 : 0;
  TRY((col<N, C_S4, TAN>(a4, B, st)));
  return 0;
}

template <int N, bool TAN>
static int run_rk4(ecs_ctx* c, double* f, double* df, int B, const double* dt, const double* ddt, int steps,
                   int masked, hipStream_t st) {
  TRY(ensure_cap(c, B));
  TRY(autotune<N>(c, B, st));
  Args a = base_args(c);
  a.masked_e = masked;
  TRY(upload_scalars(c, 0, dt, B, st));
  if (TAN) TRY(upload_scalars(c, 1, ddt, B, st));
  TRY(pack<N>(f, c->x0, B, st));
  if (TAN) TRY(pack<N>(df, c->dx0, B, st));
  if (steps > 0) {
    TRY((col<N, C_LOAD, TAN>(a, B, st)));
    // full chunks of K steps replay one captured graph; the tail runs eagerly
    constexpr int K = 16;
    int s = 0;
    if (use_graphs() && steps - 1 >= K) {
      const std::string key = graph_key(TAN ? "rk4t" : "rk4", a, B, K);
      for (; s + K <= steps - 1; s += K) {
        TRY(graph_run(c, key, st, [&](hipStream_t gs) -> int {
          for (int j = 0; j < K; ++j) TRY((rk4_step<N, TAN>(a, B, gs, false)));
          return 0;
        }));
      }
    }
    for (; s < steps; ++s) TRY((rk4_step<N, TAN>(a, B, st, s == steps - 1)));
  }
  // unmasked e keeps the modes outside the 2/3 region unchanged (L = 0 there)
  TRY(unpack<N>(c->x0, f, B, masked ? 1 : 0, st));
  if (TAN) TRY(unpack<N>(c->dx0, df, B, masked ? 1 : 0, st));
  return 0;
}

// ------------------------------------------------------------------ Lawson RK4(3)
template <int N>
static int rk43_attempt(const Args& ai, int B, hipStream_t st, bool first) {
  if (first) TRY((col<N, C_A0, false>(ai, B, st)));
  else TRY((col<N, C_A, false>(ai, B, st)));
  TRY((row<N, R_RHS, false>(ai, B, st)));
  TRY((col<N, C_S2, false>(ai, B, st)));
  TRY((row<N, R_RHS, false>(ai, B, st)));
  TRY((col<N, C_S3, false>(ai, B, st)));
  TRY((row<N, R_RHS, false>(ai, B, st)));
  TRY((col<N, C_S4, false>(ai, B, st)));
  TRY((row<N, R_RHS, false>(ai, B, st)));
  TRY((col<N, C_S5, false>(ai, B, st)));
  TRY((row<N, R_ERR, false>(ai, B, st)));
  return 0;
}

template <int N>
static int run_rk43(ecs_ctx* c, double* x, int B, const double* T, double h0, double atol, int maxs, int nck,
                    ecs_rk43_info* info, double* hs, int8_t* accf, hipStream_t st) {
  TRY(ensure_cap(c, B));
  TRY(autotune<N>(c, B, st));
  constexpr int K = 32;  // attempts per captured chunk (even: the counter parity repeats)
  const int budget = maxs * nck;
  const int hstride = budget + K + 2;
  TRY(ensure_hist(c, (size_t)B * hstride));
  HIPCHK(hipMemsetAsync(c->hist, 0, (size_t)B * hstride * sizeof(Hist), st));
  TRY(upload_scalars(c, 2, T, B, st));
  Args a = base_args(c);
  a.hist = c->hist;
  a.hstride = hstride;
  a.atol = atol;
  a.h0 = h0;
  a.budget = budget;
  TRY(pack<N>(x, c->x0, B, st));
  // k1 = t v(x) (timestepping.py:255)
  Args al = a;
  al.adaptive = AD_FIXED;
  TRY((col<N, C_LOAD, false>(al, B, st)));
  TRY((row<N, R_RHS, false>(al, B, st)));
  a.adaptive = AD_PRIMAL;
  Args a0 = a;
  a0.attempt = 0;
  TRY(rk43_attempt<N>(a0, B, st, true));
  int launched = 1;
  const bool graphs = use_graphs();
  if (graphs) TRY(set_counter(c, 1, st));
  std::vector<Hist> hh(B);
  while (launched < budget) {
    if (graphs) {
      Args ag = a;
      ag.cnt = c->cnt;
      const std::string key = graph_key("rk43", ag, B, K);
      TRY(graph_run(c, key, st, [&](hipStream_t gs) -> int {
        for (int j = 0; j < K; ++j) {
          Args aj = ag;
          aj.cslot = j & 1;
          TRY(rk43_attempt<N>(aj, B, gs, false));
        }
        return 0;
      }));
    } else {
      for (int j = 0; j < K; ++j) {
        Args aj = a;
        aj.attempt = launched + j;
        TRY(rk43_attempt<N>(aj, B, st, false));
      }
    }
    launched += K;
    // poll: stop once every candidate's last attempt was a no-op
    bool any = false;
    for (int b = 0; b < B; ++b)
      HIPCHK(hipMemcpyAsync(&hh[b], c->hist + (size_t)b * hstride + (launched - 1), sizeof(Hist),
                            hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int b = 0; b < B; ++b) any = any || hh[b].active;
    if (!any) break;
  }
  Args af = a;
  af.attempt = launched;
  TRY((col<N, C_FIN, false>(af, B, st)));
  TRY(unpack<N>(c->x0, x, B, 0, st));
  std::vector<Hist> all((size_t)B * hstride);
  HIPCHK(hipMemcpyAsync(all.data(), c->hist, (size_t)B * hstride * sizeof(Hist), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int b = 0; b < B; ++b) {
    const Hist* H = all.data() + (size_t)b * hstride;
    const Hist& fin = H[launched];
    if (info) {
      info[b].s = fin.s;
      info[b].completed = fabs(fin.s - 1.0) < 1e-10;
      info[b].accepted = fin.nacc;
      info[b].rejected = fin.nrej;
      info[b].attempts = fin.nacc + fin.nrej;
      info[b].fevals = 1 + 4 * (fin.nacc + fin.nrej);
      info[b].pad = 0;
    }
    for (int i = 0; i < budget; ++i) {
      const bool act = i < launched && H[i].active;
      if (hs) hs[(size_t)b * budget + i] = act ? H[i].h : 0.0;
      if (accf) accf[(size_t)b * budget + i] = (act && H[i + 1].prev_acc) ? 1 : 0;
    }
  }
  return 0;
}

template <int N>
static int run_rk43_jvp(ecs_ctx* c, double* x, double* dx, int B, const double* T, const double* dT,
                        const double* hs, const int8_t* accf, const int* attempts, int budget, hipStream_t st) {
  TRY(ensure_cap(c, B));
  TRY(autotune<N>(c, B, st));
  // build the accepted-step schedule with the rejections preceding each step
  std::vector<double> sh((size_t)B * budget, 0.0), srh((size_t)B * budget, 0.0);
  std::vector<int> srs((size_t)B * budget, 0), src((size_t)B * budget, 0), snacc(B, 0);
  int maxacc = 0;
  for (int b = 0; b < B; ++b) {
    int m = 0, nr = 0, rstart = 0;
    for (int i = 0; i < attempts[b]; ++i) {
      const double h = hs[(size_t)b * budget + i];
      if (accf[(size_t)b * budget + i]) {
        sh[(size_t)b * budget + m] = h;
        srs[(size_t)b * budget + m] = rstart;
        src[(size_t)b * budget + m] = nr - rstart;
        rstart = nr;
        ++m;
      } else {
        srh[(size_t)b * budget + nr] = h;
        ++nr;
      }
    }
    snacc[b] = m;
    maxacc = std::max(maxacc, m);
  }
  const size_t need = (size_t)B * budget;
  std::unique_lock<std::recursive_mutex> slk(g_alloc_mu);
  if (need > c->sched_cap) {
    if (c->s_h) { HIPCHK(hipFree(c->s_h)); HIPCHK(hipFree(c->s_rh)); HIPCHK(hipFree(c->s_rs)); HIPCHK(hipFree(c->s_rc)); }
    HIPCHK(hipMalloc((void**)&c->s_h, need * sizeof(double)));
    HIPCHK(hipMalloc((void**)&c->s_rh, need * sizeof(double)));
    HIPCHK(hipMalloc((void**)&c->s_rs, need * sizeof(int)));
    HIPCHK(hipMalloc((void**)&c->s_rc, need * sizeof(int)));
    c->sched_cap = need;
  }
  if (B > c->snacc_cap) {
    if (c->s_nacc) HIPCHK(hipFree(c->s_nacc));
    HIPCHK(hipMalloc((void**)&c->s_nacc, B * sizeof(int)));
    c->snacc_cap = B;
  }
  HIPCHK(hipMemcpyAsync(c->s_h, sh.data(), need * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->s_rh, srh.data(), need * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->s_rs, srs.data(), need * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->s_rc, src.data(), need * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->s_nacc, snacc.data(), B * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // the host vectors above go out of scope
  slk.unlock();
  TRY(upload_scalars(c, 2, T, B, st));
  TRY(upload_scalars(c, 3, dT, B, st));
  Args a = base_args(c);
  a.s_h = c->s_h; a.s_rh = c->s_rh; a.s_rs = c->s_rs; a.s_rc = c->s_rc; a.s_nacc = c->s_nacc;
  a.sstride = budget;
  a.rstride = budget;
  TRY(pack<N>(x, c->x0, B, st));
  TRY(pack<N>(dx, c->dx0, B, st));
  Args al = a;
  al.adaptive = AD_FIXED;
  TRY((col<N, C_LOAD, true>(al, B, st)));
  TRY((row<N, R_RHS, true>(al, B, st)));
  a.adaptive = AD_TANGENT;
  // one accepted step: [A] row S2 row S3 row S4 row S5; eagerly, S5 of step m
  // is fused with the first kernel (A) of step m+1 (C_S5A)
  auto step = [&](const Args& am, hipStream_t s2, bool first, bool with_a, bool fuse_next,
                  hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) -> int {
    if (with_a) {
      if (first) TRY((col<N, C_A0, true>(am, B, s2)));
      else TRY((col<N, C_A, true>(am, B, s2)));
    }
    if (e0) HIPCHK(hipEventRecord(e0, s2));
    TRY((row<N, R_RHS, true>(am, B, s2)));
    if (e1) HIPCHK(hipEventRecord(e1, s2));
    TRY((col<N, C_S2, true>(am, B, s2)));
    TRY((row<N, R_RHS, true>(am, B, s2)));
    TRY((col<N, C_S3, true>(am, B, s2)));
    TRY((row<N, R_RHS, true>(am, B, s2)));
    TRY((col<N, C_S4, true>(am, B, s2)));
    TRY((row<N, R_RHS, true>(am, B, s2)));
    if (fuse_next) TRY((col<N, C_S5A, true>(am, B, s2)));
    else TRY((col<N, C_S5, true>(am, B, s2)));
    return 0;
  };
  constexpr int K = 32;
  if (use_graphs() && maxacc > 1) {
    Args am0 = a;
    am0.attempt = 0;
    TRY(step(am0, st, true, true, false));
    TRY(set_counter(c, 1, st));
    Args ag = a;
    ag.cnt = c->cnt;
    const std::string key = graph_key("rk43t", ag, B, K);
    for (int launched = 1; launched < maxacc; launched += K) {
      TRY(graph_run(c, key, st, [&](hipStream_t gs) -> int {
        for (int j = 0; j < K; ++j) {
          Args aj = ag;
          aj.cslot = j & 1;
          TRY(step(aj, gs, false, true, false));
        }
        return 0;
      }));
    }
  } else {
    int minacc = maxacc;
    for (int b = 0; b < B; ++b) minacc = std::min(minacc, snacc[b]);
    for (int m = 0; m < maxacc; ++m) {
      Args am = a;
      am.attempt = m;
      am.run_all = m < minacc;
      am.next_all = m + 1 < minacc;
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (c->probe_every > 0 && m % c->probe_every == 0 && c->probe_used + 2 <= c->probe_ev.size()) {
        e0 = c->probe_ev[c->probe_used];
        e1 = c->probe_ev[c->probe_used + 1];
        c->probe_used += 2;
      }
      TRY(step(am, st, m == 0, m == 0, m + 1 < maxacc, e0, e1));
    }
  }
  TRY(unpack<N>(c->x0, x, B, 0, st));
  TRY(unpack<N>(c->dx0, dx, B, 0, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}


template <int N> static int rk4_p(ecs_ctx* c, double* f, int B, const double* dt, int steps, int masked, hipStream_t st) {
  return run_rk4<N, false>(c, f, nullptr, B, dt, nullptr, steps, masked, st);
}
template <int N> static int rk4_t(ecs_ctx* c, double* f, double* df, int B, const double* dt, const double* ddt, int steps, int masked, hipStream_t st) {
  return run_rk4<N, true>(c, f, df, B, dt, ddt, steps, masked, st);
}


// ------------------------------------------------------------------ adjoint
// The adjoint kernels transform with the FFT plan the autotuner fixed for the
// forward kernels of the same shape (column or row family), so they take the
// radix that suits the grid size; the plan is fixed per grid size, so the
// reverse pass, too, computes the same bits at every batch.
// Below batch 16 they take one column / row per block (the most blocks for a
// small batch); tiles do not change the arithmetic.
template <int N, int MODE, int TY>
static int acol_t(const Args& a, int B, hipStream_t st) {
  dim3 grid((Geo<N>::KC + TY - 1) / TY, 2, B);
  if (plan_of(N, K_COL) == 1) hipLaunchKernelGGL((adj_col_kernel<N, TY, NT, MODE, 8>), grid, dim3(NT), 0, st, a);
  else hipLaunchKernelGGL((adj_col_kernel<N, TY, NT, MODE, 16>), grid, dim3(NT), 0, st, a);
  LAUNCHCHK();
  return 0;
}
template <int N, int MODE>
static int acol(const Args& a, int B, hipStream_t st) {
  return B < 16 ? acol_t<N, MODE, 1>(a, B, st) : acol_t<N, MODE, Tiles<N>::TY>(a, B, st);
}
// B_OUT and the transposed stage combination `cb` (whose yb is B_OUT's output) in one launch
template <int N, int TY>
static int acol_comb_t(const Args& a, const CombAdj& cb, int B, hipStream_t st) {
  dim3 grid((Geo<N>::KC + TY - 1) / TY, 2, B);
  if (plan_of(N, K_COL) == 1) hipLaunchKernelGGL((adj_out_comb_kernel<N, TY, NT, 8>), grid, dim3(NT), 0, st, a, cb);
  else hipLaunchKernelGGL((adj_out_comb_kernel<N, TY, NT, 16>), grid, dim3(NT), 0, st, a, cb);
  LAUNCHCHK();
  return 0;
}
template <int N>
static int acol_comb(const Args& a, const CombAdj& cb, int B, hipStream_t st) {
  return B < 16 ? acol_comb_t<N, 1>(a, cb, B, st) : acol_comb_t<N, Tiles<N>::TY>(a, cb, B, st);
}
// the transposed stage combination `cb` (one of whose outputs is gin) and B_G in one launch
template <int N, int TY>
static int acol_g_comb_t(const Args& a, const CombAdj& cb, int B, hipStream_t st) {
  dim3 grid((Geo<N>::KC + TY - 1) / TY, 2, B);
  if (plan_of(N, K_COL) == 1) hipLaunchKernelGGL((adj_comb_g_kernel<N, TY, NT, 8>), grid, dim3(NT), 0, st, a, cb);
  else hipLaunchKernelGGL((adj_comb_g_kernel<N, TY, NT, 16>), grid, dim3(NT), 0, st, a, cb);
  LAUNCHCHK();
  return 0;
}
template <int N>
static int acol_g_comb(const Args& a, const CombAdj& cb, int B, hipStream_t st) {
  return B < 16 ? acol_g_comb_t<N, 1>(a, cb, B, st) : acol_g_comb_t<N, Tiles<N>::TY>(a, cb, B, st);
}
// B_OUT, the combinations `then` and `pre`, and the next pass's B_G on gnext in one launch
template <int N, int TY>
static int acol_out_g_t(const Args& a, const CombAdj& then, const CombAdj& pre, const double2* gnext, int B,
                        hipStream_t st) {
  dim3 grid((Geo<N>::KC + TY - 1) / TY, 2, B);
  if (plan_of(N, K_COL) == 1)
    hipLaunchKernelGGL((adj_out_g_kernel<N, TY, NT, 8>), grid, dim3(NT), 0, st, a, then, pre, gnext);
  else
    hipLaunchKernelGGL((adj_out_g_kernel<N, TY, NT, 16>), grid, dim3(NT), 0, st, a, then, pre, gnext);
  LAUNCHCHK();
  return 0;
}
template <int N>
static int acol_out_g(const Args& a, const CombAdj& then, const CombAdj& pre, const double2* gnext, int B,
                      hipStream_t st) {
  return B < 16 ? acol_out_g_t<N, 1>(a, then, pre, gnext, B, st)
                : acol_out_g_t<N, Tiles<N>::TY>(a, then, pre, gnext, B, st);
}
template <int N, int TX>
static int arow_t(const Args& a, int B, hipStream_t st) {
  dim3 grid(N / TX, 1, B);
  if (plan_of(N, K_ROW) == 1) hipLaunchKernelGGL((adj_row_kernel<N, TX, NT, 8>), grid, dim3(NT), 0, st, a);
  else hipLaunchKernelGGL((adj_row_kernel<N, TX, NT, 16>), grid, dim3(NT), 0, st, a);
  LAUNCHCHK();
  return 0;
}
template <int N>
static int arow(const Args& a, int B, hipStream_t st) {
  return B < 16 ? arow_t<N, 1>(a, B, st) : arow_t<N, Tiles<N>::TX>(a, B, st);
}

template <int N>
static int run_state_vel_vjp(ecs_ctx* c, const double* in, const double* cot, double* grad, int B, int diss,
                             hipStream_t st) {
  TRY(ensure_adj(c, B));
  TRY(autotune<N>(c, B, st));
  Args a = base_args(c);
  a.include_diss = diss;
  a.aout = c->AO;
  a.ain = c->AI;
  a.prim = c->I;
  a.gin = c->acc;  // compact cotangent
  TRY(pack<N>(in, c->x0, B, st));
  TRY(pack<N>(cot, c->acc, B, st));
  TRY((col<N, C_LOAD, false>(a, B, st)));
  TRY((acol<N, B_G>(a, B, st)));
  TRY(arow<N>(a, B, st));
  TRY((acol<N, B_OUT>(a, B, st)));
  TRY(unpack<N>(c->out, grad, B, 1, st));
  return 0;
}

template <int N>
static int adj_buffers(ecs_ctx* c, int B, int ckpt) {
  TRY(ensure_adj(c, B));
  using G = Geo<N>;
  const size_t cs = (size_t)B * 2 * G::R * G::KC;
  const size_t is = (size_t)B * 8 * N * G::KP;
  TRY(grow(&c->xtape, &c->vt_cap, (size_t)ckpt * 5 * cs));
  TRY(grow(&c->itape, &c->it_cap, (size_t)ckpt * 4 * is));
  return 0;
}

// The second step tape of the reverse pass, with the side stream and events
// of the overlap, when it fits REV_OVERLAP_BYTES; returns false (and the
// reverse pass runs its segments one phase after the other) otherwise, when
// the context's option is off (ecs_ctx_set_rev_overlap) or ECSMHD_REV_OVERLAP=0.
constexpr size_t REV_OVERLAP_BYTES = (size_t)4 << 30;
template <int N>
static int adj_buffers2(ecs_ctx* c, int B, int ckpt, bool* ok) {
  using G = Geo<N>;
  const size_t cs = (size_t)B * 2 * G::R * G::KC;
  const size_t is = (size_t)B * 8 * N * G::KP;
  const size_t nv = (size_t)ckpt * 5 * cs, ni = (size_t)ckpt * 4 * is;
  const char* env = getenv("ECSMHD_REV_OVERLAP");
  *ok = c->rev_overlap && (nv + ni) * sizeof(double2) <= REV_OVERLAP_BYTES && !(env && env[0] == '0');
  if (!*ok) return 0;
  TRY(grow(&c->xtape2, &c->vt2_cap, nv));
  TRY(grow(&c->itape2, &c->it2_cap, ni));
  if (!c->side) HIPCHK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
  for (hipEvent_t& e : c->ev_rev)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return 0;
}

// Forward pass of Lawson RK4 storing the compact state at the start of every
// segment of ckpt steps into ck [nseg][B][2][R][KC]; f is advanced in place.
template <int N>
static int run_rk4_ckpt(ecs_ctx* c, double* f, int B, const double* dt, int steps, int masked, int ckpt,
                        double2* ck, const int* nsteps, hipStream_t st) {
  TRY(ensure_cap(c, B));
  TRY(autotune<N>(c, B, st));
  using G = Geo<N>;
  const size_t cs = (size_t)B * 2 * G::R * G::KC;
  TRY(upload_scalars(c, 0, dt, B, st));
  Args a = base_args(c);
  a.masked_e = masked;
  if (nsteps) {
    TRY(upload_ints(c, nsteps, B, st));
    a.nstep = c->nstep;
  }
  TRY(pack<N>(f, c->x0, B, st));
  HIPCHK(hipMemcpyAsync(ck, c->x0, cs * sizeof(double2), hipMemcpyDeviceToDevice, st));
  if (steps > 0) TRY((col<N, C_LOAD, false>(a, B, st)));
  for (int s = 0; s < steps; ++s) {
    a.step = s;
    TRY((row<N, R_RHS, false>(a, B, st)));
    TRY((col<N, C_S1F, false>(a, B, st)));
    TRY((row<N, R_RHS, false>(a, B, st)));
    TRY((col<N, C_S2, false>(a, B, st)));
    TRY((row<N, R_RHS, false>(a, B, st)));
    TRY((col<N, C_S3, false>(a, B, st)));
    TRY((row<N, R_RHS, false>(a, B, st)));
    Args a4 = a;
    a4.skip_inv = (s == steps - 1);
    TRY((col<N, C_S4, false>(a4, B, st)));
    if ((s + 1) % ckpt == 0 && s + 1 < steps)
      HIPCHK(hipMemcpyAsync(ck + (size_t)((s + 1) / ckpt) * cs, c->x0, cs * sizeof(double2),
                            hipMemcpyDeviceToDevice, st));
  }
  TRY(unpack<N>(c->x0, f, B, masked ? 1 : 0, st));
  return 0;
}

// Reverse pass from the checkpoints: fbar (cotangent of the final spectra) is
// replaced by the cotangent of the initial spectra; hbar[b] = d/dh.
template <int N>
static int run_rk4_rev(ecs_ctx* c, const double2* ck, double* fbar, double* hbar, int B, const double* dt, int steps,
                       int masked, int ckpt, const int* nsteps, hipStream_t st) {
  TRY(adj_buffers<N>(c, B, ckpt));
  TRY(autotune<N>(c, B, st));
  using G = Geo<N>;
  const size_t cs = (size_t)B * 2 * G::R * G::KC;
  const size_t is = (size_t)B * 8 * N * G::KP;
  const int nseg = (steps + ckpt - 1) / ckpt;
  TRY(upload_scalars(c, 0, dt, B, st));
  Args a = base_args(c);
  a.masked_e = masked;
  a.aout = c->AO;
  a.ain = c->AI;
  a.xb = c->xb; a.k1b = c->k1b; a.k2b = c->k2b; a.k3b = c->k3b;
  a.hb = c->hb;
  a.cstride = cs;
  if (nsteps) {
    TRY(upload_ints(c, nsteps, B, st));
    a.nstep = c->nstep;
  }
  TRY(pack<N>(fbar, c->xb, B, st));
  HIPCHK(hipMemsetAsync(c->hb, 0, cs * sizeof(double), st));
  // Segment j's recomputation (forward steps from its checkpoint, storing the
  // step tape) reads and writes only the forward workspace and the tape; its
  // reverse sweep reads the tape and writes only the adjoint workspace.  With
  // two tapes, segment j - 1 is recomputed on the side stream while the
  // caller's stream sweeps segment j backward (the same kernels in the same
  // order in each chain, so the same bits): at small batches both chains are
  // launch-bound and overlap.  A tape is rewritten only after the sweep that
  // read it (events).
  bool two = false;
  TRY(adj_buffers2<N>(c, B, ckpt, &two));
  two = two && nseg > 1;
  hipStream_t fs = two ? c->side : st;
  if (two) {
    HIPCHK(hipEventRecord(c->ev_rev[0], st));
    HIPCHK(hipStreamWaitEvent(fs, c->ev_rev[0], 0));
  }
  auto itb = [&](int j) { return (two && (j & 1)) ? c->itape2 : c->itape; };
  auto xtb = [&](int j) { return (two && (j & 1)) ? c->xtape2 : c->xtape; };
  for (int j = nseg - 1; j >= 0; --j) {
    const int t0 = j * ckpt;
    const int L = std::min(ckpt, steps - t0);
    double2* itj = itb(j);
    double2* xtj = xtb(j);
    auto slot = [&](int t, int s) { return itj + ((size_t)t * 4 + s) * is; };
    auto xt = [&](int t) { return xtj + (size_t)t * cs; };
    auto vt = [&](int t) { return xtj + (size_t)ckpt * cs + (size_t)t * 4 * cs; };
    const int p = j & 1;
    // the tape this segment fills was last read by the sweep of segment j + 2
    if (two && j + 2 < nseg) HIPCHK(hipStreamWaitEvent(fs, c->ev_rev[3 + p], 0));
    HIPCHK(hipMemcpyAsync(c->x0, ck + (size_t)j * cs, cs * sizeof(double2), hipMemcpyDeviceToDevice, fs));
    Args f = a;
    f.inter = slot(0, 0);
    f.step = t0;
    TRY((col<N, C_LOAD, false>(f, B, fs)));
    for (int t = 0; t < L; ++t) {
      f.step = t0 + t;
      f.vtape = nullptr; f.xtape = nullptr;
      f.inter = slot(t, 0);
      TRY((row<N, R_RHS, false>(f, B, fs)));
      f.inter = slot(t, 1); f.vtape = vt(t); f.xtape = xt(t);
      TRY((col<N, C_S1F, false>(f, B, fs)));
      f.xtape = nullptr;
      TRY((row<N, R_RHS, false>(f, B, fs)));
      f.inter = slot(t, 2); f.vtape = vt(t) + cs;
      TRY((col<N, C_S2, false>(f, B, fs)));
      TRY((row<N, R_RHS, false>(f, B, fs)));
      f.inter = slot(t, 3); f.vtape = vt(t) + 2 * cs;
      TRY((col<N, C_S3, false>(f, B, fs)));
      TRY((row<N, R_RHS, false>(f, B, fs)));
      Args f4 = f;
      f4.vtape = vt(t) + 3 * cs;
      f4.inter = (t + 1 < L) ? slot(t + 1, 0) : slot(t, 3);
      f4.skip_inv = (t + 1 == L);
      TRY((col<N, C_S4, false>(f4, B, fs)));
    }
    if (two) {
      HIPCHK(hipEventRecord(c->ev_rev[1 + p], fs));
      HIPCHK(hipStreamWaitEvent(st, c->ev_rev[1 + p], 0));
    }
    Args b = a;
    b.tx = xt(L - 1);
    b.tv = vt(L - 1);
    b.step = t0 + L - 1;
    TRY((acol<N, B_ENTRY>(b, B, st)));
    for (int t = L - 1; t >= 0; --t) {
      b.step = t0 + t;
      b.tx = xt(t);
      b.tv = vt(t);
      b.prim = slot(t, 3);
      TRY(arow<N>(b, B, st));
      TRY((acol<N, B_S4>(b, B, st)));
      b.prim = slot(t, 2);
      TRY(arow<N>(b, B, st));
      TRY((acol<N, B_S3>(b, B, st)));
      b.prim = slot(t, 1);
      TRY(arow<N>(b, B, st));
      TRY((acol<N, B_S2>(b, B, st)));
      b.prim = slot(t, 0);
      TRY(arow<N>(b, B, st));
      if (t > 0) {
        b.txp = xt(t - 1);
        b.tvp = vt(t - 1);
        TRY((acol<N, B_S1E>(b, B, st)));
      } else {
        TRY((acol<N, B_S1>(b, B, st)));
      }
    }
    if (two) HIPCHK(hipEventRecord(c->ev_rev[3 + p], st));
  }
  TRY(unpack<N>(c->xb, fbar, B, masked ? 1 : 0, st));
  hipLaunchKernelGGL((hb_sum_kernel<N>), dim3(B), dim3(256), 0, st, c->hb, c->red);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(c->hred, c->red, B * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int b2 = 0; b2 < B; ++b2) hbar[b2] = c->hred[b2];
  return 0;
}

template <int N>
static int run_rk4_vjp(ecs_ctx* c, const double* f0, double* fbar, double* hbar, int B, const double* dt, int steps,
                       int masked, int ckpt, hipStream_t st) {
  TRY(ensure_cap(c, B));
  TRY(autotune<N>(c, B, st));
  TRY(grow(&c->ckpt, &c->ckpt_cap, ckpt_elems(N, B, steps, ckpt)));
  // the forward pass advances a private copy of f0 (held in c->out, full layout not needed)
  double* fwork = nullptr;
  const size_t full = (size_t)B * 2 * N * Geo<N>::NH * 2;
  HIPCHK(hipMallocAsync((void**)&fwork, full * sizeof(double), st));
  HIPCHK(hipMemcpyAsync(fwork, f0, full * sizeof(double), hipMemcpyDeviceToDevice, st));
  int rc = run_rk4_ckpt<N>(c, fwork, B, dt, steps, masked, ckpt, c->ckpt, nullptr, st);
  if (rc == 0) rc = run_rk4_rev<N>(c, c->ckpt, fbar, hbar, B, dt, steps, masked, ckpt, nullptr, st);
  HIPCHK(hipFreeAsync(fwork, st));
  return rc;
}

// ------------------------------------------------------------------ timing hook
// An empty kernel with the fused kernels' argument block: back to back, its
// launches measure the floor a kernel launch costs on this stream.
template <int N>
__global__ __launch_bounds__(NT) void empty_kernel(Args a) {}

// Average duration (ms) of one launch of a fused RHS kernel, timed with HIP
// events on `stream` over `reps` back-to-back launches of the same kernel on
// batch spectra already resident in the context's workspace.
// which: 0 = column kernel (stage 2 glue), 1 = row kernel, 2 = tangent column
// kernel, 3 = tangent row kernel (fixed-step instance), 4 = tangent row kernel
// of the adaptive tangent replay (the instance a GMRES matvec runs), 5 = the
// empty kernel with the tangent row kernel's grid (the launch floor).
template <int N>
static int run_time_kernel(ecs_ctx* c, int B, int which, int reps, double* ms, hipStream_t st) {
  TRY(ensure_cap(c, B));
  TRY(autotune<N>(c, B, st));
  std::vector<double> dt(B, 1e-3);
  TRY(upload_scalars(c, 0, dt.data(), B, st));
  TRY(upload_scalars(c, 1, dt.data(), B, st));
  Args a = base_args(c);
  // defined inputs: zero state, one load + row pass to fill I and O
  HIPCHK(hipMemsetAsync(c->x0, 0, c->compact_elems() * B * sizeof(double2), st));
  HIPCHK(hipMemsetAsync(c->dx0, 0, c->compact_elems() * B * sizeof(double2), st));
  HIPCHK(hipMemsetAsync(c->acc, 0, c->compact_elems() * B * sizeof(double2), st));
  HIPCHK(hipMemsetAsync(c->dacc, 0, c->compact_elems() * B * sizeof(double2), st));
  TRY((col<N, C_LOAD, true>(a, B, st)));
  TRY((row<N, R_RHS, true>(a, B, st)));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // the tangent row pass as the tangent replay of an adaptive schedule runs it
  // (stepper kind AD_TANGENT, every candidate active: no schedule is read)
  Args ar = a;
  ar.adaptive = AD_TANGENT;
  ar.run_all = 1;
  ar.attempt = 0;
  auto launch = [&]() -> int {
    switch (which) {
      case 0: return col<N, C_S2, false>(a, B, st);
      case 1: return row<N, R_RHS, false>(a, B, st);
      case 2: return col<N, C_S2, true>(a, B, st);
      case 3: return row<N, R_RHS, true>(a, B, st);
      case 4: return row<N, R_RHS, true>(ar, B, st);
      default: {
        const int tx = var_tile(tuned(N, K_ROWR, B, DefTile<N>::ROW));
        hipLaunchKernelGGL(empty_kernel<N>, dim3(N / tx, 1, B), dim3(NT), 0, st, ar);
        LAUNCHCHK();
        return 0;
      }
    }
  };
  for (int i = 0; i < 3; ++i) TRY(launch());
  HIPCHK(hipEventRecord(e0, st));
  for (int i = 0; i < reps; ++i) TRY(launch());
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  *ms = (double)t / reps;
  return 0;
}



// ------------------------------------------------------------------ generic explicit steppers
// Classical RK4 (timestepping.py:17-44), Lawson RK6 (:83-125) and two-derivative
// RK4 (:127-149): each stage a row pass and a column kernel whose glue stores
// the stage RHS and forms the next stage input from a combination table
// (comb_kernels.h) in device memory.
enum GenScheme : int { G_RK4 = 0, G_LRK6 = 1, G_TDRK4 = 2, G_LRK4 = 3 };

// Butcher form of a (Lawson) scheme: stage s reads
//   x0 E^{p_s} + h sum_j A[s][j] E^{p_s - p_j} v_j,   v_j = RHS(stage j input),
// and the step ends at x0 E^{p_S} + h sum_j b_j E^{p_S - p_j} v_j, with
// E = exp(h L / ediv).  For Lawson RK6, p counts the multiplications by
// e = exp(h/6 L) that precede each stage (after stages 1, 3, 4, 5, 6, 7).
struct Tableau {
  int S;
  double A[8][8];
  double b[8];
  int p[9];
  double ediv;
};

static const Tableau& tableau(int scheme) {
  static const Tableau rk4 = {4,
                              {{0}, {0.5}, {0, 0.5}, {0, 0, 1.0}},
                              {1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6},
                              {0, 0, 0, 0, 0},
                              1.0};
  static const Tableau lrk6 = {8,
                               {{0},
                                {1.0 / 6},
                                {1.0 / 12, 1.0 / 12},
                                {0, -4.0 / 33, 5.0 / 11},
                                {-1.0 / 4, -29.0 / 44, 31.0 / 22, 0},
                                {3.0 / 11, 8.0 / 33, -4.0 / 11, 1.0 / 11, 14.0 / 33},
                                {-17.0 / 48, -5.0 / 12, 1.0, 1.0, -13.0 / 12, 11.0 / 16},
                                {20.0 / 39, 12.0 / 39, -31.0 / 39, -1.0 / 39, 34.0 / 39, -11.0 / 39, 16.0 / 39}},
                               {13.0 / 200, 0, 4.0 / 25, 11.0 / 40, 0, 11.0 / 40, 4.0 / 25, 13.0 / 200},
                               {0, 1, 1, 2, 3, 4, 5, 6, 6},
                               6.0};
  // Lawson RK4 (timestepping.py:46-81): e = exp(h/2 L) after stages 1 and 3
  static const Tableau lrk4 = {4,
                               {{0}, {0.5}, {0, 0.5}, {0, 0, 1.0}},
                               {1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6},
                               {0, 1, 1, 2, 2},
                               2.0};
  return scheme == G_LRK6 ? lrk6 : (scheme == G_LRK4 ? lrk4 : rk4);
}

struct CombBuilder {
  Comb cb;
  CombBuilder(ecs_ctx* c, size_t elems, double ediv, int masked, double2* out, double2* dout) {
    memset(&cb, 0, sizeof(cb));
    cb.ediv = ediv; cb.masked = masked;
    cb.h = c->scal; cb.dh = c->scal + c->cap;
    cb.nu[0] = c->nu; cb.nu[1] = c->eta;
    cb.out = out; cb.dout = dout; cb.elems = elems;
  }
  CombBuilder& term(const double2* in, const double2* din, double coef, int hp, int ep) {
    const int t = cb.nterm++;
    cb.in[t] = in; cb.din[t] = din; cb.coef[t] = coef; cb.hp[t] = hp; cb.ep[t] = ep;
    return *this;
  }
};

// f (full layout [B][2][N][N/2+1]) advanced by `steps` steps of h = dt[b]
// (tangent df along ddt[b]).  The RHS includes dissipation for RK4 and TDRK4
// (loss_functions.py:53,81) and excludes it for Lawson RK6 (:68).  Unmasked
// schemes keep the modes outside the 2/3 region (their RHS is zero there).
// The two column kernels of the two-derivative RK4 tangent step (which 0: A,
// 1: B), one column per block, in FFT plan `plan` (0 wide, 1 narrow).
template <int N>
static int td_launch(int which, int plan, const Args& a, int B, hipStream_t st) {
  dim3 grid(Geo<N>::KC, 2, B);
  if (which == 0) {
    if (plan == 0) hipLaunchKernelGGL((td_col_a<N, NT, 16>), grid, dim3(NT), 0, st, a, B);
    else hipLaunchKernelGGL((td_col_a<N, NT, 8>), grid, dim3(NT), 0, st, a, B);
  } else {
    if (plan == 0) hipLaunchKernelGGL((td_col_b<N, NT, 16>), grid, dim3(NT), 0, st, a, B);
    else hipLaunchKernelGGL((td_col_b<N, NT, 8>), grid, dim3(NT), 0, st, a, B);
  }
  LAUNCHCHK();
  return 0;
}

template <int N, bool TAN>
static int run_gen(ecs_ctx* c, double* f, double* df, int B, const double* dt, const double* ddt, int steps,
                   int scheme, int masked, hipStream_t st) {
  const bool td = scheme == G_TDRK4;
  TRY(ensure_cap(c, (td && TAN) ? 3 * B : B));
  TRY(autotune<N>(c, B, st));
  if (td && TAN) TRY(autotune<N>(c, 3 * B, st));
  const size_t ce = c->compact_elems();
  const size_t cs = ce * B;
  const Tableau& tb = tableau(scheme);
  const int units = td ? (TAN ? 18 : 7) : (TAN ? 4 + 2 * tb.S : 2 + tb.S);
  TRY(grow(&c->gen, &c->gen_cap, (size_t)units * cs));
  TRY(upload_scalars(c, 0, dt, B, st));
  if (TAN) TRY(upload_scalars(c, 1, ddt, B, st));
  Args a = base_args(c);
  double2* g = c->gen;
  auto slot = [&](int u) { return g + (size_t)u * cs; };
  const int diss = scheme == G_LRK6 ? 0 : 1;
  const int msk = scheme == G_LRK6 ? masked : 0;
  double2* x = slot(0);
  TRY(pack<N>(f, x, B, st));
  if (!td) {
    // Fused form: each stage is a row pass and one C_GEN column kernel, which
    // stores the stage's RHS v_k (and tangent) and forms the next stage input
    // (or the step's result) from the tableau row in registers.  The
    // combinations are evaluated term by term in table order.
    double2* y = slot(1);
    double2* dx = TAN ? slot(2) : nullptr;
    double2* dy = TAN ? slot(3) : nullptr;
    const int v0 = TAN ? 4 : 2;
    auto v = [&](int j) { return slot(v0 + j); };
    auto dv = [&](int j) { return TAN ? slot(v0 + tb.S + j) : nullptr; };
    if (TAN) TRY(pack<N>(df, dx, B, st));
    std::vector<Comb> tabs(tb.S);
    for (int k = 0; k < tb.S; ++k) {
      const bool last = k == tb.S - 1;
      CombBuilder cb(c, ce, tb.ediv, msk, last ? x : y, last ? dx : dy);
      const int pk = last ? tb.p[tb.S] : tb.p[k + 1];
      cb.term(x, dx, 1.0, 0, pk);
      for (int j = 0; j <= k; ++j) {
        const double cj = last ? tb.b[j] : tb.A[k + 1][j];
        if (cj != 0.0) cb.term(v(j), dv(j), cj, 1, pk - tb.p[j]);
      }
      tabs[k] = cb.cb;
    }
    TRY(grow(&c->gcomb, &c->gcomb_cap, (tabs.size() * sizeof(Comb) + sizeof(double2) - 1) / sizeof(double2)));
    Comb* dtabs = (Comb*)c->gcomb;
    HIPCHK(hipMemcpyAsync(dtabs, tabs.data(), tabs.size() * sizeof(Comb), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));  // the host table goes out of scope
    if (steps > 0) {
      Args a0 = a;
      a0.x0 = x;
      a0.dx0 = dx;
      a0.adaptive = AD_FIXED;
      TRY((col<N, C_LOAD, TAN>(a0, B, st)));
    }
    for (int s = 0; s < steps; ++s) {
      for (int k = 0; k < tb.S; ++k) {
        Args ak = a;
        ak.adaptive = AD_FIXED;
        ak.include_diss = diss;
        TRY((row<N, R_RHS, TAN>(ak, B, st)));
        const bool last = k == tb.S - 1;
        ak.x0 = k == 0 ? x : y;
        ak.dx0 = k == 0 ? dx : dy;
        ak.out = v(k);
        ak.dout = dv(k);
        ak.xf = last ? x : y;
        ak.dxf = last ? dx : dy;
        ak.gc = dtabs + k;
        ak.skip_inv = (s == steps - 1 && last) ? 1 : 0;
        TRY((col<N, C_GEN, TAN>(ak, B, st)));
      }
    }
    TRY(unpack<N>(x, f, B, msk, st));
    if (TAN) TRY(unpack<N>(dx, df, B, msk, st));
    return 0;
  }
  // Two-derivative RK4 with v_j = RHS(stage input), w_j = J(stage input) v_j:
  //   y = x + h/2 v1 + h^2/8 w1,   x' = x + h v1 + h^2/6 w1 + h^2/3 w2
  // (a_j = h J k_j = h^2 w_j).  Tangents use J for the first-order terms and
  // dw = H[dz, v] + J(z) dv; J(z) v and J(z) dv share one pass of batch 2B.
  if (!TAN) {
    // Fused: per half step a row pass, C_TD1 (v = RHS(z), then the pair
    // (z, v) transformed as a tangent pass's input), the tangent row pass and
    // C_TD2 (w = J(z) v, then the combination giving the next point), so 8
    // launches per step; the combinations are evaluated term by term in table order.
    double2 *y = slot(1), *v1 = slot(2), *w1 = slot(3), *v2 = slot(4), *w2 = slot(5);
    CombBuilder cy(c, ce, 1.0, 0, y, nullptr);
    cy.term(x, nullptr, 1.0, 0, 0).term(v1, nullptr, 0.5, 1, 0).term(w1, nullptr, 0.125, 2, 0);
    CombBuilder cx(c, ce, 1.0, 0, x, nullptr);
    cx.term(x, nullptr, 1.0, 0, 0).term(v1, nullptr, 1.0, 1, 0).term(w1, nullptr, 1.0 / 6, 2, 0)
      .term(w2, nullptr, 1.0 / 3, 2, 0);
    const Comb tabs[2] = {cy.cb, cx.cb};
    TRY(grow(&c->gcomb, &c->gcomb_cap, (sizeof(tabs) + sizeof(double2) - 1) / sizeof(double2)));
    Comb* dtabs = (Comb*)c->gcomb;
    HIPCHK(hipMemcpyAsync(dtabs, tabs, sizeof(tabs), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));  // the host table goes out of scope
    Args a0 = a;
    a0.adaptive = AD_FIXED;
    a0.include_diss = diss;
    if (steps > 0) {
      Args al = a0;
      al.x0 = x;
      TRY((col<N, C_LOAD, false>(al, B, st)));
    }
    for (int s = 0; s < steps; ++s) {
      for (int half = 0; half < 2; ++half) {
        double2* z = half == 0 ? x : y;
        double2* v = half == 0 ? v1 : v2;
        double2* w = half == 0 ? w1 : w2;
        TRY((row<N, R_RHS, false>(a0, B, st)));
        Args a1 = a0;
        a1.x0 = z;
        a1.out = v;
        TRY((col<N, C_TD1, false>(a1, B, st)));
        TRY((row<N, R_RHS, true>(a0, B, st)));
        Args a2 = a0;
        a2.x0 = z;
        a2.dx0 = v;
        a2.dout = w;
        a2.gc = dtabs + half;
        a2.xf = half == 0 ? y : x;
        a2.skip_inv = (s == steps - 1 && half == 1) ? 1 : 0;
        TRY((col<N, C_TD2, true>(a2, B, st)));
      }
    }
    TRY(unpack<N>(x, f, B, 0, st));
    return 0;
  }
  // Tangent: per half step a tangent row pass at (z, dz), td_col_a (v, dv and
  // the inputs of the passes for J v, J dv and H[dz, v], batched 3B), one
  // tangent row pass over the 3B candidates (b0 = 0 for the H part) and
  // td_col_b (w = J v, J dv, H, then the combination and the next point's
  // inverse transforms): 8 launches per step.
  //   y = z + h/2 v + h^2/8 w,  dy = dz + h/2 dv + h^2/8 (J dv + H) + dh d/dh(...)
  // slots (units of B): x 0, dx 1, y 2, dy 3, v1 4, dv1 5, w1 6, Jdv1 7, H1 8,
  // v2 9, dv2 10, w2 11, Jdv2 12, H2 13
  double2 *dx = slot(1), *y = slot(2), *dy = slot(3);
  double2 *v1 = slot(4), *dv1 = slot(5), *w1 = slot(6), *jdv1 = slot(7), *h1 = slot(8);
  double2 *v2 = slot(9), *dv2 = slot(10), *w2 = slot(11), *jdv2 = slot(12), *h2 = slot(13);
  TRY(pack<N>(df, dx, B, st));
  CombBuilder cy(c, ce, 1.0, 0, y, dy);
  cy.term(x, dx, 1.0, 0, 0).term(v1, dv1, 0.5, 1, 0).term(w1, jdv1, 0.125, 2, 0).term(nullptr, h1, 0.125, 2, 0);
  CombBuilder cx(c, ce, 1.0, 0, x, dx);
  cx.term(x, dx, 1.0, 0, 0).term(v1, dv1, 1.0, 1, 0).term(w1, jdv1, 1.0 / 6, 2, 0)
    .term(nullptr, h1, 1.0 / 6, 2, 0).term(w2, jdv2, 1.0 / 3, 2, 0).term(nullptr, h2, 1.0 / 3, 2, 0);
  const Comb tabs[2] = {cy.cb, cx.cb};
  TRY(grow(&c->gcomb, &c->gcomb_cap, (sizeof(tabs) + sizeof(double2) - 1) / sizeof(double2)));
  Comb* dtabs = (Comb*)c->gcomb;
  HIPCHK(hipMemcpyAsync(dtabs, tabs, sizeof(tabs), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // the host table goes out of scope
  Args a0 = a;
  a0.adaptive = AD_FIXED;
  a0.include_diss = diss;
  if (steps > 0) {
    Args al = a0;
    al.x0 = x;
    al.dx0 = dx;
    TRY((col<N, C_LOAD, true>(al, B, st)));
  }
  const int plan = var_plan(tuned(N, K_COLT, B, DefTile<N>::COL));
  for (int s = 0; s < steps; ++s) {
    for (int half = 0; half < 2; ++half) {
      Args ah = a0;
      ah.x0 = half == 0 ? x : y;
      ah.dx0 = half == 0 ? dx : dy;
      ah.out = half == 0 ? v1 : v2;
      ah.dout = half == 0 ? dv1 : dv2;
      ah.tdw = half == 0 ? w1 : w2;
      ah.tdjdv = half == 0 ? jdv1 : jdv2;
      ah.tdh = half == 0 ? h1 : h2;
      ah.gc = dtabs + half;
      ah.xf = half == 0 ? y : x;
      ah.dxf = half == 0 ? dy : dx;
      ah.skip_inv = (s == steps - 1 && half == 1) ? 1 : 0;
      TRY((row<N, R_RHS, true>(a0, B, st)));
      TRY(td_launch<N>(0, plan, ah, B, st));
      Args a3 = a0;
      a3.b0_from = 2 * B;
      TRY((row<N, R_RHS, true>(a3, 3 * B, st)));
      TRY(td_launch<N>(1, plan, ah, B, st));
    }
  }
  TRY(unpack<N>(x, f, B, 0, st));
  TRY(unpack<N>(dx, df, B, 0, st));
  return 0;
}

// ------------------------------------------------------------------ generic steppers: reverse pass
template <int N>
static int comb_adj_launch(const CombAdj& cb, int B, hipStream_t st) {
  const size_t tot = (size_t)B * cb.elems;
  const int blocks = (int)std::min<size_t>((tot + 255) / 256, 4096);
  hipLaunchKernelGGL((comb_adj_kernel<N>), dim3(blocks), dim3(256), 0, st, cb, B);
  LAUNCHCHK();
  return 0;
}

struct CombAdjBuilder {
  CombAdj cb;
  CombAdjBuilder(ecs_ctx* c, size_t elems, double ediv, int masked, const double2* yb, double* hb) {
    memset(&cb, 0, sizeof(cb));
    cb.ediv = ediv; cb.masked = masked; cb.h = c->scal;
    cb.nu[0] = c->nu; cb.nu[1] = c->eta;
    cb.yb = yb; cb.hb = hb; cb.elems = elems;
  }
  // assign: this is the first contribution to `out` (it overwrites instead of adding)
  CombAdjBuilder& term(double2* out, const double2* in, double coef, int hp, int ep, bool assign = false) {
    const int t = cb.nterm++;
    cb.out[t] = out; cb.in[t] = in; cb.coef[t] = coef; cb.hp[t] = hp; cb.ep[t] = ep;
    if (assign) cb.assign |= 1u << t;
    return *this;
  }
};

// An RHS-VJP pass is out = J(x)^T g (with the RHS's dissipation when diss)
// on compact spectra, four kernels: the column load of x (skipped when a
// forward pass left its column output, the 8 derived fields in x-space, at
// prim), the column adjoint B_G of g, the row adjoint, and the column adjoint
// output B_OUT; hess: out = J_0(x)^T g with J_0 the Jacobian of the
// nonlinearity without b0 (the transpose of y -> H[y, x]); acc: out +=
// instead of out =.  pass_args gives one pass's arguments.  The reverse steps
// chain their passes: the transposed stage combination that completes a
// pass's input runs in its B_G (adj_comb_g_kernel), the combination that
// reads its output in its B_OUT (adj_out_comb_kernel), and the B_G of the
// next pass as well (adj_out_g_kernel), so a step of S stages takes 2 S + 1
// launches.  (These passes take no per-candidate step counts, which would end
// the adjoint column kernels early but not the combinations.)
static Args pass_args(ecs_ctx* c, const Args& base, const double2* g, double2* out, int diss, const double2* prim,
                      bool acc = false, bool hess = false) {
  Args a = base;
  a.gin = g; a.out = out; a.out_acc = acc ? 1 : 0;
  a.include_diss = hess ? 0 : diss;
  a.aout = c->AO; a.ain = c->AI; a.prim = prim;
  a.adaptive = AD_FIXED;
  if (hess) { a.b0x = 0.0; a.b0y = 0.0; }
  return a;
}

// Step tape of one segment of the generic reverse passes (the Butcher-form
// schemes, TDRK4 and the RK4(3) schedule): per step t the step input x(t),
// the stage inputs y(t, 1..S-1), the stage values v(t, 0..S-1), the raw
// first-stage value vraw(t) (RK4(3)), `ex` further stage arrays w(t, k)
// (TDRK4: w_j = J v_j), and the derived fields of the stage inputs
// d(t, 0..S-1) (the column outputs the RHS-VJP passes read).  One forward
// pass over the segment fills it, so no step is recomputed before it is
// transposed.  aux: one derived-field array after the steps, for the column
// loads of the transpose's own RHS-VJP passes (TDRK4's at points off the
// tape), so they do not share the forward pass's column output.
struct StepTape {
  double2* cm = nullptr;  // [ckpt][uc] compact arrays
  double2* dm = nullptr;  // [ckpt][S] derived-field arrays, then aux
  size_t cs = 0, is = 0;
  int uc = 0, S = 0, ckpt = 0;
  double2* x(int t) const { return cm + (size_t)t * uc * cs; }
  double2* y(int t, int k) const { return cm + ((size_t)t * uc + k) * cs; }  // k = 1 .. S-1
  double2* v(int t, int k) const { return cm + ((size_t)t * uc + S + k) * cs; }
  double2* vraw(int t) const { return cm + ((size_t)t * uc + 2 * S) * cs; }
  double2* w(int t, int k) const { return cm + ((size_t)t * uc + 2 * S + 1 + k) * cs; }
  double2* d(int t, int k) const { return dm + ((size_t)t * S + k) * is; }
  double2* aux() const { return dm + (size_t)ckpt * S * is; }
};

// The step tapes of a generic reverse pass with S stages, when one tape of
// ckpt steps fits SEG_TAPE_BYTES (and ECSMHD_SEG_TAPE is not 0): *ntape = 1,
// or 2 with the side stream and events of the overlap when the context's
// option allows it (ecs_ctx_set_rev_overlap, ECSMHD_REV_OVERLAP) and there is
// more than one segment.  *ntape = 0: the pass recomputes each step from its
// stored input before transposing it, in the shared stage arrays.
constexpr size_t SEG_TAPE_BYTES = (size_t)4 << 30;
template <int N>
static int step_tapes(ecs_ctx* c, int B, int S, int ex, int ckpt, int nseg, StepTape* tp, int* ntape) {
  using G = Geo<N>;
  *ntape = 0;
  const size_t cs = (size_t)B * 2 * G::R * G::KC;
  const size_t is = (size_t)B * 8 * N * G::KP;
  const int uc = 2 * S + 1 + ex;
  const size_t nc = (size_t)ckpt * uc * cs, ni = ((size_t)ckpt * S + 1) * is;
  const char* env = getenv("ECSMHD_SEG_TAPE");
  if ((nc + ni) * sizeof(double2) > SEG_TAPE_BYTES || (env && env[0] == '0')) return 0;
  const char* ov = getenv("ECSMHD_REV_OVERLAP");
  const int want = (c->rev_overlap && nseg > 1 && !(ov && ov[0] == '0')) ? 2 : 1;
  TRY(grow(&c->xtape, &c->vt_cap, nc));
  TRY(grow(&c->itape, &c->it_cap, ni));
  if (want == 2) {
    TRY(grow(&c->xtape2, &c->vt2_cap, nc));
    TRY(grow(&c->itape2, &c->it2_cap, ni));
    if (!c->side) HIPCHK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    for (hipEvent_t& e : c->ev_rev)
      if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  for (int p = 0; p < want; ++p) {
    tp[p].cm = p ? c->xtape2 : c->xtape;
    tp[p].dm = p ? c->itape2 : c->itape;
    tp[p].cs = cs; tp[p].is = is; tp[p].uc = uc; tp[p].S = S; tp[p].ckpt = ckpt;
  }
  *ntape = want;
  return 0;
}

// Appends the stage-combination tables of one tape's steps (Butcher form):
// stage k of step t reads x(t) and v(t, j <= k) and gives y(t, k + 1), its
// last stage x(t + 1) (`spare` for the tape's last step slot).  h: the
// candidates' steps (null: the launch's per-step table, as RK4(3) passes it).
static void tape_tables(std::vector<Comb>& tabs, ecs_ctx* c, size_t ce, const Tableau& tb, int msk,
                        const double* h, const StepTape& T, int ckpt, double2* spare) {
  const int S = tb.S;
  for (int t = 0; t < ckpt; ++t)
    for (int k = 0; k < S; ++k) {
      const bool last = k == S - 1;
      double2* out = last ? (t + 1 < ckpt ? T.x(t + 1) : spare) : T.y(t, k + 1);
      CombBuilder cb(c, ce, tb.ediv, msk, out, nullptr);
      cb.cb.h = h;
      const int pk = last ? tb.p[S] : tb.p[k + 1];
      cb.term(T.x(t), nullptr, 1.0, 0, pk);
      for (int j = 0; j <= k; ++j) {
        const double cj = last ? tb.b[j] : tb.A[k + 1][j];
        if (cj != 0.0) cb.term(T.v(t, j), nullptr, cj, 1, pk - tb.p[j]);
      }
      tabs.push_back(cb.cb);
    }
}

// The forward pass of one segment into tape T (Butcher form, C_GEN stages):
// L steps from the checkpoint ck, step m0 + t of the schedule on slot t.
// rk43: the RK4(3) reverse pass's per-step tables (dt, and the carried-k1
// factor on stage 1 with its raw value in vraw(t)).
template <int N>
static int tape_forward(const Args& af, const StepTape& T, const Comb* ttab, int S, int ckpt, int L, int m0,
                        const double2* ck, double2* spare, int B, hipStream_t fs,
                        const std::function<void(Args&, int, int)>& stage_args) {
  HIPCHK(hipMemcpyAsync(T.x(0), ck, T.cs * sizeof(double2), hipMemcpyDeviceToDevice, fs));
  Args al = af;
  al.x0 = T.x(0);
  al.inter = T.d(0, 0);
  TRY((col<N, C_LOAD, false>(al, B, fs)));
  for (int t = 0; t < L; ++t)
    for (int k = 0; k < S; ++k) {
      const bool last = k == S - 1;
      Args ar = af;
      ar.inter = T.d(t, k);
      TRY((row<N, R_RHS, false>(ar, B, fs)));
      Args ak = af;
      ak.x0 = k == 0 ? T.x(t) : T.y(t, k);
      ak.out = T.v(t, k);
      ak.xf = last ? (t + 1 < ckpt ? T.x(t + 1) : spare) : T.y(t, k + 1);
      ak.gc = ttab + (size_t)t * S + k;
      ak.skip_inv = (last && t + 1 == L) ? 1 : 0;
      ak.inter = !last ? T.d(t, k + 1) : (t + 1 < L ? T.d(t + 1, 0) : T.d(t, k));
      if (stage_args) stage_args(ak, m0 + t, k);
      TRY((col<N, C_GEN, false>(ak, B, fs)));
    }
  return 0;
}

// TDRK4's tables for one tape: per step t, y(t, 1) = x + h/2 v1 + h^2/8 w1 and
// x(t + 1) = x + h v1 + h^2/6 w1 + h^2/3 w2 (`spare` for the last step slot),
// with x = x(t), v_j = v(t, j - 1), w_j = w(t, j - 1).
static void tape_tables_td(std::vector<Comb>& tabs, ecs_ctx* c, size_t ce, const StepTape& T, int ckpt,
                           double2* spare) {
  for (int t = 0; t < ckpt; ++t) {
    CombBuilder cy(c, ce, 1.0, 0, T.y(t, 1), nullptr);
    cy.term(T.x(t), nullptr, 1.0, 0, 0).term(T.v(t, 0), nullptr, 0.5, 1, 0).term(T.w(t, 0), nullptr, 0.125, 2, 0);
    CombBuilder cx(c, ce, 1.0, 0, t + 1 < ckpt ? T.x(t + 1) : spare, nullptr);
    cx.term(T.x(t), nullptr, 1.0, 0, 0).term(T.v(t, 0), nullptr, 1.0, 1, 0).term(T.w(t, 0), nullptr, 1.0 / 6, 2, 0)
      .term(T.w(t, 1), nullptr, 1.0 / 3, 2, 0);
    tabs.push_back(cy.cb);
    tabs.push_back(cx.cb);
  }
}

// TDRK4's forward pass of one segment into tape T: per half step a row pass,
// C_TD1 (v = RHS(z)), the tangent row pass and C_TD2 (w = J(z) v, then the
// combination of table ttab[2 t + half]), as run_gen's fused steps; the
// derived fields of x(t) and y(t, 1) go to d(t, 0) and d(t, 1).
template <int N>
static int tape_forward_td(const Args& af, const StepTape& T, const Comb* ttab, int ckpt, int L, const double2* ck,
                           double2* spare, int B, hipStream_t fs) {
  HIPCHK(hipMemcpyAsync(T.x(0), ck, T.cs * sizeof(double2), hipMemcpyDeviceToDevice, fs));
  Args al = af;
  al.x0 = T.x(0);
  al.inter = T.d(0, 0);
  TRY((col<N, C_LOAD, false>(al, B, fs)));
  for (int t = 0; t < L; ++t)
    for (int half = 0; half < 2; ++half) {
      double2* z = half == 0 ? T.x(t) : T.y(t, 1);
      Args ar = af;
      ar.inter = T.d(t, half);
      TRY((row<N, R_RHS, false>(ar, B, fs)));
      Args a1 = af;
      a1.x0 = z;
      a1.out = T.v(t, half);
      TRY((col<N, C_TD1, false>(a1, B, fs)));
      TRY((row<N, R_RHS, true>(af, B, fs)));
      Args a2 = af;
      a2.x0 = z;
      a2.dx0 = T.v(t, half);
      a2.dout = T.w(t, half);
      a2.gc = ttab + 2 * (size_t)t + half;
      a2.xf = half == 0 ? T.y(t, 1) : (t + 1 < ckpt ? T.x(t + 1) : spare);
      a2.skip_inv = (half == 1 && t + 1 == L) ? 1 : 0;
      a2.inter = half == 0 ? T.d(t, 1) : (t + 1 < L ? T.d(t + 1, 0) : T.d(t, 1));
      TRY((col<N, C_TD2, true>(a2, B, fs)));
    }
  return 0;
}

// Reverse pass of run_gen: f0 (full spectra, not modified), fbar_inout (on
// entry the cotangent of the final spectra, on exit that of f0), hbar[b] =
// d/dh.  The state is checkpointed every ckpt steps; each segment runs
// forward again from its checkpoint into a step tape (StepTape; without one,
// its step inputs are stored and each step's stages recomputed from its
// input) before the stage algebra is transposed:
//   Butcher form  x' = E^{p_S} x + h sum_j b_j E^{p_S-p_j} v_j,
//                 y_s = E^{p_s} x + h sum_j A_sj E^{p_s-p_j} v_j,  v_s = f(y_s):
//     vbar_j = h b_j E^{p_S-p_j} x'bar, then for s = S..1: ybar_s = J(y_s)^T vbar_s
//     distributed to xbar and vbar_j (j < s) by the transposed combination;
//   TDRK4 (a_j = h^2 w_j, w_j = J(z_j) v_j): wbar_j gives J(z_j)^T wbar_j to
//     vbar_j and J_0(v_j)^T wbar_j to zbar_j.
// d/dh of every weight is accumulated per mode and summed at the end.
// Two-phase use: with ck_out the call only runs the forward pass over all
// steps, storing the checkpoints into ck_out ([nseg][B][2][R][KC]) and the
// final spectra into f_out (the forward entry points' result); with ck_in
// it skips the forward pass and runs the reverse from those checkpoints.
template <int N>
static int run_gen_vjp(ecs_ctx* c, const double* f0, double* fbar, double* hbar, int B, const double* dt, int steps,
                       int scheme, int masked, int ckpt, const double2* ck_in, double2* ck_out, double* f_out,
                       hipStream_t st) {
  const bool td = scheme == G_TDRK4;
  TRY(ensure_adj(c, td ? 2 * B : B));
  TRY(ensure_cap(c, td ? 2 * B : B));
  TRY(autotune<N>(c, B, st));
  if (td) TRY(autotune<N>(c, 2 * B, st));
  TRY(ensure_adj(c, td ? 2 * B : B));   // the autotuner may have grown the batch workspace
  const size_t ce = c->compact_elems();
  const size_t cs = ce * B;
  const Tableau& tb = tableau(scheme);
  const int S = td ? 2 : tb.S;
  const int nseg = (steps + ckpt - 1) / ckpt;
  // a step tape per segment when it fits (the reverse pass only)
  StepTape tp[2];
  int ntape = 0;
  if (!ck_out) TRY(step_tapes<N>(c, B, S, td ? 2 : 0, ckpt, nseg, tp, &ntape));
  // units of cs: xw (state of the fused steps), xb, xbn, yb, 3 spare | segment inputs [ckpt]
  // (without a tape) | stages
  const int stage_units = td ? 10 : 3 * S;
  const int seg_units = ntape ? 0 : ckpt;
  const int units = 7 + seg_units + stage_units;
  TRY(grow(&c->gen, &c->gen_cap, (size_t)units * cs));
  if (!ck_in && !ck_out) TRY(grow(&c->ckpt, &c->ckpt_cap, (size_t)nseg * cs));
  const double2* ckb = ck_in ? ck_in : ck_out ? ck_out : c->ckpt;
  TRY(upload_scalars(c, 0, dt, B, st));
  Args a = base_args(c);
  double2* g = c->gen;
  auto slot = [&](int u) { return g + (size_t)u * cs; };
  double2 *xw = slot(0), *xb = slot(1), *xbn = slot(2), *yb = slot(3);
  auto xt = [&](int t) { return slot(7 + t); };
  const int s0 = 7 + seg_units;
  const int diss = scheme == G_LRK6 ? 0 : 1;
  const int msk = scheme == G_LRK6 ? masked : 0;
  const double ediv = tb.ediv;
  // stage storage
  auto y = [&](int k) { return slot(s0 + k); };           // RK: y_k (k >= 1); TD: y
  auto v = [&](int k) { return slot(s0 + S + k); };       // RK: v_k;  TD: v1, v2
  auto vb = [&](int k) { return slot(s0 + 2 * S + k); };  // RK: vbar_k; TD: vbar1, vbar2
  // TDRK4 extra: w1, w2, wbar1, wbar2 at s0 + 6 .. s0 + 9
  auto w = [&](int k) { return slot(s0 + 6 + k); };
  auto wb = [&](int k) { return slot(s0 + 8 + k); };
  const size_t bytes = cs * sizeof(double2);

  // Forward steps with the fused stage kernels of run_gen (a row pass and one
  // C_GEN column kernel per stage; TDRK4: C_TD1 / C_TD2 per half step), in
  // place on the scratch slot xw: each step stores its stage inputs y(k)
  // (TDRK4: y) and stage values v(k) (TDRK4: v1, v2, w1, w2) where the
  // transposed stage algebra below reads them, and leaves its result in xw
  // (the same stages and result as the forward entry points compute).
  std::vector<Comb> tabs;
  if (!td) {
    for (int k = 0; k < tb.S; ++k) {
      const bool last = k == tb.S - 1;
      CombBuilder cb(c, ce, ediv, msk, last ? xw : y(k + 1), nullptr);
      const int pk = last ? tb.p[tb.S] : tb.p[k + 1];
      cb.term(xw, nullptr, 1.0, 0, pk);
      for (int j = 0; j <= k; ++j) {
        const double cj = last ? tb.b[j] : tb.A[k + 1][j];
        if (cj != 0.0) cb.term(v(j), nullptr, cj, 1, pk - tb.p[j]);
      }
      tabs.push_back(cb.cb);
    }
  } else {
    // y = x + h/2 v1 + h^2/8 w1,  x' = x + h v1 + h^2/6 w1 + h^2/3 w2
    CombBuilder cy(c, ce, 1.0, 0, y(0), nullptr);
    cy.term(xw, nullptr, 1.0, 0, 0).term(v(0), nullptr, 0.5, 1, 0).term(w(0), nullptr, 0.125, 2, 0);
    CombBuilder cx(c, ce, 1.0, 0, xw, nullptr);
    cx.term(xw, nullptr, 1.0, 0, 0).term(v(0), nullptr, 1.0, 1, 0).term(w(0), nullptr, 1.0 / 6, 2, 0)
      .term(w(1), nullptr, 1.0 / 3, 2, 0);
    tabs.push_back(cy.cb);
    tabs.push_back(cx.cb);
  }
  // the tapes' per-step tables (S per step) follow the forward pass's S tables
  for (int p = 0; p < ntape; ++p) {
    if (td) tape_tables_td(tabs, c, ce, tp[p], ckpt, xw);
    else tape_tables(tabs, c, ce, tb, msk, c->scal, tp[p], ckpt, xw);
  }
  TRY(grow(&c->gcomb, &c->gcomb_cap, (tabs.size() * sizeof(Comb) + sizeof(double2) - 1) / sizeof(double2)));
  Comb* dtabs = (Comb*)c->gcomb;
  HIPCHK(hipMemcpyAsync(dtabs, tabs.data(), tabs.size() * sizeof(Comb), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // the host table goes out of scope
  Args af = a;
  af.adaptive = AD_FIXED;
  af.include_diss = diss;
  // Without a tape: the derived fields (column pass output) of the stage
  // inputs of the step being transposed, kept for its RHS-VJP passes: slot k
  // for stage input k (TDRK4: 0 the step input, 1 the midpoint y).
  const size_t is = (size_t)B * 8 * N * Geo<N>::KP;
  if (!ntape) TRY(grow(&c->itape, &c->it_cap, (size_t)(td ? 2 : S + 1) * is));
  auto ti = [&](int k) { return c->itape + (size_t)k * is; };
  // n >= 1 steps from the state in xw; after(s), when given, runs once step s
  // has left its result in xw; tape (one step): the stage inputs' derived
  // fields go to the slots ti(k) instead of the shared column output
  auto fsteps = [&](int n, const std::function<int(int)>& after, bool tape) -> int {
    Args al = af;
    al.x0 = xw;
    if (tape) al.inter = ti(0);
    TRY((col<N, C_LOAD, false>(al, B, st)));
    for (int s = 0; s < n; ++s) {
      if (!td) {
        for (int k = 0; k < tb.S; ++k) {
          const bool last = k == tb.S - 1;
          Args ar = af;
          if (tape) ar.inter = ti(k);
          TRY((row<N, R_RHS, false>(ar, B, st)));
          Args ak = af;
          ak.x0 = k == 0 ? xw : y(k);
          ak.out = v(k);
          ak.xf = last ? xw : y(k + 1);
          ak.gc = dtabs + k;
          ak.skip_inv = (s == n - 1 && last) ? 1 : 0;
          if (tape) ak.inter = ti(k + 1);
          TRY((col<N, C_GEN, false>(ak, B, st)));
        }
      } else {
        for (int half = 0; half < 2; ++half) {
          double2* z = half == 0 ? xw : y(0);
          Args ar = af;
          if (tape) ar.inter = ti(half);
          TRY((row<N, R_RHS, false>(ar, B, st)));
          Args a1 = af;
          a1.x0 = z;
          a1.out = v(half);
          TRY((col<N, C_TD1, false>(a1, B, st)));
          TRY((row<N, R_RHS, true>(af, B, st)));
          Args a2 = af;
          a2.x0 = z;
          a2.dx0 = v(half);
          a2.dout = w(half);
          a2.gc = dtabs + half;
          a2.xf = half == 0 ? y(0) : xw;
          a2.skip_inv = (s == n - 1 && half == 1) ? 1 : 0;
          if (tape && half == 0) a2.inter = ti(1);
          TRY((col<N, C_TD2, true>(a2, B, st)));
        }
      }
      if (after) TRY(after(s));
    }
    return 0;
  };
  auto copy = [&](double2* dst, const double2* src) -> int {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
    return 0;
  };

  // forward pass to the last segment's start, checkpointing every ckpt steps
  // (the forward phase: over every step, leaving the result in f_out)
  if (ck_out) {
    TRY(pack<N>(f0, xw, B, st));
    TRY(copy(ck_out, xw));
    TRY(fsteps(steps, [&](int s) -> int {
      return (s + 1) % ckpt == 0 && s + 1 < steps ? copy(ck_out + (size_t)((s + 1) / ckpt) * cs, xw) : 0;
    }, false));
    TRY(unpack<N>(xw, f_out, B, msk, st));
    return 0;
  }
  if (!ck_in) {
    TRY(pack<N>(f0, xw, B, st));
    TRY(copy(c->ckpt, xw));
    if (nseg > 1)
      TRY(fsteps((nseg - 1) * ckpt, [&](int s) -> int {
        return (s + 1) % ckpt == 0 ? copy(c->ckpt + (size_t)((s + 1) / ckpt) * cs, xw) : 0;
      }, false));
  }
  // reverse
  TRY(pack<N>(fbar, xb, B, st));
  HIPCHK(hipMemsetAsync(c->hb, 0, cs * sizeof(double), st));
  // Transpose of one Butcher-form step with input xs, stage inputs ys(k),
  // stage values vs(k) and the stage inputs' derived fields ds(k).  The
  // step's cotangent accumulators start with their first contribution
  // (assign), so they need no clearing.
  using Ptr = std::function<double2*(int)>;
  // The final combination of a step with input xs and stage values vs: the
  // cotangent of its output (ybuf) to that of its input (obuf, assigned) and
  // to vbar_j (assigned; touched[j] set for each).
  auto make_fin = [&](const double2* ybuf, double2* obuf, double2* xs, const Ptr& vs, std::vector<char>& touched) {
    CombAdjBuilder fin(c, ce, ediv, msk, ybuf, c->hb);
    fin.term(obuf, xs, 1.0, 0, tb.p[S], true);
    for (int j = 0; j < S; ++j)
      if (tb.b[j] != 0.0) {
        fin.term(vb(j), vs(j), tb.b[j], 1, tb.p[S] - tb.p[j], true);
        touched[j] = 1;
      }
    return fin;
  };
  // head_done: the step's final combination and first kernel already ran in
  // the last kernel of the step transposed before it; xs_n, vs_n: the step
  // transposed next, when its stage data are resident (the same tape), whose
  // final combination and first kernel this step's last kernel runs.
  auto rev_step = [&](double2* xs, const Ptr& ys, const Ptr& vs, const Ptr& ds, bool head_done = false,
                      double2* xs_n = nullptr, const Ptr* vs_n = nullptr) -> int {
    std::vector<char> touched(S, 0);
    CombAdjBuilder fin = make_fin(xb, xbn, xs, vs, touched);
    // the stages' RHS-VJP passes, chained (pass_args): fin runs in stage
    // S - 1's first kernel, which reads vb(S - 1); stage k's combination and
    // stage k - 1's first kernel in stage k's last
    if (!head_done) {
      if (!touched[S - 1]) HIPCHK(hipMemsetAsync(vb(S - 1), 0, bytes, st));  // a stage no combination reads
      TRY(acol_g_comb<N>(pass_args(c, a, vb(S - 1), yb, diss, ds(S - 1)), fin.cb, B, st));
    }
    const CombAdj none{};
    for (int k = S - 1; k >= 0; --k) {
      CombAdjBuilder st_(c, ce, ediv, msk, yb, c->hb);
      st_.term(xbn, k == 0 ? nullptr : xs, 1.0, 0, tb.p[k]);
      for (int j = 0; j < k; ++j)
        if (tb.A[k][j] != 0.0) {
          st_.term(vb(j), vs(j), tb.A[k][j], 1, tb.p[k] - tb.p[j], !touched[j]);
          touched[j] = 1;
        }
      const Args ak = pass_args(c, a, vb(k), yb, diss, ds(k));
      TRY(arow<N>(ak, B, st));
      if (k > 0) {
        if (!touched[k - 1]) HIPCHK(hipMemsetAsync(vb(k - 1), 0, bytes, st));
        TRY(acol_out_g<N>(ak, st_.cb, none, vb(k - 1), B, st));
      } else if (xs_n && S > 1) {
        // the next step's final combination reads this step's input cotangent
        // (xbn, complete once st_ has run) and writes the other buffer
        std::vector<char> tn(S, 0);
        CombAdjBuilder fn = make_fin(xbn, xb, xs_n, *vs_n, tn);
        if (!tn[S - 1]) HIPCHK(hipMemsetAsync(vb(S - 1), 0, bytes, st));
        TRY(acol_out_g<N>(ak, st_.cb, fn.cb, vb(S - 1), B, st));
      } else {
        TRY(acol_comb<N>(ak, st_.cb, B, st));
      }
    }
    std::swap(xb, xbn);
    return 0;
  };
  // Transpose of one TDRK4 step with input xs, midpoint ym, stage values v1,
  // v2, w1 = J(xs) v1, w2 = J(ym) v2 and the derived fields d1, dm of xs and
  // ym; ab: the base arguments of the RHS-VJP passes (their column output
  // serves the passes at v1 and v2, which load their point themselves).
  // x' = x + h v1 + h^2/6 w1 + h^2/3 w2: the cotangent of x' (ybuf) to that
  // of x (obuf) and of v1, w1, w2 (all assigned; vbar2 is written whole by
  // its VJP pass)
  auto make_fin_td = [&](const double2* ybuf, double2* obuf, double2* v1, double2* w1, double2* w2) {
    CombAdjBuilder fin(c, ce, 1.0, 0, ybuf, c->hb);
    fin.term(obuf, nullptr, 1.0, 0, 0, true).term(vb(0), v1, 1.0, 1, 0, true)
      .term(wb(0), w1, 1.0 / 6, 2, 0, true).term(wb(1), w2, 1.0 / 3, 2, 0, true);
    return fin;
  };
  // head_done, next (v1, w1, w2 of the step transposed next, on the same
  // tape): as for rev_step
  auto rev_td = [&](const Args& ab, double2* xs, double2* ym, double2* v1, double2* v2, double2* w1, double2* w2,
                    double2* d1, double2* dm, bool head_done = false, double2* const* next = nullptr) -> int {
    CombAdjBuilder fin = make_fin_td(xb, xbn, v1, w1, w2);
    // y = x + h/2 v1 + h^2/8 w1 (its transposed combination follows the third pass)
    CombAdjBuilder cy(c, ce, 1.0, 0, yb, c->hb);
    cy.term(xbn, nullptr, 1.0, 0, 0).term(vb(0), v1, 0.5, 1, 0).term(wb(0), w1, 0.125, 2, 0);
    // The six RHS-VJP passes, chained (pass_args): each pass's input is
    // complete when the pass before it ends, so its B_G runs in that pass's
    // last kernel; the two passes at v1, v2 load their points (into the
    // base's column output) before that kernel.  fin runs in the first B_G.
    //   w2 = J(y) v2, v2 = f(y): vbar2 = J(y)^T wbar2; ybar = J(y)^T vbar2 + J_0(v2)^T wbar2
    //   w1 = J(x) v1: vbar1 += J(x)^T wbar1, xbar += J_0(v1)^T wbar1
    //   v1 = f(x): xbar += J(x)^T vbar1
    const Args p1 = pass_args(c, ab, wb(1), vb(1), diss, dm);
    const Args p2 = pass_args(c, ab, vb(1), yb, diss, dm);
    Args p3 = pass_args(c, ab, wb(1), yb, diss, ab.inter, true, true);
    const Args p4 = pass_args(c, ab, wb(0), vb(0), diss, d1, true);
    Args p5 = pass_args(c, ab, wb(0), xbn, diss, ab.inter, true, true);
    const Args p6 = pass_args(c, ab, vb(0), xbn, diss, d1, true);
    p3.x0 = v2;
    p5.x0 = v1;
    const CombAdj none{};
    if (!head_done) TRY(acol_g_comb<N>(p1, fin.cb, B, st));
    TRY(arow<N>(p1, B, st));
    TRY(acol_out_g<N>(p1, none, none, vb(1), B, st));
    TRY(arow<N>(p2, B, st));
    TRY((col<N, C_LOAD, false>(p3, B, st)));
    TRY(acol_out_g<N>(p2, none, none, wb(1), B, st));
    TRY(arow<N>(p3, B, st));
    TRY(acol_out_g<N>(p3, cy.cb, none, wb(0), B, st));
    TRY(arow<N>(p4, B, st));
    TRY((col<N, C_LOAD, false>(p5, B, st)));
    TRY(acol_out_g<N>(p4, none, none, wb(0), B, st));
    TRY(arow<N>(p5, B, st));
    TRY(acol_out_g<N>(p5, none, none, vb(0), B, st));
    TRY(arow<N>(p6, B, st));
    if (next) {
      // the next step's final combination (reads xbn, complete after this
      // pass) and its first pass's B_G, which reads wbar2
      CombAdjBuilder fn = make_fin_td(xbn, xb, next[0], next[1], next[2]);
      TRY(acol_out_g<N>(p6, none, fn.cb, wb(1), B, st));
    } else {
      TRY((acol<N, B_OUT>(p6, B, st)));
    }
    std::swap(xb, xbn);
    return 0;
  };
  if (ntape) {
    // Segment j's forward pass fills tape j % 2 (writing only the tape, the
    // forward workspace and the spare slot xw); its transpose reads the tape
    // and writes only the cotangent slots, the adjoint workspace and the
    // tape's aux array (TDRK4's column loads of v1, v2).  With
    // two tapes, segment j - 1 runs forward on the side stream while the
    // caller's stream transposes segment j (the same kernels in the same
    // order in each chain, so the same bits); a tape is refilled only after
    // the transpose that read it (events).
    const bool two = ntape == 2 && nseg > 1;
    hipStream_t fs = two ? c->side : st;
    if (two) {
      HIPCHK(hipEventRecord(c->ev_rev[0], st));
      HIPCHK(hipStreamWaitEvent(fs, c->ev_rev[0], 0));
    }
    for (int jseg = nseg - 1; jseg >= 0; --jseg) {
      const int t0 = jseg * ckpt;
      const int L = std::min(ckpt, steps - t0);
      const int p = two ? (jseg & 1) : 0;
      const StepTape& tap = tp[p];
      if (two && jseg + 2 < nseg) HIPCHK(hipStreamWaitEvent(fs, c->ev_rev[3 + p], 0));
      const Comb* ttab = dtabs + S + (size_t)p * ckpt * S;
      const double2* ck = ckb + (size_t)jseg * cs;
      if (td) TRY(tape_forward_td<N>(af, tap, ttab, ckpt, L, ck, xw, B, fs));
      else TRY(tape_forward<N>(af, tap, ttab, S, ckpt, L, t0, ck, xw, B, fs, nullptr));
      if (two) {
        HIPCHK(hipEventRecord(c->ev_rev[1 + p], fs));
        HIPCHK(hipStreamWaitEvent(st, c->ev_rev[1 + p], 0));
      }
      Args ab = a;
      ab.inter = tp[0].aux();
      for (int t = L - 1; t >= 0; --t) {
        if (td) {
          double2* const nx[3] = {t > 0 ? tap.v(t - 1, 0) : nullptr, t > 0 ? tap.w(t - 1, 0) : nullptr,
                                  t > 0 ? tap.w(t - 1, 1) : nullptr};
          TRY(rev_td(ab, tap.x(t), tap.y(t, 1), tap.v(t, 0), tap.v(t, 1), tap.w(t, 0), tap.w(t, 1), tap.d(t, 0),
                     tap.d(t, 1), t < L - 1, t > 0 ? nx : nullptr));
        } else {
          // consecutive steps of one tape chain across the step boundary
          const Ptr vn = [&](int k) { return tap.v(t - 1, k); };
          TRY(rev_step(tap.x(t), [&](int k) { return tap.y(t, k); }, [&](int k) { return tap.v(t, k); },
                       [&](int k) { return tap.d(t, k); }, t < L - 1 && S > 1, t > 0 ? tap.x(t - 1) : nullptr,
                       t > 0 ? &vn : nullptr));
        }
      }
      if (two) HIPCHK(hipEventRecord(c->ev_rev[3 + p], st));
    }
  }
  for (int jseg = ntape ? -1 : nseg - 1; jseg >= 0; --jseg) {
    const int t0 = jseg * ckpt;
    const int L = std::min(ckpt, steps - t0);
    // the segment's step inputs, then per step (last first) its stages
    TRY(copy(xt(0), ckb + (size_t)jseg * cs));
    if (L > 1) {
      TRY(copy(xw, xt(0)));
      TRY(fsteps(L - 1, [&](int t) -> int { return copy(xt(t + 1), xw); }, false));
    }
    for (int t = L - 1; t >= 0; --t) {
      double2* xs = xt(t);
      TRY(copy(xw, xs));
      TRY(fsteps(1, nullptr, true));
      if (td) TRY(rev_td(a, xs, y(0), v(0), v(1), w(0), w(1), ti(0), ti(1)));
      else TRY(rev_step(xs, y, v, ti));
    }
  }
  TRY(unpack<N>(xb, fbar, B, msk, st));
  hipLaunchKernelGGL((hb_sum_kernel<N>), dim3(B), dim3(256), 0, st, c->hb, c->red);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(c->hred, c->red, B * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int b2 = 0; b2 < B; ++b2) hbar[b2] = c->hred[b2];
  return 0;
}

// Reverse pass of the adaptive Lawson RK4(3) along a recorded schedule.  h is
// stop-gradient (timestepping.py:308), so the trajectory is a sequence of
// Lawson-RK4 steps of dt_m = T h_m with unmasked e = exp(dt_m/2 L), except that
// the first stage's v1 = v(x) is multiplied per mode by P_m = exp(T s_m L),
// s_m the sum of the steps of the rejections before accepted step m (each
// rejection turns the carried k1 into e e k1, timestepping.py:267,276,297).
// Candidates with fewer accepted steps take zero-length steps (dt = 0: the
// identity, with zero derivative).  d/dT is accumulated per mode with the
// weights dt_m/dT = h_m and d(T s_m)/dT = s_m.  Returns Tbar[b] on the host.
// Two-phase use: with ck_out the call only runs the forward pass to the last
// segment's start, storing the checkpoints of the accepted steps into ck_out
// ([nseg][B][2][R][KC], nseg = ceil(max(1, most accepted steps) / ckpt));
// with ck_in it runs the reverse from those checkpoints.
template <int N>
static int run_rk43_vjp(ecs_ctx* c, const double* x0f, double* xbar, double* Tbar, int B, const double* T,
                        const double* hs, const int8_t* accf, const int* attempts, int budget, int ckpt,
                        const double2* ck_in, double2* ck_out, hipStream_t st) {
  TRY(ensure_adj(c, B));
  TRY(autotune<N>(c, B, st));
  TRY(ensure_adj(c, B));
  // accepted-step tables [m][b]: dt, T s, h, s
  std::vector<int> M(B, 0);
  std::vector<std::vector<double>> hh(B), ss(B);
  int Mmax = 0;
  for (int b = 0; b < B; ++b) {
    double sig = 0.0;
    for (int i = 0; i < attempts[b]; ++i) {
      const double h = hs[(size_t)b * budget + i];
      if (accf[(size_t)b * budget + i]) {
        hh[b].push_back(h);
        ss[b].push_back(sig);
        sig = 0.0;
      } else {
        sig += h;
      }
    }
    M[b] = (int)hh[b].size();
    Mmax = std::max(Mmax, M[b]);
  }
  const int steps = std::max(Mmax, 1);
  std::vector<double> tab((size_t)4 * steps * B, 0.0);
  for (int m = 0; m < steps; ++m)
    for (int b = 0; b < B; ++b) {
      const bool on = m < M[b];
      tab[((size_t)0 * steps + m) * B + b] = on ? T[b] * hh[b][m] : 0.0;
      tab[((size_t)1 * steps + m) * B + b] = on ? T[b] * ss[b][m] : 0.0;
      tab[((size_t)2 * steps + m) * B + b] = on ? hh[b][m] : 0.0;
      tab[((size_t)3 * steps + m) * B + b] = on ? ss[b][m] : 0.0;
    }
  {
    std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
    size_t cap = c->tab_cap;
    if (tab.size() > cap) {
      if (c->tab) HIPCHK(hipFree(c->tab));
      c->tab = nullptr;
      HIPCHK(hipMalloc((void**)&c->tab, tab.size() * sizeof(double)));
      c->tab_cap = tab.size();
    }
  }
  HIPCHK(hipMemcpyAsync(c->tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));   // the host table goes out of scope
  auto DT = [&](int m) { return (const double*)c->tab + ((size_t)0 * steps + m) * B; };
  auto PDT = [&](int m) { return (const double*)c->tab + ((size_t)1 * steps + m) * B; };
  auto HSC = [&](int m) { return (const double*)c->tab + ((size_t)2 * steps + m) * B; };
  auto SIG = [&](int m) { return (const double*)c->tab + ((size_t)3 * steps + m) * B; };

  const size_t ce = c->compact_elems();
  const size_t cs = ce * B;
  const Tableau& tb = tableau(G_LRK4);
  const int S = tb.S;
  const int nseg = (steps + ckpt - 1) / ckpt;
  StepTape tp[2];
  int ntape = 0;
  if (!ck_out) TRY(step_tapes<N>(c, B, S, 0, ckpt, nseg, tp, &ntape));
  // units: xw (state of the fused steps), xb, xbn, yb, spare, vraw, vbraw | segment inputs [ckpt]
  // (without a tape) | y[S], v[S], vbar[S]
  const int seg_units = ntape ? 0 : ckpt;
  const int units = 7 + seg_units + 3 * S;
  TRY(grow(&c->gen, &c->gen_cap, (size_t)units * cs));
  if (!ck_in && !ck_out) TRY(grow(&c->ckpt, &c->ckpt_cap, (size_t)nseg * cs));
  double2* ckw = ck_out ? ck_out : c->ckpt;            // where the forward pass writes
  const double2* ckb = ck_in ? ck_in : ckw;             // where the reverse pass reads
  Args a = base_args(c);
  double2* g = c->gen;
  auto slot = [&](int u) { return g + (size_t)u * cs; };
  double2 *xw = slot(0), *xb = slot(1), *xbn = slot(2), *yb = slot(3), *vraw = slot(5);
  double2* vbraw = slot(6);
  auto xt = [&](int t) { return slot(7 + t); };
  const int s0 = 7 + seg_units;
  auto y = [&](int k) { return slot(s0 + k); };
  auto v = [&](int k) { return slot(s0 + S + k); };
  auto vb = [&](int k) { return slot(s0 + 2 * S + k); };
  const size_t bytes = cs * sizeof(double2);
  const double ediv = tb.ediv;

  // Forward steps with the fused stage kernels (a row pass and one C_GEN
  // column kernel per stage; C_GEN takes step m's dt from the launch and, on
  // stage 1, stores the raw RHS at vraw and scales it by P = exp(T s L)), in
  // place on xw; tables without a step array serve every step.  Each step
  // stores its stage inputs y(k), stage values v(k) and vraw.
  std::vector<Comb> tabs;
  for (int k = 0; k < S; ++k) {
    const bool last = k == S - 1;
    CombBuilder cb(c, ce, ediv, 0, last ? xw : y(k + 1), nullptr);
    cb.cb.h = nullptr;
    const int pk = last ? tb.p[S] : tb.p[k + 1];
    cb.term(xw, nullptr, 1.0, 0, pk);
    for (int j = 0; j <= k; ++j) {
      const double cj = last ? tb.b[j] : tb.A[k + 1][j];
      if (cj != 0.0) cb.term(v(j), nullptr, cj, 1, pk - tb.p[j]);
    }
    tabs.push_back(cb.cb);
  }
  // the tapes' per-step tables (the launch's dt) follow the forward pass's S tables
  for (int p = 0; p < ntape; ++p) tape_tables(tabs, c, ce, tb, 0, nullptr, tp[p], ckpt, xw);
  TRY(grow(&c->gcomb, &c->gcomb_cap, (tabs.size() * sizeof(Comb) + sizeof(double2) - 1) / sizeof(double2)));
  Comb* dtabs = (Comb*)c->gcomb;
  HIPCHK(hipMemcpyAsync(dtabs, tabs.data(), tabs.size() * sizeof(Comb), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // the host table goes out of scope
  // without a tape: derived fields of the stage inputs of the step being transposed
  const size_t is = (size_t)B * 8 * N * Geo<N>::KP;
  if (!ntape) TRY(grow(&c->itape, &c->it_cap, (size_t)(S + 1) * is));
  auto ti = [&](int k) { return c->itape + (size_t)k * is; };
  Args af = a;
  af.adaptive = AD_FIXED;
  af.include_diss = 0;  // v_fn of the Lawson schemes has no dissipation
  // steps m0 .. m0 + n - 1 (n >= 1) from the state in xw; after(s) runs once
  // step m0 + s has left its result in xw; tape (one step): the stage inputs'
  // derived fields go to ti(k)
  auto fsteps = [&](int m0, int n, const std::function<int(int)>& after, bool tape) -> int {
    Args al = af;
    al.x0 = xw;
    if (tape) al.inter = ti(0);
    TRY((col<N, C_LOAD, false>(al, B, st)));
    for (int s = 0; s < n; ++s) {
      const int m = m0 + s;
      for (int k = 0; k < S; ++k) {
        const bool last = k == S - 1;
        Args ar = af;
        if (tape) ar.inter = ti(k);
        TRY((row<N, R_RHS, false>(ar, B, st)));
        Args ak = af;
        ak.dt = DT(m);
        ak.x0 = k == 0 ? xw : y(k);
        ak.out = v(k);
        ak.xf = last ? xw : y(k + 1);
        ak.gc = dtabs + k;
        if (k == 0) {
          ak.pdt = PDT(m);
          ak.vraw = vraw;
        }
        ak.skip_inv = (s == n - 1 && last) ? 1 : 0;
        if (tape) ak.inter = ti(k + 1);
        TRY((col<N, C_GEN, false>(ak, B, st)));
      }
      if (after) TRY(after(s));
    }
    return 0;
  };
  auto copy = [&](double2* dst, const double2* src) -> int {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
    return 0;
  };

  // forward pass to the last segment's start, checkpointing every ckpt steps
  if (!ck_in) {
    TRY(pack<N>(x0f, xw, B, st));
    TRY(copy(ckw, xw));
    if (nseg > 1)
      TRY(fsteps(0, (nseg - 1) * ckpt, [&](int s) -> int {
        return (s + 1) % ckpt == 0 ? copy(ckw + (size_t)((s + 1) / ckpt) * cs, xw) : 0;
      }, false));
  }
  if (ck_out) return 0;
  TRY(pack<N>(xbar, xb, B, st));
  HIPCHK(hipMemsetAsync(c->hb, 0, cs * sizeof(double), st));
  // Transpose of accepted step m with input xs, stage inputs ys(k), stage
  // values vs(k), raw first-stage value vr and the stage inputs' derived
  // fields ds(k).  The step's cotangent accumulators start with their first
  // contribution.
  using Ptr = std::function<double2*(int)>;
  // the final combination of accepted step m (as make_fin of run_gen_vjp)
  auto make_fin = [&](int m, const double2* ybuf, double2* obuf, double2* xs, const Ptr& vs,
                      std::vector<char>& touched) {
    CombAdjBuilder fin(c, ce, ediv, 0, ybuf, c->hb);
    fin.cb.h = DT(m);
    fin.cb.hscale = HSC(m);
    fin.term(obuf, xs, 1.0, 0, tb.p[S], true);
    for (int j = 0; j < S; ++j)
      if (tb.b[j] != 0.0) {
        fin.term(vb(j), vs(j), tb.b[j], 1, tb.p[S] - tb.p[j], true);
        touched[j] = 1;
      }
    return fin;
  };
  // head_done, xs_n, vs_n (step m - 1, on the same tape): as in run_gen_vjp
  auto rev_step = [&](int m, double2* xs, const Ptr& ys, const Ptr& vs, double2* vr, const Ptr& ds,
                      bool head_done = false, double2* xs_n = nullptr, const Ptr* vs_n = nullptr) -> int {
    std::vector<char> touched(S, 0);
    CombAdjBuilder fin = make_fin(m, xb, xbn, xs, vs, touched);
    // v1 = P vraw: vbraw = P vbar1, d/dT through T s (the combination that
    // gives stage 0's input once vbar1 is complete)
    CombAdjBuilder pc(c, ce, 1.0, 0, vb(0), c->hb);
    pc.cb.h = PDT(m);
    pc.cb.hscale = SIG(m);
    pc.term(vbraw, vr, 1.0, 0, 1, true);
    auto gin = [&](int k) { return k == 0 ? vbraw : vb(k); };
    Args am = a;
    am.dt = DT(m);
    // the stages' RHS-VJP passes, chained (pass_args): fin runs in stage
    // S - 1's first kernel; stage k's combination sa, pc when k = 1, and
    // stage k - 1's first kernel in stage k's last
    if (!head_done) {
      if (!touched[S - 1]) HIPCHK(hipMemsetAsync(vb(S - 1), 0, bytes, st));  // a stage no combination reads
      const Args a0 = pass_args(c, am, gin(S - 1), yb, 0, ds(S - 1));
      if (S == 1) {
        TRY(comb_adj_launch<N>(fin.cb, B, st));   // a one-stage table: fin before pc
        TRY(acol_g_comb<N>(a0, pc.cb, B, st));
      } else {
        TRY(acol_g_comb<N>(a0, fin.cb, B, st));
      }
    }
    const CombAdj none{};
    for (int k = S - 1; k >= 0; --k) {
      CombAdjBuilder sa(c, ce, ediv, 0, yb, c->hb);
      sa.cb.h = DT(m);
      sa.cb.hscale = HSC(m);
      sa.term(xbn, k == 0 ? nullptr : xs, 1.0, 0, tb.p[k]);
      for (int j = 0; j < k; ++j)
        if (tb.A[k][j] != 0.0) {
          sa.term(vb(j), vs(j), tb.A[k][j], 1, tb.p[k] - tb.p[j], !touched[j]);
          touched[j] = 1;
        }
      const Args ak = pass_args(c, am, gin(k), yb, 0, ds(k));
      TRY(arow<N>(ak, B, st));
      if (k > 0) {
        if (!touched[k - 1]) HIPCHK(hipMemsetAsync(vb(k - 1), 0, bytes, st));
        TRY(acol_out_g<N>(ak, sa.cb, k == 1 ? pc.cb : none, gin(k - 1), B, st));
      } else if (xs_n && S > 1) {
        // step m - 1's final combination (reads xbn, complete once sa has run)
        // and its first kernel
        std::vector<char> tn(S, 0);
        CombAdjBuilder fn = make_fin(m - 1, xbn, xb, xs_n, *vs_n, tn);
        if (!tn[S - 1]) HIPCHK(hipMemsetAsync(vb(S - 1), 0, bytes, st));
        TRY(acol_out_g<N>(ak, sa.cb, fn.cb, vb(S - 1), B, st));
      } else {
        TRY(acol_comb<N>(ak, sa.cb, B, st));
      }
    }
    std::swap(xb, xbn);
    return 0;
  };
  if (ntape) {
    // the segments' forward passes fill the tapes, on the side stream with two
    // (as in run_gen_vjp)
    const bool two = ntape == 2 && nseg > 1;
    hipStream_t fs = two ? c->side : st;
    if (two) {
      HIPCHK(hipEventRecord(c->ev_rev[0], st));
      HIPCHK(hipStreamWaitEvent(fs, c->ev_rev[0], 0));
    }
    for (int jseg = nseg - 1; jseg >= 0; --jseg) {
      const int t0 = jseg * ckpt;
      const int L = std::min(ckpt, steps - t0);
      const int p = two ? (jseg & 1) : 0;
      const StepTape& tap = tp[p];
      if (two && jseg + 2 < nseg) HIPCHK(hipStreamWaitEvent(fs, c->ev_rev[3 + p], 0));
      TRY(tape_forward<N>(af, tap, dtabs + S + (size_t)p * ckpt * S, S, ckpt, L, t0, ckb + (size_t)jseg * cs,
                          xw, B, fs, [&](Args& ak, int m, int k) {
                            ak.dt = DT(m);
                            if (k == 0) {
                              ak.pdt = PDT(m);
                              ak.vraw = tap.vraw(m - t0);
                            }
                          }));
      if (two) {
        HIPCHK(hipEventRecord(c->ev_rev[1 + p], fs));
        HIPCHK(hipStreamWaitEvent(st, c->ev_rev[1 + p], 0));
      }
      for (int t = L - 1; t >= 0; --t) {
        // consecutive steps of one tape chain across the step boundary
        const Ptr vn = [&](int k) { return tap.v(t - 1, k); };
        TRY(rev_step(t0 + t, tap.x(t), [&](int k) { return tap.y(t, k); }, [&](int k) { return tap.v(t, k); },
                     tap.vraw(t), [&](int k) { return tap.d(t, k); }, t < L - 1 && S > 1,
                     t > 0 ? tap.x(t - 1) : nullptr, t > 0 ? &vn : nullptr));
      }
      if (two) HIPCHK(hipEventRecord(c->ev_rev[3 + p], st));
    }
  }
  for (int jseg = ntape ? -1 : nseg - 1; jseg >= 0; --jseg) {
    const int t0 = jseg * ckpt;
    const int L = std::min(ckpt, steps - t0);
    TRY(copy(xt(0), ckb + (size_t)jseg * cs));
    if (L > 1) {
      TRY(copy(xw, xt(0)));
      TRY(fsteps(t0, L - 1, [&](int t) -> int { return copy(xt(t + 1), xw); }, false));
    }
    for (int t = L - 1; t >= 0; --t) {
      const int m = t0 + t;
      double2* xs = xt(t);
      TRY(copy(xw, xs));
      TRY(fsteps(m, 1, nullptr, true));
      TRY(rev_step(m, xs, y, v, vraw, ti));
    }
  }
  TRY(unpack<N>(xb, xbar, B, 0, st));
  hipLaunchKernelGGL((hb_sum_kernel<N>), dim3(B), dim3(256), 0, st, c->hb, c->red);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(c->hred, c->red, B * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int b2 = 0; b2 < B; ++b2) Tbar[b2] = c->hred[b2];
  return 0;
}

template <int N> static int gen_p(ecs_ctx* c, double* f, int B, const double* dt, int steps, int scheme, int masked,
                                  hipStream_t st) {
  return run_gen<N, false>(c, f, nullptr, B, dt, nullptr, steps, scheme, masked, st);
}
template <int N> static int gen_t(ecs_ctx* c, double* f, double* df, int B, const double* dt, const double* ddt,
                                  int steps, int scheme, int masked, hipStream_t st) {
  return run_gen<N, true>(c, f, df, B, dt, ddt, steps, scheme, masked, st);
}

#define ECS_GRID(NV)                                                                                     \
  const GridOps* grid_ops_##NV() {                                                                       \
    static const GridOps ops = {&run_state_vel<NV>, &rk4_p<NV>,          &rk4_t<NV>,    &run_rk43<NV>,  \
                                &run_rk43_jvp<NV>,  &run_state_vel_vjp<NV>, &run_rk4_ckpt<NV>,            \
                                &run_rk4_rev<NV>,   &run_rk4_vjp<NV>,    &run_time_kernel<NV>,            \
                                &gen_p<NV>,         &gen_t<NV>,          &run_gen_vjp<NV>,                \
                                &run_rk43_vjp<NV>,  &run_state_vel_noalloc<NV>};                          \
    return &ops;                                                                                         \
  }
