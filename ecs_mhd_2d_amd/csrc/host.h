// Host-side state shared by the translation units of the HIP library:
// error reporting, the context (device workspace) and its allocation helpers,
// the autotuner's tables, hipGraph replay, and the per-grid-size dispatch
// table.  The grid-size-templated orchestration lives in engine.h, which is
// compiled once per supported N (grid_n*.hip) so the kernels build in parallel.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>
#include "kernels.h"
#include "../../include/ecsmhd.h"

using namespace ecs;

// Records msg as the calling thread's last error (ecs_last_error) and returns code.
int fail(int code, const std::string& msg);
#define HIPCHK(x)                                                                           \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return fail(2, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define LAUNCHCHK() HIPCHK(hipGetLastError())
#define TRY(x)              \
  do {                      \
    int r_ = (x);           \
    if (r_) return r_;      \
  } while (0)

// Marks a context busy for the duration of one entry point; a second,
// concurrent call on the same context fails with code 4 instead of racing on
// (or reallocating) a workspace that kernels of the first call still use.
struct CtxGuard {
  std::atomic<int>* flag;
  bool ok;
  explicit CtxGuard(std::atomic<int>* f) : flag(f), ok(f->exchange(1) == 0) {}
  ~CtxGuard() {
    if (ok) flag->store(0);
  }
};
// Makes `dev` the calling thread's current device for the scope of a call
// and restores the previous one afterwards: a context's workspace grows, and
// its kernels launch, on the context's own device whatever the caller's
// current device is.
struct DevGuard {
  int prev = -1;
  bool ok = true;
  explicit DevGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) {
      ok = false;
      return;
    }
    if (prev != dev) ok = hipSetDevice(dev) == hipSuccess;
  }
  ~DevGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};
#define GUARD(c)                                                                              \
  CtxGuard guard_(&(c)->busy);                                                                \
  if (!guard_.ok) return fail(4, "context is in use by another call (use one context per thread)"); \
  DevGuard devguard_((c)->dev);                                                               \
  if (!devguard_.ok) return fail(2, "cannot make the context's device current")

constexpr int NT = 256;

// Fixed tiles of the adjoint kernels: columns per column-kernel block (TY)
// and rows per row-kernel block (TX).
template <int N> struct Tiles;
template <> struct Tiles<16> { static constexpr int TY = 2, TX = 2; };
template <> struct Tiles<32> { static constexpr int TY = 4, TX = 2; };
template <> struct Tiles<64> { static constexpr int TY = 4, TX = 2; };
template <> struct Tiles<128> { static constexpr int TY = 4, TX = 2; };
template <> struct Tiles<256> { static constexpr int TY = 2, TX = 1; };

struct ecs_ctx {
  std::atomic<int> busy{0};    // one call at a time: the workspace is shared by all entry points
  int n = 0, dev = 0;
  double nu = 0, eta = 0, b0x = 0, b0y = 0;
  double2* tw = nullptr;
  double2* fhat = nullptr;
  int cap = 0;  // batch capacity of the workspace
  double2 *I = nullptr, *O = nullptr;
  double2 *x0 = nullptr, *acc = nullptr, *v1 = nullptr, *v4 = nullptr, *v5 = nullptr, *xf = nullptr, *out = nullptr;
  double2 *dx0 = nullptr, *dacc = nullptr, *dv1 = nullptr, *dv4 = nullptr, *dxf = nullptr, *dout = nullptr;
  double* scal = nullptr;      // device [4][cap]: dt, ddt, T, dT
  double* hscal = nullptr;     // pinned host staging [4][cap]
  hipEvent_t ev_scal[4] = {};  // completion of the last copy out of each staging slot
  int* nstep = nullptr;        // device [cap]: per-candidate step counts
  int* hnstep = nullptr;       // pinned staging [cap]
  hipEvent_t ev_nstep = nullptr;
  Hist* hist = nullptr;
  size_t hist_cap = 0;
  // adaptive tangent schedule
  double* s_h = nullptr;
  int* s_rs = nullptr;
  int* s_rc = nullptr;
  double* s_rh = nullptr;
  int* s_nacc = nullptr;
  size_t sched_cap = 0;
  int snacc_cap = 0;
  // adjoint workspace
  double2* ckpt = nullptr;     // checkpoints [nck][B][2][R][KC]
  size_t ckpt_cap = 0;
  // step tape of a reverse pass: compact arrays (Lawson RK4: the step inputs
  // [ckpt][B][2][R][KC], then the stage values [ckpt][4][...]; the generic
  // passes: StepTape's per-step arrays) and the primal column outputs
  // (Lawson RK4: [ckpt][4][B][8][N][KP]; StepTape: [ckpt][S][...] + aux)
  double2* xtape = nullptr;
  size_t vt_cap = 0;           // elements of xtape
  double2* itape = nullptr;
  size_t it_cap = 0;
  // a second step tape, so the reverse pass recomputes one segment on the
  // side stream while the caller's stream sweeps the next one backward
  double2* xtape2 = nullptr;
  size_t vt2_cap = 0;
  double2* itape2 = nullptr;
  size_t it2_cap = 0;
  hipStream_t side = nullptr;  // default priority, non-blocking
  bool rev_overlap = true;     // ecs_ctx_set_rev_overlap
  hipEvent_t ev_rev[5] = {};   // start, recomputed[2], swept[2]
  double2* gen = nullptr;      // stage arrays of the generic steppers
  size_t gen_cap = 0;
  double2* gcomb = nullptr;    // stage-combination tables of the fused generic steppers (Comb records)
  size_t gcomb_cap = 0;
  double* tab = nullptr;       // per-step tables of the RK4(3) reverse pass
  size_t tab_cap = 0;
  double2 *AO = nullptr, *AI = nullptr, *xb = nullptr, *k1b = nullptr, *k2b = nullptr, *k3b = nullptr;
  double* hb = nullptr;
  int adj_cap = 0;
  // hipGraph capture/replay of fixed chunks of stages
  hipStream_t gs = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  int* cnt = nullptr;          // device step counter [2]
  int* hcnt = nullptr;         // pinned staging
  std::map<std::string, hipGraphExec_t> graphs;
  double* red = nullptr;       // reductions [cap][64]
  double* hred = nullptr;      // pinned [cap]
  double2* noalloc = nullptr;  // device address of the library's static scratch on this context's GPU
  // measurement hook: event pairs around the first tangent row pass of every
  // probe_every-th accepted step of the tangent replay (ecs_replay_probe)
  int probe_every = 0;
  std::vector<hipEvent_t> probe_ev;
  size_t probe_used = 0;
  size_t compact_elems() const {
    int kmax = (n - 1) / 3;
    return (size_t)2 * (2 * kmax + 1) * (kmax + 1);
  }
};

// Allocation, frees and synchronous copies must not run while another thread
// captures a graph (they would join the capture's dependencies), so both take
// this lock.
extern std::recursive_mutex g_alloc_mu;

void drop_graphs(ecs_ctx* c);
// Grows the batch workspace (state arrays, column/row intermediates, scalars) to hold B candidates.
int ensure_cap(ecs_ctx* c, int B);
int ensure_hist(ecs_ctx* c, size_t need);
int ensure_adj(ecs_ctx* c, int B);
int grow(double2** p, size_t* cap, size_t need);
size_t ckpt_elems(int n, int B, int steps, int ckpt);
Args base_args(ecs_ctx* c);
// Copies B per-candidate scalars into slot (0 dt, 1 ddt, 2 T, 3 dT) of the device scalar table.
int upload_scalars(ecs_ctx* c, int slot, const double* vals, int B, hipStream_t st);
// Copies B per-candidate step counts into the device table c->nstep.
int upload_ints(ecs_ctx* c, const int* vals, int B, hipStream_t st);

// ------------------------------------------------------------------ autotuner tables
// Tile variants: column kernels take TY in {1, 2, 4} columns per block, row
// kernels TX in {1, 2, 4} rows per block; both come in a wide and a narrow
// FFT radix plan.  Variant v < 6: tile 1 << (v % 3), plan v / 3 (0: wide,
// radix 16/8; 1: narrow, radix 8/4/2).  Column families also have v = 6..9:
// one column per block with the column's output sequences split over 2 (even
// v) or 4 (odd v) blocks, plan (v - 6) / 2; the split spreads a small batch
// over more CUs and does not change the arithmetic.  For N >= 128 column
// families also have v = 10, 11: tiles of 8 columns (whole 128-byte lines of
// the padded x-space arrays), plan v - 10; they pay at large batches, where
// wider tiles move the same bytes in fewer, fuller transactions.  And v =
// 12..15: tiles of 2 columns split over 2 (even v) or 4 (odd v) blocks, plan
// (v - 12) / 2, for small batches (32-byte runs instead of 16-byte ones).  For
// N >= 128 the primal column family also has v = 16: tiles of 8 columns, narrow
// plan, with the forward transform's last pass, the glue and the inverse's
// first pass in registers (the same arithmetic as v = 11).  The autotuner picks the
// plan once per (N, kernel family), so every batch size computes with the
// same arithmetic, and the variant within that plan per (N, kernel family,
// batch bucket).
// K_COLR and K_ROWR are the tangent column and row kernels as the tangent
// replay of an adaptive schedule runs them (stepper kind AD_TANGENT): the
// same kernels and FFT plan as K_COLT and K_ROWT, with a variant table of
// their own, timed in that instance (at N = 128 and 256, batch 1 the replay
// instance and the fixed-step one rank the variants differently).
enum TuneKind : int { K_COL = 0, K_COLT = 1, K_ROW = 2, K_ROWT = 3, K_COLR = 4, K_ROWR = 5 };
constexpr int NKIND = 6;
inline int base_kind(int kind) { return kind == K_COLR ? K_COLT : kind == K_ROWR ? K_ROWT : kind; }
constexpr int VMAX = 17;  // variant indices of one family lie in [0, VMAX)
inline bool var_ok(int kind, int n, int v) {
  if (v < 0 || v >= VMAX) return false;
  kind = base_kind(kind);
  if (kind == K_ROW || kind == K_ROWT) return v < 6;
  if (v == 16) return kind == K_COL && n >= 128;
  return v < 10 || v >= 12 || n >= 128;
}
inline int var_plan(int v) {
  return v < 6 ? v / 3 : v < 10 ? (v - 6) / 2 : v < 12 ? v - 10 : v < 16 ? (v - 12) / 2 : 1;
}
inline int var_tile(int v) { return v < 6 ? 1 << (v % 3) : v < 10 ? 1 : v < 12 ? 8 : v < 16 ? 2 : 8; }
inline int var_split(int v) { return (v < 6 || v == 10 || v == 11 || v == 16) ? 1 : (v % 2 == 0 ? 2 : 4); }
// padded row stride of the x-space arrays of grid n (Geo<N>::KP)
inline int kp_of(int n) { return (((n - 1) / 3 + 1) + 7) & ~7; }
// the variants of family `kind` of grid n in plan p (every plan for p < 0);
// returns their number
inline int var_list(int kind, int n, int p, int* out) {
  int k = 0;
  for (int v = 0; v < VMAX; ++v)
    if (var_ok(kind, n, v) && (p < 0 || var_plan(v) == p)) out[k++] = v;
  return k;
}
extern std::mutex g_tune_mu;
extern std::map<long long, int> g_tune;  // (n, kind, bucket) -> variant
extern std::map<long long, int> g_plan;  // (n, kind) -> plan index
long long tune_key(int n, int kind, int B);
long long plan_key(int n, int kind);
int tuned(int n, int kind, int B, int dflt);
// the FFT plan fixed for (n, kind): 0 wide, 1 narrow; the plan table's entry
// until the autotuner (or ECSMHD_PLAN / ECSMHD_VARIANTS) has fixed it
int plan_of(int n, int kind);
// FFT plan (0 wide, 1 narrow) of kernel family `kind` for grid size n: the
// plan that timed fastest on MI355X at batch 1 and 16, fixed so that every
// process computes the same bits
int plan_table(int n, int kind);

// ------------------------------------------------------------------ hipGraph chunks
// Captures the launches `record` makes on the context's own stream once per
// key (the bytes of the kernels' arguments plus a tag), then replays the
// instantiated graph; the caller's stream is joined before and after.
// Opt-in (ECSMHD_GRAPH=1): on this ROCm stack, eager launches of the same
// chunk measured faster, so eager is the default.
bool use_graphs();
std::string graph_key(const char* tag, const Args& a, int B, int K);
int set_counter(ecs_ctx* c, int v, hipStream_t st);

template <class F>
inline int graph_run(ecs_ctx* c, const std::string& key, hipStream_t st, F&& record) {
  if (!c->gs) {
    HIPCHK(hipStreamCreateWithFlags(&c->gs, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming));
  }
  HIPCHK(hipEventRecord(c->ev_in, st));
  HIPCHK(hipStreamWaitEvent(c->gs, c->ev_in, 0));
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    std::lock_guard<std::recursive_mutex> lk(g_alloc_mu);
    HIPCHK(hipStreamBeginCapture(c->gs, hipStreamCaptureModeThreadLocal));
    const int rc = record(c->gs);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->gs, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(2, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) return fail(2, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
    it = c->graphs.emplace(key, ex).first;
  }
  HIPCHK(hipGraphLaunch(it->second, c->gs));
  HIPCHK(hipEventRecord(c->ev_out, c->gs));
  HIPCHK(hipStreamWaitEvent(st, c->ev_out, 0));
  return 0;
}

// ------------------------------------------------------------------ per-N dispatch
// Entry points of one grid size; grid_ops(n) returns null for unsupported n.
struct GridOps {
  int (*state_vel)(ecs_ctx*, const double*, const double*, double*, double*, int, int, hipStream_t);
  int (*rk4)(ecs_ctx*, double*, int, const double*, int, int, hipStream_t);
  int (*rk4_jvp)(ecs_ctx*, double*, double*, int, const double*, const double*, int, int, hipStream_t);
  int (*rk43)(ecs_ctx*, double*, int, const double*, double, double, int, int, ecs_rk43_info*, double*, int8_t*,
              hipStream_t);
  int (*rk43_jvp)(ecs_ctx*, double*, double*, int, const double*, const double*, const double*, const int8_t*,
                  const int*, int, hipStream_t);
  int (*state_vel_vjp)(ecs_ctx*, const double*, const double*, double*, int, int, hipStream_t);
  int (*rk4_ckpt)(ecs_ctx*, double*, int, const double*, int, int, int, double2*, const int*, hipStream_t);
  int (*rk4_rev)(ecs_ctx*, const double2*, double*, double*, int, const double*, int, int, int, const int*,
                 hipStream_t);
  int (*rk4_vjp)(ecs_ctx*, const double*, double*, double*, int, const double*, int, int, int, hipStream_t);
  int (*time_kernel)(ecs_ctx*, int, int, int, double*, hipStream_t);
  int (*gen)(ecs_ctx*, double*, int, const double*, int, int, int, hipStream_t);
  int (*gen_jvp)(ecs_ctx*, double*, double*, int, const double*, const double*, int, int, int, hipStream_t);
  // reverse passes: one call, or (checkpoints in / checkpoints out) one phase of the two-phase form
  int (*gen_vjp)(ecs_ctx*, const double*, double*, double*, int, const double*, int, int, int, int,
                 const double2*, double2*, double*, hipStream_t);
  int (*rk43_vjp)(ecs_ctx*, const double*, double*, double*, int, const double*, const double*, const int8_t*,
                  const int*, int, int, const double2*, double2*, hipStream_t);
  int (*state_vel_noalloc)(ecs_ctx*, const double*, double*, int, int, double2*, size_t, hipStream_t);
};
const GridOps* grid_ops(int n);
const GridOps* grid_ops_16();
const GridOps* grid_ops_32();
const GridOps* grid_ops_64();
const GridOps* grid_ops_128();
const GridOps* grid_ops_256();
